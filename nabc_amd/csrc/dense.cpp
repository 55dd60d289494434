/*
 * dense.cpp -- dense symmetric linear algebra on the device for the
 * second-derivative drivers (newton, nmode): the Newton-Raphson system of
 * newton.c:825 (dposv_) and the normal-mode eigenproblem of nmode.c:2419
 * (dsyevd_) / :2208 (dsyev_), solved with rocSOLVER on the GPU instead of
 * host LAPACK.  Matrices are n x n, symmetric, in host memory; both
 * functions copy in, solve on the null stream and copy back.
 */
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstddef>

#include "host_internal.h"

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

int dense_fail(const char *what)
{
    nabc_seterr("%s", what);
    return -1;
}

} // namespace

/*
 * Eigenvalues (ascending, into w[n]) of the symmetric matrix A[n*n]; with
 * vectors != 0 the eigenvectors overwrite A, one per row of the row-major
 * array (column j of the column-major result = eigenvector j).  Returns 0,
 * or -1 with a message (mm_get_error).
 */
extern "C" int nabc_dense_eigh(int n, double *A, double *w, int vectors)
{
    if (n < 1 || !A || !w) return dense_fail("dense eigensolver: bad arguments");
    const size_t nn = (size_t)n * (size_t)n;
    DevBuf dA, dw, de, dinfo;
    if (hipMalloc(&dA.p, nn * sizeof(double)) != hipSuccess || hipMalloc(&dw.p, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc(&de.p, (size_t)n * sizeof(double)) != hipSuccess || hipMalloc(&dinfo.p, sizeof(rocblas_int)) != hipSuccess) {
        (void)hipGetLastError();
        return dense_fail("dense eigensolver: device allocation failed");
    }
    if (hipMemcpy(dA.p, A, nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return dense_fail("dense eigensolver: copy to the device failed");
    rocblas_handle hd = nullptr;
    if (rocblas_create_handle(&hd) != rocblas_status_success) return dense_fail("rocblas_create_handle failed");
    /* symmetric: the row-major array is its own column-major transpose */
    const rocblas_status st = rocsolver_dsyevd(hd, vectors ? rocblas_evect_original : rocblas_evect_none,
                                               rocblas_fill_upper, n, (double *)dA.p, n, (double *)dw.p,
                                               (double *)de.p, (rocblas_int *)dinfo.p);
    rocblas_int info = 0;
    const bool ok = st == rocblas_status_success && hipStreamSynchronize(nullptr) == hipSuccess &&
                    hipMemcpy(&info, dinfo.p, sizeof(info), hipMemcpyDeviceToHost) == hipSuccess &&
                    hipMemcpy(w, dw.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
                    (!vectors || hipMemcpy(A, dA.p, nn * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess);
    rocblas_destroy_handle(hd);
    if (!ok) return dense_fail("rocsolver_dsyevd failed");
    if (info != 0) {
        nabc_seterr("rocsolver_dsyevd: %d off-diagonal elements did not converge", (int)info);
        return -1;
    }
    return 0;
}

/*
 * Solve A x = b for symmetric positive definite A[n*n] by Cholesky
 * factorisation (lower triangle); x overwrites b.  Returns 0, the order
 * (> 0) of the leading minor that is not positive definite (as dposv_'s
 * info), or -1 with a message.
 */
extern "C" int nabc_dense_posv(int n, const double *A, double *b)
{
    if (n < 1 || !A || !b) return dense_fail("dense solver: bad arguments");
    const size_t nn = (size_t)n * (size_t)n;
    DevBuf dA, db, dinfo;
    if (hipMalloc(&dA.p, nn * sizeof(double)) != hipSuccess || hipMalloc(&db.p, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dinfo.p, sizeof(rocblas_int)) != hipSuccess) {
        (void)hipGetLastError();
        return dense_fail("dense solver: device allocation failed");
    }
    if (hipMemcpy(dA.p, A, nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db.p, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return dense_fail("dense solver: copy to the device failed");
    rocblas_handle hd = nullptr;
    if (rocblas_create_handle(&hd) != rocblas_status_success) return dense_fail("rocblas_create_handle failed");
    const rocblas_status st = rocsolver_dposv(hd, rocblas_fill_lower, n, 1, (double *)dA.p, n, (double *)db.p, n,
                                              (rocblas_int *)dinfo.p);
    rocblas_int info = 0;
    const bool ok = st == rocblas_status_success && hipStreamSynchronize(nullptr) == hipSuccess &&
                    hipMemcpy(&info, dinfo.p, sizeof(info), hipMemcpyDeviceToHost) == hipSuccess &&
                    hipMemcpy(b, db.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
    rocblas_destroy_handle(hd);
    if (!ok) return dense_fail("rocsolver_dposv failed");
    return (int)info;
}
