/*
 * dense.hip -- dense symmetric linear algebra for the second-derivative
 * drivers (newton, nmode): the Newton-Raphson system of newton.c:825
 * (dposv_) and the normal-mode eigenproblem of nmode.c:2419 (dsyevd_) /
 * :2208 (dsyev_), which the reference hands to host LAPACK.
 *
 *   nabc_dense_posv   Cholesky factorisation A = L L^T and two triangular
 *                     solves, on the host.
 *   nabc_dense_eigh   Householder reduction to tridiagonal form (Golub and
 *                     Van Loan, Matrix Computations, 8.3.1), on the GPU for
 *                     n >= DEVICE_MIN_N (one reflector per step: norm, the
 *                     product p = beta A v, w = p - (beta/2)(v.p) v, and the
 *                     rank-2 update A -= v w^T + w v^T of the trailing
 *                     block), on the host below; then the implicit
 *                     symmetric QR iteration with Wilkinson shifts on the
 *                     tridiagonal matrix (ibid. 8.3.2-8.3.3) on the host,
 *                     with the eigenvectors accumulated and transformed back
 *                     through the reflectors when asked for.
 *
 * No vendor library: the library must coexist in one process with other
 * users of the ROCm math libraries (which may bring their own copies).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "gbhip.h"
#include "host_internal.h"

#define DEVICE_MIN_N 1024

namespace {

int dense_fail(const char *what)
{
    nabc_seterr("%s", what);
    return -1;
}

/* ------------------------------------------------ tridiagonal reduction */

/*
 * Reflector k on the host: x = A[k+1.., k]; alpha = -sign(x0) |x|;
 * v = x - alpha e0; beta = 2 / v.v (0 when x = 0); the trailing block
 * A22 <- H A22 H.  d[k] = A[k][k], e[k] = alpha.  V row k keeps v (at
 * columns k+1..n-1), betas[k] its beta.
 */
void tridiag_host(int n, std::vector<double> &A, double *d, double *e, std::vector<double> &V,
                  std::vector<double> &betas)
{
    std::vector<double> p(n), w(n);
    for (int k = 0; k + 2 < n; k++) {
        const int m = n - k - 1;
        double *v = &V[(size_t)k * n + k + 1];
        double sigma = 0.0;
        for (int i = 0; i < m; i++) {
            v[i] = A[(size_t)(k + 1 + i) * n + k];
            sigma += v[i] * v[i];
        }
        d[k] = A[(size_t)k * n + k];
        if (sigma == 0.0) {
            e[k] = 0.0;
            betas[k] = 0.0;
            continue;
        }
        const double nx = std::sqrt(sigma), x0 = v[0], alpha = x0 > 0.0 ? -nx : nx;
        v[0] = x0 - alpha;
        const double beta = 2.0 / (sigma - x0 * x0 + v[0] * v[0]);
        betas[k] = beta;
        e[k] = alpha;
        double vp = 0.0;
        for (int i = 0; i < m; i++) {
            const double *row = &A[(size_t)(k + 1 + i) * n + k + 1];
            double s = 0.0;
            for (int j = 0; j < m; j++) s += row[j] * v[j];
            p[i] = beta * s;
            vp += v[i] * p[i];
        }
        const double K = 0.5 * beta * vp;
        for (int i = 0; i < m; i++) w[i] = p[i] - K * v[i];
        for (int i = 0; i < m; i++) {
            double *row = &A[(size_t)(k + 1 + i) * n + k + 1];
            for (int j = 0; j < m; j++) row[j] -= v[i] * w[j] + w[i] * v[j];
        }
    }
    d[n - 2] = A[(size_t)(n - 2) * n + n - 2];
    d[n - 1] = A[(size_t)(n - 1) * n + n - 1];
    e[n - 2] = A[(size_t)(n - 1) * n + n - 2];
}

/* device: reflector k from column k (one workgroup of 1024) */
__global__ __launch_bounds__(1024) void k_hh_reflector(double *A, int n, int k, double *V, double *betas, double *e)
{
    __shared__ double red[1024];
    const int m = n - k - 1, t = threadIdx.x;
    double *v = V + (size_t)k * n + k + 1;
    double s = 0.0;
    for (int i = t; i < m; i += 1024) {
        const double x = A[(size_t)(k + 1 + i) * n + k];
        v[i] = x;
        s += x * x;
    }
    red[t] = s;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) {
        const double sigma = red[0];
        if (sigma == 0.0) {
            e[k] = 0.0;
            betas[k] = 0.0;
        } else {
            const double nx = sqrt(sigma), x0 = v[0], alpha = x0 > 0.0 ? -nx : nx;
            v[0] = x0 - alpha;
            betas[k] = 2.0 / (sigma - x0 * x0 + v[0] * v[0]);
            e[k] = alpha;
        }
    }
}

/* p_i = beta (A22 v)_i, one workgroup of 256 per row */
__global__ __launch_bounds__(256) void k_hh_matvec(const double *A, int n, int k, const double *V, const double *betas,
                                                   double *p)
{
    __shared__ double red[256];
    const int m = n - k - 1, i = blockIdx.x, t = threadIdx.x;
    const double beta = betas[k];
    const double *row = A + (size_t)(k + 1 + i) * n + k + 1, *v = V + (size_t)k * n + k + 1;
    double s = 0.0;
    if (beta != 0.0)
        for (int j = t; j < m; j += 256) s += row[j] * v[j];
    red[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) p[i] = beta * red[0];
}

/* w = p - (beta/2)(v.p) v, in place (one workgroup of 1024) */
__global__ __launch_bounds__(1024) void k_hh_w(double *p, int n, int k, const double *V, const double *betas)
{
    __shared__ double red[1024];
    const int m = n - k - 1, t = threadIdx.x;
    const double *v = V + (size_t)k * n + k + 1;
    double s = 0.0;
    for (int i = t; i < m; i += 1024) s += v[i] * p[i];
    red[t] = s;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    const double K = 0.5 * betas[k] * red[0];
    for (int i = t; i < m; i += 1024) p[i] -= K * v[i];
}

/* A22 -= v w^T + w v^T over 16 x 16 tiles */
__global__ __launch_bounds__(256) void k_hh_update(double *A, int n, int k, const double *V, const double *w)
{
    const int m = n - k - 1;
    const int i = blockIdx.y * 16 + threadIdx.y, j = blockIdx.x * 16 + threadIdx.x;
    if (i < m && j < m) {
        const double *v = V + (size_t)k * n + k + 1;
        A[(size_t)(k + 1 + i) * n + k + 1 + j] -= v[i] * w[j] + w[i] * v[j];
    }
}

/* diagonal and last off-diagonal element after the reduction */
__global__ void k_hh_diag(const double *A, int n, double *d, double *e)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) d[i] = A[(size_t)i * n + i];
    if (i == 0) e[n - 2] = A[(size_t)(n - 1) * n + n - 2];
}

struct DevMem {
    void *p = nullptr;
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
};

int tridiag_device(int n, const std::vector<double> &A, double *d, double *e, std::vector<double> *V,
                   std::vector<double> *betas)
{
    const size_t nn = (size_t)n * (size_t)n;
    DevMem dA, dV, dB, dE, dP, dD;
    if (hipMalloc(&dA.p, nn * sizeof(double)) != hipSuccess || hipMalloc(&dV.p, nn * sizeof(double)) != hipSuccess ||
        hipMalloc(&dB.p, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dE.p, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dP.p, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dD.p, (size_t)n * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return dense_fail("dense eigensolver: device allocation failed");
    }
    double *a = (double *)dA.p, *vv = (double *)dV.p, *bb = (double *)dB.p, *ee = (double *)dE.p,
           *pp = (double *)dP.p, *dd = (double *)dD.p;
    if (hipMemcpy(a, A.data(), nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return dense_fail("dense eigensolver: copy to the device failed");
    for (int k = 0; k + 2 < n; k++) {
        const int m = n - k - 1;
        hipLaunchKernelGGL(k_hh_reflector, dim3(1), dim3(1024), 0, nullptr, a, n, k, vv, bb, ee);
        hipLaunchKernelGGL(k_hh_matvec, dim3(m), dim3(256), 0, nullptr, a, n, k, vv, bb, pp);
        hipLaunchKernelGGL(k_hh_w, dim3(1), dim3(1024), 0, nullptr, pp, n, k, vv, bb);
        const unsigned g = (unsigned)((m + 15) / 16);
        hipLaunchKernelGGL(k_hh_update, dim3(g, g), dim3(16, 16), 0, nullptr, a, n, k, vv, pp);
    }
    hipLaunchKernelGGL(k_hh_diag, dim3((n + 255) / 256), dim3(256), 0, nullptr, a, n, dd, ee);
    if (hipGetLastError() != hipSuccess) return dense_fail("dense eigensolver: kernel launch failed");
    if (hipMemcpy(d, dd, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(e, ee, (size_t)(n - 1) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return dense_fail("dense eigensolver: copy from the device failed");
    if (V && (hipMemcpy(V->data(), vv, nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(betas->data(), bb, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
        return dense_fail("dense eigensolver: copy from the device failed");
    return 0;
}

/* ------------------------------------------ tridiagonal QR iteration */

/*
 * Eigenvalues of the symmetric tridiagonal (d, e) into d; with Z (n x n,
 * row-major, identity on entry) the rotations are accumulated so that
 * column j of Z is the eigenvector of d[j] in the tridiagonal basis.  An
 * off-diagonal element below eps (|d_i| + |d_i+1|) is set to zero; the
 * largest unreduced block [lo, hi] then gets one implicit QR step with the
 * Wilkinson shift of its trailing 2 x 2 (the bulge chased down by Givens
 * rotations).  Returns 0, or -1 when 30 n steps do not suffice.
 */
int tridiag_qr(int n, double *d, double *e, double *Z)
{
    const double eps = DBL_EPSILON;
    long long steps = 0;
    int hi = n - 1;
    while (hi > 0) {
        if (std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
            e[hi - 1] = 0.0;
            hi--;
            continue;
        }
        int lo = hi - 1;
        while (lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo]))) lo--;
        if (lo > 0) e[lo - 1] = 0.0;
        if (++steps > 30LL * n) return -1;
        /* Wilkinson shift: the eigenvalue of the trailing 2 x 2 nearer d[hi] */
        const double t = 0.5 * (d[hi - 1] - d[hi]), b2 = e[hi - 1] * e[hi - 1];
        const double mu = d[hi] - b2 / (t + (t >= 0.0 ? 1.0 : -1.0) * std::hypot(t, e[hi - 1]));
        double x = d[lo] - mu, z = e[lo], bulge = 0.0;
        for (int k = lo; k < hi; k++) {
            if (k > lo) {
                x = e[k - 1];
                z = bulge;
            }
            const double r = std::hypot(x, z);
            const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
            if (k > lo) e[k - 1] = r;
            /* rows and columns k, k+1 turned by [c s; -s c] */
            const double a = d[k], b = e[k], dd = d[k + 1];
            d[k] = c * c * a + 2.0 * c * s * b + s * s * dd;
            d[k + 1] = s * s * a - 2.0 * c * s * b + c * c * dd;
            e[k] = c * s * (dd - a) + (c * c - s * s) * b;
            if (k + 1 < hi) {
                bulge = s * e[k + 1];
                e[k + 1] = c * e[k + 1];
            }
            if (Z) {
                for (int i = 0; i < n; i++) {
                    const double zk = Z[(size_t)i * n + k], zk1 = Z[(size_t)i * n + k + 1];
                    Z[(size_t)i * n + k] = c * zk + s * zk1;
                    Z[(size_t)i * n + k + 1] = -s * zk + c * zk1;
                }
            }
        }
    }
    return 0;
}

} // namespace

/*
 * Eigenvalues (ascending, into w[n]) of the symmetric matrix A[n*n]; with
 * vectors != 0 the eigenvectors overwrite A, one per row (row j belongs to
 * w[j]).  Returns 0, or -1 with a message (mm_get_error).
 */
extern "C" int nabc_dense_eigh(int n, double *A, double *w, int vectors)
{
    if (n < 1 || !A || !w) return dense_fail("dense eigensolver: bad arguments");
    if (n == 1) {
        w[0] = A[0];
        if (vectors) A[0] = 1.0;
        return 0;
    }
    const size_t nn = (size_t)n * (size_t)n;
    std::vector<double> a(A, A + nn), d(n), e(n, 0.0), V, betas;
    if (vectors) {
        V.assign(nn, 0.0);
        betas.assign(n, 0.0);
    }
    if (n >= DEVICE_MIN_N) {
        if (gbhip_device_count() < 1) return dense_fail("dense eigensolver: no GPU is visible to the HIP runtime");
        if (tridiag_device(n, a, d.data(), e.data(), vectors ? &V : nullptr, vectors ? &betas : nullptr)) return -1;
    } else {
        if (!vectors) {
            V.assign(nn, 0.0);
            betas.assign(n, 0.0);
        }
        tridiag_host(n, a, d.data(), e.data(), V, betas);
    }
    std::vector<double> Z;
    if (vectors) {
        Z.assign(nn, 0.0);
        for (int i = 0; i < n; i++) Z[(size_t)i * n + i] = 1.0;
    }
    if (tridiag_qr(n, d.data(), e.data(), vectors ? Z.data() : nullptr))
        return dense_fail("dense eigensolver: the tridiagonal QR iteration did not converge");
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&d](int a_, int b_) { return d[a_] < d[b_]; });
    for (int i = 0; i < n; i++) w[i] = d[order[i]];
    if (vectors) {
        /* eigenvector j of A = H_0 H_1 ... H_{n-3} z_j, H_k = I - beta_k v_k v_k^T */
        for (int jj = 0; jj < n; jj++) {
            const int j = order[jj];
            std::vector<double> y(n);
            for (int i = 0; i < n; i++) y[i] = Z[(size_t)i * n + j];
            for (int k = n - 3; k >= 0; k--) {
                if (betas[k] == 0.0) continue;
                const double *v = &V[(size_t)k * n + k + 1];
                double s = 0.0;
                for (int i = 0; i < n - k - 1; i++) s += v[i] * y[k + 1 + i];
                s *= betas[k];
                for (int i = 0; i < n - k - 1; i++) y[k + 1 + i] -= s * v[i];
            }
            std::memcpy(A + (size_t)jj * n, y.data(), sizeof(double) * n);
        }
    }
    return 0;
}

/*
 * Solve A x = b for symmetric positive definite A[n*n] (the lower
 * triangle is read) by Cholesky factorisation; x overwrites b.  Returns 0,
 * the order (> 0) of the leading minor that is not positive definite (as
 * dposv_'s info; b is then unchanged), or -1 with a message.
 */
extern "C" int nabc_dense_posv(int n, const double *A, double *b)
{
    if (n < 1 || !A || !b) return dense_fail("dense solver: bad arguments");
    std::vector<double> L((size_t)n * (size_t)n, 0.0);
    for (int j = 0; j < n; j++) {
        double s = A[(size_t)j * n + j];
        const double *lj = &L[(size_t)j * n];
        for (int k = 0; k < j; k++) s -= lj[k] * lj[k];
        if (!(s > 0.0)) return j + 1;
        const double ljj = std::sqrt(s);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; i++) {
            const double *li = &L[(size_t)i * n];
            double t = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) t -= li[k] * lj[k];
            L[(size_t)i * n + j] = t / ljj;
        }
    }
    std::vector<double> y(n);
    for (int i = 0; i < n; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= L[(size_t)i * n + k] * y[k];
        y[i] = t / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double t = y[i];
        for (int k = i + 1; k < n; k++) t -= L[(size_t)k * n + i] * b[k];
        b[i] = t / L[(size_t)i * n + i];
    }
    return 0;
}
