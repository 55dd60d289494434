/*
 * nmode.c -- second-derivative drivers of libsff over the device Hessian.
 *
 *   newton  newton.c:357-1371 (non-ScaLAPACK path): Newton-Raphson with the
 *           translation/rotation level shift (level, newton.c:135-226) and
 *           the cubic line search of linemin (newton.c:233-340)
 *   nmode   nmode.c:1596-2781, ntrun 0, 1 and 2: mass-weighted Hessian,
 *           eigenvalues as frequencies, thermochemistry (thermo,
 *           nmode.c:224-600, with mofi :150-221 and symnum :131-146)
 *
 * The Hessian comes from the caller's mme2-like function (normally the
 * library's mme2, the device's central differences); the dense Cholesky
 * solve and the symmetric eigenproblem are dense.hip's (Householder
 * reduction on the GPU for large matrices, QR iteration on the host) where
 * the reference calls LAPACK's dposv_/dsyev_/dsyevd_.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"
#include "sff_gb.h"

static FILE *nm_out(void) { return nabout ? nabout : stdout; }

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    int i;
    for (i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static double norm2(int n, const double *a) { return sqrt(dot(n, a, a)); }

/*
 * The six mass-weighted unit vectors of overall translation (x, y, z) and
 * rotation about z, x and y (Eckart's conditions; newton.c:135-226), as the
 * rows of d[6][n]: a translation along axis k has sqrt(m_a) / sqrt(M) in
 * coordinate k of every atom; the rotations pair -y/x, -z/y and z/-x
 * components normalised by the mass-weighted sums of squares of the
 * coordinates they involve.
 */
static void eckart_vectors(const double *x, const double *m, int natom, double *d)
{
    const int n = 3 * natom;
    double sx = 0.0, sy = 0.0, sz = 0.0, sm = 0.0, nz, nxr, nyr, nt;
    int a;
    for (a = 0; a < natom; a++) {
        sx += x[3 * a] * x[3 * a] * m[a];
        sy += x[3 * a + 1] * x[3 * a + 1] * m[a];
        sz += x[3 * a + 2] * x[3 * a + 2] * m[a];
        sm += m[a];
    }
    nz = sqrt(sx + sy);   /* rotation about z */
    nxr = sqrt(sy + sz);  /* rotation about x */
    nyr = sqrt(sx + sz);  /* rotation about y */
    nt = sqrt(sm);
    memset(d, 0, sizeof(double) * 6 * (size_t)n);
    for (a = 0; a < natom; a++) {
        const double w = sqrt(m[a]), xa = x[3 * a], ya = x[3 * a + 1], za = x[3 * a + 2];
        const int i = 3 * a;
        d[0 * n + i] = w / nt;
        d[1 * n + i + 1] = w / nt;
        d[2 * n + i + 2] = w / nt;
        d[3 * n + i] = -w * ya / nz;
        d[3 * n + i + 1] = w * xa / nz;
        d[4 * n + i + 1] = -w * za / nxr;
        d[4 * n + i + 2] = w * ya / nxr;
        d[5 * n + i] = w * za / nyr;
        d[5 * n + i + 2] = -w * xa / nyr;
    }
}

/*
 * Minimum of func along x + alpha d (newton.c:233-340, after H. T. Lau's
 * routine): two evaluations bracket the slope's sign change, then at most
 * evlmax cubic-interpolation or halving steps, each ending when the step
 * falls below the tolerance (relative |x| tol[0] plus absolute tol[1],
 * divided by |d|).  x ends at the last point evaluated.
 */
static void line_minimum(int n, double *x, const double *d, double alpha, double (*func)(double *, double *, int *),
                         int evlmax, const double tol[2], double *work)
{
    double *x0 = work, *g = work + n;
    const double nd = norm2(n, d);
    int evl = 0, inside = 0, it = -2, i;
    double alpha0 = 0.0, f0, df0, oldf, olddf, y, f1, df1, eps;
    f0 = func(x, g, &it);
    printf("For alpha = %10.5f energy = %20.10f\n", alpha0, f0);
    df0 = dot(n, d, g);
    oldf = f0;
    olddf = df0;
    y = alpha;
    for (i = 0; i < n; i++) x0[i] = x[i] + alpha * d[i];
    it = -2;
    f1 = func(x0, g, &it);
    printf("For alpha = %10.5f energy = %20.10f\n", y, f1);
    df1 = dot(n, d, g);
    memcpy(x0, x, sizeof(double) * n);
    eps = (norm2(n, x) * tol[0] + tol[1]) / nd;
    for (;;) {
        double prev, f, df;
        /* "inside" turns on once the slope at the far point is >= 0: the
         * minimum is bracketed */
        if (!inside) inside = !(df1 < 0.0);
        prev = alpha;
        if (df1 >= 0.0) {
            const double z = 3.0 * (oldf - f1) / alpha + olddf + df1;
            const double w = sqrt(z * z - olddf * df1);
            alpha *= 1.0 - (df1 + w - z) / (df1 - olddf + w + w);
            if (alpha < eps)
                alpha = eps;
            else if (prev - alpha < eps)
                alpha = prev - eps;
        } else if (!inside) {
            alpha0 = alpha = y;
            olddf = df1;
            oldf = f1;
        } else {
            alpha *= 0.5;
        }
        y = alpha + alpha0;
        for (i = 0; i < n; i++) x[i] = x0[i] + y * d[i];
        eps = (norm2(n, x) * tol[0] + tol[1]) / nd;
        it = -2;
        f = func(x, g, &it);
        printf("For alpha = %10.5f energy = %20.10f\n", y, f);
        df = dot(n, d, g);
        if (++evl > evlmax) break;
        if (!inside || df > 0.0) {
            df1 = df;
            f1 = f;
        } else {
            alpha0 = y;
            alpha = prev - alpha;
            olddf = df;
            oldf = f;
        }
        if (alpha < 2.0 * eps) break;
    }
    fflush(stdout);
}

int newton(double x[], int *n_, double *f, double (*func1)(double *, double *, int *), nabc_mme2_fn func2,
           double *rmsgrad, double *nradd, int *maxiter)
{
    const int n = *n_;
    const size_t nn = (size_t)n * (size_t)n;
    double *g = NULL, *grad = NULL, *h = NULL, *m = NULL, *d = NULL, *work = NULL;
    char *name = NULL;
    int niter = 0, ret = -1, natom = 0, it, i, j, k;
    const double sqrn = sqrt((double)n), dgrad = (*rmsgrad) * (*rmsgrad) * n;
    int dummy[9] = {0};
    if (n < 3 || n % 3) {
        nabc_seterr("newton: n must be 3 * natom");
        return -2;
    }
    g = (double *)malloc(sizeof(double) * n);
    grad = (double *)malloc(sizeof(double) * n);
    h = (double *)malloc(sizeof(double) * nn);
    m = (double *)malloc(sizeof(double) * n);
    d = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    work = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    name = (char *)malloc((size_t)n);
    if (!g || !grad || !h || !m || !d || !work || !name) {
        nabc_seterr("newton: host allocation failed");
        ret = -2;
        goto done;
    }
    do {
        double e, alphal;
        ++niter;
        it = niter;
        *f = func1(x, grad, &it);
        if (dot(n, grad, grad) < dgrad) break;
        it = niter;
        e = func2(x, g, h, m, grad, &dummy[0], &dummy[1], &dummy[2], &dummy[3], &dummy[4], &dummy[5], &dummy[6],
                  &natom, &it, name);
        if (isnan(e)) {
            ret = -2;
            goto done;
        }
        *f = e;
        /* the translations and rotations moved to 3500 (newton.c:775-800),
         * then nradd on the diagonal (newton.c:807-818) */
        eckart_vectors(x, m, natom, d);
        for (i = 0; i < n; i++)
            for (j = 0; j < n; j++) {
                double s = 0.0;
                for (k = 0; k < 6; k++) s += d[k * n + i] * d[k * n + j];
                h[(size_t)i * n + j] += 3500.0 * s;
            }
        printf(" adding %10.5f to diagonal of the hessian\n", *nradd);
        for (i = 0; i < n; i++) h[(size_t)i * n + i] += *nradd;
        /* the Newton step H^-1 g into g (newton.c:825) */
        {
            const int info = nabc_dense_posv(n, h, g);
            if (info < 0) {
                printf("dposv returns %d; giving up...\n\n", info);
                ret = -2;
                goto done;
            }
            /* info > 0 (not positive definite): the reference, built
             * without its DSYEV option, carries on with the solver's
             * output (newton.c:847-933) */
        }
        memcpy(grad, g, sizeof(double) * n);
        alphal = norm2(n, grad) / sqrn;
        printf("rms of search direction: %10.7f\n", alphal);
        if (alphal < 0.0001) {
            /* near a stationary point: the full Newton step */
            for (i = 0; i < n; i++) x[i] -= grad[i];
        } else {
            const double tol[2] = {1.0e-8, 1.0e-8};
            for (i = 0; i < n; i++) grad[i] = -grad[i];
            line_minimum(n, x, grad, 1.0, func1, 5, tol, work);
        }
    } while (niter < *maxiter);
    ret = niter >= *maxiter ? -1 : niter;
done:
    it = -3;
    if (func2 && g && h && m && grad && name)
        (void)func2(x, g, h, m, grad, &dummy[0], &dummy[1], &dummy[2], &dummy[3], &dummy[4], &dummy[5], &dummy[6],
                    &natom, &it, name);
    free(g);
    free(grad);
    free(h);
    free(m);
    free(d);
    free(work);
    free(name);
    return ret;
}

/* ------------------------------------------------------------- nmode */

/* principal moments of inertia (amu A^2, ascending) about the centre of
 * mass (nmode.c:150-221) */
static int moments(int natom, const double *x, const double *m, double mom[3])
{
    double c[3] = {0.0, 0.0, 0.0}, mt = 0.0, t[9] = {0};
    int a, k;
    for (a = 0; a < natom; a++) {
        mt += m[a];
        for (k = 0; k < 3; k++) c[k] += m[a] * x[3 * a + k];
    }
    for (k = 0; k < 3; k++) c[k] /= mt;
    for (a = 0; a < natom; a++) {
        const double dx = x[3 * a] - c[0], dy = x[3 * a + 1] - c[1], dz = x[3 * a + 2] - c[2];
        t[0] += m[a] * (dy * dy + dz * dz);
        t[4] += m[a] * (dx * dx + dz * dz);
        t[8] += m[a] * (dx * dx + dy * dy);
        t[1] -= m[a] * dx * dy;
        t[2] -= m[a] * dx * dz;
        t[5] -= m[a] * dy * dz;
    }
    t[3] = t[1];
    t[6] = t[2];
    t[7] = t[5];
    return nabc_dense_eigh(3, t, mom, 0);
}

/*
 * Ideal-gas thermochemistry from the frequencies (ascending, cm^-1) at
 * temperature t and pressure patm (nmode.c:224-600; McQuarrie,
 * Statistical Thermodynamics, ch. 5, 6 and 8): translation, rotation
 * (symmetry number 1, 2 for a homonuclear diatomic), and every real
 * vibration past the first 6 (5 if linear) plus the imaginary ones;
 * printed as the reference prints it, per mode E (kcal/mol), Cv and S
 * (cal/mol-K).  The physical constants are the reference's.
 */
static void thermo(int natom, const double *x, const double *m, const double *freq, double t, double patm,
                   double energy)
{
    const double tokg = 1.660531e-27, boltz = 1.380622e-23, planck = 6.626196e-34, avog = 6.022169e23;
    const double jpcal = 4.18674, tomet = 1e-10, pstd = 101325.0;
    const double gas = avog * boltz, pi = 4.0 * atan(1.0), euler = exp(1.0);
    const double tocal = 1.0 / jpcal, tokcal = tocal / 1.0e3, rt = gas * t, p = pstd * patm;
    const int n3 = 3 * natom;
    double weight = 0.0, arg, stran, etran, ctran, erot, crot, srot, mom[3] = {0, 0, 0}, sn = 1.0, con;
    double evib = 0.0, cvib = 0.0, svib = 0.0, ezpe = 0.0, rtemp1 = 0, rtemp2 = 0, rtemp3 = 0, rtemp = 0;
    int linear = 0, nimag = 0, iff, ndof, i, a;
    double *ev = (double *)calloc((size_t)n3 + 1, sizeof(double)), *cv = (double *)calloc((size_t)n3 + 1, sizeof(double));
    double *sv = (double *)calloc((size_t)n3 + 1, sizeof(double));
    FILE *o = nm_out();
    if (!ev || !cv || !sv) goto done;
    fprintf(o, "\n                - Thermochemistry -\n\n");
    fprintf(o, "Temperature:  %8.3f\n   Pressure:  %8.3f\n", t, patm);
    for (a = 0; a < natom; a++) weight += m[a];
    fprintf(o, "       Mass:  %8.3f\n", weight);
    weight *= tokg;
    /* translation: Sackur-Tetrode */
    {
        const double kt = boltz * t;
        arg = pow(kt, 1.5) / planck;
        arg = arg / p * (kt / planck);
        arg = arg * pow(2.0 * pi, 1.5) * (weight / planck);
        arg = arg * sqrt(weight) * pow(euler, 2.5);
        stran = gas * log(arg);
        etran = 1.5 * rt;
        ctran = 1.5 * gas;
    }
    if (natom <= 1) goto done;
    if (moments(natom, x, m, mom) != 0) goto done;
    fprintf(o, "Principal moments of inertia in amu-A**2:\n");
    fprintf(o, "     %12.2f%12.2f%12.2f\n", mom[0], mom[1], mom[2]);
    if (natom <= 2) {
        linear = 1;
        if (m[0] == m[1]) sn = 2.0;
    }
    fprintf(o, "Rotational symmetry number is %2.0f\n", sn);
    con = planck / (boltz * 8.0 * pi * pi);
    con = con / tokg * (planck / (tomet * tomet));
    if (linear) {
        rtemp = con / mom[2];
        if (rtemp < 0.2) fprintf(o, "Assuming classical behavior for rotation\n");
        fprintf(o, "Rotational temperature: %8.3f\n", rtemp);
        erot = rt;
        crot = gas;
        srot = gas * log(t / rtemp * (euler / sn));
    } else {
        rtemp1 = con / mom[0];
        rtemp2 = con / mom[1];
        rtemp3 = con / mom[2];
        if (rtemp1 < 0.2) fprintf(o, "Assuming classical behavior for rotation\n");
        fprintf(o, "Rotational temperatures: %8.3f  %8.3f  %8.3f\n", rtemp1, rtemp2, rtemp3);
        erot = 1.5 * rt;
        crot = 1.5 * gas;
        arg = sqrt(pi * euler * euler * euler) / sn * sqrt(t / rtemp1 * (t / rtemp2) * (t / rtemp3));
        srot = gas * log(arg);
    }
    for (i = 0; i < n3; i++)
        if (freq[i] < -0.5) nimag++;
    if (nimag > 0) fprintf(o, "found %d imaginary frequencies\n", nimag);
    iff = (linear ? 5 : 6) + nimag;
    ndof = n3 - iff;
    con = planck / boltz;
    for (i = 0; i < ndof; i++) ezpe += freq[i + iff] * 3e10;
    ezpe = 0.5 * planck * ezpe;
    fprintf(o, "Zero-point vibrational energy: %10.3f\n", ezpe * tokcal * avog);
    for (i = 0; i < ndof; i++) {
        /* vibrational temperature over t, per mode */
        const double u = freq[i + iff] * con * 3e10 / t, eu = exp(u), em1 = eu - 1.0, q = u / em1;
        const double ec = u * (0.5 + 1.0 / em1), cc = eu * q * q, argd = 1.0 - 1.0 / eu;
        double sc = 0.0;
        if (argd > 1e-7)
            sc = q - log(argd);
        else
            fprintf(o, "Warning: setting vibrational entropy to zero for mode %d with vtemp = %8.3f\n", i + 1, u * t);
        ev[i] = ec * rt * tokcal;
        cv[i] = cc * gas * tocal;
        sv[i] = sc * gas * tocal;
        evib += ec;
        cvib += cc;
        svib += sc;
    }
    evib *= rt * tokcal;
    cvib *= gas * tocal;
    svib *= gas * tocal;
    etran *= tokcal;
    ctran *= tocal;
    stran *= tocal;
    erot *= tokcal;
    crot *= tocal;
    srot *= tocal;
    fprintf(o, "\n             freq.         E               Cv              S\n");
    fprintf(o, "            cm**-1       kcal/mol       cal/mol-K     cal/mol-K\n");
    fprintf(o, "Total:               %11.3f    %11.3f    %11.3f\n", etran + erot + evib + energy, ctran + crot + cvib,
            stran + srot + svib);
    fprintf(o, "translational:       %11.3f    %11.3f    %11.3f\n", etran, ctran, stran);
    fprintf(o, "rotational:          %11.3f    %11.3f    %11.3f\n", erot, crot, srot);
    fprintf(o, "vibrational:         %11.3f    %11.3f    %11.3f\n", evib, cvib, svib);
    fprintf(o, "ff   energy:         %11.3f\n", energy);
    for (i = 0; i < iff; i++) fprintf(o, " %5d   %10.3f\n", i + 1, freq[i]);
    for (i = 0; i < ndof; i++)
        fprintf(o, " %5d   %10.3f  %11.3f    %11.3f    %11.3f\n", i + iff + 1, freq[i + iff], ev[i], cv[i], sv[i]);
    fflush(o);
done:
    free(ev);
    free(cv);
    free(sv);
}

/* eigenvalue (mass-weighted Hessian) -> signed frequency in cm^-1 */
static double to_wavenumber(double v) { return v >= 0.0 ? 108.587 * sqrt(v) : -108.587 * sqrt(-v); }

/* the modes file of nmode.c:2446-2468: coordinates, then per mode its
 * number, frequency and un-mass-weighted vector, 7 numbers a line */
static int write_vecs(int n, const double *x, const double *freq, const double *vec, int eigp)
{
    FILE *fp = fopen("vecs", "w");
    int i, j;
    if (!fp) {
        nabc_seterr("nmode: unable to open file vecs");
        return -1;
    }
    fprintf(fp, " modes from NAB\n%5d\n", n);
    for (i = 0; i < n; i++) {
        fprintf(fp, "%11.5f", x[i]);
        if (i % 7 == 6) fprintf(fp, "\n");
    }
    if (n % 7 != 0) fprintf(fp, "\n");
    for (j = 0; j < eigp && j < n; j++) {
        fprintf(fp, "****\n%5d%12.5f\n", j + 1, freq[j]);
        for (i = 0; i < n; i++) {
            fprintf(fp, "%11.5f", vec[(size_t)j * n + i]);
            if (i % 7 == 6) fprintf(fp, "\n");
        }
        if (n % 7 != 0) fprintf(fp, "\n");
    }
    fclose(fp);
    return 0;
}

int nmode(double *x, int n, nabc_mme2_fn func, int eigp, int ntrun, double eta, double hrmax, int ioseen)
{
    const size_t nn = (size_t)n * (size_t)n;
    double *g = NULL, *h = NULL, *m = NULL, *grad = NULL, *w = NULL, *v = NULL, energy, sumg = 0.0;
    char *name = NULL;
    int natom = 0, niter = 1, ret = -1, i, j;
    int dummy[9] = {0};
    (void)eta;
    (void)hrmax;
    (void)ioseen;
    if (ntrun == 3) {
        nabc_seterr("nmode: Langevin modes (ntrun=3) are not provided by this library");
        return -1;
    }
    if (ntrun < 0 || ntrun > 3 || n < 3 || n % 3 || !func) {
        nabc_seterr("nmode: bad arguments");
        return -1;
    }
    g = (double *)malloc(sizeof(double) * n);
    grad = (double *)malloc(sizeof(double) * n);
    h = (double *)malloc(sizeof(double) * nn);
    m = (double *)malloc(sizeof(double) * n);
    w = (double *)malloc(sizeof(double) * n);
    v = (double *)malloc(sizeof(double) * n);
    name = (char *)malloc((size_t)n);
    if (!g || !grad || !h || !m || !w || !v || !name) {
        nabc_seterr("nmode: host allocation failed");
        goto done;
    }
    energy = func(x, g, h, m, grad, &dummy[0], &dummy[1], &dummy[2], &dummy[3], &dummy[4], &dummy[5], &dummy[6],
                  &natom, &niter, name);
    if (isnan(energy)) goto done; /* the energy function has set the message */
    if (natom * 3 != n) {
        nabc_seterr("nmode: n must be 3 * natom");
        goto done;
    }
    for (i = 0; i < n; i++) sumg += g[i] * g[i];
    sumg = sqrt(sumg / n);
    fprintf(nm_out(), "\nEnergy       = %20.10e\n", energy);
    fprintf(nm_out(), "RMS gradient = %20.10e\n", sumg);
    /* mass weighting: H_ij / sqrt(m_i m_j) */
    for (i = 0; i < natom; i++) g[3 * i] = g[3 * i + 1] = g[3 * i + 2] = 1.0 / sqrt(m[i]);
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) h[(size_t)i * n + j] *= g[i] * g[j];
    if (nabc_dense_eigh(n, h, w, eigp > 0) != 0) goto done;
    for (i = 0; i < n; i++) v[i] = to_wavenumber(w[i]);
    if (eigp > 0) {
        /* un-mass-weight the eigenvectors (one per row of h) */
        for (j = 0; j < n; j++)
            for (i = 0; i < n; i++) h[(size_t)j * n + i] *= g[i];
        if (write_vecs(n, x, v, h, eigp) != 0) goto done;
    }
    if (ntrun == 2) {
        /* the lowest eigp frequencies with the mass and moments
         * (nmode.c:2284-2393) */
        double mom[3] = {0, 0, 0}, weight = 0.0;
        FILE *o = nm_out();
        fprintf(o, "\n                - Thermochemistry -\n\n");
        fprintf(o, "Temperature:  %8.3f\n   Pressure:  %8.3f\n", 298.15, 1.0);
        for (i = 0; i < natom; i++) weight += m[i];
        fprintf(o, "       Mass:  %8.3f\n", weight);
        if (moments(natom, x, m, mom) != 0) goto done;
        fprintf(o, "Principal moments of inertia in amu-A**2:\n");
        fprintf(o, "     %12.2f%12.2f%12.2f\n", mom[0], mom[1], mom[2]);
        fflush(o);
        printf("\nVibrational frequencies in cm**-1 \n\n");
        for (i = 0; i < eigp && i < n; i++) printf("%5d  %12.5f \n", i + 1, v[i]);
        fflush(stdout);
    } else {
        thermo(natom, x, m, v, 298.15, 1.0, energy);
    }
    ret = 0;
done:
    niter = -3;
    if (g && h && m && grad && name)
        (void)func(x, g, h, m, grad, &dummy[0], &dummy[1], &dummy[2], &dummy[3], &dummy[4], &dummy[5], &dummy[6],
                   &natom, &niter, name);
    free(g);
    free(h);
    free(m);
    free(grad);
    free(w);
    free(v);
    free(name);
    return ret;
}
