"""Second-derivative drivers of the oracle (test infrastructure only).

A CPU restatement, in numpy over the oracle's mme34 (oracle.c), of the
reference's Newton-Raphson minimiser and normal-mode analysis, used by the
tests as the checker of the library's newton/nmode and of its device
Hessian; nothing in the library imports it.

  hessian    the Hessian of mme34 by central differences of its analytic
             gradient (step h), symmetrised, with the pair list held; the
             reference forms it analytically (sff2.c:2271 egb2, :628
             ebond2, :950 eangl2, :1193 ephi2).  Pinned by the normal modes
             of test/nab/asp_nm1.out.check (all 66 rows, 4 columns, at
             their printed precision) and asp_nm2.out.check (modes 7-20 at
             1e-4) in tests/test_oracle_nmode.py.
  level      newton.c:135-226  translation/rotation unit vectors (Eckart)
  linemin    newton.c:233-340  cubic line search (after H. T. Lau)
  newton     newton.c:357-1371 (non-ScaLAPACK path, DSYEV undefined)
  frequencies, thermo
             nmode.c:1920-1945, :2394-2481 (ntrun 1: mass weighting,
             eigenvalues -> 108.587 sqrt(lambda) cm^-1) and thermo(),
             nmode.c:224-600 (McQuarrie's ideal-gas formulas, the
             reference's constants), per-mode E, Cv, S.
"""
import numpy as np

HESS_STEP = 1.0e-4


class MME:
    """mme34 of an oracle.Oracle with its pair list held (built at the first call)."""

    def __init__(self, orc):
        self.orc = orc
        self.built = False

    def __call__(self, x):
        ene, frms, g = self.orc.mme(np.ascontiguousarray(x, dtype=np.float64), rebuild=not self.built)
        self.built = True
        return ene[0], g


def hessian(mme, x, h=HESS_STEP):
    n = len(x)
    H = np.zeros((n, n))
    for c in range(n):
        xp = x.copy()
        xp[c] += h
        xm = x.copy()
        xm[c] -= h
        H[:, c] = (mme(xp)[1] - mme(xm)[1]) / (2.0 * h)
    return 0.5 * (H + H.T)


def level(x, m):
    """The six Eckart vectors (rows) of newton.c:135-226."""
    X = x.reshape(-1, 3)
    n = len(x)
    sx, sy, sz = ((X[:, k] ** 2 * m).sum() for k in range(3))
    nz, nxr, nyr, nt = np.sqrt(sx + sy), np.sqrt(sy + sz), np.sqrt(sx + sz), np.sqrt(m.sum())
    d = np.zeros((6, n))
    w = np.sqrt(m)
    idx = np.arange(len(m)) * 3
    d[0, idx] = w / nt
    d[1, idx + 1] = w / nt
    d[2, idx + 2] = w / nt
    d[3, idx] = -w * X[:, 1] / nz
    d[3, idx + 1] = w * X[:, 0] / nz
    d[4, idx + 1] = -w * X[:, 2] / nxr
    d[4, idx + 2] = w * X[:, 1] / nxr
    d[5, idx] = w * X[:, 2] / nyr
    d[5, idx + 2] = -w * X[:, 0] / nyr
    return d


def linemin(mme, x, d, alpha=1.0, evlmax=5, tol=(1e-8, 1e-8)):
    nd = np.linalg.norm(d)
    evl, alpha0 = 0, 0.0
    f0, g = mme(x)
    oldf, olddf = f0, d @ g
    y = alpha
    f1, g = mme(x + alpha * d)
    df1 = d @ g
    inside = False
    x0 = x.copy()
    eps = (np.linalg.norm(x) * tol[0] + tol[1]) / nd
    while True:
        if not inside:
            inside = not (df1 < 0.0)
        prev = alpha
        if df1 >= 0.0:
            z = 3.0 * (oldf - f1) / alpha + olddf + df1
            w = np.sqrt(z * z - olddf * df1)
            alpha *= 1.0 - (df1 + w - z) / (df1 - olddf + w + w)
            if alpha < eps:
                alpha = eps
            elif prev - alpha < eps:
                alpha = prev - eps
        elif not inside:
            alpha0 = alpha = y
            olddf, oldf = df1, f1
        else:
            alpha *= 0.5
        y = alpha + alpha0
        x = x0 + y * d
        eps = (np.linalg.norm(x) * tol[0] + tol[1]) / nd
        f, g = mme(x)
        df = d @ g
        evl += 1
        if evl > evlmax:
            break
        if not inside or df > 0.0:
            df1, f1 = df, f
        else:
            alpha0 = y
            alpha = prev - alpha
            olddf, oldf = df, f
        if alpha < 2.0 * eps:
            break
    return x


def newton(mme, x, m, rmsgrad, nradd=0.0, maxiter=15, h=HESS_STEP):
    """Returns (x, return code as newton(), energy)."""
    n = len(x)
    dgrad = rmsgrad * rmsgrad * n
    x = x.copy()
    f = None
    for it in range(1, maxiter + 1):
        f, g = mme(x)
        if g @ g < dgrad:
            return x, it, f
        H = hessian(mme, x, h)
        d = level(x, m)
        H = H + 3500.0 * d.T @ d + nradd * np.eye(n)
        dx = np.linalg.solve(H, g)
        if np.linalg.norm(dx) / np.sqrt(n) < 1e-4:
            x = x - dx
        else:
            x = linemin(mme, x, -dx)
    return x, -1, f


def frequencies(H, m):
    """Signed frequencies (cm^-1, ascending) of the mass-weighted Hessian."""
    w = np.repeat(1.0 / np.sqrt(m), 3)
    lam = np.linalg.eigvalsh(w[:, None] * H * w[None, :])
    return np.where(lam >= 0.0, 108.587 * np.sqrt(np.abs(lam)), -108.587 * np.sqrt(np.abs(lam)))


def thermo_modes(freq, t=298.15, linear=False):
    """Per-mode (mode number, frequency, E kcal/mol, Cv, S cal/mol-K) of the
    vibrations thermo() includes."""
    boltz, planck, avog, jpcal = 1.380622e-23, 6.626196e-34, 6.022169e23, 4.18674
    gas = avog * boltz
    rt = gas * t
    nimag = int(np.sum(freq < -0.5))
    iff = (5 if linear else 6) + nimag
    u = freq[iff:] * planck / boltz * 3e10 / t
    eu = np.exp(u)
    em1 = eu - 1.0
    e = u * (0.5 + 1.0 / em1) * rt / jpcal / 1e3
    cv = eu * (u / em1) ** 2 * gas / jpcal
    s = (u / em1 - np.log(1.0 - 1.0 / eu)) * gas / jpcal
    return [(iff + k + 1, freq[iff + k], e[k], cv[k], s[k]) for k in range(len(u))]
