"""Independent brute-force evaluation of primitive Cartesian two-electron
integrals, used to pin the oracle's Hermite (McMurchie-Davidson) machinery.

1/r12 = (2/sqrt(pi)) int_0^inf exp(-t^2 r12^2) dt turns [ab|cd] into a
t-integral of products over x, y, z of 2-D Gaussian integrals
    int int P_bra(x1) P_ket(x2) exp(-p(x1-P)^2 - q(x2-Q)^2 - t^2 (x1-x2)^2)
which are evaluated exactly by tensor Gauss-Hermite quadrature after
whitening the quadratic form.  The t-integral is mapped to u in [0, 1) by
t = sqrt(rho) u / sqrt(1 - u^2) (rho = pq/(p+q)), where the integrand is
smooth, and done by Gauss-Legendre.  No recurrence relation is involved.
"""

import numpy as np

_GH_X, _GH_W = np.polynomial.hermite.hermgauss(12)
_GL_X, _GL_W = np.polynomial.legendre.leggauss(400)


def cart_list(L):
    out = []
    for lx in range(L, -1, -1):
        for ly in range(L - lx, -1, -1):
            out.append((lx, ly, L - lx - ly))
    return out


def _dir_integral(t2, p, q, P, Q, pows_bra, pows_ket):
    """2-D integral for one direction; pows_* are lists of (centre, power)."""
    M = np.array([[p + t2, -t2], [-t2, q + t2]])
    D = P - Q
    b = t2 * D * np.array([1.0, -1.0])
    Minv = np.linalg.inv(M)
    mu = -Minv @ b
    const = b @ Minv @ b - t2 * D * D
    R = np.linalg.cholesky(M).T  # M = R^T R
    Rinv = np.linalg.inv(R)
    W1, W2 = np.meshgrid(_GH_X, _GH_X, indexing="ij")
    WW = np.outer(_GH_W, _GH_W)
    z1 = mu[0] + Rinv[0, 0] * W1 + Rinv[0, 1] * W2
    z2 = mu[1] + Rinv[1, 0] * W1 + Rinv[1, 1] * W2
    x1 = z1 + P
    x2 = z2 + Q
    g = np.ones_like(x1)
    for c, k in pows_bra:
        g = g * (x1 - c) ** k
    for c, k in pows_ket:
        g = g * (x2 - c) ** k
    return np.exp(const) * np.sum(WW * g) / np.linalg.det(R)


def prim_cart_eri(la, lb, lc, ld, a, b, c, d, A, B, C, D):
    A, B, C, D = (np.asarray(x, dtype=float) for x in (A, B, C, D))
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    Kab = np.exp(-a * b / p * np.sum((A - B) ** 2))
    Kcd = np.exp(-c * d / q * np.sum((C - D) ** 2)) if q > 0 else 1.0
    rho = p * q / (p + q)
    u = 0.5 * (_GL_X + 1.0)
    wu = 0.5 * _GL_W
    t = np.sqrt(rho) * u / np.sqrt(1.0 - u * u)
    dtdu = np.sqrt(rho) / (1.0 - u * u) ** 1.5
    ca, cb, cc, cd_ = cart_list(la), cart_list(lb), cart_list(lc), cart_list(ld)
    out = np.zeros((len(ca), len(cb), len(cc), len(cd_)))
    for iu in range(u.size):
        t2 = t[iu] ** 2
        f = wu[iu] * dtdu[iu] * 2.0 / np.sqrt(np.pi) * Kab * Kcd
        for i, ea in enumerate(ca):
            for j, eb in enumerate(cb):
                for k, ec in enumerate(cc):
                    for l, ed in enumerate(cd_):
                        val = f
                        for x in range(3):
                            val *= _dir_integral(
                                t2, p, q, P[x], Q[x],
                                [(A[x], ea[x]), (B[x], eb[x])],
                                [(C[x], ec[x]), (D[x], ed[x])],
                            )
                        out[i, j, k, l] += val
    return out
