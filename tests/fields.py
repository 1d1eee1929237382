"""numpy helpers for closed-form test fields (test infrastructure).

Quaternions are (..., 4) float64 arrays {c0, c1, c2, c3} with
U = c0 + i(c1 s1 + c2 s2 + c3 s3) (LGT_Code/include/gauge.h:21-24); fields are
(VOL, 4, 4) in the reference's site-major order, site n = x3 + L3(x2 + L2(x1 + L1 x0)).
"""
import numpy as np


def qmul(v, w):
    v0, v1, v2, v3 = np.moveaxis(v, -1, 0)
    w0, w1, w2, w3 = np.moveaxis(w, -1, 0)
    return np.stack([
        v0 * w0 - v1 * w1 - v2 * w2 - v3 * w3,
        v0 * w1 + v1 * w0 - v2 * w3 + v3 * w2,
        v0 * w2 + v1 * w3 + v2 * w0 - v3 * w1,
        v0 * w3 - v1 * w2 + v2 * w1 + v3 * w0,
    ], axis=-1)


def qdag(v):
    return v * np.array([1.0, -1.0, -1.0, -1.0])


def to_matrix(q):
    """2x2 complex form of the quaternion (gauge.h:190-198)."""
    c0, c1, c2, c3 = np.moveaxis(q, -1, 0)
    m = np.empty(q.shape[:-1] + (2, 2), dtype=np.complex128)
    m[..., 0, 0] = c0 + 1j * c3
    m[..., 0, 1] = c2 + 1j * c1
    m[..., 1, 0] = -c2 + 1j * c1
    m[..., 1, 1] = c0 - 1j * c3
    return m


def coords(L):
    n = np.arange(int(np.prod(L)))
    x3 = n % L[3]
    r = n // L[3]
    x2 = r % L[2]
    r //= L[2]
    x1 = r % L[1]
    x0 = r // L[1]
    return np.stack([x0, x1, x2, x3], axis=0)


def lex(L, x):
    return ((x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3]


def shift_index(L, mu, step):
    x = coords(L).copy()
    x[mu] = (x[mu] + step) % L[mu]
    return lex(L, x)


def haar(rng, shape):
    q = rng.normal(size=shape + (4,))
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def gauge_transform(U, g, L):
    """U'_mu(x) = g(x) U_mu(x) g(x+mu)^+"""
    V = np.empty_like(U)
    for mu in range(4):
        up = shift_index(L, mu, 1)
        V[:, mu] = qmul(qmul(g, U[:, mu]), qdag(g[up]))
    return V


def abelian_field(theta):
    """U_mu(x) = exp(i s3 theta_mu(x)) for theta of shape (VOL, 4)."""
    U = np.zeros(theta.shape + (4,))
    U[..., 0] = np.cos(theta)
    U[..., 3] = np.sin(theta)
    return U


def flux_angles(L, a, b, k):
    """Angles of a uniform abelian flux 2 pi k / (L_a L_b) through every
    (a, b) plaquette (orientation a then b), periodic via a twist at x_a = L_a - 1."""
    x = coords(L)
    phi = 2.0 * np.pi * k / (L[a] * L[b])
    th = np.zeros((int(np.prod(L)), 4))
    th[:, b] = phi * x[a]
    th[:, a] = np.where(x[a] == L[a] - 1, -phi * L[a] * x[b], 0.0)
    return th, phi


def abelian_plaquette(theta, L):
    tot = 0.0
    for mu in range(3):
        for nu in range(mu + 1, 4):
            th = theta[:, mu] + theta[shift_index(L, mu, 1), nu] - theta[shift_index(L, nu, 1), mu] - theta[:, nu]
            tot += np.cos(th).sum()
    return tot / (6 * theta.shape[0])


def abelian_wilson(theta, L, R, T):
    vol = theta.shape[0]
    acc = 0.0
    for d in (1, 2, 3):
        ang = np.zeros(vol)
        for i in range(T):
            idx = lex(L, _shifted(L, [(0, i)]))
            ang += theta[idx, 0]
            idx = lex(L, _shifted(L, [(d, R), (0, i)]))
            ang -= theta[idx, 0]
        for j in range(R):
            idx = lex(L, _shifted(L, [(0, T), (d, j)]))
            ang += theta[idx, d]
            idx = lex(L, _shifted(L, [(d, j)]))
            ang -= theta[idx, d]
        acc += np.cos(ang).sum()
    return acc / (3 * vol)


def _shifted(L, moves):
    x = coords(L).copy()
    for mu, s in moves:
        x[mu] = (x[mu] + s) % L[mu]
    return x


def improved_f(phi):
    """clover field strength of a uniform abelian flux phi per plaquette:
    1.5 sin(phi) - 0.15 sin(4 phi) + sin(9 phi)/90 (LGT_Code/modules/obs/topologicalcharge.c:123-162)."""
    return 1.5 * np.sin(phi) - 0.15 * np.sin(4 * phi) + np.sin(9 * phi) / 90.0
