"""Pins the CPU chain rule of N-step APE smearing (spatial, all directions,
temporal) + Wilson-loop objective (oracle/su2_oracle.c, or_ape_vjp) with
central finite differences along su(2) directions: U -> exp(eps X) U,
dL/deps = <dL/dU, X U>."""
import numpy as np
import pytest

import fields as F
from oracle_lib import Oracle


def su2_exp(X, eps):
    """exp(eps X) for a pure-imaginary quaternion X = (0, x)."""
    n = np.linalg.norm(X[1:])
    out = np.zeros(4)
    out[0] = np.cos(eps * n)
    out[1:] = np.sin(eps * n) * X[1:] / n
    return out


def objective_after(o, U, nsteps, sp, R, T, mode=0):
    V = U.copy()
    o.ape(V, mode, nsteps, sp)
    return o.wloop_objective(V, R, T)


@pytest.mark.parametrize("mode,sp", [(0, 0.25), (1, 1.0 / 6.0), (2, 0.25)])
@pytest.mark.parametrize("L,nsteps,R,T", [((4, 4, 4, 4), 0, 2, 2), ((6, 4, 4, 8), 3, 2, 3), ((4, 4, 4, 4), 4, 1, 2)])
def test_ape_vjp_matches_finite_differences(L, nsteps, R, T, mode, sp):
    o = Oracle(L)
    U = o.hot(3)
    o.engine_sweep(U, 3.09, 1, 3, 0)
    Lval, G = o.ape_vjp(U, nsteps, sp, R, T, mode)
    assert Lval == pytest.approx(objective_after(o, U, nsteps, sp, R, T, mode), rel=1e-14)
    rng = np.random.default_rng(7)
    eps = 1e-5
    # every direction at least twice: the smeared and the unsmeared ones
    for i in range(10):
        n = int(rng.integers(o.vol))
        mu = i % 4
        X = np.concatenate([[0.0], rng.normal(size=3)])
        Up, Um = U.copy(), U.copy()
        Up[n, mu] = F.qmul(su2_exp(X, eps), U[n, mu])
        Um[n, mu] = F.qmul(su2_exp(X, -eps), U[n, mu])
        fd = (objective_after(o, Up, nsteps, sp, R, T, mode) - objective_after(o, Um, nsteps, sp, R, T, mode)) / (2 * eps)
        an = float(np.dot(G[n, mu], F.qmul(X, U[n, mu])))
        assert an == pytest.approx(fd, rel=1e-6, abs=1e-8), (mode, n, mu, an, fd)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ape_vjp_gradient_is_gauge_covariant(mode):
    """L is gauge invariant and smearing is covariant, so the gradient
    transforms like a link: G'_mu(x) = g(x) G_mu(x) g(x+mu)^+."""
    L = (6, 4, 4, 8)
    o = Oracle(L)
    U = o.hot(5)
    rng = np.random.default_rng(2)
    g = F.haar(rng, (o.vol,))
    V = F.gauge_transform(U, g, L)
    la, Ga = o.ape_vjp(U, 3, 0.25, 2, 2, mode)
    lb, Gb = o.ape_vjp(V, 3, 0.25, 2, 2, mode)
    assert la == pytest.approx(lb, rel=1e-12)
    np.testing.assert_allclose(F.gauge_transform(Ga, g, L), Gb, atol=1e-11)
