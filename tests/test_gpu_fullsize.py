"""Size-independent properties at the bench's full size, 48^4, in both
precisions: where the oracle would take minutes, the mathematics fixes the
answer.

* Overrelaxation is an involution: the same (mu, parity) sub-sweep applied
  twice returns every link (its staple does not involve the links it
  changes).
* An overrelaxation sub-sweep preserves Re tr(U A) of every link it changes,
  so it leaves the total action, i.e. the plaquette, unchanged.
* Spatial APE smearing leaves the temporal links untouched and every smeared
  link a unit quaternion.
* Results do not depend on the tile order: sweeps, FP32 spatial smearing
  (the site-pair kernel) and the charge give the same bits under the plain
  linear order as under the tuned one.
* The plaquette, Wilson loops and both charges are gauge invariant, and APE
  smearing is gauge covariant.
* A heatbath sub-sweep does not read the links it replaces.
* A heatbath-only sweep replaces every link and leaves each a unit
  quaternion (FP64: a projection pass after the sweep; FP32: each link
  projected as the heatbath stores it).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = (48, 48, 48, 48)


@pytest.fixture(scope="module")
def lgt():
    import masterarbeit_amd as m
    m.load_library()
    return m


@pytest.fixture(scope="module", params=[0, 1], ids=["fp64", "fp32"])
def thermal(lgt, request):
    """(precision, field) after 5 heatbath + overrelaxation sweeps from a hot start."""
    prec = request.param
    with lgt.Engine(L, prec, 7) as e:
        e.init_field(1)
        e.update(5, 1, 3.09)
        return prec, e.download()


def test_overrelax_is_an_involution_and_keeps_the_action(lgt, thermal):
    prec, U = thermal
    tol_link = 1e-14 if prec == 0 else 1e-6     # measured: 5e-16, 2.4e-7
    tol_plaq = 1e-14 if prec == 0 else 1e-9     # measured: 0, 1.1e-11
    with lgt.Engine(L, prec, 7) as e:
        e.upload(U)
        P0 = e.plaquette()
        for mu, p in ((0, 0), (2, 1)):
            e.overrelax_subsweep(mu, p)
            P1 = e.plaquette()
            print(f"prec {prec} OR({mu},{p}): dP = {P1 - P0:.2e}")
            assert abs(P1 - P0) <= tol_plaq
            e.overrelax_subsweep(mu, p)
        d = np.max(np.abs(e.download() - U))
        print(f"prec {prec}: OR twice, max |dU| = {d:.2e}")
        assert d <= tol_link


def test_spatial_smearing_keeps_temporal_links_and_unit_norm(lgt, thermal):
    prec, U = thermal
    with lgt.Engine(L, prec, 7) as e:
        e.upload(U)
        e.ape(lgt.APE_SPATIAL, 3, 0.25)
        V = e.download()
    np.testing.assert_array_equal(V[:, 0], U[:, 0])
    np.testing.assert_allclose(np.linalg.norm(V[:, 1:], axis=-1), 1.0, atol=1e-14 if prec == 0 else 1e-6)
    assert np.max(np.abs(V[:, 1:] - U[:, 1:])) > 1e-3


def test_tile_order_does_not_change_results(lgt, thermal):
    prec, U = thermal
    res = []
    for order in (-1, 0):
        with lgt.Engine(L, prec, 7) as e:
            e.set_option(lgt.engine.OPT_TILE_ORDER, order)
            e.upload(U)
            e.update(1, 1, 3.09)
            e.ape(lgt.APE_SPATIAL, 2, 0.25)
            res.append((e.download(), e.qtop(lgt.Q_IMPROVED)))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1]


def test_observables_are_gauge_invariant(lgt, thermal):
    """The plaquette, Wilson loops and both charges of a random gauge
    transform U'_mu(x) = g(x) U_mu(x) g(x+mu)^+ of the 48^4 field equal the
    untransformed ones (FP64: to the rounding of the transform; FP32: to the
    rounding of the transformed links to FP32)."""
    import fields as F
    prec, U = thermal
    g = F.haar(np.random.default_rng(5), (U.shape[0],))
    V = F.gauge_transform(U, g, L)
    R, T = [1, 3, 5], [1, 4, 5]
    out = []
    for W in (U, V):
        with lgt.Engine(L, prec, 7) as e:
            e.upload(W)
            out.append((e.plaquette(), e.wilson_loops(R, T), e.qtop(lgt.Q_IMPROVED), e.qtop(lgt.Q_CLOVER)))
    (p0, w0, q0, c0), (p1, w1, q1, c1) = out
    tol = 1e-13 if prec == 0 else 1e-8   # measured: 0 and 1.5e-11 (P), 7e-11 (W)
    print(f"prec {prec}: dP {p1 - p0:.1e} dW {np.max(np.abs(np.subtract(w1, w0))):.1e} "
          f"dQ {q1[0] - q0[0]:.1e} (sum |q| {q0[1]:.1f}) dQc {c1[0] - c0[0]:.1e}")
    assert abs(p1 - p0) <= tol
    np.testing.assert_allclose(w1, w0, rtol=0, atol=tol)
    assert abs(q1[0] - q0[0]) <= (1e-12 if prec == 0 else 1e-8) * q0[1]   # measured 3e-19, 1e-10
    assert abs(c1[0] - c0[0]) <= (1e-12 if prec == 0 else 1e-8) * c0[1]


@pytest.mark.parametrize("mode,sp", [(0, 0.25), (1, 1.0 / 6.0), (2, 0.25)], ids=["spatial", "4d", "temporal"])
def test_smearing_is_gauge_covariant(lgt, thermal, mode, sp):
    """Smearing the gauge transform of the 48^4 field gives the gauge transform
    of the smeared field (FP32 spatial smearing through the site-pair kernel)."""
    import fields as F
    prec, U = thermal
    g = F.haar(np.random.default_rng(8), (U.shape[0],))
    out = []
    for W in (U, F.gauge_transform(U, g, L)):
        with lgt.Engine(L, prec, 7) as e:
            e.upload(W)
            e.ape(mode, 2, sp)
            out.append(e.download())
    d = np.max(np.abs(F.gauge_transform(out[0], g, L) - out[1]))
    print(f"prec {prec} mode {mode}: max |g S(U) g+ - S(g U g+)| = {d:.1e}")
    assert d <= (1e-14 if prec == 0 else 1e-6)   # measured: 8e-16, 2.5e-7


def test_heatbath_ignores_the_links_it_replaces(lgt, thermal):
    """A heatbath sub-sweep draws each new link from its staple and its own
    Philox stream only: two fields that differ only in the links the sub-sweep
    replaces give the same result bit for bit."""
    import fields as F
    prec, U = thermal
    mu, p = 1, 1
    sites = np.flatnonzero(F.coords(L).sum(axis=0) % 2 == p)
    U2 = U.copy()
    U2[sites, mu] = F.haar(np.random.default_rng(2), (sites.size,))
    out = []
    for W in (U, U2):
        with lgt.Engine(L, prec, 13) as e:
            e.upload(W)
            e.heatbath_subsweep(mu, p, 3.09)
            out.append(e.download())
    np.testing.assert_array_equal(out[0], out[1])
    assert np.max(np.abs(out[0][sites, mu] - U[sites, mu])) > 0.1


def test_heatbath_only_sweep_leaves_unit_links(lgt, thermal):
    prec, U = thermal
    with lgt.Engine(L, prec, 21) as e:
        e.upload(U)
        e.update(1, 0, 3.09)
        V = e.download()
        P = e.plaquette()
    d = np.max(np.abs(np.linalg.norm(V, axis=-1) - 1.0))
    print(f"prec {prec}: heatbath-only sweep, max ||U| - 1| = {d:.1e}, plaquette {P:.6f}")
    assert d <= (1e-15 if prec == 0 else 2e-7)      # measured: 3.3e-16, 1.4e-7
    assert np.all(np.any(V != U, axis=-1))
    assert 0.6 < P < 0.8
