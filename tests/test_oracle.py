"""Pins the FP64 oracle (oracle/su2_oracle.c): against the reference's golden
fixtures (tests/golden, written from the reference's own C code), against the
reference build itself where it exists (oracle/_ref), and against closed forms
that a dropped term, a wrong sign, index or orientation would violate."""
import os

import numpy as np
import pytest

import fields as F
from oracle_lib import REF_SIZES, Oracle, RefLib, heatbath_batch, philox

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


# ---------------------------------------------------------------------------
# Philox4x32-10: known-answer vectors of the Random123 distribution (kat_vectors)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expect):
    assert philox(ctr, key) == expect


def test_u53_range():
    lib = Oracle.lib()
    assert lib.or_u53(0, 0) == 0.0
    top = lib.or_u53(0xffffffff, 0xffffffff)
    assert top < 1.0 and top == 1.0 - 2.0 ** -53


# ---------------------------------------------------------------------------
# quaternion algebra against 2x2 complex matrices (gauge.h:83-102,190-198)
# ---------------------------------------------------------------------------
def test_quaternion_products_match_matrices():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(64, 4))
    b = rng.normal(size=(64, 4))
    lib = Oracle.lib()
    import ctypes as C
    for i in range(64):
        o = np.zeros(4)
        od = np.zeros(4)
        pa, pb = a[i].copy(), b[i].copy()
        lib.or_qmul(pa.ctypes.data_as(C.POINTER(C.c_double)), pb.ctypes.data_as(C.POINTER(C.c_double)),
                    o.ctypes.data_as(C.POINTER(C.c_double)))
        lib.or_qmul_dag(pa.ctypes.data_as(C.POINTER(C.c_double)), pb.ctypes.data_as(C.POINTER(C.c_double)),
                        od.ctypes.data_as(C.POINTER(C.c_double)))
        Ma, Mb = F.to_matrix(a[i]), F.to_matrix(b[i])
        np.testing.assert_allclose(F.to_matrix(o), Ma @ Mb, atol=1e-13)
        np.testing.assert_allclose(F.to_matrix(od), Ma @ Mb.conj().T, atol=1e-13)


# ---------------------------------------------------------------------------
# golden fixtures written from the reference's own code
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["ref_l4.npz", "ref_a6448.npz"])
def test_oracle_matches_reference_fixture(name):
    fx = load(name)
    L = tuple(int(x) for x in fx["L"])
    o = Oracle(L)
    U0 = fx["U0"]
    assert o.plaquette(U0) == fx["plaq"]                       # sequential sum, bit-exact
    np.testing.assert_array_equal(o.wilson(U0, fx["wR"], fx["wT"], seq=True), fx["w"])
    assert o.qtop(U0, "improved", seq=True)[0] == fx["qimp"]
    assert o.qtop(U0, "clover", seq=True)[0] == fx["qclov"]
    for tag in ["ape_spatial", "ape_all", "ape_temporal"]:
        mode, nst, sp = fx[tag + "_args"]
        U = U0.copy()
        o.ape(U, int(mode), int(nst), float(sp))
        if tag in fx:
            np.testing.assert_array_equal(U, fx[tag])
        assert o.plaquette(U) == fx[tag + "_plaq"]
        np.testing.assert_array_equal(o.wilson(U, fx[tag + "_wR"], fx[tag + "_wT"], seq=True), fx[tag + "_w"])
        assert o.qtop(U, "improved", seq=True)[0] == fx[tag + "_qimp"]
        assert o.qtop(U, "clover", seq=True)[0] == fx[tag + "_qclov"]
    # one lexicographic heatbath sweep with the reference's own RANLUX uniforms
    U = U0.copy()
    acc, used = o.hb_sweep_lex(U, float(fx["beta"]), fx["hb_uniforms"])
    assert used == fx["hb_uniforms"].size
    if "hb_U1" in fx:
        np.testing.assert_array_equal(U, fx["hb_U1"])
    assert o.plaquette(U) == fx["hb_plaq"]


def test_oracle_heatbath_links_match_reference():
    fx = load("hb_links.npz")
    for beta in np.unique(fx["beta"]):
        sel = np.nonzero(fx["beta"] == beta)[0]
        uni = np.concatenate([fx["uniforms"][fx["offsets"][i]:fx["offsets"][i] + fx["lens"][i]] for i in sel])
        links, counts, used = heatbath_batch(fx["staples"][sel], float(beta), uni)
        assert used == uni.size
        np.testing.assert_array_equal(counts, fx["counts"][sel])
        np.testing.assert_array_equal(links, fx["links"][sel])
    assert fx["counts"].max() > 1   # the fixture exercises rejections


def test_oracle_random_matrix_matches_reference():
    fx = load("su2_random.npz")
    lib = Oracle.lib()
    import ctypes as C

    class UB(C.Structure):
        _fields_ = [("buf", C.POINTER(C.c_double)), ("n", C.c_int64), ("pos", C.c_int64), ("exhausted", C.c_int)]

    lib.or_su2_random_matrix.argtypes = [C.POINTER(UB), C.POINTER(C.c_double)]
    for u, link in zip(fx["uniforms"], fx["links"]):
        u = np.ascontiguousarray(u)
        ub = UB(u.ctypes.data_as(C.POINTER(C.c_double)), 3, 0, 0)
        out = np.zeros(4)
        lib.or_su2_random_matrix(C.byref(ub), out.ctypes.data_as(C.POINTER(C.c_double)))
        np.testing.assert_array_equal(out, link)


# ---------------------------------------------------------------------------
# the live reference build (present only where the reference tree was)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["l8", "a6448"])
def test_oracle_matches_reference_build(name):
    if not RefLib.available(name):
        pytest.skip("reference build absent (oracle/_ref)")
    ref = RefLib(name)
    ref.setup(2.3, 21)
    ref.lib.ref_init_field(1)
    ref.lib.ref_update(2)
    U = ref.get_field()
    o = Oracle(REF_SIZES[name])
    rng = np.random.default_rng(5)
    for n in rng.integers(0, o.vol, 64):
        for mu in range(4):
            np.testing.assert_array_equal(o.staple(U, int(n), mu), ref.staple(int(n), mu))
    tmax = REF_SIZES[name][0] - 1
    R, T = np.meshgrid(np.arange(1, 5), np.arange(1, tmax + 1), indexing="ij")
    np.testing.assert_array_equal(o.wilson(U, R.ravel(), T.ravel(), seq=True), ref.wilson(R.ravel(), T.ravel()))
    assert o.qtop(U, "improved", seq=True)[0] == ref.lib.ref_qtop_improved()
    for mode, sp in [(0, 0.25), (1, 1 / 6), (2, 0.4)]:
        ref.set_field(U)
        ref.lib.ref_ape(mode, 4, sp)
        V = U.copy()
        o.ape(V, mode, 4, sp)
        np.testing.assert_array_equal(V, ref.get_field())
    ref.set_field(U)
    st = ref.rng_state()
    ref.lib.ref_update(1)
    U1 = ref.get_field()
    ref.rng_reset(st)
    buf = ref.rng_draw(40 * 4 * o.vol)
    V = U.copy()
    o.hb_sweep_lex(V, 2.3, buf)
    np.testing.assert_array_equal(V, U1)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
LATS = [(4, 4, 4, 4), (6, 4, 4, 8), (4, 6, 8, 4)]


@pytest.mark.parametrize("L", LATS)
def test_cold_field(L):
    o = Oracle(L)
    U = o.cold()
    assert o.plaquette(U) == 1.0
    np.testing.assert_array_equal(o.staple(U, 3, 1), [6.0, 0, 0, 0])
    np.testing.assert_array_equal(o.wilson(U, [1, 2], [1, 3]), [1.0, 1.0])
    assert o.qtop(U)[0] == 0.0 and o.qtop(U, "clover")[0] == 0.0
    V = U.copy()
    o.ape(V, 1, 3, 0.3)
    np.testing.assert_array_equal(V, U)    # the cold field is a fixed point of APE


@pytest.mark.parametrize("L", LATS)
def test_abelian_plaquette_and_loops(L):
    rng = np.random.default_rng(3)
    o = Oracle(L)
    th = rng.uniform(-np.pi, np.pi, size=(o.vol, 4))
    U = F.abelian_field(th)
    assert o.plaquette(U, acc=True) == pytest.approx(F.abelian_plaquette(th, L), abs=1e-14)
    R, T = [1, 2, 3, 2], [1, 1, 2, 3]
    expect = [F.abelian_wilson(th, L, r, t) for r, t in zip(R, T)]
    np.testing.assert_allclose(o.wilson(U, R, T), expect, atol=1e-14)


@pytest.mark.parametrize("L,planes", [
    ((6, 4, 4, 8), ((0, 1), (2, 3))),
    ((6, 4, 4, 8), ((0, 2), (3, 1))),
    ((6, 4, 4, 8), ((0, 3), (1, 2))),
    ((8, 8, 8, 8), ((0, 1), (2, 3))),
])
def test_qtop_uniform_abelian_flux(L, planes):
    """Uniform fluxes phi in plane (a,b), psi in plane (c,d): every m x m leaf
    carries m^2 phi, so q(x) = 8 * (-2 f(phi) f(psi)) with f = 1.5 sin x -
    0.15 sin 4x + sin(9x)/90 for the improved clover, -2 sin phi sin psi per
    site for the 1x1 clover."""
    o = Oracle(L)
    (a, b), (c, d) = planes
    th1, phi = F.flux_angles(L, a, b, 1)
    th2, psi = F.flux_angles(L, c, d, 2)
    U = F.abelian_field(th1 + th2)
    q_imp = o.vol * 8 * (-2 * F.improved_f(phi) * F.improved_f(psi)) / (32 * np.pi ** 2)
    assert o.qtop(U, "improved")[0] == pytest.approx(q_imp, rel=1e-12)
    q_clov = o.vol * (-2 * np.sin(phi) * np.sin(psi)) / (4 * np.pi ** 2)
    assert o.qtop(U, "clover")[0] == pytest.approx(q_clov, rel=1e-12)


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (6, 4, 4, 8)])
def test_gauge_invariance(L):
    rng = np.random.default_rng(8)
    o = Oracle(L)
    U = o.hot(99)
    o.ape(U, 1, 2, 0.2)          # smoother than hot
    g = F.haar(rng, (o.vol,))
    V = F.gauge_transform(U, g, L)
    assert o.plaquette(V, acc=True) == pytest.approx(o.plaquette(U, acc=True), abs=1e-14)
    np.testing.assert_allclose(o.wilson(V, [1, 2, 3], [1, 2, 3]), o.wilson(U, [1, 2, 3], [1, 2, 3]), atol=1e-14)
    qu, su = o.qtop(U)
    qv, _ = o.qtop(V)
    assert abs(qu - qv) <= 1e-13 * su
    # APE is covariant: smearing commutes with the gauge transformation
    U2, V2 = U.copy(), V.copy()
    o.ape(U2, 0, 3, 0.25)
    o.ape(V2, 0, 3, 0.25)
    np.testing.assert_allclose(F.gauge_transform(U2, g, L), V2, atol=1e-13)


def test_hot_start_is_unit_and_isotropic():
    o = Oracle((8, 8, 8, 8))
    U = o.hot(1)
    np.testing.assert_allclose(np.linalg.norm(U, axis=-1), 1.0, atol=1e-15)
    # Haar: E[c_i] = 0, E[c_i^2] = 1/4, E[c0^4] = 1/8
    c = U.reshape(-1, 4)
    np.testing.assert_allclose(c.mean(axis=0), 0.0, atol=0.01)
    np.testing.assert_allclose((c ** 2).mean(axis=0), 0.25, atol=0.005)
    assert abs((c[:, 0] ** 4).mean() - 0.125) < 0.004


# ---------------------------------------------------------------------------
# heatbath sampler distribution and the overrelaxation map
# ---------------------------------------------------------------------------
def test_kp_sampler_distribution():
    """x0 of the KP draw at alpha=7 against p(x0) ~ sqrt(1-x0^2) exp(7 x0) and
    against the reference's own 10^8-draw histogram (LGT_Code/main_test/improved.dat)."""
    rng = np.random.default_rng(11)
    n = 400_000
    alpha = 7.0
    staples = np.tile([alpha, 0.0, 0.0, 0.0], (n, 1))     # |A| beta = alpha with beta = 1, V = 1
    links, counts, used = heatbath_batch(staples, 1.0, rng.random(12 * n))
    x0 = links[:, 0]
    edges = np.linspace(-1, 1, 201)
    hist, _ = np.histogram(x0, bins=edges)
    xs = np.linspace(-1, 1, 200001)
    dens = np.sqrt(1 - xs ** 2) * np.exp(alpha * xs)
    cdf = np.concatenate([[0], np.cumsum((dens[1:] + dens[:-1]) / 2)])
    cdf /= cdf[-1]
    expect = n * np.diff(np.interp(edges, xs, cdf))
    sel = expect > 50
    chi2 = (((hist - expect) ** 2) / expect)[sel].sum()
    dof = sel.sum()
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
    ref = np.loadtxt(os.path.join(GOLDEN, "kp_improved_alpha7.dat"))
    ref_counts = ref[:, 1] / 200.0 * n                   # value = count * M / N
    sel = ref_counts > 50
    chi2r = (((hist - ref_counts) ** 2) / ref_counts)[sel].sum()
    assert chi2r < sel.sum() + 5 * np.sqrt(2 * sel.sum()), chi2r
    # the direction of x is isotropic
    v = links[:, 1:] / np.linalg.norm(links[:, 1:], axis=1, keepdims=True)
    np.testing.assert_allclose(v.mean(axis=0), 0.0, atol=0.01)


def test_overrelaxation_closed_form():
    rng = np.random.default_rng(4)
    lib = Oracle.lib()
    import ctypes as C
    P = C.POINTER(C.c_double)
    for _ in range(200):
        A = rng.normal(size=4) * rng.uniform(0.1, 6)
        U = F.haar(rng, ())
        out = np.zeros(4)
        back = np.zeros(4)
        lib.or_or_link(A.ctypes.data_as(P), U.ctypes.data_as(P), out.ctypes.data_as(P))
        lib.or_or_link(A.ctypes.data_as(P), out.ctypes.data_as(P), back.ctypes.data_as(P))
        # keeps Re tr(U A), stays on S^3, is an involution
        assert F.qmul(out, A)[0] == pytest.approx(F.qmul(U, A)[0], abs=1e-14)
        assert np.linalg.norm(out) == pytest.approx(1.0, abs=1e-15)
        np.testing.assert_allclose(back, U, atol=1e-14)


def link_uniforms(seed, site, tag, sweep, n):
    """FP64 uniform stream of one link: Philox4x32-10 with key = seed and
    counter (cb site, mu | parity<<2 | kind<<3, sweep, draw/2)."""
    lib = Oracle.lib()
    out = []
    j = 0
    while len(out) < n:
        w = philox((site, tag, sweep, j), (seed & 0xffffffff, seed >> 32))
        out += [lib.or_u53(w[0], w[1]), lib.or_u53(w[2], w[3])]
        j += 1
    return np.array(out[:n])


@pytest.mark.parametrize("L,mu,p", [((4, 4, 4, 4), 2, 1), ((6, 4, 4, 8), 0, 0), ((6, 4, 4, 8), 3, 1)])
def test_checkerboard_subsweep_is_order_free(L, mu, p):
    """Links of one (mu, parity) never read each other: taking every staple
    from the untouched input (any update order) and drawing each link's
    uniforms from its own Philox counter reproduces the sequential sub-sweep."""
    o = Oracle(L)
    seed, sweep, beta = (7 << 32) | 77, 5, 3.09
    U = o.hot(5)
    A = U.copy()
    o.hb_subsweep(A, mu, p, beta, seed, sweep)
    x = F.coords(L)
    par = x.sum(axis=0) & 1
    sites = np.nonzero(par == p)[0]
    staples = np.array([o.staple(U, int(n), mu) for n in sites])
    B = U.copy()
    for n, S in zip(sites[::-1], staples[::-1]):
        uni = link_uniforms(seed, int(n) // 2, mu | (p << 2), sweep, 200)
        link, cnt, _ = heatbath_batch(S[None, :], beta, uni)
        B[n, mu] = link[0]
    np.testing.assert_array_equal(A, B)
    changed = np.nonzero(np.any(A != U, axis=(1, 2)))[0]
    assert set(changed) == set(sites)


# ---------------------------------------------------------------------------
# configuration file format (config_IO.c:40-117,537-621)
# ---------------------------------------------------------------------------
@pytest.mark.skipif(not RefLib.available("a6448"), reason="reference build missing")
def test_reference_config_file_layout(tmp_path):
    """The reference's writeConfig output parses with the test reader: header
    (DIM, SUN, extents t first, plaquette) and links site-major, c0..c3."""
    import config_format as CF
    ref = RefLib("a6448")
    ref.setup(3.09, 5)
    ref.lib.ref_init_field(1)
    U = ref.get_field()
    ref.write_config(tmp_path / "conf")
    dim, sun, L, plaq, V = CF.read(tmp_path / "conf")
    assert (dim, sun, L) == (4, 2, (6, 4, 4, 8))
    assert plaq == ref.lib.ref_plaquette()
    np.testing.assert_array_equal(V, U)
    # and the reference reads back what the test writer produced
    W = U.copy()
    W[:, :, :] = W[::-1]
    CF.write(tmp_path / "conf2", L, 0.0, W)
    ref.set_field(W)
    p = ref.lib.ref_plaquette()
    CF.write(tmp_path / "conf2", L, p, W)
    ref.lib.ref_init_field(0)
    ref.read_config(tmp_path / "conf2")
    np.testing.assert_array_equal(ref.get_field(), W)


# ---------------------------------------------------------------------------
# Metropolis (metro.c:65-109, exp_fct.c:137-151)
# ---------------------------------------------------------------------------
@pytest.mark.skipif(not RefLib.available("l4"), reason="reference build missing")
def test_metropolis_link_matches_reference():
    """The oracle's Metropolis hit, fed the RANLUX uniforms the reference
    consumes (replayed), gives the reference's new link bit for bit, accepted
    or not, for several proposal widths."""
    from oracle_lib import metropolis_batch
    ref = RefLib("l4")
    ref.setup(2.3, 11)
    ref.lib.ref_init_field(1)
    nacc = 0
    for k, (n, mu, eps) in enumerate([(0, 0, 0.2), (7, 1, 0.5), (33, 2, 1.0), (101, 3, 2.0), (200, 0, 0.05),
                                       (250, 3, 1.5), (17, 2, 0.8), (90, 1, 0.3)]):
        A = ref.staple(n, mu)
        U = ref.get_link(n, mu)
        st = ref.rng_state()
        r = ref.rng_draw(4)
        ref.rng_reset(st)
        a = ref.lib.ref_metropolis_link(n, mu, eps)
        new = ref.get_link(n, mu)
        out, acc = metropolis_batch(A[None], U[None], 2.3, eps, r[None])
        np.testing.assert_array_equal(out[0], new)
        assert acc[0] == int(a)
        nacc += acc[0]
    assert 0 < nacc < 8      # both branches are exercised


def test_metropolis_detailed_balance_two_state():
    """A hit with proposal X and its inverse -X (the proposal is symmetric):
    P(U -> U') exp(S-weight U) equals P(U' -> U) exp(S-weight U') at fixed
    staple (detailed balance of metro.c:80-108)."""
    from oracle_lib import metropolis_batch
    rng = np.random.default_rng(3)
    A = F.haar(rng, (1,))[0] * 2.7
    U = F.haar(rng, (1,))[0]
    beta, eps = 2.3, 0.7
    r = np.array([0.2, 0.9, 0.35, 0.0])        # accept for sure: r4 = 0
    out, acc = metropolis_batch(A[None], U[None], beta, eps, r[None])
    assert acc[0] == 1
    Up = out[0]
    rinv = np.array([1 - 0.2, 1 - 0.9, 1 - 0.35, 0.0])   # X -> -X
    back, _ = metropolis_batch(A[None], Up[None], beta, eps, rinv[None])
    np.testing.assert_allclose(back[0], U, atol=1e-14)   # expx(-X) = expx(X)^+
    w = lambda V: np.exp(beta / 2 * 2 * F.qmul(V, A)[0])
    p_fwd = min(1.0, w(Up) / w(U))
    p_bwd = min(1.0, w(U) / w(Up))
    assert p_fwd * w(U) == pytest.approx(p_bwd * w(Up), rel=1e-12)
