"""Parity of the HIP engine with the FP64 oracle and the reference fixtures,
through the C ABI (masterarbeit_amd/liblgt_su2.so).

Tolerances: FP64 per-link and per-site results are expected bit-identical to
the oracle (same operation order, no FMA contraction, IEEE division and
square root); global sums are double-double on the GPU and compensated on the
CPU, compared at 1e-12 relative (|dQ| <= 1e-13 sum|q| for the charge).  The
heatbath's transcendentals (log, cos, acos, sin) come from the device math
library, so heatbath links agree with glibc's to a few ulp with identical
trial counts.  FP32-storage results are compared with FP64 at the stated
FP32 tolerances."""
import os

import numpy as np
import pytest

import fields as F
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LATS = [(4, 4, 4, 4), (6, 4, 4, 8), (8, 8, 8, 8)]


@pytest.fixture(scope="module")
def lgt():
    import masterarbeit_amd as m
    m.load_library()
    return m


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def engine(lgt, L, prec=0, seed=1, U=None):
    e = lgt.Engine(L, prec, seed)
    if U is not None:
        e.upload(U)
    return e


def thermal_field(L, seed=3, nsweeps=2):
    o = Oracle(L)
    U = o.hot(seed)
    for s in range(nsweeps):
        o.engine_sweep(U, 3.09, 1, seed, s)
    return o, U


def ulp_diff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


# ---------------------------------------------------------------------------
# field life cycle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("L", LATS)
def test_upload_download_roundtrip(lgt, L):
    o = Oracle(L)
    U = o.hot(4)
    with engine(lgt, L, 0, U=U) as e:
        np.testing.assert_array_equal(e.download(), U)
    with engine(lgt, L, 1, U=U) as e:
        np.testing.assert_array_equal(e.download(), U.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("L", [(24, 24, 24, 24), (16, 32, 20, 22)])
def test_multi_chunk_transfers(lgt, L):
    """Host arrays larger than one 32 MB pinned chunk (42 and 59 MB, the last
    chunk partial) survive upload and download bit for bit, and so do the
    charge density (VOL doubles) and a gradient download."""
    rng = np.random.default_rng(8)
    vol = int(np.prod(L))
    U = rng.normal(size=(vol, 4, 4))
    U /= np.linalg.norm(U, axis=-1, keepdims=True)
    with engine(lgt, L, 0, U=U) as e:
        np.testing.assert_array_equal(e.download(), U)
        V = U[::-1].copy()
        e.upload(V)
        np.testing.assert_array_equal(e.download(), V)
        q, s, qx = e.qtop(1, density=True)
        assert qx.shape == (vol,) and np.isfinite(qx).all()
        # the density is q(x) before the 1/(4 pi^2) of the 1x1 clover charge
        assert np.sum(qx) == pytest.approx(q * 4 * np.pi ** 2, rel=1e-9, abs=1e-9 * s * 4 * np.pi ** 2)
        e.ape_vjp_forward(2, 0.25)
        cot = rng.normal(size=(vol, 4, 4))
        G = e.ape_vjp_backward(cot)
        np.testing.assert_array_equal(G[:, 0], cot[:, 0])    # temporal links pass through


@pytest.mark.parametrize("prec", [0, 1])
def test_cold_start(lgt, prec):
    L = (6, 4, 4, 8)
    with engine(lgt, L, prec) as e:
        U = e.download()
        np.testing.assert_array_equal(U, Oracle(L).cold())
        assert e.plaquette() == 1.0
        assert e.qtop()[0] == 0.0


@pytest.mark.parametrize("L", LATS)
def test_hot_start_matches_oracle(lgt, L):
    with engine(lgt, L, 0, seed=12345) as e:
        e.init_field(1)
        G = e.download()
    C = Oracle(L).hot(12345)
    assert np.max(np.abs(G - C)) <= 4e-16


# ---------------------------------------------------------------------------
# observables
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("L", LATS)
def test_plaquette_parity(lgt, L):
    o, U = thermal_field(L)
    with engine(lgt, L, 0, U=U) as e:
        p = e.plaquette()
    assert p == pytest.approx(o.plaquette(U, acc=True), rel=1e-14)
    assert p == pytest.approx(o.plaquette(U), rel=1e-12)


def test_plaquette_reference_fixture(lgt):
    for name in ["ref_l4.npz", "ref_a6448.npz"]:
        fx = load(name)
        L = tuple(int(x) for x in fx["L"])
        with engine(lgt, L, 0, U=fx["U0"]) as e:
            assert e.plaquette() == pytest.approx(float(fx["plaq"]), rel=1e-13)


@pytest.mark.parametrize("L", LATS)
@pytest.mark.parametrize("mode,sp", [(0, 0.25), (1, 1.0 / 6.0), (2, 0.3)])
def test_ape_parity(lgt, L, mode, sp):
    o, U = thermal_field(L)
    for nsteps in (1, 7):
        V = U.copy()
        o.ape(V, mode, nsteps, sp)
        with engine(lgt, L, 0, U=U) as e:
            e.ape(mode, nsteps, sp)
            G = e.download()
        d = np.max(np.abs(G - V))
        print(f"APE L={L} mode={mode} n={nsteps}: max|dU| = {d:.3g} exact={np.array_equal(G, V)}")
        assert d <= 1e-14


def test_ape_reference_fixture(lgt):
    fx = load("ref_l4.npz")
    L = tuple(int(x) for x in fx["L"])
    for tag in ["ape_spatial", "ape_all", "ape_temporal"]:
        mode, nst, sp = fx[tag + "_args"]
        with engine(lgt, L, 0, U=fx["U0"]) as e:
            e.ape(int(mode), int(nst), float(sp))
            G = e.download()
            assert np.max(np.abs(G - fx[tag])) <= 1e-14
            w = e.wilson_loops(fx[tag + "_wR"], fx[tag + "_wT"])
            np.testing.assert_allclose(w, fx[tag + "_w"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("L", LATS)
def test_wilson_loop_parity(lgt, L):
    o, U = thermal_field(L)
    o.ape(U, 0, 4, 0.25)     # smeared: loops well above rounding noise
    tmax = min(L[0] - 1, 6)
    R, T = np.meshgrid(np.arange(1, 5), np.arange(1, tmax + 1), indexing="ij")
    R, T = R.ravel(), T.ravel()
    with engine(lgt, L, 0, U=U) as e:
        w = e.wilson_loops(R, T)
        w_perm = e.wilson_loops(R[::-1], T[::-1])       # order of the list does not matter
    ref = o.wilson(U, R, T)
    np.testing.assert_allclose(w, ref, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(w_perm[::-1], w)


@pytest.mark.parametrize("prec", [0, 1])
def test_wilson_loop_unsorted_repeated_and_wrapping(lgt, prec):
    """Loops listed in no order, repeated, with R both below and beyond the
    spatial extent 4 (the line wraps around the lattice): each value equals
    the oracle's for its own (R, T), and a repeated loop gets one value."""
    L = (6, 4, 4, 8)
    o, U = thermal_field(L)
    o.ape(U, 0, 4, 0.25)
    R = np.array([5, 1, 3, 5, 2, 4, 1, 6, 3])
    T = np.array([2, 1, 5, 2, 3, 5, 5, 1, 1])
    ref = o.wilson(U, R, T)
    with engine(lgt, L, prec, U=U) as e:
        w = e.wilson_loops(R, T)
        single = np.array([e.wilson_loops([r], [t])[0] for r, t in zip(R, T)])
    tol = dict(rtol=1e-12, atol=1e-15) if prec == 0 else dict(rtol=0, atol=2e-6)
    np.testing.assert_allclose(w, ref, **tol)
    np.testing.assert_array_equal(single, w)
    assert w[0] == w[3]


@pytest.mark.parametrize("prec", [0, 1])
def test_wilson_loop_long_t(lgt, prec):
    """T beyond 8: the loop kernel sums eight T at a time across a wave, so
    this covers a full batch, a partial one (T up to 11), T values missing
    from a group, and spatial lines advanced by several links between groups
    (R = 1, 2, 5, 6)."""
    L = (12, 4, 4, 6)
    o, U = thermal_field(L)
    o.ape(U, 0, 4, 0.25)
    R = np.array([1] * 11 + [2, 2, 2, 5, 5, 6])
    T = np.array(list(range(1, 12)) + [2, 9, 11, 10, 3, 8])
    ref = o.wilson(U, R, T)
    with engine(lgt, L, prec, U=U) as e:
        w = e.wilson_loops(R, T)
        single = np.array([e.wilson_loops([r], [t])[0] for r, t in zip(R, T)])
    tol = dict(rtol=1e-12, atol=1e-15) if prec == 0 else dict(rtol=0, atol=2e-6)
    np.testing.assert_allclose(w, ref, **tol)
    np.testing.assert_array_equal(single, w)


def test_wilson_loop_reference_fixture(lgt):
    for name in ["ref_l4.npz", "ref_a6448.npz"]:
        fx = load(name)
        L = tuple(int(x) for x in fx["L"])
        with engine(lgt, L, 0, U=fx["U0"]) as e:
            np.testing.assert_allclose(e.wilson_loops(fx["wR"], fx["wT"]), fx["w"], rtol=1e-12, atol=1e-15)


def test_wilson_loop_rejects_long_t(lgt):
    with engine(lgt, (4, 4, 4, 4), 0) as e:
        with pytest.raises(lgt.LgtError):
            e.wilson_loops([1], [4])                       # wilson.c:290-291: T <= L0 - 1
        for R in (0, 4097, 1 << 30):                       # spatial lines of 1..4096 links
            with pytest.raises(lgt.LgtError):
                e.wilson_loops([1, R], [1, 1])
        assert np.isfinite(e.wilson_loops([4096], [1])[0])


@pytest.mark.parametrize("L", LATS)
@pytest.mark.parametrize("kind", [0, 1])
def test_qtop_parity(lgt, L, kind):
    o, U = thermal_field(L)
    o.ape(U, 1, 3, 1.0 / 6.0)
    q_ref, s_ref, qx_ref = o.qtop(U, "improved" if kind == 0 else "clover", density=True)
    with engine(lgt, L, 0, U=U) as e:
        q, s, qx = e.qtop(kind, density=True)
    d = np.max(np.abs(qx - qx_ref))
    print(f"Q L={L} kind={kind}: Q={q:.6g} ref={q_ref:.6g} max|dq(x)|={d:.3g} exact={np.array_equal(qx, qx_ref)}")
    assert d <= 1e-15 * max(1.0, np.max(np.abs(qx_ref)))
    assert abs(q - q_ref) <= 1e-13 * s_ref
    assert s == pytest.approx(s_ref, rel=1e-12)


@pytest.mark.parametrize("L", LATS)
def test_fp32_storage_against_oracle(lgt, L):
    """FP32 link storage against the FP64 oracle on the same field (rounded to
    FP32 on upload): the plaquette (FP64 products and sums of the FP32 links),
    7 steps of every APE mode, and both charges of the 4D-smeared field."""
    o, U = thermal_field(L)
    U32 = U.astype(np.float32).astype(np.float64)
    with engine(lgt, L, 1, U=U) as e:
        dp = abs(e.plaquette() - o.plaquette(U32, acc=True))
        dU = []
        for mode, sp in ((0, 0.25), (1, 1.0 / 6.0), (2, 0.3)):
            e.upload(U)
            V = U32.copy()
            e.ape(mode, 7, sp)
            o.ape(V, mode, 7, sp)
            dU.append(np.max(np.abs(e.download() - V)))
        e.upload(U)
        e.ape(1, 3, 1.0 / 6.0)
        W = U32.copy()
        o.ape(W, 1, 3, 1.0 / 6.0)
        dq = []
        for kind in (0, 1):
            q, s = e.qtop(kind)[:2]
            q_ref, s_ref = o.qtop(W, "improved" if kind == 0 else "clover")[:2]
            dq.append(abs(q - q_ref) / s_ref)
    print(f"FP32 L={L}: |dP| {dp:.1e}, APE max|dU| {max(dU):.1e}, |dQ|/sum|q| {max(dq):.1e}")
    assert dp <= 1e-14          # FP64 products of the same FP32 links; measured: 0
    assert max(dU) <= 1e-6      # measured: 1.7e-7
    assert max(dq) <= 1e-6      # measured: 6.2e-8


def test_qtop_reference_fixture(lgt):
    for name in ["ref_l4.npz", "ref_a6448.npz"]:
        fx = load(name)
        L = tuple(int(x) for x in fx["L"])
        with engine(lgt, L, 0, U=fx["U0"]) as e:
            q, s = e.qtop(0)
            assert abs(q - float(fx["qimp"])) <= 1e-13 * s
            q, s = e.qtop(1)
            assert abs(q - float(fx["qclov"])) <= 1e-13 * s


def test_qtop_abelian_flux_closed_form(lgt):
    L = (6, 4, 4, 8)
    th1, phi = F.flux_angles(L, 0, 1, 1)
    th2, psi = F.flux_angles(L, 2, 3, 2)
    U = F.abelian_field(th1 + th2)
    vol = int(np.prod(L))
    with engine(lgt, L, 0, U=U) as e:
        q, _ = e.qtop(0)
        qc, _ = e.qtop(1)
    assert q == pytest.approx(vol * 8 * (-2 * F.improved_f(phi) * F.improved_f(psi)) / (32 * np.pi ** 2), rel=1e-12)
    assert qc == pytest.approx(vol * (-2 * np.sin(phi) * np.sin(psi)) / (4 * np.pi ** 2), rel=1e-12)


# ---------------------------------------------------------------------------
# update: per-link harnesses, sub-sweep replays, full sweeps
# ---------------------------------------------------------------------------
def test_heatbath_links_match_reference(lgt):
    """The reference's localHeatbathUpdate, fed the RANLUX uniforms it consumed,
    against the engine's FP64 per-link heatbath fed the same uniforms."""
    fx = load("hb_links.npz")
    links = np.zeros_like(fx["links"])
    counts = np.zeros_like(fx["counts"])
    for beta in np.unique(fx["beta"]):
        sel = np.nonzero(fx["beta"] == beta)[0]
        l, c = lgt.heatbath_harness(fx["staples"][sel], float(beta), fx["uniforms"], fx["offsets"][sel],
                                    fx["lens"][sel])
        links[sel], counts[sel] = l, c
    np.testing.assert_array_equal(counts, fx["counts"])
    u = ulp_diff(links, fx["links"])
    exact = np.mean(np.all(links == fx["links"], axis=1))
    print(f"heatbath per link: {exact * 100:.1f}% bit-identical, max {u.max():.0f} ulp")
    assert np.max(np.abs(links - fx["links"])) <= 1e-14


@pytest.mark.parametrize("prec", [0, 1])
def test_heatbath_draw_distribution(lgt, prec):
    """On a cold field every staple is 6 * 1, so a heatbath sub-sweep draws
    each new link straight from the Kennedy-Pendleton density at alpha = 6 beta
    (heatbath.c:76-123): at beta = 7/6, x0 = U.c0 of 663 552 links (four
    sub-sweeps of a 24^4 lattice) follows p(x0) ~ sqrt(1 - x0^2) exp(7 x0) and
    the reference's own 10^8-draw histogram (LGT_Code/main_test/improved.dat)
    by chi^2, and the direction of (c1, c2, c3) is isotropic, in both storages
    (FP32: hardware log2/cos and 24-bit uniforms)."""
    L = (24, 24, 24, 24)
    alpha = 7.0
    even = F.coords(L).sum(axis=0) % 2 == 0
    links = []
    with engine(lgt, L, prec, seed=77) as e:
        for k in range(4):
            e.init_field(0)
            e.sweep_index = k
            e.heatbath_subsweep(0, 0, alpha / 6.0)
            links.append(e.download()[even, 0])
    links = np.concatenate(links)
    n = len(links)
    np.testing.assert_allclose(np.linalg.norm(links, axis=1), 1.0, atol=1e-15 if prec == 0 else 3e-7)
    edges = np.linspace(-1, 1, 201)
    hist, _ = np.histogram(links[:, 0], bins=edges)
    xs = np.linspace(-1, 1, 200001)
    dens = np.sqrt(1 - xs ** 2) * np.exp(alpha * xs)
    cdf = np.concatenate([[0], np.cumsum((dens[1:] + dens[:-1]) / 2)])
    cdf /= cdf[-1]
    expect = n * np.diff(np.interp(edges, xs, cdf))
    sel = expect > 50
    chi2 = (((hist - expect) ** 2) / expect)[sel].sum()
    dof = sel.sum()
    ref = np.loadtxt(os.path.join(GOLDEN, "kp_improved_alpha7.dat"))
    ref_counts = ref[:, 1] / 200.0 * n                    # value = count * M / N
    selr = ref_counts > 50
    chi2r = (((hist - ref_counts) ** 2) / ref_counts)[selr].sum()
    v = links[:, 1:] / np.linalg.norm(links[:, 1:], axis=1, keepdims=True)
    print(f"prec {prec}: chi2 {chi2:.1f} / {dof} bins, against the reference histogram {chi2r:.1f} / "
          f"{selr.sum()}, direction mean {np.abs(v.mean(axis=0)).max():.1e}")
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
    assert chi2r < selr.sum() + 5 * np.sqrt(2 * selr.sum()), chi2r
    np.testing.assert_allclose(v.mean(axis=0), 0.0, atol=0.01)
    np.testing.assert_allclose((v ** 2).mean(axis=0), 1.0 / 3.0, atol=0.01)


def test_heatbath_degenerate_branch(lgt):
    fx = load("su2_random.npz")
    n = len(fx["links"])
    staples = np.zeros((n, 4))
    lens = np.full(n, 3, dtype=np.int32)
    offsets = np.arange(n, dtype=np.int32) * 3
    links, counts = lgt.heatbath_harness(staples, 3.09, fx["uniforms"].ravel(), offsets, lens)
    assert np.all(counts == 0)
    assert np.max(np.abs(links - fx["links"])) <= 1e-15


def test_overrelax_harness(lgt):
    rng = np.random.default_rng(2)
    n = 512
    A = rng.normal(size=(n, 4)) * rng.uniform(0.1, 6, size=(n, 1))
    U = F.haar(rng, (n,))
    out = lgt.overrelax_harness(A, U)
    import ctypes as C
    lib = Oracle.lib()
    P = C.POINTER(C.c_double)
    ref = np.zeros_like(out)
    for i in range(n):
        a, u, r = A[i].copy(), U[i].copy(), np.zeros(4)
        lib.or_or_link(a.ctypes.data_as(P), u.ctypes.data_as(P), r.ctypes.data_as(P))
        ref[i] = r
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (6, 4, 4, 8)])
def test_heatbath_subsweep_replay(lgt, L):
    """Each parallel (mu, parity) sub-sweep equals the oracle's sequential
    replay with the same Philox streams (checkerboard independence)."""
    seed = 99
    o, U = thermal_field(L)
    with engine(lgt, L, 0, seed=seed, U=U) as e:
        V = U.copy()
        for mu in range(4):
            for p in range(2):
                o.hb_subsweep(V, mu, p, 3.09, seed, 0)
                e.heatbath_subsweep(mu, p, 3.09)
                G = e.download()
                d = np.max(np.abs(G - V))
                assert d <= 1e-13, (mu, p, d)
                V = G.copy()                 # continue from the engine's field


@pytest.mark.parametrize("prec", [0, 1])
def test_overrelax_subsweep_replay(lgt, prec):
    """Each (mu, parity) overrelaxation sub-sweep equals the oracle's
    sequential one (FP64: to rounding of the closed form; FP32: to FP32
    rounding, unit links)."""
    L = (6, 4, 4, 8)
    o, U = thermal_field(L)
    with engine(lgt, L, prec, seed=3, U=U) as e:
        V = e.download()
        for mu in range(4):
            for p in range(2):
                o.or_subsweep(V, mu, p)
                e.overrelax_subsweep(mu, p)
                G = e.download()
                d = np.max(np.abs(G - V))
                assert d <= (1e-14 if prec == 0 else 2e-6), (mu, p, d)
                V = G.copy()
        np.testing.assert_allclose(np.linalg.norm(V, axis=-1), 1.0, atol=1e-14 if prec == 0 else 1e-6)


@pytest.mark.parametrize("prec", [0, 1])
def test_chain_resumes_from_sweep_index(lgt, prec):
    """The RNG is keyed by the sweep index: a chain stopped after one sweep and
    resumed on a fresh engine from its field and sweep index reproduces the
    uninterrupted chain bit for bit."""
    L = (6, 4, 4, 8)
    o, U = thermal_field(L)
    with engine(lgt, L, prec, seed=21, U=U) as e:
        assert e.extents() == (L, prec)
        assert e.device_bytes() >= 4 * np.prod(L) * (32 if prec == 0 else 16)
        e.update(3, 1, 3.09)
        full = e.download()
        assert e.sweep_index == 3
    with engine(lgt, L, prec, seed=21, U=U) as e:
        e.update(1, 1, 3.09)
        mid = e.download()
    with engine(lgt, L, prec, seed=21, U=mid) as e:
        assert e.sweep_index == 0
        e.sweep_index = 1
        e.update(2, 1, 3.09)
        np.testing.assert_array_equal(e.download(), full)
    with engine(lgt, L, prec, seed=21, U=mid) as e:
        e.update(2, 1, 3.09)   # resumed at index 0: a different chain
        assert not np.array_equal(e.download(), full)


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (6, 4, 4, 8)])
@pytest.mark.parametrize("n_or", [0, 1, 3])
def test_update_matches_oracle_chain(lgt, L, n_or):
    seed = 2024
    o, U = thermal_field(L, seed=5, nsweeps=1)
    V = U.copy()
    with engine(lgt, L, 0, seed=seed, U=U) as e:
        acc = e.update(2, n_or, 3.09, acceptance=True)
        G = e.download()
    acc_ref = 0.0
    for s in range(2):
        acc_ref += o.engine_sweep(V, 3.09, n_or, seed, s) / 2
    d = np.max(np.abs(G - V))
    print(f"update L={L} n_or={n_or}: max|dU| = {d:.3g}, acceptance {acc:.6f} vs {acc_ref:.6f}")
    assert d <= 1e-12
    assert acc == pytest.approx(acc_ref, rel=1e-12)
    np.testing.assert_allclose(np.linalg.norm(G, axis=-1), 1.0, atol=1e-15)


@pytest.mark.parametrize("L", [(2, 2, 2, 4), (4, 2, 6, 4), (2, 6, 2, 4)])
def test_smallest_extents_match_oracle(lgt, L):
    """Extent 2 (x+1 and x-1 are the same site) and lattices smaller than one
    256-site tile: sweeps, every APE mode, plaquette, loops up to T = L0 - 1
    with wrapping R, and both charges against the oracle; FP32 smearing (the
    site-pair kernel in spatial smearing) and charges within FP32 rounding."""
    seed = 11
    o, U = thermal_field(L, seed=4, nsweeps=1)
    V = U.copy()
    with engine(lgt, L, 0, seed=seed, U=U) as e:
        acc = e.update(2, 1, 3.09, acceptance=True)
        G = e.download()
        acc_ref = sum(o.engine_sweep(V, 3.09, 1, seed, s) for s in range(2)) / 2
        assert np.max(np.abs(G - V)) <= 1e-12
        assert acc == pytest.approx(acc_ref, rel=1e-12)
        assert e.plaquette() == pytest.approx(o.plaquette(V, acc=True), abs=1e-14)
        R = [1, 2, 3, 1, 3]
        T = [1, 1, 1, L[0] - 1, L[0] - 1]
        np.testing.assert_allclose(e.wilson_loops(R, T), o.wilson(V, R, T), rtol=1e-12, atol=1e-15)
        for mode, sp in ((0, 0.25), (1, 1.0 / 6.0), (2, 0.25)):
            e.upload(V)
            W = V.copy()
            e.ape(mode, 3, sp)
            o.ape(W, mode, 3, sp)
            assert np.max(np.abs(e.download() - W)) <= 1e-14
        for kind in (0, 1):
            q, s_abs = e.qtop(kind)
            q_ref, s_ref = o.qtop(W, "improved" if kind == 0 else "clover")[:2]
            assert abs(q - q_ref) <= 1e-13 * s_ref
    with engine(lgt, L, 1, seed=seed, U=V) as e32:
        assert abs(e32.plaquette() - o.plaquette(V, acc=True)) < 1e-6
        for mode, sp in ((0, 0.25), (1, 1.0 / 6.0), (2, 0.25)):
            e32.upload(V)
            W = V.copy()
            e32.ape(mode, 3, sp)
            o.ape(W, mode, 3, sp)
            assert np.max(np.abs(e32.download() - W)) <= 2e-6
        for kind in (0, 1):
            q, s_abs = e32.qtop(kind)
            q_ref, s_ref = o.qtop(W, "improved" if kind == 0 else "clover")[:2]
            assert abs(q - q_ref) <= 1e-5 * s_ref
        e32.update(2, 1, 3.09)
        np.testing.assert_allclose(np.linalg.norm(e32.download(), axis=-1), 1.0, atol=1e-6)


@pytest.mark.parametrize("prec", [0, 1])
def test_results_independent_of_tile_order(lgt, prec):
    """Every tile order (XCD-contiguous, time-fastest, both, neither, each also
    reversed, and update launches alternating or not) gives bit-identical
    sweeps, acceptance, smearing and observables; repeated runs too."""
    L = (8, 16, 16, 16)    # a time slice holds 8 whole tiles: every order applies
    res = []
    for order, group, alt in ((-1, 0, 1), (0, 0, 1), (1, 0, 1), (2, 0, 1), (3, 0, 1), (4, 0, 1), (5, 0, 1),
                              (7, 0, 1), (5, 2, 1), (5, 4, 1), (5, 8, 1), (-1, 0, 0), (8, 0, 0), (13, 0, 0),
                              (13, 4, 1), (15, 0, 1), (-1, 0, 1)):
        with engine(lgt, L, prec, seed=7) as e:
            e.set_option(0, order)
            e.set_option(lgt.engine.OPT_GROUP, group)
            e.set_option(lgt.engine.OPT_ALTERNATE, alt)
            e.init_field(1)
            a = e.update(3, 1, 3.09, acceptance=True)
            e.update(1, 0, 3.09)
            e.metropolis_update(1, 3.09, 0.5)
            e.ape(0, 2, 0.25)
            res.append((e.download(), a, e.plaquette(), e.wilson_loops([1, 2], [2, 3]).tolist(), e.qtop(0)))
    for r in res[1:]:
        np.testing.assert_array_equal(r[0], res[0][0])
        assert r[1:] == res[0][1:]


def test_group_size_must_divide_time_slice(lgt):
    with engine(lgt, (8, 16, 16, 16), 0, seed=7) as e:
        with pytest.raises(lgt.LgtError):
            e.set_option(lgt.engine.OPT_GROUP, 3)
        e.set_option(lgt.engine.OPT_GROUP, 4)
        for opt, bad in ((lgt.engine.OPT_ALTERNATE, 2), (lgt.engine.OPT_TILE_ORDER, 16),
                         (lgt.engine.OPT_TILE_ORDER, -2)):
            with pytest.raises(lgt.LgtError):
                e.set_option(opt, bad)


# ---------------------------------------------------------------------------
# FP32 storage against FP64
# ---------------------------------------------------------------------------
def test_fp32_observables_track_fp64(lgt):
    L = (8, 8, 8, 8)
    o, U = thermal_field(L)
    with engine(lgt, L, 0, U=U) as e64, engine(lgt, L, 1, U=U) as e32:
        assert e32.plaquette() == pytest.approx(e64.plaquette(), abs=2e-7)
        e64.ape(0, 20, 0.25)
        e32.ape(0, 20, 0.25)
        d = np.max(np.abs(e32.download() - e64.download()))
        assert d < 2e-5, d
        R, T = [1, 2, 3, 4], [1, 2, 3, 4]
        np.testing.assert_allclose(e32.wilson_loops(R, T), e64.wilson_loops(R, T), atol=2e-6)
        e64.ape(1, 5, 1.0 / 6.0)
        e32.ape(1, 5, 1.0 / 6.0)
        q64, s64 = e64.qtop(0)
        q32, s32 = e32.qtop(0)
        assert abs(q32 - q64) < 1e-4 * s64


def test_fp32_update_stays_unitary(lgt):
    L = (8, 8, 8, 8)
    with engine(lgt, L, 1, seed=3) as e:
        e.init_field(1)
        acc = e.update(5, 1, 3.09, acceptance=True)
        U = e.download()
        p = e.plaquette()
    np.testing.assert_allclose(np.linalg.norm(U, axis=-1), 1.0, atol=3e-7)
    assert 0.5 < acc <= 1.0
    assert 0.3 < p < 0.8


def test_acceptance_over_many_sweeps(lgt):
    """The per-sweep acceptance is read back in batches of 1024 sweeps: one
    call of 1030 sweeps gives the field of 1030 calls of one sweep bit for bit
    and the mean of their acceptances; Metropolis sweeps give the same field
    with and without acceptance."""
    L = (4, 4, 4, 4)
    n = 1030
    with engine(lgt, L, 1, seed=21) as a, engine(lgt, L, 1, seed=21) as b:
        a.init_field(1)
        b.init_field(1)
        acc = a.update(n, 1, 3.09, acceptance=True)
        each = [b.update(1, 1, 3.09, acceptance=True) for _ in range(n)]
        assert acc == pytest.approx(np.mean(each), rel=1e-12)
        np.testing.assert_array_equal(a.download(), b.download())
        acc = a.metropolis_update(n, 3.09, 0.5, acceptance=True)
        b.metropolis_update(n, 3.09, 0.5)
        np.testing.assert_array_equal(a.download(), b.download())
        assert 0.3 < acc < 1.0


def test_fp32_heatbath_only_projects_on_store(lgt):
    """An FP32 heatbath-only sweep stores every link projected onto SU(2)
    instead of projecting the field after the sweep: it equals a replay of
    its eight sub-sweeps, each followed by an FP64 projection of the links it
    updated, to FP32 rounding; its acceptance equals the replay's."""
    L = (6, 4, 4, 8)
    _, U = thermal_field(L)
    parity = F.coords(L).sum(axis=0) % 2
    with engine(lgt, L, 1, seed=8, U=U) as e:
        acc = e.update(1, 0, 3.09, acceptance=True)
        G = e.download()
    acc_rep = 0.0
    with engine(lgt, L, 1, seed=8, U=U) as e:
        for mu in range(4):
            for p in range(2):
                acc_rep += e.heatbath_subsweep(mu, p, 3.09, acceptance=True) / (4 * np.prod(L))
                V = e.download()
                s = parity == p
                V[s, mu] /= np.linalg.norm(V[s, mu], axis=-1, keepdims=True)
                e.upload(V)
        R = e.download()
    d = np.max(np.abs(G - R))
    print(f"FP32 heatbath-only: max|fused - replay| = {d:.3g}, acceptance {acc:.6f} vs {acc_rep:.6f}")
    assert d <= 2e-6
    assert acc == pytest.approx(acc_rep, abs=1e-6)
    np.testing.assert_allclose(np.linalg.norm(G, axis=-1), 1.0, atol=2e-7)


def test_fp32_heatbath_only_equilibrium(lgt):
    """FP32 and FP64 heatbath-only chains sample the same Wilson action: their
    mean plaquettes at beta = 3.09 agree within 4 sigma (8^4, independent
    chains of 600 measured sweeps after 200 of thermalisation)."""
    from masterarbeit_amd.studies import tau_int
    means = []
    for prec, seed in ((0, 41), (1, 42)):
        with engine(lgt, (8, 8, 8, 8), prec, seed=seed) as e:
            e.init_field(1)
            e.update(200, 0, 3.09)
            p = []
            for _ in range(600):
                e.update(1, 0, 3.09)
                p.append(e.plaquette())
        p = np.array(p)
        means.append((p.mean(), tau_int(p)[1]))
    (m1, e1), (m2, e2) = means
    assert abs(m1 - m2) <= 4 * np.hypot(e1, e2), means


# ---------------------------------------------------------------------------
# Metropolis (metro.c:65-109, exp_fct.c:137-151)
# ---------------------------------------------------------------------------
def test_metropolis_harness_matches_oracle(lgt):
    """One hit per link with injected uniforms: the GPU's FP64 links and
    decisions equal the oracle's (which equals the reference's)."""
    from oracle_lib import metropolis_batch
    rng = np.random.default_rng(12)
    n = 4096
    A = F.haar(rng, (n,)) * rng.uniform(0.5, 3.5, size=(n, 1))
    U = F.haar(rng, (n,))
    r = rng.uniform(size=(n, 4))
    for eps in (0.1, 0.6, 2.0):
        out, acc = lgt.metropolis_harness(A, U, 2.3, eps, r)
        ref, racc = metropolis_batch(A, U, 2.3, eps, r)
        np.testing.assert_array_equal(acc, racc)
        assert np.max(np.abs(out - ref)) <= 1e-15
        assert 0.05 < acc.mean() < 0.999


@pytest.mark.parametrize("prec", [0, 1])
def test_metropolis_update_matches_oracle(lgt, prec):
    L = (6, 4, 4, 8)
    o, U = thermal_field(L)
    V = U.copy()
    acc_ref = o.metropolis_update(V, 2, 3.09, 0.4, 5, 0)
    with engine(lgt, L, prec, seed=5, U=U) as e:
        acc = e.metropolis_update(2, 3.09, 0.4, acceptance=True)
        G = e.download()
        assert e._lib.lgt_get_sweep_index(e._h) == 2
    if prec == 0:
        assert np.max(np.abs(G - V)) <= 1e-12
        assert acc == pytest.approx(acc_ref, abs=1e-15)
    else:
        np.testing.assert_allclose(np.linalg.norm(G, axis=-1), 1.0, atol=1e-6)
        assert abs(acc - acc_ref) < 0.02


@pytest.mark.parametrize("prec", [0, 1])
def test_metropolis_equilibrium_matches_heatbath(lgt, prec):
    """Metropolis and heatbath chains sample the same Wilson action: their
    mean plaquettes at beta = 3.09 agree within 4 sigma (8^4, independent
    chains of 600 measured sweeps after 200 of thermalisation), in both
    storages."""
    L = (8, 8, 8, 8)
    means = []
    for algo in ("hb", "metro"):
        with engine(lgt, L, prec, seed=31 if algo == "hb" else 32) as e:
            e.init_field(1)
            p = []
            for k in range(800):
                if algo == "hb":
                    e.update(1, 0, 3.09)
                else:
                    e.metropolis_update(1, 3.09, 0.5)
                if k >= 200:
                    p.append(e.plaquette())
        p = np.array(p)
        from masterarbeit_amd.studies import tau_int
        tau, err = tau_int(p)
        means.append((p.mean(), err))
    (m1, e1), (m2, e2) = means
    assert abs(m1 - m2) <= 4 * np.hypot(e1, e2), (means)
