"""Adjoint of N-step APE smearing (spatial, all directions, temporal) with
the Wilson-loop objective L = sum_x sum_d 2 Re[W_{RxT}]_0, through the C ABI.

Checked against the FP64 CPU chain rule (oracle/su2_oracle.c, or_ape_vjp,
which scatters each staple term's factor gradients while the engine gathers
them per link), against central finite differences along su(2) directions
computed by the engine itself (double-double loop sums), and across
checkpoint segmentations (sqrt(N) default, uneven segments, one segment)."""
import numpy as np
import pytest

import fields as F
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lgt():
    import masterarbeit_amd as m
    m.load_library()
    return m


def su2_exp(X, eps):
    n = np.linalg.norm(X[1:])
    return np.concatenate([[np.cos(eps * n)], np.sin(eps * n) * X[1:] / n])


def thermal(L, seed=3):
    o = Oracle(L)
    U = o.hot(seed)
    for s in range(2):
        o.engine_sweep(U, 3.09, 1, seed, s)
    return o, U


@pytest.mark.parametrize("L", [(6, 4, 4, 8), (8, 8, 8, 8)])
def test_wloop_sum_and_gradient(lgt, L):
    o, U = thermal(L)
    R, T = 2, 3
    with lgt.Engine(L, 0, 1) as e:
        e.upload(U)
        hi, lo = e.wloop_sum(R, T)
        G = e.wloop_grad(R, T)
    Lref, Gref = o.ape_vjp(U, 0, 0.25, R, T)
    assert hi + lo == pytest.approx(Lref, rel=1e-14, abs=1e-12)
    np.testing.assert_allclose(G, Gref, rtol=0, atol=1e-12 * np.abs(Gref).max())


MODES = [(0, 0.25), (1, 1.0 / 6.0), (2, 0.25)]     # (mode, sp): spatial, all directions, temporal


@pytest.mark.parametrize("mode,sp", MODES)
@pytest.mark.parametrize("L,nsteps,segment", [((6, 4, 4, 8), 7, 0), ((6, 4, 4, 8), 7, 3), ((8, 8, 8, 8), 5, 5),
                                              ((8, 8, 8, 8), 9, 2), ((2, 2, 2, 4), 3, 0), ((4, 2, 6, 4), 4, 2)])
def test_vjp_matches_cpu_chain_rule(lgt, L, nsteps, segment, mode, sp):
    """Extents 2 (x + 1 and x - 1 the same site, loops of R = 2 closing on
    themselves; L0 = 2 allows T = 1 only) and lattices smaller than one
    256-site tile included."""
    o, U = thermal(L)
    R, T = 2, min(2, L[0] - 1)
    Lref, Gref = o.ape_vjp(U, nsteps, sp, R, T, mode)
    with lgt.Engine(L, 0, 1) as e:
        e.upload(U)
        e.ape_vjp_forward(nsteps, sp, segment, mode=mode)
        hi, lo = e.wloop_sum(R, T)
        e.wloop_grad(R, T, host=False)
        G = e.ape_vjp_backward()
        V = e.download()
    W = U.copy()
    o.ape(W, mode, nsteps, sp)
    np.testing.assert_array_equal(V, W)              # the forward pass is the plain smearing, bit for bit
    assert hi + lo == pytest.approx(Lref, rel=1e-13)
    scale = np.abs(Gref).max()
    err = np.abs(G - Gref).max() / scale
    print(f"VJP mode={mode} L={L} N={nsteps} seg={segment}: max rel err {err:.2e}")
    assert err < 1e-11


def test_vjp_modes_share_an_engine(lgt):
    """Alternating modes on one engine (each re-lays out the buffers) gives
    the same gradients as fresh engines."""
    L = (6, 4, 4, 8)
    o, U = thermal(L)
    res = {}
    with lgt.Engine(L, 0, 1) as e:
        for mode, sp in MODES + MODES[:1]:
            e.upload(U)
            e.ape_vjp_forward(5, sp, mode=mode)
            e.wloop_grad(2, 2, host=False)
            res.setdefault(mode, []).append(e.ape_vjp_backward())
    for mode, sp in MODES:
        with lgt.Engine(L, 0, 1) as e:
            e.upload(U)
            e.ape_vjp_forward(5, sp, mode=mode)
            e.wloop_grad(2, 2, host=False)
            G = e.ape_vjp_backward()
        for Gm in res[mode]:
            np.testing.assert_array_equal(Gm, G)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("mode,sp", MODES)
def test_vjp_segmentations_agree_bitwise(lgt, prec, mode, sp):
    """Every checkpoint segmentation runs the same kernels on the same inputs:
    sqrt(N) checkpoints, uneven segments, every step kept (segment = N) and a
    repeated backward give the same gradient bit for bit, in both precisions
    (FP32 spatial smearing through the site-pair kernel)."""
    L = (8, 8, 8, 8)
    o, U = thermal(L)
    N = 9
    res = []
    with lgt.Engine(L, prec, 1) as e:
        for seg in (0, 2, 4, N, 1):
            e.upload(U)
            e.ape_vjp_forward(N, sp, seg, mode=mode)
            e.wloop_grad(2, 3, host=False)
            res.append(e.ape_vjp_backward())
        # a second backward of the last forward recomputes its top segment
        e.wloop_grad(2, 3, host=False)
        res.append(e.ape_vjp_backward())
    for G in res[1:]:
        np.testing.assert_array_equal(G, res[0])


@pytest.mark.parametrize("mode,sp", MODES)
def test_vjp_fp32_tracks_fp64(lgt, mode, sp):
    """The FP32-storage adjoint (FP32 links, cotangents and products) against
    the FP64 one on the same thin field: 20 steps, the 3x3 loop objective."""
    L = (8, 8, 8, 8)
    o, U = thermal(L)
    G = []
    for prec in (0, 1):
        with lgt.Engine(L, prec, 1) as e:
            e.upload(U)
            e.ape_vjp_forward(20, sp, mode=mode)
            e.wloop_grad(3, 3, host=False)
            G.append(e.ape_vjp_backward())
    scale = np.abs(G[0]).max()
    d = np.abs(G[1] - G[0]).max() / scale
    print(f"mode {mode}: FP32 adjoint max deviation {d:.2e} of max |G| = {scale:.3g}")
    assert d <= 5e-6   # measured: 8.9e-7 (spatial), 3.1e-7 (4D), 2.8e-7 (temporal)


@pytest.mark.parametrize("mode,sp", MODES)
def test_vjp_explicit_cotangent(lgt, mode, sp):
    L = (6, 4, 4, 8)
    o, U = thermal(L)
    rng = np.random.default_rng(4)
    cot = rng.normal(size=(o.vol, 4, 4))
    with lgt.Engine(L, 0, 1) as e:
        e.upload(U)
        e.ape_vjp_forward(4, sp, mode=mode)
        G1 = e.ape_vjp_backward(cot)
        G2 = e.ape_vjp_backward(cot)                 # repeatable, the forward state is kept
    np.testing.assert_array_equal(G1, G2)
    if mode == 0:
        np.testing.assert_array_equal(G1[:, 0], cot[:, 0])   # temporal links pass through
    # linearity in the cotangent
    with lgt.Engine(L, 0, 1) as e:
        e.upload(U)
        e.ape_vjp_forward(4, sp, mode=mode)
        G3 = e.ape_vjp_backward(2.5 * cot)
    np.testing.assert_allclose(G3, 2.5 * G1, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("mode,sp", MODES)
@pytest.mark.parametrize("prec,rtol,eps", [(0, 2e-6, 1e-4), (1, 2e-2, 1e-2)])
def test_vjp_matches_gpu_finite_differences(lgt, prec, rtol, eps, mode, sp):
    L = (8, 8, 8, 8)
    o, U = thermal(L)
    R, T, nsteps = 3, 3, 12
    with lgt.Engine(L, prec, 1) as e:
        e.upload(U)
        e.ape_vjp_forward(nsteps, sp, mode=mode)
        e.wloop_grad(R, T, host=False)
        G = e.ape_vjp_backward()
        rng = np.random.default_rng(9)
        for i in range(8):
            n, mu = int(rng.integers(o.vol)), i % 4
            X = np.concatenate([[0.0], rng.normal(size=3)])
            vals = []
            for sgn in (1, -1):
                V = U.copy()
                V[n, mu] = F.qmul(su2_exp(X, sgn * eps), U[n, mu])
                e.upload(V)
                e.ape(mode, nsteps, sp)
                hi, lo = e.wloop_sum(R, T)
                vals.append((hi, lo))
            fd = ((vals[0][0] - vals[1][0]) + (vals[0][1] - vals[1][1])) / (2 * eps)
            an = float(np.dot(G[n, mu], F.qmul(X, U[n, mu])))
            assert an == pytest.approx(fd, rel=rtol, abs=1e-7 if prec == 0 else 1e-3), (mode, n, mu, an, fd)


def test_vjp_400_steps_48_finite_differences(lgt):
    """The 48^4 case: FP64, thermalised from seed 11, 400 spatial APE steps
    (alpha = 0.5), L = sum of 5x5 loops, sqrt(N) = 20 checkpoints; central
    differences (eps = 1e-4) along 16 random su(2) directions on random links,
    with the loop sums kept in double-double."""
    L = (48, 48, 48, 48)
    R = T = 5
    with lgt.Engine(L, 0, 11) as e:
        e.init_field(1)
        e.update(100, 1, 3.09)
        U = e.download()
        e.ape_vjp_forward(400, 0.25)
        e.wloop_grad(R, T, host=False)
        G = e.ape_vjp_backward()
        rng = np.random.default_rng(11)
        worst = 0.0
        for _ in range(16):
            n, mu = int(rng.integers(U.shape[0])), int(rng.integers(4))
            X = np.concatenate([[0.0], rng.normal(size=3)])
            vals = []
            for sgn in (1, -1):
                old = U[n, mu].copy()
                U[n, mu] = F.qmul(su2_exp(X, sgn * 1e-4), old)
                e.upload(U)
                U[n, mu] = old
                e.ape(0, 400, 0.25)
                vals.append(e.wloop_sum(R, T))
            fd = ((vals[0][0] - vals[1][0]) + (vals[0][1] - vals[1][1])) / 2e-4
            an = float(np.dot(G[n, mu], F.qmul(X, U[n, mu])))
            rel = abs(an - fd) / max(abs(fd), 1e-30)
            worst = max(worst, rel)
            assert rel < 1e-6 or abs(an - fd) < 1e-9, (n, mu, an, fd)
        print(f"48^4 APE-400 adjoint vs finite differences: worst relative deviation {worst:.2e}")
    # the same gradient in FP32 storage (FP32 links and arithmetic, FP64
    # checkpoint-boundary and objective sums) against the FP64 one
    with lgt.Engine(L, 1, 11) as e:
        e.upload(U)
        e.ape_vjp_forward(400, 0.25)
        e.wloop_grad(R, T, host=False)
        G32 = e.ape_vjp_backward()
    scale = float(np.max(np.abs(G)))
    dmax = float(np.max(np.abs(G32 - G))) / scale
    drms = float(np.sqrt(np.mean((G32 - G) ** 2) / np.mean(G ** 2)))
    print(f"48^4 APE-400 adjoint, FP32 storage - FP64: max {dmax:.2e} of the largest component, rms {drms:.2e} "
          f"relative")
    assert dmax <= 1e-4 and drms <= 1e-5


def test_vjp_rejects_bad_calls(lgt):
    """Argument and order errors come back as errors, not as results."""
    L = (4, 4, 4, 4)
    with lgt.Engine(L, 0, 1) as e:
        e.init_field(1)
        with pytest.raises(lgt.LgtError):
            e.ape_vjp_backward(np.zeros((e.vol, 4, 4)))          # no forward pass yet
        for bad in (dict(nsteps=0, sp=0.25), dict(nsteps=3, sp=0.25, mode=3), dict(nsteps=3, sp=0.25, segment=-1)):
            with pytest.raises(lgt.LgtError):
                e.ape_vjp_forward(**bad)
        e.ape_vjp_forward(3, 0.25)
        with pytest.raises(lgt.LgtError):
            e.ape_vjp_backward()                                  # no cotangent: neither passed nor left by wloop_grad
        with pytest.raises(lgt.LgtError):
            e.wloop_grad(1, L[0])                                 # T must stay below the time extent
        e.wloop_grad(1, 2, host=False)
        G = e.ape_vjp_backward()
        assert np.isfinite(G).all()
