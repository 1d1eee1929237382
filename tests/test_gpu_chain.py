"""Statistical parity of the engine's Markov chain with the reference's:
the reference's own 12^4 heatbath chain at beta = 3.09 (tests/golden/
ref_chain_l12.npz, written by gen_chain_fixture.py from the reference's
gaugefieldUpdate) against the engine's checkerboard heatbath + overrelaxation
chain, FP64 and FP32 storage: <plaquette> over 900 thermalised sweeps, and
W(R,T) (R,T = 1..6) after 20 spatial APE steps on every 10th configuration.
Errors of the means come from binning (bins of 50 sweeps, 5 measurements),
which absorbs the autocorrelation; the chains are seeded, so the outcome of
each comparison is fixed."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def binned(x, nbin):
    x = np.asarray(x)
    n = len(x) // nbin * nbin
    b = x[:n].reshape(-1, nbin, *x.shape[1:]).mean(axis=1)
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(len(b))


@pytest.fixture(scope="module")
def ref_chain():
    return dict(np.load(os.path.join(GOLDEN, "ref_chain_l12.npz"), allow_pickle=False))


@pytest.mark.parametrize("prec", [0, 1])
def test_chain_matches_reference_statistically(ref_chain, prec):
    import masterarbeit_amd as lgt
    fx = ref_chain
    beta, therm, every = float(fx["beta"]), int(fx["therm"]), int(fx["every"])
    nsweep = len(fx["plaq"])
    R, T = fx["R"], fx["T"]
    plaq, w = [], []
    with lgt.Engine((12, 12, 12, 12), prec, 2024) as e:
        for s in range(nsweep):
            e.update(1, 1, beta)
            plaq.append(e.plaquette())
            if s + 1 >= therm and (s + 1) % every == 0:
                U = e.download()
                e.ape(lgt.APE_SPATIAL, 20, 0.25)
                w.append(e.wilson_loops(R, T))
                e.upload(U)
    p_gpu, dp_gpu = binned(plaq[therm:], 50)
    p_ref, dp_ref = binned(fx["plaq"][therm:], 50)
    z = (p_gpu - p_ref) / np.hypot(dp_gpu, dp_ref)
    print(f"prec={prec}: <P> engine {p_gpu:.6f}({dp_gpu:.6f}) reference {p_ref:.6f}({dp_ref:.6f}) -> {z:+.2f} sigma")
    assert abs(z) < 2.0
    w_gpu, dw_gpu = binned(np.array(w), 5)
    w_ref, dw_ref = binned(fx["w"], 5)
    zw = (w_gpu - w_ref) / np.hypot(dw_gpu, dw_ref)
    print(f"prec={prec}: W(R,T) deviations in sigma: max |z| = {np.abs(zw).max():.2f}, "
          f"rms = {np.sqrt(np.mean(zw ** 2)):.2f}")
    assert np.sqrt(np.mean(zw ** 2)) < 2.0 and np.abs(zw).max() < 3.5
