"""The engine's chains against the reference's own equilibrated chains at 24^4,
beta = 3.09: the means of W(1,1), W(3,3), W(5,5), W(7,7), W(9,9), W(10,16)
over configurations 10 000..49 999 of the reference's heatbath chains from a
cold and from a hot start, and of W(10,16) after 100 spatial APE steps over
its configurations 1000..9999 (Thermalisation/24x24x24x24_b3.09_WL_*.dat;
statistics in tests/golden/ref_therm24.json, written by
tests/golden/gen_therm24_fixture.py, which states why the first 10 000
configurations are left out; reference_values() states how the two starts
are combined).

Each precision runs four independent chains (own seeds, cold and hot starts
alternating) of heatbath + 1 overrelaxation sweeps: 2000 thermalisation
sweeps, then the six loops every second sweep (6000 measurements) and the
smeared W(10,16) (sp = 0.25, alpha = 0.5) on every 20th measured
configuration.  Errors follow LGT_Code/Autocorrelation/t_int.c (integrated
autocorrelation time, summed up to the first negative autocorrelation).

The bound: |z| < 2 per loop, z = (engine - reference) / combined error, for
7 loops x 2 precisions.  The chains are deterministic (counter-based RNG), so
the z values are fixed numbers, printed by the test and quoted in the design
notes."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LOOPS = [(1, 1), (3, 3), (5, 5), (7, 7), (9, 9), (10, 16)]


def reference_values():
    """(mean, error) per loop: the mean of the cold and hot means (over the
    configurations the fixture keeps), the error from the means of blocks of
    1000 configurations of each start; for the smeared loop at least half the
    cold - hot difference (those configurations include the hot chain's
    thermalisation: 2 tau_int = 28)."""
    with open(os.path.join(HERE, "golden", "ref_therm24.json")) as f:
        d = json.load(f)
    out = {}
    for key, src in [(f"{R},{T}", d["unsmeared"]) for R, T in LOOPS] + [("ape100", d["ape100"])]:
        stats = src["10,16"] if key == "ape100" else src[key]
        c, h = stats["cold"], stats["hot"]
        mean = 0.5 * (c["mean"] + h["mean"])
        be = [np.std(s["block_means"], ddof=1) / np.sqrt(len(s["block_means"])) for s in (c, h)]
        err = 0.5 * float(np.hypot(*be))
        if key == "ape100":
            err = max(err, 0.5 * abs(c["mean"] - h["mean"]))
        out[key] = (mean, err)
    return out


@pytest.mark.parametrize("prec,seeds", [(0, (51, 52, 53, 54)), (1, (61, 62, 63, 64))], ids=["fp64", "fp32"])
def test_chains_match_reference_equilibrium_24(prec, seeds):
    import masterarbeit_amd as lgt
    from masterarbeit_amd.studies import tau_int
    lgt.load_library()
    L = (24, 24, 24, 24)
    R = [r for r, _ in LOOPS]
    T = [t for _, t in LOOPS]
    ref = reference_values()
    means, errs = [], []
    for k, seed in enumerate(seeds):
        with lgt.Engine(L, prec, seed) as e, lgt.Engine(L, prec, seed) as m:
            e.init_field(k % 2)
            e.update(2000, 1, 3.09)
            w, wa = [], []
            for i in range(6000):
                e.update(2, 1, 3.09)
                w.append(e.wilson_loops(R, T))
                if i % 20 == 0:
                    m.upload(e.download())
                    m.ape(lgt.APE_SPATIAL, 100, 0.25)
                    wa.append(m.wilson_loops([10], [16])[0])
        series = [np.array(w)[:, j] for j in range(len(LOOPS))] + [np.array(wa)]
        means.append([x.mean() for x in series])
        errs.append([tau_int(x)[1] for x in series])
        print(f"chain {seed} ({'hot' if k % 2 else 'cold'} start): W = {np.array(means[-1])}", flush=True)
    mean = np.mean(means, axis=0)
    err = np.sqrt(np.sum(np.array(errs) ** 2, axis=0)) / len(seeds)
    bad = []
    for j, key in enumerate([f"{r},{t}" for r, t in LOOPS] + ["ape100"]):
        r, re = ref[key]
        z = (mean[j] - r) / np.hypot(err[j], re)
        print(f"{'W(' + key + ')' if key != 'ape100' else 'APE-100 W(10,16)':>18}: engine {mean[j]:.7g} +- {err[j]:.2g}"
              f"  reference {r:.7g} +- {re:.2g}  z = {z:+.2f}")
        if abs(z) >= 2.0:
            bad.append((key, z))
    assert not bad, bad
