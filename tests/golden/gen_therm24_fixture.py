"""Writes tests/golden/ref_therm24.json: the equilibrium statistics of the
reference's own 24^4 chains at beta = 3.09 (container-only; reads the
reference's data files as text, runs nothing of the reference).

Inputs: /root/reference/Thermalisation/24x24x24x24_b3.09_WL_<R>_<T>_<start>.dat,
one Wilson loop W(R,T) (measureWilsonLoop normalisation: mean of
(1/2) Re tr over sites and the three (t, d) planes, LGT_Code/modules/obs/
wilson.c:268-324) per stored configuration of a heatbath chain of 50 000
configurations, from a cold and from a hot start; and ..._10_16_<start>_NAPE_100.dat,
W(10,16) after 100 APE smearing steps on the first 10 000 configurations of
the same chains (24x24x24x24_SU2_b3.090000_id1_<start>_NAPE_100.meas).

For each series the first `skip` configurations are dropped, then the
mean, the spread per configuration, the integrated autocorrelation time and
the error of the mean follow LGT_Code/Autocorrelation/t_int.c:31-69
(tau = 1/2 + sum_k Gamma_k/Gamma_0 up to the first negative Gamma_k;
error = sqrt(2 tau Gamma_0/(N-1))).  Means of blocks of 1000 configurations
are kept too, so that a blocking error can be formed without the series.

Skip: the reference's own analysis drops 1000 configurations
(Thermalisation/test2.py), but its hot chain is still thermalising after them
for the larger loops: over configurations 1000..9999 its W(5,5) averages
0.044163 against 0.04428-0.04429 over later stretches (W(7,7): 0.009419
against 0.00950), and its cold chain approaches from above (W(5,5) 0.044314).
The unsmeared series therefore drop 10 000 of their 50 000 configurations.
The smeared series hold only configurations 1..10 000 and drop 1000; the
two starts then differ by 1.4e-4, which the test takes into its error.

Usage: python tests/golden/gen_therm24_fixture.py
"""
import json
import os

import numpy as np

REF = "/root/reference/Thermalisation"
HERE = os.path.dirname(os.path.abspath(__file__))
LOOPS = [(1, 1), (3, 3), (5, 5), (7, 7), (9, 9), (10, 16)]
BLOCK = 1000


def tau_int(x):
    x = np.asarray(x, dtype=np.float64)
    n = len(x)

    def gamma(k):
        a, b = x[:n - k], x[k:]
        return float(np.mean((a - a.mean()) * (b - b.mean())))

    g0 = gamma(0)
    t = 0.5 * g0
    for k in range(1, n):
        gk = gamma(k)
        if gk < 0.0:
            break
        t += gk
    tau = t / g0
    return tau, float(np.sqrt(2.0 * tau * g0 / (n - 1)))


def stats(path, skip):
    x = np.loadtxt(path)
    n_all = len(x)
    x = x[skip:]
    tau, err = tau_int(x)
    nb = len(x) // BLOCK
    blocks = x[:nb * BLOCK].reshape(nb, BLOCK).mean(axis=1)
    return {"file": "Thermalisation/" + os.path.basename(path), "n_configs": n_all, "skip": skip, "n": len(x),
            "mean": float(x.mean()), "sigma": float(x.std(ddof=1)), "tau_int": tau, "err": err,
            "block_means": [float(b) for b in blocks]}


def main():
    out = {"lattice": [24, 24, 24, 24], "beta": 3.09, "block": BLOCK, "unsmeared": {}, "ape100": {}}
    for R, T in LOOPS:
        for start in ("cold", "hot"):
            s = stats(os.path.join(REF, f"24x24x24x24_b3.09_WL_{R}_{T}_{start}.dat"), 10000)
            out["unsmeared"].setdefault(f"{R},{T}", {})[start] = s
            print(f"W({R},{T}) {start}: {s['mean']:.6e} +- {s['err']:.1e} (2 tau_int {2 * s['tau_int']:.2f})")
    for start in ("cold", "hot"):
        s = stats(os.path.join(REF, f"24x24x24x24_b3.09_WL_10_16_{start}_NAPE_100.dat"), 1000)
        out["ape100"].setdefault("10,16", {})[start] = s
        print(f"APE-100 W(10,16) {start}: {s['mean']:.6e} +- {s['err']:.1e} (2 tau_int {2 * s['tau_int']:.2f})")
    with open(os.path.join(HERE, "ref_therm24.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
