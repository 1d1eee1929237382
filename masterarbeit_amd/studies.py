#!/usr/bin/env python3
"""Physics studies on the GPU engine: thermalisation, FP32-storage precision,
topology and a 96^4 production chain.

    python -m masterarbeit_amd.studies c1|c2|c3|c4|c5|c6|c7|c8|all [--quick]

Each study prints one JSON line on stdout (progress on stderr).  The field is
resident on the device throughout; host transfers happen only where a
configuration is handed from one engine to another (same configuration in FP32
and FP64 storage) or restored after a destructive measurement.

  c1  12^4, beta 3.09, cold start: 100 x (heatbath + overrelaxation) sweeps
      with the plaquette after every sweep; 20 spatial APE steps (alpha 0.5,
      sp 0.25); W(R,T) for R,T in 1..6.
  c2  24^4, hot start (seed 1): 1000 sweeps in FP32 storage and in FP64;
      <P> of both chains with autocorrelation-corrected errors; on a common
      configuration every 100 sweeps, spatial APE to N = 100..500 and W for
      R 1..10 x T 1..16 in both storages: the FP32 - FP64 differences.
  c3  48^4, seed 7: 500 sweeps (FP64); 4D APE (sp 1/6) for 15 and 30 steps and
      the improved charge Q, on the same configuration in FP64 and FP32
      storage; |dQ|.
  c4  48^4, FP64, seed 11, 100 sweeps: the adjoint of 400 spatial APE steps
      (sqrt(N) = 20 checkpoints) for L = sum of the 5x5 loops, against central
      differences (eps = 1e-4, loop sums in double-double) along 16 random
      su(2) directions on random links.
  c5  96^4, FP32 storage, hot start (seed 3): 200 sweeps; 400 spatial APE
      steps and W at R = T in {5, 7, 9}; 30 steps of 4D APE and Q; device
      memory against the 288 GB of one MI355X.
  c6  48^4: <P> of FP32-storage and FP64 chains from 8 seeds (7..14, or
      --seeds N), 200 thermalisation + 1500 measured sweeps each; the
      differences against their autocorrelation-corrected errors.
  c7  48^4: unsmeared W(1,1), W(5,5), W(7,7), W(9,9) of FP64 and FP32-storage
      chains (4 seeds, hot and cold starts, 3000 thermalisation sweeps, 1000
      measurements 4 sweeps apart) against the reference's own chain at
      beta 3.09 (W_REF).
  c8  24^4: W(R,T) for (1,1), (3,3), (5,5), (7,7), (9,9), (10,16) of FP64 and FP32-storage
      chains (cold and hot starts, 2000 thermalisation sweeps, --nmeas measurements --every
      sweeps apart) and W(10,16) after 100 spatial APE steps, against the reference's own
      equilibrated 24^4 chains (W_REF24).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import engine as E

BETA = 3.09
SP_SPATIAL = 0.25          # alpha = 0.5 in the reference's input convention
SP_4D = 1.0 / 6.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tau_int(x):
    """Integrated autocorrelation time and error of the mean, summing the
    normalised autocorrelation function until its first negative value
    (LGT_Code/Autocorrelation/t_int.c:31-69): tau = 1/2 + sum_k Gamma_k/Gamma_0,
    error = sqrt(2 tau) * sqrt(Gamma_0 / (N - 1))."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)

    def gamma(k):
        a, b = x[:n - k], x[k:]
        return float(np.mean((a - a.mean()) * (b - b.mean())))

    g0 = gamma(0)
    # a constant series (up to the rounding of its mean) has no fluctuation
    if g0 <= 1e-24 * float(np.mean(x * x)):
        return 0.5, 0.0
    t = 0.5 * g0
    for k in range(1, n):
        gk = gamma(k)
        if gk < 0.0:
            break
        t += gk
    tau = t / g0
    return tau, float(np.sqrt(2.0 * tau) * np.sqrt(g0 / (n - 1)))


def loops(Rs, Ts):
    R, T = np.meshgrid(np.asarray(Rs), np.asarray(Ts), indexing="ij")
    return R.ravel(), T.ravel()


def c1(quick=False):
    L = (12, 12, 12, 12)
    t0 = time.perf_counter()
    with E.Engine(L, E.FP64, 1) as e:
        e.init_field(0)
        plaq = []
        for _ in range(100):
            e.update(1, 1, BETA)
            plaq.append(e.plaquette())
        R, T = loops(range(1, 7), range(1, 7))
        e.ape(E.APE_SPATIAL, 20, SP_SPATIAL)
        w = e.wilson_loops(R, T)
    tail = np.array(plaq[50:])
    tau, err = tau_int(tail)
    return {"study": "c1", "lattice": L, "beta": BETA, "start": "cold", "precision": "fp64",
            "plaquette_per_sweep": plaq, "plaquette_mean_50_100": float(tail.mean()), "err": err, "tau_int": tau,
            "ape": {"steps": 20, "sp": SP_SPATIAL},
            "wloops": {f"{r},{t}": float(v) for r, t, v in zip(R, T, w)}, "seconds": time.perf_counter() - t0}


def c2(quick=False):
    L = (24, 24, 24, 24)
    nsweeps = 200 if quick else 1000
    therm, every = 100, 100
    R, T = loops(range(1, 11), range(1, 17))
    Ns = [100, 200, 300, 400, 500]
    t0 = time.perf_counter()
    e64 = E.Engine(L, E.FP64, 1)
    e32 = E.Engine(L, E.FP32, 1)
    m64 = E.Engine(L, E.FP64, 2)      # measurement engines (APE is destructive)
    m32 = E.Engine(L, E.FP32, 2)
    e64.init_field(1)
    e32.init_field(1)
    p64, p32, dW = [], [], []
    W64s = []
    for s in range(1, nsweeps + 1):
        e64.update(1, 1, BETA)
        e32.update(1, 1, BETA)
        p64.append(e64.plaquette())
        p32.append(e32.plaquette())
        if s > therm and s % every == 0:
            U = e64.download()
            m64.upload(U)
            m32.upload(U)
            d_n, w_n = [], []
            done = 0
            for n in Ns:
                m64.ape(E.APE_SPATIAL, n - done, SP_SPATIAL)
                m32.ape(E.APE_SPATIAL, n - done, SP_SPATIAL)
                done = n
                w64 = m64.wilson_loops(R, T)
                w32 = m32.wilson_loops(R, T)
                d_n.append(w32 - w64)
                w_n.append(w64)
            dW.append(np.array(d_n))
            W64s.append(np.array(w_n))
            log(f"[c2] sweep {s}: P64 {p64[-1]:.6f} P32 {p32[-1]:.6f} max|dW| {np.abs(dW[-1]).max():.3g}")
    for x in (e64, e32, m64, m32):
        x.close()
    a64, a32 = np.array(p64[therm:]), np.array(p32[therm:])
    t64, s64 = tau_int(a64)
    t32, s32 = tau_int(a32)
    dW = np.array(dW)        # [conf, N, loop]
    W64s = np.array(W64s)
    rel = np.abs(dW) / np.maximum(np.abs(W64s), 1e-300)
    per_n = {}
    for j, n in enumerate(Ns):
        per_n[str(n)] = {
            "max_abs_dW": float(np.abs(dW[:, j]).max()),
            "max_rel_dW": float(rel[:, j].max()),
            "rel_dW_W55": float(rel[:, j, list(zip(R, T)).index((5, 5))].max()),
            "W55_mean_fp64": float(W64s[:, j, list(zip(R, T)).index((5, 5))].mean()),
        }
    dP = float(a32.mean() - a64.mean())
    sig = float(np.hypot(s32, s64))
    return {"study": "c2", "lattice": L, "beta": BETA, "sweeps": nsweeps, "thermalisation": therm,
            "plaquette": {"fp64": float(a64.mean()), "err64": s64, "tau64": t64,
                          "fp32": float(a32.mean()), "err32": s32, "tau32": t32,
                          "dP": dP, "dP_over_sigma": dP / sig if sig > 0 else None},
            "ape_wloops_same_config": {"configs": int(dW.shape[0]), "loops": "R 1..10 x T 1..16", "per_N": per_n},
            "seconds": time.perf_counter() - t0}


def c3(quick=False):
    L = (48, 48, 48, 48)
    nsweeps = 50 if quick else 500
    t0 = time.perf_counter()
    with E.Engine(L, E.FP64, 7) as e:
        e.init_field(1)
        for s in range(0, nsweeps, 50):
            e.update(min(50, nsweeps - s), 1, BETA)
            log(f"[c3] sweep {s + 50}: P {e.plaquette():.6f}")
        plaq = e.plaquette()
        U = e.download()
    out = {"study": "c3", "lattice": L, "beta": BETA, "sweeps": nsweeps, "plaquette": plaq, "steps": {}}
    res = {}
    for prec, name in ((E.FP64, "fp64"), (E.FP32, "fp32")):
        with E.Engine(L, prec, 7) as e:
            e.upload(U)
            done = 0
            for n in (15, 30):
                e.ape(E.APE_ALL, n - done, SP_4D)
                done = n
                q, qa = e.qtop(E.Q_IMPROVED)
                res[(name, n)] = (q, qa)
    del U
    for n in (15, 30):
        q64, qa64 = res[("fp64", n)]
        q32, _ = res[("fp32", n)]
        out["steps"][str(n)] = {"Q_fp64": q64, "Q_fp32": q32, "abs_dQ": abs(q32 - q64), "sum_abs_q_fp64": qa64}
    out["seconds"] = time.perf_counter() - t0
    return out


def su2_exp(X, eps):
    """exp(eps X) for a pure-imaginary quaternion X = (0, x)."""
    n = np.linalg.norm(X[1:])
    return np.concatenate([[np.cos(eps * n)], np.sin(eps * n) * X[1:] / n])


def qmul(a, b):
    """Quaternion product a b (the engine's link algebra)."""
    return np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                     a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2],
                     a[0] * b[2] + a[2] * b[0] - a[3] * b[1] + a[1] * b[3],
                     a[0] * b[3] + a[3] * b[0] - a[1] * b[2] + a[2] * b[1]])


def c4(quick=False):
    L = (48, 48, 48, 48)
    R = T = 5
    N = 100 if quick else 400
    ndir = 4 if quick else 16
    t0 = time.perf_counter()
    with E.Engine(L, E.FP64, 11) as e:
        e.init_field(1)
        e.update(100, 1, BETA)
        U = e.download()
        t1 = time.perf_counter()
        e.ape_vjp_forward(N, SP_SPATIAL)
        e.sync()
        t_fwd = time.perf_counter() - t1
        e.wloop_grad(R, T, host=False)
        t1 = time.perf_counter()
        G = e.ape_vjp_backward()
        t_bwd = time.perf_counter() - t1
        rng = np.random.default_rng(11)
        dev = []
        for _ in range(ndir):
            n, mu = int(rng.integers(U.shape[0])), int(rng.integers(4))
            X = np.concatenate([[0.0], rng.normal(size=3)])
            vals = []
            for sgn in (1, -1):
                old = U[n, mu].copy()
                U[n, mu] = qmul(su2_exp(X, sgn * 1e-4), old)
                e.upload(U)
                U[n, mu] = old
                e.ape(E.APE_SPATIAL, N, SP_SPATIAL)
                vals.append(e.wloop_sum(R, T))
            fd = ((vals[0][0] - vals[1][0]) + (vals[0][1] - vals[1][1])) / 2e-4
            an = float(np.dot(G[n, mu], qmul(X, U[n, mu])))
            dev.append(abs(an - fd) / max(abs(fd), 1e-30))
            log(f"[c4] link ({n},{mu}): adjoint {an:.10g} differences {fd:.10g}")
    return {"study": "c4", "lattice": L, "beta": BETA, "precision": "fp64", "ape_steps": N, "sp": SP_SPATIAL,
            "loop": [R, T], "directions": ndir, "worst_relative_deviation": float(max(dev)),
            "median_relative_deviation": float(np.median(dev)), "forward_s": t_fwd, "backward_s": t_bwd,
            "seconds": time.perf_counter() - t0}


def c6(quick=False, nseeds=8):
    L = (48, 48, 48, 48)
    seeds = (7, 8) if quick else tuple(range(7, 7 + nseeds))
    nmeas = 300 if quick else 1500
    t0 = time.perf_counter()
    rows = []
    for seed in seeds:
        r = {"seed": seed}
        for prec, name in ((E.FP32, "fp32"), (E.FP64, "fp64")):
            with E.Engine(L, prec, seed) as e:
                e.init_field(1)
                e.update(200, 1, BETA)
                p = []
                for _ in range(nmeas):
                    e.update(1, 1, BETA)
                    p.append(e.plaquette())
            p = np.array(p)
            tau, err = tau_int(p)
            r[name] = {"mean": float(p.mean()), "err": err, "tau_int": tau}
        r["diff"] = r["fp32"]["mean"] - r["fp64"]["mean"]
        r["diff_over_sigma"] = r["diff"] / float(np.hypot(r["fp32"]["err"], r["fp64"]["err"]))
        log(f"[c6] seed {seed}: dP = {r['diff']:.3g} = {r['diff_over_sigma']:+.2f} sigma")
        rows.append(r)
    d = np.array([r["diff"] for r in rows])
    sig = np.array([np.hypot(r["fp32"]["err"], r["fp64"]["err"]) for r in rows])
    mean_sig = float(np.sqrt(np.sum(sig ** 2)) / len(d))
    return {"study": "c6", "lattice": L, "beta": BETA, "thermalisation": 200, "measured_sweeps": nmeas,
            "per_seed": rows, "mean_diff": float(d.mean()), "mean_diff_err": mean_sig,
            "mean_diff_over_sigma": float(d.mean()) / mean_sig,
            "rms_z": float(np.sqrt(np.mean((d / sig) ** 2))), "seconds": time.perf_counter() - t0}


# the reference's unsmeared loops at beta 3.09, 48^4: mean and autocorrelation-
# corrected error of its configurations 200..999 (of 1000) in
# measurements/Therm/48x48x48x48_SU2_b3.090000_id1.meas (the same tau_int
# method as here)
W_REF = {1: (0.732540, 3.7e-6), 5: (0.044296, 1.5e-5), 7: (0.009507, 9.3e-6), 9: (0.001973, 4.9e-6)}


def c7(quick=False):
    # Large loops relax slowly under local updates (W(9,9) drifts for ~2000
    # sweeps from a hot start, from below, and from a cold start from above),
    # so the chains start alternately hot and cold, thermalise 3000 sweeps and
    # measure every 4th of the next 4000.
    L = (48, 48, 48, 48)
    seeds = (7,) if quick else (7, 8, 9, 10)
    therm, nmeas, every = (100, 100, 1) if quick else (3000, 1000, 4)
    Rs = sorted(W_REF)
    t0 = time.perf_counter()
    out = {}
    for prec, name in ((E.FP64, "fp64"), (E.FP32, "fp32")):
        means, errs = [], []
        for k, seed in enumerate(seeds):
            with E.Engine(L, prec, seed) as e:
                e.init_field(1 - k % 2)
                e.update(therm, 1, BETA)
                w = []
                for _ in range(nmeas):
                    e.update(every, 1, BETA)
                    w.append(e.wilson_loops(Rs, Rs))
            w = np.array(w)
            means.append(w.mean(axis=0))
            errs.append([tau_int(w[:, k])[1] for k in range(len(Rs))])
            log(f"[c7] {name} seed {seed}: W = {np.array2string(means[-1], precision=6)}")
        m = np.mean(means, axis=0)
        err = np.sqrt(np.sum(np.array(errs) ** 2, axis=0)) / len(seeds)
        out[name] = {f"W{R}x{R}": {"mean": float(m[k]), "err": float(err[k]), "reference": W_REF[R][0],
                                   "reference_err": W_REF[R][1], "diff": float(m[k] - W_REF[R][0]),
                                   "z": float((m[k] - W_REF[R][0]) / np.hypot(err[k], W_REF[R][1]))}
                     for k, R in enumerate(Rs)}
    return {"study": "c7", "lattice": L, "beta": BETA, "seeds": list(seeds), "starts": "hot, cold alternately",
            "thermalisation": therm, "measurements": nmeas, "sweeps_between": every, "loops": out,
            "seconds": time.perf_counter() - t0}


# The reference's own 24^4 chains at beta 3.09 (heatbath, 50 000 stored
# configurations each from a cold and from a hot start;
# Thermalisation/24x24x24x24_b3.09_WL_<R>_<T>_{cold,hot}.dat): mean over
# configurations 10000..49999 of the two starts and error from the means of
# blocks of 1000 configurations (its hot chain still drifts upwards in the
# larger loops over configurations 1000..9999).  W(10,16) after 100 spatial
# APE steps: configurations 1000..9999 (..._WL_10_16_{cold,hot}_NAPE_100.dat,
# the only ones smeared; the smearing parameter is not archived with them),
# whose two starts differ by 1.4e-4, five times their block error, so the
# error here is half that difference.  The same figures with
# autocorrelation-time errors per start are in tests/golden/ref_therm24.json.
W_REF24 = {(1, 1): (0.73253802, 1.4e-6), (3, 3): (0.19556128, 6.4e-6), (5, 5): (0.044281381, 7.3e-6),
           (7, 7): (0.0095079745, 5.7e-6), (9, 9): (0.0019805584, 3.4e-6), (10, 16): (8.4210018e-05, 2.1e-6)}
W_REF24_APE100 = (0.0017639191, 6.8e-5)


def c8(quick=False, nmeas=5000, every=2, therm=2000, ape_every=10):
    """24^4 chains (FP64 and FP32 storage, cold and hot starts, heatbath + 1
    overrelaxation per sweep): W(R,T) for the reference's six loops after
    every `every` sweeps, and W(10,16) after 100 spatial APE steps (sp 0.25)
    on every `ape_every`-th measured configuration, against W_REF24."""
    L = (24, 24, 24, 24)
    if quick:
        nmeas, therm = 300, 200
    keys = sorted(W_REF24)
    R = [k[0] for k in keys]
    T = [k[1] for k in keys]
    t0 = time.perf_counter()
    out = {}
    for prec, name in ((E.FP64, "fp64"), (E.FP32, "fp32")):
        chains = {}
        for start, sname, seed in ((0, "cold", 21), (1, "hot", 22)):
            t1 = time.perf_counter()
            with E.Engine(L, prec, seed) as e, E.Engine(L, prec, seed) as m:
                e.init_field(start)
                e.update(therm, 1, BETA)
                w, wa = [], []
                for i in range(nmeas):
                    e.update(every, 1, BETA)
                    w.append(e.wilson_loops(R, T))
                    if i % ape_every == 0:
                        m.upload(e.download())
                        m.ape(E.APE_SPATIAL, 100, SP_SPATIAL)
                        wa.append(m.wilson_loops([10], [16])[0])
                    if (i + 1) % 1000 == 0:
                        log(f"[c8] {name} {sname}: {i + 1} measurements, {time.perf_counter() - t1:.1f} s")
            w = np.array(w)
            chains[sname] = {"w": w, "wa": np.array(wa)}
            log(f"[c8] {name} {sname}: {time.perf_counter() - t1:.1f} s, W = {np.array2string(w.mean(axis=0), precision=7)}"
                f", APE-100 W(10,16) = {np.mean(wa):.6g}")
        res = {}
        for k, key in enumerate(keys + [("ape100", 10, 16)]):
            series = [c["wa"] if key[0] == "ape100" else c["w"][:, k] for c in chains.values()]
            stats = [tau_int(x) for x in series]
            m_ = float(np.mean([x.mean() for x in series]))
            err = float(np.sqrt(sum(s_[1] ** 2 for s_ in stats)) / len(series))
            ref, ref_err = W_REF24_APE100 if key[0] == "ape100" else W_REF24[key]
            label = "APE100 W10x16" if key[0] == "ape100" else f"W{key[0]}x{key[1]}"
            res[label] = {"mean": m_, "err": err, "tau_int": [s_[0] for s_ in stats],
                          "per_start": [float(x.mean()) for x in series], "reference": ref,
                          "reference_err": ref_err, "z": (m_ - ref) / float(np.hypot(err, ref_err))}
        out[name] = res
    return {"study": "c8", "lattice": L, "beta": BETA, "thermalisation": therm, "measurements": nmeas,
            "sweeps_between": every, "ape_every": ape_every, "loops": out, "seconds": time.perf_counter() - t0}


def c5(quick=False):
    L = (96, 96, 96, 96)
    nsweeps = 20 if quick else 200
    t0 = time.perf_counter()
    with E.Engine(L, E.FP32, 3) as e:
        e.init_field(1)
        e.sync()
        t1 = time.perf_counter()
        for s in range(0, nsweeps, 20):
            e.update(min(20, nsweeps - s), 1, BETA)
            log(f"[c5] sweep {s + 20}: P {e.plaquette():.6f}")
        e.sync()
        t_sweeps = time.perf_counter() - t1
        plaq = e.plaquette()
        U = e.download()
        t1 = time.perf_counter()
        e.ape(E.APE_SPATIAL, 400, SP_SPATIAL)
        e.sync()
        t_ape = time.perf_counter() - t1
        w = e.wilson_loops([5, 7, 9], [5, 7, 9])
        log(f"[c5] W(5,5) {w[0]:.6g} W(7,7) {w[1]:.6g} W(9,9) {w[2]:.6g}")
        e.upload(U)
        del U
        t1 = time.perf_counter()
        e.ape(E.APE_ALL, 30, SP_4D)
        e.sync()
        t_ape4 = time.perf_counter() - t1
        t1 = time.perf_counter()
        q, qa = e.qtop(E.Q_IMPROVED)
        t_q = time.perf_counter() - t1
        dev = e.device_bytes()
    vol = int(np.prod(L))
    return {"study": "c5", "lattice": L, "beta": BETA, "precision": "fp32-store", "sweeps": nsweeps,
            "plaquette": plaq, "sweep_s": t_sweeps / nsweeps,
            "link_updates_per_s": 4.0 * vol * nsweeps / t_sweeps,
            "ape400_s": t_ape, "wloops": {"5,5": w[0], "7,7": w[1], "9,9": w[2]},
            "ape4d_30_s": t_ape4, "Q_improved": q, "sum_abs_q": qa, "qtop_s": t_q,
            "device_bytes": dev, "device_fraction_of_288GB": dev / 288e9,
            "seconds": time.perf_counter() - t0}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("study", choices=["c1", "c2", "c3", "c4", "c5", "c6", "c7", "c8", "all"])
    ap.add_argument("--quick", action="store_true", help="fewer sweeps (a functional check)")
    ap.add_argument("--seeds", type=int, default=8, help="C6: independent chains per precision")
    ap.add_argument("--nmeas", type=int, default=5000, help="C8: measured configurations per chain")
    ap.add_argument("--every", type=int, default=2, help="C8: sweeps between measurements")
    a = ap.parse_args()
    E.load_library()
    todo = ["c1", "c2", "c3", "c4", "c5", "c6", "c7", "c8"] if a.study == "all" else [a.study]
    for s in todo:
        log(f"[{s}] start")
        if s == "c6":
            r = c6(a.quick, a.seeds)
        elif s == "c8":
            r = c8(a.quick, a.nmeas, a.every)
        else:
            r = globals()[s](a.quick)
        print(json.dumps(r), flush=True)
        log(f"[{s}] done in {r['seconds']:.1f} s")


if __name__ == "__main__":
    main()
