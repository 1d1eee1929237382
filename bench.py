#!/usr/bin/env python3
"""Headline benchmark: SU(2) heatbath + overrelaxation sweeps at 48^4, beta=3.09.

A step is one sweep over the whole lattice: one checkerboard Kennedy-Pendleton
heatbath sweep plus one overrelaxation sweep (16 kernel launches), counted as
4*VOL link updates.  The FP32-storage chain is the headline `value`; the
all-FP64 chain is timed the same way for the FP32/FP64 ratio, and both chains'
plaquettes are compared over a post-timing series.  The field is resident in
HBM for the whole run (hot start on the device, no host transfers inside the
timed region).  Kernel durations come from HIP events on the engine's stream
recorded around every launch of a second run of the same K steps (the event
records would otherwise add gaps to the timed steps).

Also reported (extra keys): spatial APE steps/s, Wilson-loop and charge times,
the APE-400 adjoint/forward time ratio, the FP64 field's host round trip, and
the reference CPU code timed on this host on one sweep of the same 48^4
configuration (one core; then --cpu-procs independent processes at once).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("link updates/s at 48^4 FP32-store vs FP64, same ⟨plaq⟩ ±1σ; "
          "APE VJP/fwd time")
HBM_PEAK_GBS = 8000.0
# algorithmic (compulsory) bytes per link update: every field element moves once
# per kernel pass; heatbath reads 7 half-planes and writes 1, overrelaxation
# also reads the old link
BYTES_PER_UPDATE = {"heatbath": {1: 128, 0: 256}, "overrelax": {1: 144, 0: 288}}
LINK_BYTES = {1: 16, 0: 32}   # FP32 storage, FP64


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_chain(lgt, L, prec, seed, beta, n_or, warmup, steps, nplaq):
    vol = int(np.prod(L))
    e = lgt.Engine(L, prec, seed)
    e.init_field(1)
    e.update(warmup, n_or, beta)
    e.sync()
    t0 = time.perf_counter()
    e.update(steps, n_or, beta)
    e.sync()
    t1 = time.perf_counter()
    # the same number of steps again with a HIP event pair around every launch
    # (on the engine's stream): per-kernel durations, kept out of the wall
    # time above because the event records add gaps between launches
    e.set_option(lgt.engine.OPT_PROFILE, 1)
    e.profile_reset()
    e.sync()
    tp = time.perf_counter()
    e.update(steps, n_or, beta)
    e.sync()
    tp = time.perf_counter() - tp
    kern = {}
    for cls, name in enumerate(lgt.engine.KERNEL_CLASSES):
        ms, n = e.profile_read(cls)
        if n:
            kern[name] = {"ms_total": ms, "launches": n, "ms_avg": ms / n}
    e.set_option(lgt.engine.OPT_PROFILE, 0)
    # the reference main's loop (qcd.c:53-59): one sweep call, then the
    # plaquette, per configuration
    plaqs = []
    e.sync()
    tq = time.perf_counter()
    for _ in range(nplaq):
        e.update(1, n_or, beta)
        plaqs.append(e.plaquette())
    tq = time.perf_counter() - tq
    dt = t1 - t0
    return e, {
        "seconds": dt,
        "ms_per_step": 1e3 * dt / steps,
        "ms_per_step_profiled": 1e3 * tp / steps,
        "link_updates_per_s": 4.0 * vol * steps / dt,
        "kernels": kern,
        "plaq": plaqs,
        "qcd_loop_ms_per_step": 1e3 * tq / nplaq if nplaq else None,
        "device_bytes": e.device_bytes(),
    }


# Compulsory link transfers of one APE step, per smeared link: spatial
# smearing reads and writes the 3 VOL spatial links (its staples use only
# spatial links, smearing.c:107-138), 4D smearing the 4 VOL links
# (smearing.c:74-105); temporal smearing writes the VOL temporal links but its
# staples read all 4 VOL links (smearing.c:140-168), i.e. 5 transfers per
# smeared link.
APE_SMEARED = {0: 3, 1: 4, 2: 1}
APE_TRANSFERS = {0: 2, 1: 2, 2: 5}


def time_heatbath_only(e, L, beta, steps):
    """Heatbath-only sweeps, the reference's algorithm (update.c:102-176: one
    heatbath sweep, then one projection of the field): the GPU figure that
    stands beside the heatbath-only CPU baseline."""
    vol = int(np.prod(L))
    e.sync()
    t0 = time.perf_counter()
    e.update(steps, 0, beta)
    e.sync()
    dt = time.perf_counter() - t0
    return {"link_updates_per_s": 4.0 * vol * steps / dt, "ms_per_step": 1e3 * dt / steps, "steps": steps}


def cpu_model():
    """The host CPU as /proc/cpuinfo names it, with its family, model and
    clock (some hosts give only a generic model name)."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if fields:
                        break
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                fields.setdefault(k, v)
    except OSError:
        return None
    name = fields.get("model name")
    if name is None:
        return None
    extra = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    if "cpu MHz" in fields:
        extra.append(f"{float(fields['cpu MHz']):.0f} MHz")
    return f"{name} ({', '.join(extra)})" if extra else name


def observables_24(lgt, beta, seed=5, reps=5):
    """One spatial APE step (sp = 0.25), the Wilson loops R,T in 1..6 and the
    improved charge at 24^4 in FP64: the reference's own functions
    (oracle/_ref, one host core) and the engine on the same configuration,
    timed in the same run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import RefLib
    if not RefLib.available("l24"):
        return None
    L = (24, 24, 24, 24)
    R6, T6 = np.meshgrid(np.arange(1, 7), np.arange(1, 7), indexing="ij")
    R6, T6 = R6.ravel().astype(np.int32), T6.ravel().astype(np.int32)

    def timed(f, n=1):
        t0 = time.perf_counter()
        for _ in range(n):
            f()
        return (time.perf_counter() - t0) / n

    with lgt.Engine(L, lgt.FP64, seed) as e:
        e.init_field(1)
        e.update(50, 1, beta)
        U = e.download()

        def gpu(f):
            f()
            e.sync()

        e.ape(0, 1, 0.25)       # untimed first calls allocate the engine's work buffers
        e.wilson_loops(R6, T6)
        e.qtop(0)
        e.upload(U)
        gpu_ape = timed(lambda: gpu(lambda: e.ape(0, 1, 0.25)), reps)
        e.upload(U)
        gpu_w = timed(lambda: e.wilson_loops(R6, T6), reps)
        gpu_q = timed(lambda: e.qtop(0), reps)
    ref = RefLib("l24")
    ref.setup(beta, 1)
    ref.set_field(U)
    cpu_ape = timed(lambda: ref.lib.ref_ape(0, 1, 0.25))
    ref.set_field(U)
    cpu_w = timed(lambda: ref.wilson(R6, T6))
    cpu_q = timed(lambda: ref.lib.ref_qtop_improved())
    rows = []
    for op, c, g in (("spatial APE step", cpu_ape, gpu_ape), ("Wilson loops R,T 1..6", cpu_w, gpu_w),
                     ("improved charge", cpu_q, gpu_q)):
        rows.append({"op": op, "reference_cpu_s": c, "gpu_fp64_s": g, "speedup": c / g})
    return {"lattice": list(L), "cores": 1, "cpu_model": cpu_model(), "rows": rows}


def time_ape(e, prec, L, nsteps=20, mode=0, sp=0.25):
    """One APE step per launch; algorithmic bytes: APE_TRANSFERS link
    transfers per smeared link (links of 16 or 32 bytes)."""
    vol = int(np.prod(L))
    nlinks = APE_SMEARED[mode] * vol
    e.sync()
    t0 = time.perf_counter()
    e.ape(mode, nsteps, sp)
    e.sync()
    dt = time.perf_counter() - t0
    e.set_option(1, 1)
    e.profile_reset()
    e.ape(mode, nsteps, sp)
    ms, n = e.profile_read(4)
    e.set_option(1, 0)
    avg = ms / n
    nbytes = nlinks * APE_TRANSFERS[mode] * LINK_BYTES[prec]
    gbs = nbytes / (avg * 1e-3) / 1e9
    return {"steps_per_s": nsteps / dt, "kernel_ms_avg": avg, "algorithmic_bytes": nbytes, "achieved_GBs": gbs,
            "frac_hbm": gbs / HBM_PEAK_GBS}


def time_plaquette(e, prec, L, reps=20):
    """Average plaquette: every link read once (4 links per site)."""
    vol = int(np.prod(L))
    e.set_option(1, 1)
    e.profile_reset()
    for _ in range(reps):
        e.plaquette()
    ms, n = e.profile_read(3)
    e.set_option(1, 0)
    avg = ms / n
    gbs = vol * 4 * LINK_BYTES[prec] / (avg * 1e-3) / 1e9
    return {"kernel_ms_avg": avg, "achieved_GBs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}


def time_stream(e, prec, L, reps=10):
    """The projection kernel reads and writes every link once: a plain
    streaming pass of known size, the achievable-bandwidth reference and the
    calibration kernel of the HBM byte counters."""
    vol = int(np.prod(L))
    e.set_option(1, 1)
    e.profile_reset()
    for _ in range(reps):
        e.project()
    ms, n = e.profile_read(2)
    e.set_option(1, 0)
    nbytes = 2 * 4 * vol * (16 if prec == 1 else 32)
    return {"bytes_per_launch": nbytes, "kernel_ms_avg": ms / n, "achieved_GBs": nbytes / (ms / n * 1e-3) / 1e9}


def time_vjp(lgt, L, prec, seed, beta, nsteps=400, sp=0.25, R=5, T=5, therm=50, mode=0, segment=0):
    """Adjoint of N APE steps (spatial: alpha = 0.5, sp = 0.25; mode 1: all
    directions) with the loop objective L = sum_x sum_d 2 Re W_{RxT}: forward
    with sqrt(N) checkpoints (segment 0) or checkpoints every `segment` steps
    (segment = N: every step's output kept, nothing recomputed), then the
    backward pass (recompute + adjoint steps) for the objective's gradient.
    The ratio is backward / forward wall time."""
    with lgt.Engine(L, prec, seed) as e:
        return _time_vjp(e, nsteps, sp, R, T, therm, mode, segment, beta)


def _time_vjp(e, nsteps, sp, R, T, therm, mode, segment, beta):
    e.init_field(1)
    e.update(therm, 1, beta)
    U0 = e.download()
    # one untimed pass allocates the checkpoint and segment buffers (kept by
    # the engine for later passes of the same layout)
    e.ape_vjp_forward(nsteps, sp, segment, mode=mode)
    e.wloop_grad(R, T, host=False)
    e.ape_vjp_backward(grad=False)
    e.upload(U0)
    del U0
    e.sync()
    t0 = time.perf_counter()
    e.ape_vjp_forward(nsteps, sp, segment, mode=mode)
    e.sync()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    e.wloop_grad(R, T, host=False)
    e.sync()
    t_obj = time.perf_counter() - t0
    t0 = time.perf_counter()
    e.ape_vjp_backward(grad=False)
    e.sync()
    t_bwd = time.perf_counter() - t0
    out = {"mode": ("spatial", "4d", "temporal")[mode], "nsteps": nsteps, "segment": segment or int(np.ceil(np.sqrt(nsteps))), "forward_s": t_fwd, "backward_s": t_bwd,
           "objective_grad_s": t_obj, "vjp_over_forward": t_bwd / t_fwd, "device_bytes": e.device_bytes()}
    return out


# quaternion product = 16 multiplications + 12 additions
FLOP_QMUL = 28
# improved clover charge: ~576 products per site plus the field-strength sums
FLOP_QTOP_SITE = 16100


def time_observables(e, L):
    """Wilson loops R 1..10 x T 1..16 and the improved charge, with their rate
    in reference-equivalent operations: the flops the reference's algorithm
    would spend (a loop costs 3 directions x 2(R+T) products per site,
    wilson.c:21-56; the charge ~16 100 flops per site,
    topologicalcharge.c:112-199) divided by the GPU time.  The kernels here
    perform far fewer (shared line products), so this is not an executed
    FLOP rate and is not comparable with a FLOP/s peak."""
    vol = int(np.prod(L))
    R, T = np.meshgrid(np.arange(1, 11), np.arange(1, 17), indexing="ij")
    R, T = R.ravel(), T.ravel()
    # untimed first calls: they allocate the loops' two line fields (a second
    # one only when the loops span more than one R) and the charge's two line
    # fields, which later calls reuse
    e.wilson_loops([1, 2], [1, 1])
    e.qtop(0)
    e.sync()
    t0 = time.perf_counter()
    e.wilson_loops(R, T)
    tw = time.perf_counter() - t0
    t0 = time.perf_counter()
    e.qtop(0)
    tq = time.perf_counter() - t0
    fw = float(vol * 3 * np.sum(2 * (R + T)) * FLOP_QMUL)
    fq = float(vol * FLOP_QTOP_SITE)
    return {
        "wloops_R1-10_T1-16": {"s": tw, "reference_flops": fw, "ref_equiv_Tops_per_s": fw / tw / 1e12},
        "qtop_improved": {"s": tq, "reference_flops": fq, "ref_equiv_Tops_per_s": fq / tq / 1e12},
    }


# one reference sweep in a fresh interpreter (no GPU): prints its own duration
_CPU_WORKER = """
import sys, time, numpy as np
sys.path.insert(0, {tests!r})
from oracle_lib import RefLib
ref = RefLib({name!r}); ref.setup({beta!r}, 1)
ref.set_field(np.load({path!r}, mmap_mode="r"))
t0 = time.perf_counter(); ref.lib.ref_update(1); print(time.perf_counter() - t0, flush=True)
"""


def cpu_baseline(U, L, beta, procs):
    """The reference's own heatbath sweep (LGT_Code/modules/update/update.c:102-176,
    compiled from its sources into oracle/_ref) on the thermalised FP64
    configuration of the GPU run: once on one core (the reference is
    single-threaded), then as `procs` independent processes at once, the
    aggregate over the host's cores."""
    import subprocess
    import tempfile
    tests = os.path.join(ROOT, "tests")
    sys.path.insert(0, tests)
    from oracle_lib import REF_SIZES, RefLib
    name = {v: k for k, v in REF_SIZES.items()}.get(tuple(L))
    if name is None or not RefLib.available(name):
        return None
    vol = int(np.prod(L))
    ref = RefLib(name)
    ref.setup(beta, 1)
    ref.set_field(U)
    t0 = time.perf_counter()
    ref.lib.ref_update(1)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref.lib.ref_ape(0, 1, 0.25)     # one spatial APE step (smearing.c:107-138) of the same field
    dt_ape = time.perf_counter() - t0
    out = {
        "value": 4.0 * vol / dt,
        "unit": "link updates/s",
        "cores": 1,
        "kind": "reference",
        "sample": f"1 lexicographic heatbath sweep (gaugefieldUpdate, no overrelaxation) of the thermalised "
                  f"{'x'.join(map(str, L))} FP64 configuration, {dt:.1f} s on 1 host core",
        "spatial_ape_step_s": dt_ape,
    }
    if procs > 1:
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "field.npy")
            np.save(path, U)
            del U
            code = _CPU_WORKER.format(tests=tests, name=name, beta=beta, path=path)
            t0 = time.perf_counter()
            ps = [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, text=True)
                  for _ in range(procs)]
            times = [float(p.communicate()[0].split()[-1]) for p in ps]
            wall = time.perf_counter() - t0
        out["all_cores"] = {
            "value": 4.0 * vol * procs / wall,
            "unit": "link updates/s",
            "processes": procs,
            "host_cpus": os.cpu_count(),
            "sweep_s_per_process": [round(t, 2) for t in times],
            "wall_s": wall,
        }
    return out


def cpu_table(beta=3.09):
    """The reference's own functions (oracle/_ref, compiled from its sources)
    and the C restatement of them (oracle/su2_oracle.c) on one core: heatbath
    sweep at 24^4 and 48^4, one spatial APE step, Wilson loops R,T in 1..6 and
    the improved charge at 24^4.  The reference sweeps lexicographically with
    RANLUX (update.c:102-176); the restatement runs the engine's checkerboard
    sweep with Philox (the same per-link work).  No GPU is used."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, RefLib
    R6, T6 = np.meshgrid(np.arange(1, 7), np.arange(1, 7), indexing="ij")
    R6, T6 = R6.ravel().astype(np.int32), T6.ravel().astype(np.int32)
    rows = []

    def timed(f):
        t0 = time.perf_counter()
        f()
        return time.perf_counter() - t0

    for name, L in (("l24", (24, 24, 24, 24)), ("l48", (48, 48, 48, 48))):
        if not RefLib.available(name):
            continue
        vol = int(np.prod(L))
        o = Oracle(L)
        U = o.hot(5)
        o.engine_sweep(U, beta, 0, 5, 0)
        ref = RefLib(name)
        ref.setup(beta, 1)
        ref.set_field(U)
        n = 2 if L[0] == 24 else 1
        ref.lib.ref_update(1)
        t_ref = timed(lambda: ref.lib.ref_update(n)) / n
        V = U.copy()
        t_or = timed(lambda: [o.engine_sweep(V, beta, 0, 5, 1 + k) for k in range(n)]) / n
        rows.append({"op": "heatbath sweep", "lattice": L, "reference_s": t_ref, "restatement_s": t_or,
                     "reference_link_updates_per_s": 4 * vol / t_ref})
        ref.set_field(U)
        V = U.copy()
        rows.append({"op": "spatial APE step", "lattice": L, "reference_s": timed(lambda: ref.lib.ref_ape(0, 1, 0.25)),
                     "restatement_s": timed(lambda: o.ape(V, 0, 1, 0.25))})
        if L[0] == 24:
            ref.set_field(U)
            rows.append({"op": "Wilson loops R,T 1..6", "lattice": L,
                         "reference_s": timed(lambda: ref.wilson(R6, T6)),
                         "restatement_s": timed(lambda: o.wilson(U, R6, T6, seq=True))})
            rows.append({"op": "improved charge", "lattice": L,
                         "reference_s": timed(lambda: ref.lib.ref_qtop_improved()),
                         "restatement_s": timed(lambda: o.qtop(U, seq=True))})
        log(f"[cpu-table] {name} done")
    for r in rows:
        r["restatement_over_reference"] = r["restatement_s"] / r["reference_s"]
    return {"cpu_table": rows, "cores": 1}


def time_host_transfer(e64):
    """Host <-> device round trip of the FP64 field in the reference's layout
    (what the drop-in entry points pay per call when the field lives in the
    program's host array)."""
    U = e64.download()
    t0 = time.perf_counter()
    e64.upload(U)
    e64.sync()
    t_up = time.perf_counter() - t0
    t0 = time.perf_counter()
    U = e64.download()
    t_down = time.perf_counter() - t0
    return {"bytes": int(U.nbytes), "upload_s": t_up, "download_s": t_down}


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tab = json.load(f)
    return tab.get(workload, {}).get(kernel)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--lattice", type=int, nargs=4, default=[48, 48, 48, 48])
    ap.add_argument("--beta", type=float, default=3.09)
    ap.add_argument("--n-or", type=int, default=1)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--nplaq", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-table", action="store_true",
                    help="only time the reference and the restatement on the CPU (no GPU) and exit")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="independent reference processes for the all-cores CPU figure (1: skip)")
    ap.add_argument("--vjp-steps", type=int, default=400)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported (the engine has no "
              f"multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    if args.cpu_table:
        print(json.dumps(cpu_table(args.beta)), flush=True)
        return

    import masterarbeit_amd as lgt
    lgt.load_library()
    L = tuple(args.lattice)
    vol = int(np.prod(L))
    lname = "x".join(map(str, L))
    log(f"[bench] lattice {lname} beta={args.beta} n_or={args.n_or} warmup={args.warmup} steps={args.steps}")

    e32, r32 = run_chain(lgt, L, lgt.FP32, args.seed, args.beta, args.n_or, args.warmup, args.steps, args.nplaq)
    log(f"[bench] fp32 sweep: {r32['ms_per_step']:.3f} ms/step, {r32['link_updates_per_s']:.4g} link updates/s")
    hb32 = time_heatbath_only(e32, L, args.beta, args.steps)
    log(f"[bench] fp32 heatbath-only: {hb32['link_updates_per_s']:.4g} link updates/s")
    stream32 = time_stream(e32, 1, L)
    log(f"[bench] fp32 streaming pass (project): {stream32['achieved_GBs']:.0f} GB/s")
    plaq32 = time_plaquette(e32, 1, L)
    ape32 = time_ape(e32, 1, L)
    log(f"[bench] fp32 spatial APE: {ape32['steps_per_s']:.1f} steps/s")
    ape4d32 = time_ape(e32, 1, L, mode=1, sp=1.0 / 6.0)
    apet32 = time_ape(e32, 1, L, mode=2)
    obs32 = time_observables(e32, L)
    e32.close()

    e64, r64 = run_chain(lgt, L, lgt.FP64, args.seed, args.beta, args.n_or, args.warmup, args.steps, args.nplaq)
    log(f"[bench] fp64 sweep: {r64['ms_per_step']:.3f} ms/step, {r64['link_updates_per_s']:.4g} link updates/s")
    U64 = e64.download()   # the thermalised configuration: input of the CPU baseline
    hb64 = time_heatbath_only(e64, L, args.beta, args.steps)
    e64.upload(U64)
    plaq64 = time_plaquette(e64, 0, L)
    ape64 = time_ape(e64, 0, L)
    ape4d64 = time_ape(e64, 0, L, mode=1, sp=1.0 / 6.0)
    apet64 = time_ape(e64, 0, L, mode=2)
    obs64 = time_observables(e64, L)
    log("[bench] APE-400 adjoint (fp64) ...")
    vjp64 = time_vjp(lgt, L, lgt.FP64, args.seed, args.beta, nsteps=args.vjp_steps)
    log(f"[bench] fp64 VJP/forward = {vjp64['vjp_over_forward']:.2f} (forward {vjp64['forward_s']:.3f} s)")
    vjp32 = time_vjp(lgt, L, lgt.FP32, args.seed, args.beta, nsteps=args.vjp_steps)
    # the same adjoint with every step's output kept in HBM (48^4 FP32, 400
    # steps: 128 GB), so that the backward recomputes nothing
    try:
        vjp32_all = time_vjp(lgt, L, lgt.FP32, args.seed, args.beta, nsteps=args.vjp_steps, segment=args.vjp_steps)
        log(f"[bench] fp32 VJP/forward = {vjp32['vjp_over_forward']:.2f}, all states kept: "
            f"{vjp32_all['vjp_over_forward']:.2f} ({vjp32_all['device_bytes'] / 1e9:.0f} GB)")
    except lgt.LgtError as err:   # more steps or a larger lattice than the device holds
        vjp32_all = {"error": str(err)}
        log(f"[bench] all-states adjoint skipped: {err}")
    vjp4d64 = time_vjp(lgt, L, lgt.FP64, args.seed, args.beta, nsteps=args.vjp_steps, sp=1.0 / 6.0, mode=1)
    log(f"[bench] fp64 4D APE VJP/forward = {vjp4d64['vjp_over_forward']:.2f}")

    # dominant kernel of the headline chain
    k = r32["kernels"]
    dom = max(("heatbath", "overrelax"), key=lambda n: k.get(n, {}).get("ms_total", 0.0))
    per_launch_bytes = (vol // 2) * BYTES_PER_UPDATE[dom][1]
    achieved = per_launch_bytes / (k[dom]["ms_avg"] * 1e-3) / 1e9
    traffic = load_traffic(dom, f"{lname}_fp32")

    p32, p64 = np.array(r32["plaq"]), np.array(r64["plaq"])
    sigma = float(np.sqrt(0.5 * (p32.var(ddof=1) + p64.var(ddof=1)))) if len(p32) > 1 else None
    diff = float(p32.mean() - p64.mean()) if len(p32) else None
    # errors of the two means with the integrated autocorrelation time
    # (LGT_Code/Autocorrelation/t_int.c), the chains being independent
    from masterarbeit_amd.studies import tau_int
    err_mean = float(np.hypot(tau_int(p32)[1], tau_int(p64)[1])) if len(p32) > 2 else None

    base = None
    xfer = time_host_transfer(e64)
    if not args.no_cpu_baseline:
        log("[bench] timing the reference CPU sweep (one core, then all cores) ...")
        base = cpu_baseline(U64, L, args.beta, args.cpu_procs)
        if base is not None:
            base["cpu_model"] = cpu_model()
    del U64
    e64.close()
    obs24 = None
    if not args.no_cpu_baseline:
        log("[bench] 24^4 observables: reference CPU functions and the engine ...")
        obs24 = observables_24(lgt, args.beta)

    out = {
        "metric": METRIC,
        "value": r32["link_updates_per_s"],
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"SU(2) pure gauge, {lname} lattice, beta={args.beta}: 1 checkerboard heatbath + "
                        f"{args.n_or} overrelaxation sweep per step, FP32 link storage with FP64 global sums, "
                        f"hot start (Haar, seed {args.seed}) thermalised by the warm-up sweeps",
            "lattice": list(L),
            "beta": args.beta,
            "n_or": args.n_or,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": per_launch_bytes,
            "avg_launch_ms": k[dom]["ms_avg"],
        },
        "cpu_baseline": base,
        "fp64": {
            "link_updates_per_s": r64["link_updates_per_s"],
            "ms_per_step": r64["ms_per_step"],
            "ms_per_step_profiled": r64["ms_per_step_profiled"],
            "qcd_loop_ms_per_step": r64["qcd_loop_ms_per_step"],
            "kernels": r64["kernels"],
        },
        "heatbath_only": {"fp32": hb32, "fp64": hb64,
                          "note": "one heatbath sweep + one projection of the field per step, the reference's "
                                  "algorithm (FP32: links projected as the heatbath stores them; FP64: a separate "
                                  "projection pass); compare with cpu_baseline"},
        "observables_24_cpu_vs_gpu": obs24,
        "fp32_qcd_loop_ms_per_step": r32["qcd_loop_ms_per_step"],
        "fp32_over_fp64": r32["link_updates_per_s"] / r64["link_updates_per_s"],
        "fp32_kernels": r32["kernels"],
        "fp32_ms_per_step_profiled": r32["ms_per_step_profiled"],
        "plaquette": {
            "fp32_mean": float(p32.mean()) if len(p32) else None,
            "fp64_mean": float(p64.mean()) if len(p64) else None,
            "sigma_per_config": sigma,
            "diff": diff,
            "within_1sigma": (abs(diff) <= sigma) if sigma else None,
            "sigma_of_diff_of_means": err_mean,
            "diff_over_sigma_of_means": (diff / err_mean) if err_mean else None,
            "n": len(p32),
        },
        "ape_spatial": {"fp32": ape32, "fp64": ape64},
        "ape_4d": {"fp32": ape4d32, "fp64": ape4d64},
        "ape_temporal": {"fp32": apet32, "fp64": apet64},
        "plaquette_kernel": {"fp32": plaq32, "fp64": plaq64},
        "ape_vjp": {"fp64": vjp64, "fp32": vjp32},
        "ape_vjp_4d": {"fp64": vjp4d64},
        "ape_vjp_all_states": {"fp32": vjp32_all},
        "stream_pass": stream32,
        "observables_s": {"fp32": obs32, "fp64": obs64},
        "device_bytes": {"fp32": r32["device_bytes"], "fp64": r64["device_bytes"]},
        "host_transfer_fp64": xfer,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
