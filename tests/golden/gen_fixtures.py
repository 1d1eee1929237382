"""Writes the golden fixtures of this directory from the reference's own C code.

Run in a container that holds the reference tree, after ``make -C oracle ref``
(it loads oracle/_ref/<size>/libref.so, the reference modules compiled from
their sources).  The outputs are data only: inputs and the reference's outputs.

* ref_l4.npz     -- 4^4 lattice: a configuration made by the reference (its hot
                    start, su2RandomMatrix, then two heatbath sweeps at beta=3.09),
                    its plaquette, Wilson loops, both charges; the fields after
                    spatial / 4D / temporal APE; one lexicographic heatbath sweep
                    with the RANLUX uniforms it consumed and the field it produced.
* ref_a6448.npz  -- the same observables on a 6x4x4x8 lattice (asymmetric extents
                    pin the index order), fields only for the input configuration.
* hb_links.npz   -- per-link heatbath vectors: staple, beta, the uniforms the
                    reference's localHeatbathUpdate consumed, the link it wrote and
                    the trial count (LGT_Code/modules/update/heatbath.c:124-149).
* su2_random.npz -- su2RandomMatrix outputs with their three uniforms
                    (LGT_Code/modules/random/random_su3.c:141-155).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import RefLib  # noqa: E402

BETA = 3.09


def loop_grid(rmax, tmax):
    R, T = np.meshgrid(np.arange(1, rmax + 1), np.arange(1, tmax + 1), indexing="ij")
    return R.ravel().astype(np.int32), T.ravel().astype(np.int32)


def observables(ref, prefix, out, rmax, tmax):
    R, T = loop_grid(rmax, tmax)
    out[prefix + "plaq"] = ref.lib.ref_plaquette()
    out[prefix + "wR"] = R
    out[prefix + "wT"] = T
    out[prefix + "w"] = ref.wilson(R, T)
    out[prefix + "qimp"] = ref.lib.ref_qtop_improved()
    out[prefix + "qclov"] = ref.lib.ref_qtop_clover()


def make_config(ref, seed, nsweeps):
    ref.setup(BETA, seed)
    ref.lib.ref_init_field(1)
    for _ in range(nsweeps):
        ref.lib.ref_update(1)
    return ref.get_field()


def gen_lattice(name, rmax, tmax, keep_fields):
    ref = RefLib(name)
    out = {"L": np.array(ref.L, dtype=np.int32), "beta": BETA}
    U0 = make_config(ref, seed=11, nsweeps=2)
    out["U0"] = U0
    observables(ref, "", out, rmax, tmax)
    for mode, nst, sp, tag in [(0, 5, 0.25, "ape_spatial"), (1, 3, 1.0 / 6.0, "ape_all"), (2, 2, 0.3, "ape_temporal")]:
        ref.set_field(U0)
        ref.lib.ref_ape(mode, nst, sp)
        out[tag + "_args"] = np.array([mode, nst, sp])
        if keep_fields:
            out[tag] = ref.get_field()
        observables(ref, tag + "_", out, rmax, tmax)
    # one lexicographic heatbath sweep (update.c:102-176) with its uniforms
    ref.set_field(U0)
    st = ref.rng_state()
    ref.lib.ref_update(1)
    U1 = ref.get_field()
    st_after = ref.rng_state()
    ref.rng_reset(st)
    used = 0
    chunk = []
    while True:  # replay until the generator state matches the one after the sweep
        chunk.append(ref.rng_draw(1)[0])
        used += 1
        if np.array_equal(ref.rng_state(), st_after):
            break
        if used > 64 * 4 * ref.vol:
            raise RuntimeError("could not replay the sweep's uniforms")
    out["hb_uniforms"] = np.array(chunk)
    if keep_fields:
        out["hb_U1"] = U1
    out["hb_plaq"] = ref.lib.ref_plaquette()
    return out


def gen_hb_links():
    ref = RefLib("l4")
    rows = []
    for seed, beta, presweeps in [(3, 3.09, 0), (5, 3.09, 3), (7, 1.0, 1), (9, 0.3, 0)]:
        ref.setup(beta, seed)
        ref.lib.ref_init_field(1)
        for _ in range(presweeps):
            ref.lib.ref_update(1)
        for n in range(0, ref.vol, 2):
            mu = n % 4
            A = ref.staple(n, mu)
            st = ref.rng_state()
            r = ref.lib.ref_heatbath_link(n, mu)
            link = ref.get_link(n, mu)
            count = 0 if r == 0.0 else int(round(1.0 / r))
            nuse = 3 if count == 0 else 4 * count + 2
            ref.rng_reset(st)
            u = ref.rng_draw(nuse)
            rows.append((A, beta, u, link, count))
    lens = np.array([len(r[2]) for r in rows], dtype=np.int32)
    offsets = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int32)
    return {
        "staples": np.array([r[0] for r in rows]),
        "beta": np.array([r[1] for r in rows]),
        "uniforms": np.concatenate([r[2] for r in rows]),
        "offsets": offsets,
        "lens": lens,
        "links": np.array([r[3] for r in rows]),
        "counts": np.array([r[4] for r in rows], dtype=np.int32),
    }


def gen_su2_random():
    ref = RefLib("l4")
    ref.setup(BETA, 13)
    us, mats = [], []
    for _ in range(64):
        st = ref.rng_state()
        mats.append(ref.su2_random())
        ref.rng_reset(st)
        us.append(ref.rng_draw(3))
    return {"uniforms": np.array(us), "links": np.array(mats)}


def main():
    for name in ["l4", "a6448"]:
        if not RefLib.available(name):
            raise SystemExit(f"reference build {name} missing: run make -C oracle ref")
    np.savez_compressed(os.path.join(HERE, "ref_l4.npz"), **gen_lattice("l4", 3, 3, keep_fields=True))
    np.savez_compressed(os.path.join(HERE, "ref_a6448.npz"), **gen_lattice("a6448", 4, 5, keep_fields=False))
    np.savez_compressed(os.path.join(HERE, "hb_links.npz"), **gen_hb_links())
    np.savez_compressed(os.path.join(HERE, "su2_random.npz"), **gen_su2_random())
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
