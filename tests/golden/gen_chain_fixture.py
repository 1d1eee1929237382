"""Writes tests/golden/ref_chain_l12.npz: the reference's own heatbath chain
(gaugefieldUpdate, LGT_Code/modules/update/update.c:102-176, compiled from its
sources into oracle/_ref/l12) on a 12^4 lattice at beta = 3.09 from a cold
start, seeded with rlxd_init(2, seed):

* plaq[s]    -- plaquette() after sweep s+1 (LGT_Code/main/qcd.c:57-60)
* w[m, :]    -- measureWilsonLoop on configurations 100, 110, ..., after
                smearing_APE_spatial(20, 0.25) (LGT_Code/main_meas/meas.c:105-109),
                for the loop grid R, T = 1..6 (R outer, T inner)

Data only; the statistical tests compare the engine's chain with it."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import RefLib  # noqa: E402

BETA, SEED, NSWEEP, THERM, EVERY = 3.09, 5, 1000, 100, 10


def main():
    ref = RefLib("l12")
    ref.setup(BETA, SEED)
    ref.lib.ref_init_field(0)
    R, T = np.meshgrid(np.arange(1, 7), np.arange(1, 7), indexing="ij")
    R, T = R.ravel().astype(np.int32), T.ravel().astype(np.int32)
    plaq, w = [], []
    for s in range(NSWEEP):
        ref.lib.ref_update(1)
        plaq.append(ref.lib.ref_plaquette())
        if s + 1 >= THERM and (s + 1) % EVERY == 0:
            U = ref.get_field()
            ref.lib.ref_ape(0, 20, 0.25)
            w.append(ref.wilson(R, T))
            ref.set_field(U)
    np.savez_compressed(os.path.join(HERE, "ref_chain_l12.npz"), beta=BETA, seed=SEED, plaq=np.array(plaq),
                        w=np.array(w), R=R, T=T, therm=THERM, every=EVERY)
    print("written", len(plaq), len(w))


if __name__ == "__main__":
    main()
