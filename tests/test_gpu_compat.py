"""The reference's own entry points (include/lgt_compat.h) served by the GPU
engine, driven by a reference-style C host program (tests/compat/
compat_driver.c) that owns pu/latParams/runParams/measParams exactly as
LGT_Code/main/qcd.c does, compared with the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "compat", "_build", "compat_driver")
DRIVER_LAZY = os.path.join(ROOT, "tests", "compat", "_build", "compat_driver_lazy")
SEED = ((1 << 30) << 32) | (1 << 31)     # drawn from the driver's ranlxd: 0.25, 0.5


def read_summary(path):
    out = {}
    for line in open(path):
        parts = line.split()
        if parts[0] == "WIL":
            out[("WIL", int(parts[1]), int(parts[2]))] = float(parts[3])
        else:
            out[parts[0]] = float(parts[1])
    return out


@pytest.mark.parametrize("prec,sync", [("fp64", "eager"), ("fp32", "eager"), ("fp64", "lazy"), ("fp32", "lazy")])
def test_reference_entry_points_through_compat(tmp_path, prec, sync):
    """Both coherence modes give the same results; the lazy build (the
    program's lifecycle functions wrapped at link time) moves the field only
    when the program itself reads or writes it."""
    driver = DRIVER if sync == "eager" else DRIVER_LAZY
    if not os.path.exists(driver):
        pytest.fail("compat driver not built (run __graft_entry__.build())")
    L = (6, 4, 4, 8)
    o = Oracle(L)
    U = o.hot(21)
    o.engine_sweep(U, 3.09, 1, 21, 0)
    U.tofile(tmp_path / "field.bin")
    env = dict(os.environ, LGT_N_OR="1", LGT_PRECISION=prec)
    r = subprocess.run([driver, *map(str, L), str(tmp_path / "field.bin"), str(tmp_path)], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    s = read_summary(tmp_path / "summary.txt")
    vol = o.vol
    load = lambda n: np.fromfile(tmp_path / n, dtype=np.float64).reshape(vol, 4, 4)
    tol = 1e-13 if prec == "fp64" else 3e-7
    assert s["plaq"] == pytest.approx(o.plaquette(U, acc=True), abs=tol)
    assert s["action"] == pytest.approx(3.09 * 6 * vol * (1 - s["plaq"]), rel=1e-15)
    R, T = [1, 1, 1, 2, 2, 2], [1, 2, 3, 1, 2, 3]
    w_ref = o.wilson(U, R, T)
    for r_, t_, w in zip(R, T, w_ref):
        assert s[("WIL", r_, t_)] == pytest.approx(w, abs=10 * tol)
    q, qs = o.qtop(U)
    assert abs(s["QTOP"] - q) <= (1e-13 if prec == "fp64" else 1e-5) * qs
    qc, qcs = o.qtop(U, "clover")
    assert abs(s["QCLOV"] - qc) <= (1e-13 if prec == "fp64" else 1e-5) * qcs
    V = U.copy()
    for name, mode, n, sp in [("ape_spatial.bin", 0, 3, 0.25), ("ape_all.bin", 1, 2, 1 / 6), ("ape_temporal.bin", 2, 1, 0.3)]:
        o.ape(V, mode, n, sp)
        d = np.max(np.abs(load(name) - V))
        assert d <= (0.0 if prec == "fp64" else 2e-6), (name, d)
    V = U.copy()
    for sw in range(2):
        o.engine_sweep(V, 3.09, 1, SEED, sw)
    if prec == "fp64":
        assert np.max(np.abs(load("update.bin") - V)) <= 1e-12
    else:
        G = load("update.bin")
        np.testing.assert_allclose(np.linalg.norm(G, axis=-1), 1.0, atol=3e-7)
    # Metropolis after the 2 heatbath sweeps: sweep indices 2 and 3
    V = U.copy()
    o.metropolis_update(V, 2, 3.09, 0.4, SEED, 2)
    if prec == "fp64":
        assert np.max(np.abs(load("metro.bin") - V)) <= 1e-12
    else:
        np.testing.assert_allclose(np.linalg.norm(load("metro.bin"), axis=-1), 1.0, atol=3e-7)
    if sync == "lazy":
        # two uploads (after each readConfig), a download per writeConfig
        assert (s["uploads"], s["downloads"]) == (2, 5)
    else:
        assert (s["uploads"], s["downloads"]) == (11, 5)
    log = open(tmp_path / "log.txt").read()
    assert "Update 1 done:\n(algorithm 1; sweep 0; acc=" in log and "update took" in log
    assert "calculating topological charge" in log
