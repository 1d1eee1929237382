"""The reference's own entry points (include/lgt_compat.h) served by the GPU
engine, driven by a reference-style C host program (tests/compat/
compat_driver.c) that owns pu/latParams/runParams/measParams exactly as
LGT_Code/main/qcd.c does, compared with the FP64 oracle."""
import os

import numpy as np
import pytest

import reflogs
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "compat", "_build", "compat_driver")
DRIVER_LAZY = os.path.join(ROOT, "tests", "compat", "_build", "compat_driver_lazy")
SEED = ((1 << 30) << 32) | (1 << 31)     # drawn from compat_driver's ranlxd: 0.25, 0.5


def read_summary(path):
    out = {}
    for line in open(path):
        parts = line.split()
        if parts[0] == "WIL":
            out[("WIL", int(parts[1]), int(parts[2]))] = float(parts[3])
        else:
            out[parts[0]] = float(parts[1])
    return out


@pytest.mark.parametrize("prec,sync,order", [("fp64", "eager", "contiguous"), ("fp32", "eager", "contiguous"),
                                             ("fp64", "lazy", "contiguous"), ("fp32", "lazy", "contiguous"),
                                             ("fp64", "eager", "reversed"), ("fp64", "lazy", "reversed")])
def test_reference_entry_points_through_compat(tmp_path, prec, sync, order):
    """Both coherence modes give the same results; the lazy build (the
    program's lifecycle functions wrapped at link time) moves the field only
    when the program itself reads or writes it.  A field in the reference's
    contiguous layout goes to the engine in place, any other pointer table
    (here: sites reversed in the allocation) through the packing buffer."""
    driver = DRIVER if sync == "eager" else DRIVER_LAZY
    if not os.path.exists(driver):
        pytest.fail("compat driver not built (run __graft_entry__.build())")
    L = (6, 4, 4, 8)
    o = Oracle(L)
    U = o.hot(21)
    o.engine_sweep(U, 3.09, 1, 21, 0)
    U.tofile(tmp_path / "field.bin")
    env = dict(os.environ, LGT_N_OR="1", LGT_PRECISION=prec, COMPAT_SITE_ORDER=order)
    r = reflogs.run_program([driver, *map(str, L), str(tmp_path / "field.bin"), str(tmp_path)], env=env,
                            timeout=120)
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
    log = reflogs.read_log(tmp_path / "log.txt")
    assert "Update 1 done:\n(algorithm 1; sweep 0; acc=" in log and "update took" in log
    assert "calculating topological charge" in log


def lcg_uniforms(n, x=12345):
    """The compat driver's stand-in stream: x <- 6364136223846793005 x +
    1442695040888963407 mod 2^64, r = (x >> 11) / 2^53."""
    out = np.empty(n)
    for i in range(n):
        x = (6364136223846793005 * x + 1442695040888963407) & ((1 << 64) - 1)
        out[i] = (x >> 11) / 9007199254740992.0
    return out


@pytest.mark.parametrize("prec,sync", [("fp64", "eager"), ("fp64", "lazy"), ("fp32", "lazy")])
def test_reference_sweep_loop_through_per_link_entry_points(tmp_path, prec, sync):
    """A program that runs the reference's own sweep loop (update.c:122-166)
    with the exported per-link entry points: staples() of some links, then
    localHeatbathUpdate() for every site and direction in lexicographic order,
    its uniforms drawn from the program's stream, and project_gfield_to_sun.
    FP64: equal to the oracle's lexicographic sweep (the reference's
    algorithm, pinned bit for bit against it) on the same uniforms, with the
    same number of uniforms drawn and the same acceptance.  FP32 storage:
    unit links and a plausible acceptance (its FP32 staples move the
    acceptance decisions, so the draws diverge)."""
    driver = DRIVER if sync == "eager" else DRIVER_LAZY
    L = (6, 4, 4, 8)
    o = Oracle(L)
    U = o.hot(33)
    o.engine_sweep(U, 3.09, 1, 33, 0)
    U.tofile(tmp_path / "field.bin")
    env = dict(os.environ, LGT_PRECISION=prec)
    r = reflogs.run_program([driver, *map(str, L), str(tmp_path / "field.bin"), str(tmp_path), "perlink"], env=env,
                            timeout=120)
    assert r.returncode == 0, r.stderr
    s = read_summary(tmp_path / "summary.txt")
    stap = np.fromfile(tmp_path / "staples.bin", dtype=np.float64).reshape(-1, 4)
    ref_stap = np.array([o.staple(U, n, d) for n in range(0, o.vol, 37) for d in range(4)])
    got = np.fromfile(tmp_path / "perlink_sweep.bin", dtype=np.float64).reshape(o.vol, 4, 4)
    if prec == "fp64":
        assert np.array_equal(stap, ref_stap)
        V = U.copy()
        uni = lcg_uniforms(int(s["uniforms"]) + 64)
        acc, used = o.hb_sweep_lex(V, 3.09, uni)
        assert used == s["uniforms"]
        assert s["acc"] == pytest.approx(acc, rel=1e-14)
        d = float(np.max(np.abs(got - V)))
        print(f"per-link sweep vs oracle: max |dU| = {d:.1e}, {used} uniforms")
        assert d <= 1e-12
    else:
        assert np.max(np.abs(stap - ref_stap)) <= 1e-5 * np.max(np.abs(ref_stap))
        np.testing.assert_allclose(np.linalg.norm(got, axis=-1), 1.0, atol=3e-7)
        assert 0.7 < s["acc"] <= 1.0 and 4 * o.vol * 6 <= s["uniforms"] < 4 * o.vol * 12


def test_local_heatbath_degenerate_staple():
    """A link whose staple is exactly zero (on a cold field, U_mu(n+k) = -1 for
    the three k != mu makes the three forward terms -1 and leaves the three
    backward terms +1): lgt_staple returns 0, and lgt_local_heatbath asks for
    3 uniforms, then draws su2RandomMatrix from them (heatbath.c:133-138),
    equal to the oracle's degenerate branch."""
    import fields as F
    import masterarbeit_amd as lgt
    from oracle_lib import heatbath_batch
    lgt.load_library()
    L = (4, 4, 4, 4)
    n, mu = 85, 2
    U = np.zeros((256, 4, 4))
    U[..., 0] = 1.0
    for k in range(4):
        if k != mu:
            U[F.shift_index(L, k, 1)[n], mu, 0] = -1.0
    with lgt.Engine(L, lgt.FP64, 1) as e:
        e.upload(U)
        assert np.array_equal(e.staple(n, mu), np.zeros(4))
        assert e.local_heatbath(n, mu, 3.09, [])[:2] == (-1, 3)
        u = [0.3, 0.6, 0.9]
        trials, used, link = e.local_heatbath(n, mu, 3.09, u)
        G = e.download()
    ref, counts, consumed = heatbath_batch(np.zeros((1, 4)), 3.09, np.array(u))
    assert (trials, used, consumed, counts[0]) == (0, 3, 3, 0)
    assert np.array_equal(link, ref[0]) and np.array_equal(G[n, mu], ref[0])
    G[n, mu] = U[n, mu]
    assert np.array_equal(G, U)     # nothing else changed


def test_per_link_entry_points_reject_bad_links():
    """lgt_staple and lgt_local_heatbath return LGT_EINVAL for a site outside
    the lattice, a direction outside 0..3 and a negative or non-finite beta,
    and leave the field as it was."""
    import masterarbeit_amd as lgt
    lgt.load_library()
    with lgt.Engine((4, 4, 4, 4), lgt.FP64, 3) as e:
        e.init_field(1)
        U = e.download()
        for site, mu in [(-1, 0), (256, 0), (0, -1), (0, 4)]:
            with pytest.raises(lgt.LgtError, match="outside the lattice"):
                e.staple(site, mu)
            with pytest.raises(lgt.LgtError, match="outside the lattice"):
                e.local_heatbath(site, mu, 3.09, [0.5] * 6)
        for beta in (-1.0, float("nan"), float("inf")):
            with pytest.raises(lgt.LgtError, match="beta"):
                e.local_heatbath(5, 1, beta, [0.5] * 6)
        assert np.array_equal(e.download(), U)


QCD = os.path.join(ROOT, "oracle", "_ref", "qcd_l8", "qcd")


@pytest.mark.skipif(not os.path.exists(QCD), reason="reference program not built (oracle/Makefile ref)")
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reference_qcd_program_on_the_engine(tmp_path, prec):
    """The reference's own main/qcd.c (with its input, I/O and admin modules,
    compiled unchanged at 8^4) linked against liblgt_compat.so and the
    lifecycle interceptors: it reads its input file, thermalises from a hot
    start on the GPU, logs its plaquettes and writes configurations that the
    engine reads back with the reference's plaquette test."""
    import config_format as CF
    import masterarbeit_amd as lgt
    out = tmp_path / "out"
    out.mkdir()
    (tmp_path / "run.in").write_text(
        f"outFileId     1\noutputDir     {out}\nstart         1\nbeta          3.09\nepsilon       0.15\n"
        "nconf         12\nntherm        4\nndecorr       1\nwriteConfFreq 4\nseed          1337\n")
    env = dict(os.environ, LGT_N_OR="1", LGT_PRECISION=prec)
    r = reflogs.run_program([QCD, "-i", str(tmp_path / "run.in")], cwd=tmp_path, env=env, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    logs = list(out.glob("*.log"))
    assert len(logs) == 1
    # the startup block holds qcd.c's uninitialised cnfg_file (qcd.c:31,42,
    # inp_IO.c:244): arbitrary bytes, hence the byte-safe reader
    log = reflogs.read_log(logs[0])
    assert log.count("(algorithm 1; sweep 0; acc=") == 12
    plaq = reflogs.plaquettes(log)
    assert len(plaq) == 12 and plaq[0] < plaq[-1] and 0.65 < plaq[-1] < 0.8
    confs = sorted(out.glob("*_n*"))
    assert [c.name.rsplit("_", 1)[1] for c in confs] == ["n1", "n2"]
    lgt.load_library()
    for c, step in zip(confs, (7, 11)):
        dim, sun, L, p, U = CF.read(c)
        assert (dim, sun, L) == (4, 2, (8, 8, 8, 8))
        assert p == pytest.approx(plaq[step], abs=1e-8)
        # the FP32 engine evaluates the plaquette in FP64 too: the files of
        # either storage pass the FP64 engine's (the reference's) test
        with lgt.Engine(L, 0, 1) as e:
            p0, p1 = e.read_config(c)
        assert p0 == p


QTOP = os.path.join(ROOT, "oracle", "_ref", "qtop_l8", "qtop")


@pytest.mark.skipif(not os.path.exists(QTOP), reason="reference program not built (oracle/Makefile ref)")
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reference_qtop_program_on_the_engine(tmp_path, prec):
    """The reference's main_qtop/qtop.c compiled unchanged against the drop-in
    library: it reads a configuration, alternates 4D APE steps (sp = 0.1) with
    the improved charge 21 times and logs QTOP lines; the charges equal the
    FP64 oracle's on the same configuration.  In FP32 storage the program's
    readConfig plaquette test (config_IO.c:609-614, sqrt(6 VOL) DBL_EPSILON)
    still passes on the FP64-written file: the compat plaquette() evaluates
    pu's FP64 links as uploaded."""
    import masterarbeit_amd as lgt
    lgt.load_library()
    L = (8, 8, 8, 8)
    with lgt.Engine(L, 0, 9) as e:
        e.init_field(1)
        e.update(30, 1, 2.3)
        U = e.download()
        e.write_config(tmp_path / "8x8x8x8_SU2_b2.300000_id1_n1")
    env = dict(os.environ, LGT_PRECISION=prec)
    r = reflogs.run_program([QTOP, str(tmp_path)], cwd=tmp_path, env=env, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = reflogs.qtop_scan(reflogs.read_log(tmp_path / "8x8x8x8_SU2_b2.300000_id1.meas"))
    assert [n for n, _ in got] == list(range(21))
    o = Oracle(L)
    V = U.copy()
    for n, q in got:
        qo, qs = o.qtop(V)
        tol = max(1e-8 * abs(qo), 1e-13 * qs) if prec == "fp64" else 1e-5 * qs
        assert q == pytest.approx(qo, abs=tol)
        o.ape(V, 1, 1, 0.1)
    assert abs(got[-1][1]) > 1e-3     # smearing leaves a non-trivial charge density


MEAS = os.path.join(ROOT, "oracle", "_ref", "meas_l8", "meas")


@pytest.mark.skipif(not os.path.exists(MEAS), reason="reference program not built (oracle/Makefile ref)")
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reference_meas_program_on_the_engine(tmp_path, prec):
    """The reference's main_meas/meas.c compiled unchanged against the drop-in
    library (2-pt functions off): for two configurations it logs the charge
    after 4D APE and the Wilson loops after spatial APE, equal to the FP64
    oracle's on the same configurations."""
    import masterarbeit_amd as lgt
    lgt.load_library()
    L = (8, 8, 8, 8)
    base = tmp_path / "8x8x8x8_SU2_b2.300000_id1"
    fields = []
    with lgt.Engine(L, 0, 13) as e:
        e.init_field(1)
        for n in (1, 2):
            e.update(15, 1, 2.3)
            fields.append(e.download())
            e.write_config(f"{base}_n{n}")
    (tmp_path / "meas.in").write_text(
        f"outFileId      1\nodir           {tmp_path}\nbeta           2.3\nseed           1\nsconf          1\n"
        "fconf          2\neconf          1\nmtopcharge     1\nsmear_tc       10 0.1\nmwil           1 0\n"
        "tLoopExtent    1 4 1\ntSmear         0 0.0\nrLoopExtent    1 3 1\nrSmear         5 0.25\n"
        "mcorrs         0\nkappa          0.12\nsource         0 0 0 0\nstype          0\nnmax           100\n"
        "prec           1e-8\nn2pt           0\n")
    env = dict(os.environ, LGT_PRECISION=prec)
    r = reflogs.run_program([MEAS, "-i", str(tmp_path / "meas.in")], cwd=tmp_path, env=env, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    meas = reflogs.read_log(tmp_path / "8x8x8x8_SU2_b2.300000_id1.meas")
    qs = reflogs.meas_charges(meas)
    ws = reflogs.meas_loops(meas)
    assert len(qs) == 2 and len(ws) == 2 * 12
    o = Oracle(L)
    R = [r_ for r_ in (1, 2, 3) for _ in (1, 2, 3, 4)]
    T = [t_ for _ in (1, 2, 3) for t_ in (1, 2, 3, 4)]
    for k, U in enumerate(fields):
        V = U.copy()
        o.ape(V, 1, 10, 0.1)
        q, qsum = o.qtop(V)
        tol = max(1e-8 * abs(q), 1e-13 * qsum) if prec == "fp64" else 1e-5 * qsum
        assert qs[k] == pytest.approx(q, abs=tol)
        V = U.copy()
        o.ape(V, 0, 5, 0.25)
        w = o.wilson(V, R, T)
        got = ws[12 * k:12 * (k + 1)]
        assert [(a, b) for a, b, _ in got] == list(zip(R, T))
        np.testing.assert_allclose([c for _, _, c in got], w, rtol=1e-8 if prec == "fp64" else 1e-5)
