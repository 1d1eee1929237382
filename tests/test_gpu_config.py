"""Configuration files in the reference's binary format through the engine
(lgt_write_config / lgt_read_config) against the reference's own writeConfig
and readConfig (LGT_Code/modules/io/config_IO.c:40-117,537-621)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import config_format as CF
from oracle_lib import Oracle, RefLib

pytestmark = pytest.mark.gpu

L = (6, 4, 4, 8)     # asymmetric: the extents' order is checked


@pytest.fixture(scope="module")
def lgt():
    import masterarbeit_amd as m
    m.load_library()
    return m


def engine(lgt, prec=0, seed=3):
    return lgt.Engine(L, prec, seed)


def test_engine_writes_reference_format(lgt, tmp_path):
    with engine(lgt) as e:
        e.init_field(1)
        e.update(2, 1, 3.09)
        U = e.download()
        p = e.plaquette()
        e.write_config(tmp_path / "conf")
    dim, sun, Lf, plaq, V = CF.read(tmp_path / "conf")
    assert (dim, sun, Lf) == (4, 2, L)
    assert plaq == p
    np.testing.assert_array_equal(V, U)
    assert plaq == pytest.approx(Oracle(L).plaquette(U, acc=True), abs=1e-14)


@pytest.mark.skipif(not RefLib.available("a6448"), reason="reference build missing")
def test_reference_reads_engine_file(lgt, tmp_path):
    """The reference's readConfig accepts the engine's file (its plaquette
    test passes); run in a child process because readConfig exits on failure."""
    with engine(lgt) as e:
        e.init_field(1)
        e.update(1, 1, 3.09)
        U = e.download()
        e.write_config(tmp_path / "conf")
    np.save(tmp_path / "U.npy", U)
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {os.path.dirname(__file__)!r})\n"
        "from oracle_lib import RefLib\n"
        "r = RefLib('a6448'); r.setup(3.09, 1); r.lib.ref_init_field(0)\n"
        f"r.read_config({str(tmp_path / 'conf')!r})\n"
        f"assert (r.get_field() == np.load({str(tmp_path / 'U.npy')!r})).all()\n"
        "print('ok')\n")
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "ok" in res.stdout, res.stdout + res.stderr


@pytest.mark.skipif(not RefLib.available("a6448"), reason="reference build missing")
@pytest.mark.parametrize("prec", [0, 1])
def test_engine_reads_reference_file(lgt, tmp_path, prec):
    ref = RefLib("a6448")
    ref.setup(3.09, 9)
    ref.lib.ref_init_field(1)
    ref.lib.ref_update(2)
    U = ref.get_field()
    ref.write_config(tmp_path / "conf")
    with engine(lgt, prec) as e:
        p0, p1 = e.read_config(tmp_path / "conf")
        V = e.download()
    assert p0 == ref.lib.ref_plaquette()
    # the reference's own test in both storages: the plaquette of the links as
    # read, before FP32 storage rounds them
    assert abs(p1 - p0) <= np.sqrt(6 * np.prod(L)) * np.finfo(np.float64).eps
    if prec == 0:
        np.testing.assert_array_equal(V, U)
    else:
        np.testing.assert_allclose(V, U, atol=1e-7)


@pytest.mark.parametrize("prec", [0, 1])
def test_upload_plaquette_is_that_of_the_fp64_links(lgt, prec):
    """lgt_upload_plaquette returns the plaquette of the FP64 links as given
    (the oracle's, to the rounding of the sum); for FP32 storage the plaquette
    of the rounded links on the device differs from it by far more than
    readConfig's tolerance sqrt(6 VOL) DBL_EPSILON."""
    o = Oracle(L)
    U = o.hot(41)
    o.engine_sweep(U, 3.09, 1, 41, 0)
    ref = o.plaquette(U, acc=True)
    with engine(lgt, prec) as e:
        p = e.upload_plaquette(U)
        p_dev = e.plaquette()
    assert p == pytest.approx(ref, abs=1e-15)
    tol = np.sqrt(6 * np.prod(L)) * np.finfo(np.float64).eps
    if prec == 0:
        assert abs(p_dev - ref) <= 1e-15
    else:
        print(f"FP32 storage: plaquette of the rounded links - FP64 plaquette = {p_dev - ref:.2e} (tolerance {tol:.1e})")
        assert abs(p_dev - ref) > tol


def test_read_config_rejects_mismatch(lgt, tmp_path):
    U = np.zeros((int(np.prod(L)), 4, 4))
    U[:, :, 0] = 1.0
    CF.write(tmp_path / "wrong_extent", (6, 4, 8, 4), 1.0, U)
    CF.write(tmp_path / "wrong_sun", L, 1.0, U, sun=3)
    CF.write(tmp_path / "bad_plaq", L, 0.99, U)
    CF.write(tmp_path / "good", L, 1.0, U)
    with engine(lgt) as e:
        for name, msg in (("wrong_extent", b"lattice"), ("wrong_sun", b"SUN"), ("bad_plaq", b"plaquette test")):
            with pytest.raises(lgt.LgtError):
                e.read_config(tmp_path / name)
            assert msg in e._lib.lgt_last_error()
        assert e.read_config(tmp_path / "good") == (1.0, 1.0)
        with pytest.raises(lgt.LgtError):
            e.read_config(tmp_path / "missing")


@pytest.mark.parametrize("small,big", [("l4", "l8"), ("l12", "l24")])
def test_reflection_doubling_matches_reference(lgt, tmp_path, small, big):
    """lgt_read_config_doubled against the reference's
    readConfigForThermalisation (config_IO.c:348-535) on the same file: the
    doubled fields are identical and both pass the plaquette test."""
    if not (RefLib.available(small) and RefLib.available(big)):
        pytest.skip("reference build missing")
    ref = RefLib(small)
    ref.setup(3.09, 4)
    ref.lib.ref_init_field(1)
    ref.lib.ref_update(3)
    ref.write_config(tmp_path / "small")
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {os.path.dirname(__file__)!r})\n"
        "from oracle_lib import RefLib\n"
        f"r = RefLib({big!r}); r.setup(3.09, 1); r.lib.ref_init_field(0)\n"
        f"r.read_config_doubled({str(tmp_path / 'small')!r})\n"
        f"np.save({str(tmp_path / 'big.npy')!r}, r.get_field())\n"
        "print('ok')\n")
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "ok" in res.stdout, res.stdout + res.stderr
    Lbig = RefLib(big).L
    with lgt.Engine(Lbig, 0, 1) as e:
        p0, p1 = e.read_config_doubled(tmp_path / "small")
        V = e.download()
    np.testing.assert_array_equal(V, np.load(tmp_path / "big.npy"))
    assert p0 == ref.lib.ref_plaquette()
    assert abs(p1 - p0) <= np.sqrt(6 * np.prod(Lbig)) * np.finfo(np.float64).eps


def test_reflection_doubling_rejects_wrong_size(lgt, tmp_path):
    U = np.zeros((4 ** 4, 4, 4))
    U[:, :, 0] = 1.0
    CF.write(tmp_path / "c4", (4, 4, 4, 4), 1.0, U)
    with lgt.Engine((6, 8, 8, 8), 0, 1) as e:
        with pytest.raises(lgt.LgtError):
            e.read_config_doubled(tmp_path / "c4")
    with lgt.Engine((8, 8, 8, 8), 0, 1) as e:
        assert e.read_config_doubled(tmp_path / "c4") == (1.0, 1.0)
        assert np.all(e.download()[:, :, 0] == 1.0)
