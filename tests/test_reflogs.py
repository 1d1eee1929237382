"""The parsers of the reference programs' logs (tests/reflogs.py) on logs
whose startup block holds bytes that are not UTF-8: main/qcd.c prints its
uninitialised configuration path there when it starts without -c or -t
(LGT_Code/main/qcd.c:31,42; LGT_Code/modules/io/inp_IO.c:244)."""
import os
import sys

import reflogs

HEADER = (b"*******************************\nLATTICE QCD CODE\n*******************************\n"
          b"Global parameters:\n*******************************\nUsing a four -dimensional lattice:\n"
          b"latt  :\t8x8x8x8\nbeta :\t3.090000\nepsilon :\t0.150000\nnconf :\t12\nndecorr :\t1\nntherm :\t4\n"
          b"writeConfFreq :\t4\nseed :\t1337\n*******************************\nInline measurements:\n"
          b"None at the moment!\n*******************************\n")


def qcd_log(garbage):
    body = b"".join(b"\nUpdate %d done:\n(algorithm 1; sweep 0; acc=1.000)\nupdate took 0.000123456 seconds\n"
                    b"plaq\t%.8f\n" % (i + 1, 0.5 + 0.01 * i) for i in range(3))
    return HEADER + b'Start from "' + garbage + b'" configuration:\n*******************************\n' + body


def test_log_with_stack_garbage_parses(tmp_path):
    garbage = bytes([0x10, 0xfe, 0xff, 0x80, 0xc3, 0x28, 0x7f])   # none of it valid UTF-8
    raw = qcd_log(garbage)
    try:
        raw.decode("utf-8")
        raise AssertionError("the fixture is meant not to be UTF-8")
    except UnicodeDecodeError:
        pass
    p = tmp_path / "8x8x8x8_SU2_b3.090000_id1.log"
    p.write_bytes(raw)
    log = reflogs.read_log(p)
    assert log.count("(algorithm 1; sweep 0; acc=") == 3
    assert reflogs.plaquettes(log) == [0.5, 0.51, 0.52]
    assert log.encode("latin-1") == raw       # nothing lost or replaced


def test_program_output_with_stack_garbage_decodes(tmp_path):
    prog = tmp_path / "emit.py"
    prog.write_text("import sys\nsys.stdout.buffer.write(bytes([0x41, 0xfe, 0x42]))\n"
                    "sys.stderr.buffer.write(bytes([0xff]))\n")
    r = reflogs.run_program([sys.executable, str(prog)], timeout=60)
    assert r.returncode == 0
    assert r.stdout == "A\xfeB" and r.stderr == "\xff"


def test_meas_line_parsers():
    meas = "Configuration imported (took 1e-3 sec)!\nQTOP 3.25000000e-01\nWIL 1 2\t4.5e-01\nWIL 2 1\t4.4e-01\n"
    assert reflogs.meas_charges(meas) == [0.325]
    assert reflogs.meas_loops(meas) == [(1, 2, 0.45), (2, 1, 0.44)]
    scan = "QTOP 1 0 1.000000e-01 +1.50000000e-02\nQTOP 1 1 1.000000e-01 -2.50000000e-02\n"
    assert reflogs.qtop_scan(scan) == [(0, 0.015), (1, -0.025)]
