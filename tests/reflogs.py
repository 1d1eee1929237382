"""Reading what the reference's programs write: their .log and .meas files and
their stdout/stderr.

The text is decoded as latin-1, which maps every byte to one character and
cannot fail.  The reference's main/qcd.c passes the path buffer cnfg_file to
printStartupInfo (LGT_Code/main/qcd.c:31,42) without initialising it when
neither -c nor -t is given, and inp_IO.c:244 writes it into the log with %s:
the log then holds whatever bytes that stack buffer held, which need not be
valid UTF-8.  Nothing parsed below comes from that part of the log."""
import subprocess


def decode(raw):
    return raw.decode("latin-1")


def read_log(path):
    with open(path, "rb") as f:
        return decode(f.read())


def run_program(args, **kw):
    """subprocess.run with stdout/stderr captured and decoded as latin-1."""
    r = subprocess.run(args, capture_output=True, **kw)
    r.stdout = decode(r.stdout)
    r.stderr = decode(r.stderr)
    return r


def plaquettes(log):
    """The 'plaq\\t%.8f' lines of main/qcd.c:60."""
    return [float(l.split("\t")[1]) for l in log.splitlines() if l.startswith("plaq\t")]


def qtop_scan(meas):
    """The 'QTOP %d %d %e %+.8e' lines of main_qtop/qtop.c:85: (N_APE, Q)."""
    return [(int(l.split()[2]), float(l.split()[4])) for l in meas.splitlines() if l.startswith("QTOP ")]


def meas_charges(meas):
    """The 'QTOP %.8e' lines of main_meas/meas.c:98."""
    return [float(l.split()[1]) for l in meas.splitlines() if l.startswith("QTOP ")]


def meas_loops(meas):
    """The 'WIL %d %d\\t%.8e' lines of main_meas/meas.c:113: (R, T, W)."""
    return [(int(l.split()[1]), int(l.split()[2]), float(l.split()[3])) for l in meas.splitlines()
            if l.startswith("WIL ")]
