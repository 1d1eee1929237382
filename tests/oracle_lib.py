"""ctypes bindings for the CPU checkers (test infrastructure only).

* :class:`Oracle` wraps ``oracle/_build/libsu2oracle.so``, the FP64 restatement of
  the reference algorithm (runtime lattice size, reference AoS layout).
* :class:`RefLib` wraps ``oracle/_ref/<size>/libref.so``, the reference's own C
  modules compiled from source with a fixed lattice size.  It exists only where
  the reference tree was present at build time; callers skip without it.

Fields are numpy float64 arrays of shape (VOL, 4, 4): site, direction, {c0..c3},
the reference's memory order (LGT_Code/modules/admin/init.c:842-868).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libsu2oracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

REF_SIZES = {
    "l4": (4, 4, 4, 4),
    "a6448": (6, 4, 4, 8),
    "l8": (8, 8, 8, 8),
    "l12": (12, 12, 12, 12),
    "l24": (24, 24, 24, 24),
    "l48": (48, 48, 48, 48),
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class OLat(C.Structure):
    _fields_ = [("L", C.c_int * 4), ("vol", C.c_int64), ("nb", C.POINTER(C.c_int64))]


class Oracle:
    """The FP64 restatement, bound to one lattice."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run __graft_entry__.build())")
            lib = C.CDLL(ORACLE_SO)
            P = C.POINTER(OLat)
            sig = {
                "or_lat_init": (C.c_int, [P, _ip]),
                "or_lat_free": (None, [P]),
                "or_qmul": (None, [_dp, _dp, _dp]),
                "or_qmul_dag": (None, [_dp, _dp, _dp]),
                "or_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
                "or_u53": (C.c_double, [C.c_uint32, C.c_uint32]),
                "or_heatbath_batch": (C.c_int, [C.c_int64, _dp, C.c_double, _dp, C.c_int64, _dp, _ip,
                                                C.POINTER(C.c_int64)]),
                "or_staple": (None, [P, _dp, C.c_int64, C.c_int, _dp]),
                "or_project_field": (None, [P, _dp]),
                "or_hb_sweep_lex": (C.c_double, [P, _dp, C.c_double, _dp, C.c_int64, C.POINTER(C.c_int64)]),
                "or_hb_subsweep_philox": (C.c_double, [P, _dp, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                                        C.c_uint32]),
                "or_or_subsweep": (None, [P, _dp, C.c_int, C.c_int]),
                "or_or_link": (None, [_dp, _dp, _dp]),
                "or_engine_sweep": (C.c_double, [P, _dp, C.c_double, C.c_int, C.c_uint64, C.c_uint32]),
                "or_hot_start": (None, [P, _dp, C.c_uint64]),
                "or_cold_start": (None, [P, _dp]),
                "or_plaquette": (C.c_double, [P, _dp]),
                "or_plaquette_acc": (C.c_double, [P, _dp]),
                "or_ape": (None, [P, _dp, C.c_int, C.c_int, C.c_double]),
                "or_wilson_loops": (C.c_int, [P, _dp, C.c_int, _ip, _ip, C.c_int, _dp]),
                "or_wloop_site": (C.c_double, [P, _dp, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int]),
                "or_qtop_improved": (C.c_double, [P, _dp, C.c_int, _dp, _dp]),
                "or_qtop_clover": (C.c_double, [P, _dp, C.c_int, _dp, _dp]),
                "or_wloop_objective": (C.c_double, [P, _dp, C.c_int, C.c_int]),
                "or_ape_vjp": (C.c_double, [P, _dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _dp]),
                "or_metropolis_batch": (C.c_int, [C.c_int64, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _ip]),
                "or_metro_subsweep_philox": (C.c_int64, [P, _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                                          C.c_uint64, C.c_uint32]),
                "or_metropolis_update": (C.c_double, [P, _dp, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                                       C.c_uint32]),
            }
            for name, (res, args) in sig.items():
                f = getattr(lib, name)
                f.restype = res
                f.argtypes = args
            cls._lib = lib
        return cls._lib

    def __init__(self, L):
        self.L = tuple(int(x) for x in L)
        self.vol = int(np.prod(self.L))
        self.lat = OLat()
        Larr = (C.c_int * 4)(*self.L)
        rc = self.lib().or_lat_init(C.byref(self.lat), Larr)
        if rc != 0:
            raise ValueError(f"bad lattice {L}")

    def __del__(self):
        try:
            self.lib().or_lat_free(C.byref(self.lat))
        except Exception:
            pass

    def field(self):
        return np.zeros((self.vol, 4, 4), dtype=np.float64)

    # --- field generation
    def cold(self):
        U = self.field()
        self.lib().or_cold_start(C.byref(self.lat), _ptr(U))
        return U

    def hot(self, seed):
        U = self.field()
        self.lib().or_hot_start(C.byref(self.lat), _ptr(U), C.c_uint64(seed))
        return U

    # --- updates
    def staple(self, U, n, mu):
        A = np.zeros(4)
        self.lib().or_staple(C.byref(self.lat), _ptr(U), n, mu, _ptr(A))
        return A

    def hb_sweep_lex(self, U, beta, uniforms):
        uniforms = np.ascontiguousarray(uniforms, dtype=np.float64)
        used = C.c_int64(0)
        acc = self.lib().or_hb_sweep_lex(C.byref(self.lat), _ptr(U), beta, _ptr(uniforms), uniforms.size,
                                          C.byref(used))
        return acc, used.value

    def hb_subsweep(self, U, mu, parity, beta, seed, sweep):
        return self.lib().or_hb_subsweep_philox(C.byref(self.lat), _ptr(U), mu, parity, beta,
                                                C.c_uint64(seed), sweep)

    def or_subsweep(self, U, mu, parity):
        self.lib().or_or_subsweep(C.byref(self.lat), _ptr(U), mu, parity)

    def engine_sweep(self, U, beta, n_or, seed, sweep):
        return self.lib().or_engine_sweep(C.byref(self.lat), _ptr(U), beta, n_or, C.c_uint64(seed), sweep)

    def metro_subsweep(self, U, mu, parity, beta, eps, seed, sweep):
        return self.lib().or_metro_subsweep_philox(C.byref(self.lat), _ptr(U), mu, parity, beta, eps,
                                                   C.c_uint64(seed), sweep)

    def metropolis_update(self, U, nsweeps, beta, eps, seed, sweep0=0):
        return self.lib().or_metropolis_update(C.byref(self.lat), _ptr(U), nsweeps, beta, eps, C.c_uint64(seed),
                                               sweep0)

    def project(self, U):
        self.lib().or_project_field(C.byref(self.lat), _ptr(U))

    # --- observables
    def plaquette(self, U, acc=False):
        f = self.lib().or_plaquette_acc if acc else self.lib().or_plaquette
        return f(C.byref(self.lat), _ptr(U))

    def ape(self, U, mode, nsteps, sp):
        self.lib().or_ape(C.byref(self.lat), _ptr(U), mode, nsteps, sp)

    def wilson(self, U, R, T, seq=False):
        R = np.ascontiguousarray(R, dtype=np.int32)
        T = np.ascontiguousarray(T, dtype=np.int32)
        w = np.zeros(len(R))
        rc = self.lib().or_wilson_loops(C.byref(self.lat), _ptr(U), len(R), _ptr(R, C.c_int), _ptr(T, C.c_int),
                                        1 if seq else 0, _ptr(w))
        if rc != 0:
            raise ValueError("wilson loop extents")
        return w

    def wloop_site(self, U, n, d1, d2, l1, l2):
        return self.lib().or_wloop_site(C.byref(self.lat), _ptr(U), n, d1, d2, l1, l2)

    def qtop(self, U, kind="improved", seq=False, density=False):
        qx = np.zeros(self.vol) if density else None
        s = C.c_double(0)
        f = self.lib().or_qtop_improved if kind == "improved" else self.lib().or_qtop_clover
        q = f(C.byref(self.lat), _ptr(U), 1 if seq else 0, _ptr(qx) if density else None, C.byref(s))
        return (q, s.value, qx) if density else (q, s.value)


    # --- adjoint of spatial APE
    def wloop_objective(self, U, R, T):
        return self.lib().or_wloop_objective(C.byref(self.lat), _ptr(U), R, T)

    def ape_vjp(self, U, nsteps, sp, R, T, mode=0):
        G = self.field()
        L = self.lib().or_ape_vjp(C.byref(self.lat), _ptr(U), mode, nsteps, sp, R, T, _ptr(G))
        return L, G


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    Oracle.lib().or_philox4x32_10(c, k, o)
    return tuple(o)


def metropolis_batch(staples, links, beta, eps, uniforms):
    """Oracle Metropolis hit per link (or_metropolis_batch): 4 uniforms per link."""
    staples = np.ascontiguousarray(staples, dtype=np.float64)
    links = np.ascontiguousarray(links, dtype=np.float64)
    uniforms = np.ascontiguousarray(uniforms, dtype=np.float64)
    n = len(staples)
    out = np.zeros((n, 4))
    acc = np.zeros(n, dtype=np.int32)
    Oracle.lib().or_metropolis_batch(n, _ptr(staples), _ptr(links), beta, eps, _ptr(uniforms), _ptr(out),
                                     _ptr(acc, C.c_int))
    return out, acc


def heatbath_batch(staples, beta, uniforms):
    staples = np.ascontiguousarray(staples, dtype=np.float64)
    uniforms = np.ascontiguousarray(uniforms, dtype=np.float64)
    n = staples.shape[0]
    links = np.zeros((n, 4))
    counts = np.zeros(n, dtype=np.int32)
    used = C.c_int64(0)
    rc = Oracle.lib().or_heatbath_batch(n, _ptr(staples), beta, _ptr(uniforms), uniforms.size, _ptr(links),
                                        _ptr(counts, C.c_int), C.byref(used))
    if rc != 0:
        raise ValueError("uniform buffer exhausted")
    return links, counts, used.value


class RefLib:
    """The reference's own modules for one compile-time lattice size."""

    _cache = {}

    @staticmethod
    def available(name):
        return os.path.exists(os.path.join(REF_DIR, name, "libref.so"))

    def __init__(self, name):
        self.name = name
        self.L = REF_SIZES[name]
        self.vol = int(np.prod(self.L))
        path = os.path.join(REF_DIR, name, "libref.so")
        if name not in RefLib._cache:
            lib = C.CDLL(path)
            sig = {
                "ref_setup": (C.c_int, [C.c_double, C.c_int]),
                "ref_set_beta": (None, [C.c_double]),
                "ref_set_field": (None, [_dp]),
                "ref_get_field": (None, [_dp]),
                "ref_init_field": (None, [C.c_int]),
                "ref_get_link": (None, [C.c_long, C.c_int, _dp]),
                "ref_su2_random": (None, [_dp]),
                "ref_rng_state_size": (C.c_int, []),
                "ref_rng_get": (None, [_ip]),
                "ref_rng_reset": (None, [_ip]),
                "ref_rng_draw": (None, [_dp, C.c_int]),
                "ref_staple": (None, [C.c_long, C.c_int, _dp]),
                "ref_heatbath_link": (C.c_double, [C.c_long, C.c_int]),
                "ref_metropolis_link": (C.c_double, [C.c_long, C.c_int, C.c_double]),
                "ref_update": (None, [C.c_int]),
                "ref_project": (None, []),
                "ref_plaquette": (C.c_double, []),
                "ref_gauge_action": (C.c_double, []),
                "ref_ape": (None, [C.c_int, C.c_int, C.c_double]),
                "ref_wilson": (None, [C.c_int, _ip, _ip, _dp]),
                "ref_qtop_improved": (C.c_double, []),
                "ref_qtop_clover": (C.c_double, []),
                "ref_write_config": (None, [C.c_char_p]),
                "ref_read_config": (None, [C.c_char_p]),
                "ref_read_config_doubled": (None, [C.c_char_p]),
            }
            for fname, (res, args) in sig.items():
                f = getattr(lib, fname)
                f.restype = res
                f.argtypes = args
            RefLib._cache[name] = lib
        self.lib = RefLib._cache[name]

    def setup(self, beta=3.09, seed=1):
        self.lib.ref_setup(beta, seed)

    def set_field(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self.lib.ref_set_field(_ptr(U))

    def write_config(self, path):
        self.lib.ref_write_config(str(path).encode())

    def read_config(self, path):
        """The reference's readConfig: on a header mismatch or a failed
        plaquette test it calls error(), which ends the process."""
        self.lib.ref_read_config(str(path).encode())

    def read_config_doubled(self, path):
        """readConfigForThermalisation: a configuration of half the extents,
        doubled by reflection; exits the process on a failed check."""
        self.lib.ref_read_config_doubled(str(path).encode())

    def get_field(self):
        U = np.zeros((self.vol, 4, 4))
        self.lib.ref_get_field(_ptr(U))
        return U

    def rng_state(self):
        st = np.zeros(self.lib.ref_rng_state_size(), dtype=np.int32)
        self.lib.ref_rng_get(_ptr(st, C.c_int))
        return st

    def rng_reset(self, st):
        st = np.ascontiguousarray(st, dtype=np.int32)
        self.lib.ref_rng_reset(_ptr(st, C.c_int))

    def rng_draw(self, n):
        r = np.zeros(n)
        self.lib.ref_rng_draw(_ptr(r), n)
        return r

    def staple(self, n, mu):
        A = np.zeros(4)
        self.lib.ref_staple(n, mu, _ptr(A))
        return A

    def get_link(self, n, mu):
        a = np.zeros(4)
        self.lib.ref_get_link(n, mu, _ptr(a))
        return a

    def su2_random(self):
        a = np.zeros(4)
        self.lib.ref_su2_random(_ptr(a))
        return a

    def wilson(self, R, T):
        R = np.ascontiguousarray(R, dtype=np.int32)
        T = np.ascontiguousarray(T, dtype=np.int32)
        w = np.zeros(len(R))
        self.lib.ref_wilson(len(R), _ptr(R, C.c_int), _ptr(T, C.c_int), _ptr(w))
        return w
