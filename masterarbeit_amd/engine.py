"""Python binding of the engine's C ABI (include/lgt_su2.h).

The engine itself is ``masterarbeit_amd/liblgt_su2.so`` (hand-written HIP for
gfx950).  This module only loads it with ctypes and converts numpy arrays; it
has no compute path of its own, and it raises if the library is missing.

Fields cross the boundary in the reference's layout: float64 arrays of shape
(VOL, 4, 4) = site, direction, {c0, c1, c2, c3}
(LGT_Code/modules/admin/init.c:842-868).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblgt_su2.so")

FP64, FP32 = 0, 1
APE_SPATIAL, APE_ALL, APE_TEMPORAL = 0, 1, 2
Q_IMPROVED, Q_CLOVER = 0, 1
OPT_TILE_ORDER, OPT_PROFILE, OPT_GROUP, OPT_ALTERNATE = 0, 1, 2, 3
(K_HEATBATH, K_OVERRELAX, K_PROJECT, K_PLAQUETTE, K_APE, K_WLOOP, K_QTOP, K_REDUCE, K_FIELD, K_APE_SAVE, K_VJP,
 K_WLOOP_GRAD, K_METROPOLIS) = range(13)
KERNEL_CLASSES = ["heatbath", "overrelax", "project", "plaquette", "ape", "wloop", "qtop", "reduce", "field",
                  "ape_save", "vjp", "wloop_grad", "metropolis"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# exported symbols and their signatures (include/lgt_su2.h)
SIGNATURES = {
    "lgt_create": (C.c_int, [_ip, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "lgt_destroy": (C.c_int, [C.c_void_p]),
    "lgt_strerror": (C.c_char_p, [C.c_int]),
    "lgt_last_error": (C.c_char_p, []),
    "lgt_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "lgt_get_extents": (C.c_int, [C.c_void_p, _ip, _ip]),
    "lgt_device_bytes": (C.c_int64, [C.c_void_p]),
    "lgt_init_field": (C.c_int, [C.c_void_p, C.c_int]),
    "lgt_upload": (C.c_int, [C.c_void_p, _dp]),
    "lgt_download": (C.c_int, [C.c_void_p, _dp]),
    "lgt_upload_plaquette": (C.c_int, [C.c_void_p, _dp, _dp]),
    "lgt_staple": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, _dp]),
    "lgt_local_heatbath": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_double, _dp, C.c_int, _dp, _ip, _ip]),
    "lgt_write_config": (C.c_int, [C.c_void_p, C.c_char_p]),
    "lgt_read_config": (C.c_int, [C.c_void_p, C.c_char_p, _dp, _dp]),
    "lgt_read_config_doubled": (C.c_int, [C.c_void_p, C.c_char_p, _dp, _dp]),
    "lgt_update": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp]),
    "lgt_heatbath_subsweep": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp]),
    "lgt_overrelax_subsweep": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "lgt_project": (C.c_int, [C.c_void_p]),
    "lgt_set_sweep_index": (C.c_int, [C.c_void_p, C.c_uint32]),
    "lgt_gauge_transform": (C.c_int, [C.c_void_p, C.c_uint64, _dp]),
    "lgt_get_sweep_index": (C.c_uint32, [C.c_void_p]),
    "lgt_plaquette": (C.c_int, [C.c_void_p, _dp]),
    "lgt_gauge_action": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "lgt_ape": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
    "lgt_wilson_loops": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _dp]),
    "lgt_qtop": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "lgt_ape_vjp_forward": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int]),
    "lgt_ape_vjp_backward": (C.c_int, [C.c_void_p, _dp, _dp]),
    "lgt_wloop_sum": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "lgt_wloop_grad": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
    "lgt_heatbath_harness": (C.c_int, [C.c_int, _dp, C.c_double, _dp, _ip, _ip, _dp, _ip]),
    "lgt_overrelax_harness": (C.c_int, [C.c_int, _dp, _dp, _dp]),
    "lgt_metropolis_update": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp]),
    "lgt_metropolis_harness": (C.c_int, [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _ip]),
    "lgt_plaquette_aos_noalloc": (C.c_int, [_ip, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lgt_profile_read": (C.c_int, [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_int64)]),
    "lgt_profile_reset": (C.c_int, [C.c_void_p]),
    "lgt_sync": (C.c_int, [C.c_void_p]),
}

_lib = None


class LgtError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Loads the HIP engine; raises if it was not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise LgtError(f"HIP engine library not built: {path} (run __graft_entry__.build())")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _check(rc, what):
    if rc != 0:
        lib = load_library()
        msg = lib.lgt_last_error().decode()
        raise LgtError(f"{what}: {lib.lgt_strerror(rc).decode()} ({msg})")


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Engine:
    """One lattice resident on the current HIP device."""

    def __init__(self, extents, precision=FP64, seed=1):
        lib = load_library()
        self.L = tuple(int(x) for x in extents)
        self.vol = int(np.prod(self.L))
        self.precision = precision
        h = C.c_void_p()
        ext = (C.c_int * 4)(*self.L)
        _check(lib.lgt_create(ext, precision, C.c_uint64(seed), C.byref(h)), "lgt_create")
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lgt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- options / profiling
    def set_option(self, opt, value):
        _check(self._lib.lgt_set_option(self._h, opt, int(value)), "lgt_set_option")

    def profile_read(self, cls):
        ms = C.c_double(0)
        n = C.c_int64(0)
        _check(self._lib.lgt_profile_read(self._h, cls, C.byref(ms), C.byref(n)), "lgt_profile_read")
        return ms.value, n.value

    def profile_reset(self):
        _check(self._lib.lgt_profile_reset(self._h), "lgt_profile_reset")

    def sync(self):
        _check(self._lib.lgt_sync(self._h), "lgt_sync")

    def device_bytes(self):
        return self._lib.lgt_device_bytes(self._h)

    def extents(self):
        """(lattice extents, precision) the engine was created with."""
        L = (C.c_int * 4)()
        prec = C.c_int(0)
        _check(self._lib.lgt_get_extents(self._h, L, C.byref(prec)), "lgt_get_extents")
        return tuple(L), prec.value

    # ---- field
    def init_field(self, start):
        _check(self._lib.lgt_init_field(self._h, start), "lgt_init_field")

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != 16 * self.vol:
            raise ValueError("field size mismatch")
        _check(self._lib.lgt_upload(self._h, _ptr(U)), "lgt_upload")

    def upload_plaquette(self, U):
        """Uploads U and returns the FP64 plaquette of its links as given."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != 16 * self.vol:
            raise ValueError("field size mismatch")
        p = C.c_double(0)
        _check(self._lib.lgt_upload_plaquette(self._h, _ptr(U), C.byref(p)), "lgt_upload_plaquette")
        return p.value

    def staple(self, site, mu):
        """staples(n, mu) of the current field at lexicographic site n."""
        A = np.zeros(4)
        _check(self._lib.lgt_staple(self._h, int(site), int(mu), _ptr(A)), "lgt_staple")
        return A

    def local_heatbath(self, site, mu, beta, uniforms):
        """localHeatbathUpdate of one link with the given uniforms: returns
        (trials, needed, new link); trials = -1 (and no change) when the
        uniforms run out, needed = how many the next call needs."""
        u = np.ascontiguousarray(uniforms, dtype=np.float64).ravel()
        link = np.zeros(4)
        trials = C.c_int(0)
        needed = C.c_int(0)
        _check(self._lib.lgt_local_heatbath(self._h, int(site), int(mu), beta, _ptr(u) if u.size else None,
                                            int(u.size), _ptr(link), C.byref(trials), C.byref(needed)),
               "lgt_local_heatbath")
        return trials.value, needed.value, (link if trials.value >= 0 else None)

    def write_config(self, path):
        _check(self._lib.lgt_write_config(self._h, os.fsencode(path)), "lgt_write_config")

    def read_config(self, path):
        """Loads a reference configuration file; returns (file plaquette,
        recomputed plaquette).  Raises on a header mismatch or a failed
        plaquette test (the field then stays loaded)."""
        p0 = C.c_double(0)
        p1 = C.c_double(0)
        _check(self._lib.lgt_read_config(self._h, os.fsencode(path), C.byref(p0), C.byref(p1)), "lgt_read_config")
        return p0.value, p1.value

    def read_config_doubled(self, path):
        """Loads a configuration of half the extents, doubled by reflection
        (readConfigForThermalisation); returns (file, recomputed) plaquette."""
        p0 = C.c_double(0)
        p1 = C.c_double(0)
        _check(self._lib.lgt_read_config_doubled(self._h, os.fsencode(path), C.byref(p0), C.byref(p1)),
               "lgt_read_config_doubled")
        return p0.value, p1.value

    def download(self):
        U = np.empty((self.vol, 4, 4), dtype=np.float64)
        _check(self._lib.lgt_download(self._h, _ptr(U)), "lgt_download")
        return U

    # ---- update
    def update(self, nsweeps, n_or, beta, acceptance=False):
        acc = C.c_double(0)
        _check(self._lib.lgt_update(self._h, nsweeps, n_or, beta, C.byref(acc) if acceptance else None),
               "lgt_update")
        return acc.value if acceptance else None

    def metropolis_update(self, nsweeps, beta, eps, acceptance=False):
        acc = C.c_double(0)
        _check(self._lib.lgt_metropolis_update(self._h, nsweeps, beta, eps, C.byref(acc) if acceptance else None),
               "lgt_metropolis_update")
        return acc.value if acceptance else None

    def heatbath_subsweep(self, mu, parity, beta, acceptance=False):
        acc = C.c_double(0)
        _check(self._lib.lgt_heatbath_subsweep(self._h, mu, parity, beta, C.byref(acc) if acceptance else None),
               "lgt_heatbath_subsweep")
        return acc.value if acceptance else None

    def overrelax_subsweep(self, mu, parity):
        _check(self._lib.lgt_overrelax_subsweep(self._h, mu, parity), "lgt_overrelax_subsweep")

    def gauge_transform(self, seed, elements=False):
        """Random gauge transformation of the field in place; returns g(x)
        per lexicographic site if elements is set."""
        G = np.empty((self.vol, 4)) if elements else None
        _check(self._lib.lgt_gauge_transform(self._h, C.c_uint64(seed), _ptr(G) if elements else None),
               "lgt_gauge_transform")
        return G

    def project(self):
        _check(self._lib.lgt_project(self._h), "lgt_project")

    @property
    def sweep_index(self):
        return self._lib.lgt_get_sweep_index(self._h)

    @sweep_index.setter
    def sweep_index(self, v):
        _check(self._lib.lgt_set_sweep_index(self._h, v), "lgt_set_sweep_index")

    # ---- observables
    def plaquette(self):
        p = C.c_double(0)
        _check(self._lib.lgt_plaquette(self._h, C.byref(p)), "lgt_plaquette")
        return p.value

    def gauge_action(self, beta):
        a = C.c_double(0)
        _check(self._lib.lgt_gauge_action(self._h, beta, C.byref(a)), "lgt_gauge_action")
        return a.value

    def ape(self, mode, nsteps, sp):
        _check(self._lib.lgt_ape(self._h, mode, nsteps, sp), "lgt_ape")

    def wilson_loops(self, R, T):
        R = np.ascontiguousarray(R, dtype=np.int32).ravel()
        T = np.ascontiguousarray(T, dtype=np.int32).ravel()
        if len(R) != len(T):
            raise ValueError(f"wilson_loops: {len(R)} spatial extents but {len(T)} temporal extents")
        w = np.zeros(len(R))
        _check(self._lib.lgt_wilson_loops(self._h, len(R), _ptr(R, C.c_int), _ptr(T, C.c_int), _ptr(w)),
               "lgt_wilson_loops")
        return w

    def qtop(self, kind=Q_IMPROVED, density=False):
        q = C.c_double(0)
        s = C.c_double(0)
        qx = np.zeros(self.vol) if density else None
        _check(self._lib.lgt_qtop(self._h, kind, C.byref(q), C.byref(s), _ptr(qx) if density else None),
               "lgt_qtop")
        return (q.value, s.value, qx) if density else (q.value, s.value)


    # ---- adjoint of APE smearing (mode: APE_SPATIAL, APE_ALL, APE_TEMPORAL)
    def ape_vjp_forward(self, nsteps, sp, segment=0, mode=APE_SPATIAL):
        _check(self._lib.lgt_ape_vjp_forward(self._h, mode, nsteps, sp, segment), "lgt_ape_vjp_forward")

    def ape_vjp_backward(self, cotangent=None, grad=True):
        cot = None if cotangent is None else np.ascontiguousarray(cotangent, dtype=np.float64)
        if cot is not None and cot.size != 16 * self.vol:
            raise ValueError(f"ape_vjp_backward: cotangent has {cot.size} values, the field {16 * self.vol}")
        G = np.empty((self.vol, 4, 4), dtype=np.float64) if grad else None
        _check(self._lib.lgt_ape_vjp_backward(self._h, _ptr(cot) if cot is not None else None,
                                              _ptr(G) if grad else None), "lgt_ape_vjp_backward")
        return G

    def wloop_sum(self, R, T):
        """(hi, lo) of the unnormalised loop sum."""
        out = np.zeros(2)
        _check(self._lib.lgt_wloop_sum(self._h, R, T, _ptr(out)), "lgt_wloop_sum")
        return out[0], out[1]

    def wloop_grad(self, R, T, host=True):
        G = np.empty((self.vol, 4, 4), dtype=np.float64) if host else None
        _check(self._lib.lgt_wloop_grad(self._h, R, T, _ptr(G) if host else None), "lgt_wloop_grad")
        return G


def heatbath_harness(staples, beta, uniforms, offsets, lens):
    lib = load_library()
    staples = np.ascontiguousarray(staples, dtype=np.float64)
    uniforms = np.ascontiguousarray(uniforms, dtype=np.float64).ravel()
    offsets = np.ascontiguousarray(offsets, dtype=np.int32).ravel()
    lens = np.ascontiguousarray(lens, dtype=np.int32).ravel()
    n = staples.shape[0]
    if staples.shape != (n, 4) or offsets.size != n or lens.size != n:
        raise ValueError("heatbath_harness: shapes")
    if n and (offsets.min() < 0 or lens.min() < 0 or (offsets.astype(np.int64) + lens).max() > uniforms.size):
        raise ValueError("heatbath_harness: uniform ranges outside the uniform array")
    links = np.zeros((n, 4))
    counts = np.zeros(n, dtype=np.int32)
    _check(lib.lgt_heatbath_harness(n, _ptr(staples), beta, _ptr(uniforms), _ptr(offsets, C.c_int),
                                    _ptr(lens, C.c_int), _ptr(links), _ptr(counts, C.c_int)), "lgt_heatbath_harness")
    return links, counts


def overrelax_harness(staples, links_in):
    lib = load_library()
    staples = np.ascontiguousarray(staples, dtype=np.float64)
    links_in = np.ascontiguousarray(links_in, dtype=np.float64)
    n = staples.shape[0]
    if staples.shape != (n, 4) or links_in.shape != (n, 4):
        raise ValueError("overrelax_harness: shapes")
    out = np.zeros((n, 4))
    _check(lib.lgt_overrelax_harness(n, _ptr(staples), _ptr(links_in), _ptr(out)), "lgt_overrelax_harness")
    return out


def metropolis_harness(staples, links_in, beta, eps, uniforms):
    """One Metropolis hit per link on the GPU (FP64) with injected uniforms
    (4 per link); returns (new links, accepted flags)."""
    lib = load_library()
    staples = np.ascontiguousarray(staples, dtype=np.float64)
    links_in = np.ascontiguousarray(links_in, dtype=np.float64)
    uniforms = np.ascontiguousarray(uniforms, dtype=np.float64)
    n = staples.shape[0]
    if links_in.shape != (n, 4) or uniforms.shape != (n, 4):
        raise ValueError("metropolis_harness: shapes")
    out = np.zeros((n, 4))
    acc = np.zeros(n, dtype=np.int32)
    _check(lib.lgt_metropolis_harness(n, _ptr(staples), _ptr(links_in), beta, eps, _ptr(uniforms), _ptr(out),
                                      _ptr(acc, C.c_int)), "lgt_metropolis_harness")
    return out, acc
