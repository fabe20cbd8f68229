"""Python front end of liburbansph.so (the C ABI in include/urbansph.h).

Loads the in-tree HIP library and raises if it is missing: there is no CPU
fallback.  ``Case`` wraps one context with numpy arrays in and out.
"""
import ctypes as C
import os

import numpy as np

import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liburbansph.so")

_lib = None


def lib():
    """Load the HIP library (once).  Raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.sph_params_default.argtypes = [C.POINTER(abi.Params)]
    L.sph_create.argtypes = [C.POINTER(abi.Params), C.c_int, C.POINTER(vp)]
    L.sph_upload.argtypes = [vp, vp, vp, vp, C.c_int]
    L.sph_step.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    L.sph_download.argtypes = [vp, vp, vp, vp, C.c_int]
    L.sph_get_field_f64.argtypes = [vp, C.c_int, vp, C.c_int]
    L.sph_get_field_i32.argtypes = [vp, C.c_int, vp, C.c_int]
    L.sph_nnps_debug.argtypes = [vp, vp, vp, vp, vp]
    L.sph_neighbors.argtypes = [vp, C.c_int, vp, vp, C.POINTER(C.c_longlong)]
    L.sph_domain_info.argtypes = [vp, C.POINTER(abi.Domain)]
    L.sph_ppe_info.argtypes = [vp, C.POINTER(abi.PPEReport)]
    L.sph_stats.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.sph_synchronize.argtypes = [vp]
    L.sph_profile_enable.argtypes = [vp, C.c_int]
    L.sph_profile_report.argtypes = [vp, C.POINTER(abi.KernelTime), C.c_int]
    L.sph_device_bytes.restype = C.c_longlong
    L.sph_device_bytes.argtypes = [vp]
    L.sph_last_error.restype = C.c_char_p
    L.sph_last_error.argtypes = [vp]
    L.sph_destroy.argtypes = [vp]
    _lib = L
    return L


EXPORTED = [
    "sph_params_default", "sph_create", "sph_upload", "sph_step", "sph_download",
    "sph_get_field_f64", "sph_get_field_i32", "sph_nnps_debug", "sph_neighbors",
    "sph_domain_info", "sph_ppe_info", "sph_stats", "sph_synchronize",
    "sph_profile_enable", "sph_profile_report", "sph_device_bytes",
    "sph_last_error", "sph_destroy",
]


def default_params():
    p = abi.Params()
    lib().sph_params_default(C.byref(p))
    return p


class SPHError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"rc={rc}: {msg}")
        self.rc = rc


class Case:
    """One simulation case on the GPU (one sph_ctx)."""

    def __init__(self, params, n_part2):
        self.L = lib()
        self.params = params
        h = C.c_void_p()
        rc = self.L.sph_create(C.byref(params), int(n_part2), C.byref(h))
        if rc:
            raise SPHError(rc, "sph_create failed")
        self.h = h
        self.n = 0

    def _check(self, rc, what):
        if rc:
            raise SPHError(rc, f"{what}: {self.L.sph_last_error(self.h).decode()}")

    def upload(self, P1, P2=None, normals=None):
        P1 = np.ascontiguousarray(P1)
        assert P1.dtype == abi.PART1
        if P2 is not None:
            P2 = np.ascontiguousarray(P2)
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32)
        self._check(self.L.sph_upload(self.h, abi.ptr(P1), abi.ptr(P2), abi.ptr(nrm), len(P1)), "sph_upload")
        self.n = len(P1)

    def step(self, count, dt, time=None):
        if time is None:
            time = count * dt
        self._check(self.L.sph_step(self.h, int(count), float(dt), float(time)), "sph_step")

    def sync(self):
        self._check(self.L.sph_synchronize(self.h), "sph_synchronize")

    def field(self, name):
        out = np.empty(self.n, np.float64)
        self._check(self.L.sph_get_field_f64(self.h, abi.F[name], abi.ptr(out), self.n), "get_field")
        return out

    def ifield(self, name):
        out = np.empty(self.n, np.int32)
        self._check(self.L.sph_get_field_i32(self.h, abi.I[name], abi.ptr(out), self.n), "get_field_i32")
        return out

    def download(self):
        P1 = np.zeros(self.n, abi.PART1)
        P2 = np.zeros(self.n, abi.PART2)
        P3 = np.zeros(self.n, abi.PART3)
        self._check(self.L.sph_download(self.h, abi.ptr(P1), abi.ptr(P2), abi.ptr(P3), self.n), "download")
        return P1, P2, P3

    def domain(self):
        d = abi.Domain()
        self._check(self.L.sph_domain_info(self.h, C.byref(d)), "domain")
        return d

    def nnps(self):
        d = self.domain()
        keys = np.empty(self.n, np.int32)
        perm = np.empty(self.n, np.int32)
        gs = np.empty(d.num_cells, np.int32)
        ge = np.empty(d.num_cells, np.int32)
        self._check(self.L.sph_nnps_debug(self.h, abi.ptr(keys), abi.ptr(perm), abi.ptr(gs), abi.ptr(ge)), "nnps")
        return keys, perm, gs, ge

    def neighbors(self, kind):
        tot = C.c_longlong(0)
        self._check(self.L.sph_neighbors(self.h, kind, None, None, C.byref(tot)), "neighbors")
        off = np.empty(self.n + 1, np.int32)
        nbr = np.empty(max(tot.value, 1), np.int32)
        self._check(self.L.sph_neighbors(self.h, kind, abi.ptr(off), abi.ptr(nbr), C.byref(tot)), "neighbors")
        return off, nbr[:tot.value]

    def ppe_report(self):
        r = abi.PPEReport()
        self.L.sph_ppe_info(self.h, C.byref(r))
        return r

    def stats(self):
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        self._check(self.L.sph_stats(self.h, C.byref(a), C.byref(b), C.byref(c)), "stats")
        return a.value, b.value, c.value

    def profile(self, on=True):
        self._check(self.L.sph_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_report(self):
        buf = (abi.KernelTime * 128)()
        k = self.L.sph_profile_report(self.h, buf, 128)
        if k < 0:
            raise SPHError(k, "profile_report")
        return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms, buf[i].alg_bytes) for i in range(k)}

    def device_bytes(self):
        return self.L.sph_device_bytes(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.L.sph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
