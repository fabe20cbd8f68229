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
    L.sph_stats_noalloc.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.sph_synchronize.argtypes = [vp]
    L.sph_profile_enable.argtypes = [vp, C.c_int]
    L.sph_profile_report.argtypes = [vp, C.POINTER(abi.KernelTime), C.c_int]
    L.sph_device_bytes.restype = C.c_longlong
    L.sph_device_bytes.argtypes = [vp]
    L.sph_last_error.restype = C.c_char_p
    L.sph_last_error.argtypes = [vp]
    L.sph_destroy.argtypes = [vp]
    L.sph_set_engine_options.argtypes = [vp, C.c_int, C.c_int]
    L.sph_apr_refine.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
    L.sph_tuning.argtypes = [vp, vp, C.c_int]
    L.sph_batch_create.argtypes = [C.POINTER(vp)]
    L.sph_batch_step.argtypes = [vp, vp, C.c_int, vp, vp, vp]
    L.sph_batch_synchronize.argtypes = [vp]
    L.sph_batch_set_options.argtypes = [vp, C.c_int, C.c_int]
    L.sph_batch_profile_enable.argtypes = [vp, C.c_int]
    L.sph_batch_profile_report.argtypes = [vp, C.POINTER(abi.KernelTime), C.c_int]
    L.sph_batch_tuning.argtypes = [vp, vp, C.c_int]
    L.sph_batch_last_error.restype = C.c_char_p
    L.sph_batch_last_error.argtypes = [vp]
    L.sph_batch_destroy.argtypes = [vp]
    L.sph_estimate_bytes.restype = C.c_longlong
    L.sph_estimate_bytes.argtypes = [C.POINTER(abi.Params), C.c_int]
    L.sph_server_create.argtypes = [C.c_longlong, C.c_int, C.c_double, C.POINTER(vp)]
    L.sph_server_submit.argtypes = [vp, C.POINTER(abi.Params), vp, vp, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.POINTER(C.c_longlong)]
    L.sph_server_tick.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
    L.sph_server_idle_until.argtypes = [vp, C.c_double]
    L.sph_server_clock.restype = C.c_double
    L.sph_server_clock.argtypes = [vp]
    L.sph_server_status.argtypes = [vp, C.POINTER(abi.ServerInfo)]
    L.sph_server_request.argtypes = [vp, C.c_longlong, C.POINTER(abi.RequestInfo)]
    L.sph_server_result.argtypes = [vp, C.c_longlong, C.c_int, vp, C.c_int]
    L.sph_server_result_part1.argtypes = [vp, C.c_longlong, vp, C.c_int]
    L.sph_server_release.argtypes = [vp, C.c_longlong]
    L.sph_server_last_error.restype = C.c_char_p
    L.sph_server_last_error.argtypes = [vp]
    L.sph_server_destroy.argtypes = [vp]
    L.sph_read_solv.argtypes = [C.c_char_p, C.POINTER(abi.Params), C.POINTER(abi.RunConfig)]
    L.sph_read_deck.argtypes = [C.c_char_p, vp, C.c_int, C.POINTER(C.c_int)]
    L.sph_write_vtk.argtypes = [C.c_char_p, vp, vp, vp, C.c_int, C.POINTER(C.c_char_p), C.c_int]
    L.sph_read_tables.argtypes = [C.c_char_p, C.POINTER(abi.PropTables)]
    L.sph_table_interp.restype = C.c_float
    L.sph_table_interp.argtypes = [vp, vp, C.c_int, C.c_float]
    _lib = L
    return L


EXPORTED = [
    "sph_params_default", "sph_create", "sph_upload", "sph_step", "sph_download",
    "sph_get_field_f64", "sph_get_field_i32", "sph_nnps_debug", "sph_neighbors",
    "sph_domain_info", "sph_ppe_info", "sph_stats", "sph_stats_noalloc", "sph_synchronize",
    "sph_profile_enable", "sph_profile_report", "sph_device_bytes",
    "sph_last_error", "sph_destroy", "sph_set_engine_options", "sph_tuning", "sph_apr_refine",
    "sph_batch_create", "sph_batch_step", "sph_batch_synchronize", "sph_batch_set_options",
    "sph_batch_profile_enable", "sph_batch_profile_report", "sph_batch_tuning",
    "sph_batch_last_error", "sph_batch_destroy", "sph_estimate_bytes", "sph_server_create",
    "sph_server_submit", "sph_server_tick", "sph_server_idle_until", "sph_server_clock",
    "sph_server_status", "sph_server_request", "sph_server_result", "sph_server_result_part1",
    "sph_server_release", "sph_server_last_error", "sph_server_destroy",
    "sph_read_solv", "sph_read_deck", "sph_write_vtk", "sph_read_tables", "sph_table_interp",
]

TUNE_PASSES = ["grad_corr", "prep", "boundary3d", "neumann", "advection", "freeslip", "pressure_force",
               "ibm_force", "ibm_spread", "ppe_assemble"]


def _tuning_rows(fn, h):
    out = np.zeros((64, 1 + len(TUNE_PASSES)), np.int32)
    k = fn(h, abi.ptr(out), 64)
    return {int(r[0]): dict(zip(TUNE_PASSES, (int(v) for v in r[1:]))) for r in out[:max(k, 0)]}


def _profile_rows(buf, k):
    if k < 0:
        raise SPHError(k, "profile_report")
    return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms, buf[i].alg_bytes) for i in range(k)}


def default_params():
    p = abi.Params()
    lib().sph_params_default(C.byref(p))
    return p


class SPHError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"rc={rc}: {msg}")
        self.rc = rc


def read_solv(path, params=None):
    """solv.txt -> (sph_params, run settings dict) (function_init.cuh:166-463).
    Members the file does not name keep their values from `params`
    (default: the reference defaults)."""
    prm = abi.Params()
    C.pointer(prm)[0] = params if params is not None else default_params()
    run = abi.RunConfig()
    rc = lib().sph_read_solv(os.fsencode(path), C.byref(prm), C.byref(run))
    if rc:
        raise SPHError(rc, f"cannot read {path}")
    return prm, run.as_dict()


def read_deck(path):
    """Particle deck -> part1 records (function_init.cuh:550-676)."""
    L = lib()
    n = C.c_int()
    rc = L.sph_read_deck(os.fsencode(path), None, 0, C.byref(n))
    if rc:
        raise SPHError(rc, f"cannot read deck {path}")
    P = np.zeros(n.value, abi.PART1)
    rc = L.sph_read_deck(os.fsencode(path), abi.ptr(P), len(P), C.byref(n))
    if rc:
        raise SPHError(rc, f"cannot read deck {path}")
    return P[:n.value]


def read_tables(path):
    """Property tables -> list of per-p_type dicts of float32 arrays
    T, h, k, cp, mu (function_init.cuh:724-970)."""
    t = abi.PropTables()
    rc = lib().sph_read_tables(os.fsencode(path), C.byref(t))
    if rc:
        raise SPHError(rc, f"cannot read tables {path}")
    out = []
    for i in range(t.ntables):
        a, n = t.index[i], t.size[i]
        out.append({k: np.array(getattr(t, k)[a:a + n], np.float32) for k in ("T", "h", "k", "cp", "mu")})
    return out


def table_interp(x, y, v):
    """interp2 (functions_PROP.cuh:8-40) on float32 knots."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return np.float32(lib().sph_table_interp(abi.ptr(x), abi.ptr(y), len(x), float(v)))


def write_vtk(path, P1, labels, P2=None, P3=None):
    """Legacy binary VTK of the i_type 1/2 particles (functions_OUTPUT.cuh:121-616)."""
    P1 = np.ascontiguousarray(P1, abi.PART1)
    P2 = None if P2 is None else np.ascontiguousarray(P2, abi.PART2)
    P3 = None if P3 is None else np.ascontiguousarray(P3, abi.PART3)
    arr = (C.c_char_p * max(len(labels), 1))(*[l.encode() for l in labels])
    rc = lib().sph_write_vtk(os.fsencode(path), abi.ptr(P1), abi.ptr(P2), abi.ptr(P3), len(P1), arr, len(labels))
    if rc:
        raise SPHError(rc, f"cannot write {path}")


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

    def stats_noalloc(self):
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        self._check(self.L.sph_stats_noalloc(self.h, C.byref(a), C.byref(b), C.byref(c)), "stats_noalloc")
        return a.value, b.value, c.value

    def profile(self, on=True):
        self._check(self.L.sph_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_report(self):
        buf = (abi.KernelTime * 128)()
        return _profile_rows(buf, self.L.sph_profile_report(self.h, buf, 128))

    def options(self, graphs=True, autotune=True):
        self._check(self.L.sph_set_engine_options(self.h, int(graphs), int(autotune)), "engine_options")

    def apr_refine(self, target=None, max_passes=1):
        """Refinement passes (sph_apr_refine); target: per-particle target
        level in device order, or None to keep the current targets."""
        t = None if target is None else np.ascontiguousarray(target, np.float64)
        n_out = C.c_int(0)
        self._check(self.L.sph_apr_refine(self.h, abi.ptr(t), int(max_passes), C.byref(n_out)), "apr_refine")
        self.n = n_out.value
        return self.n

    def tuning(self):
        return _tuning_rows(self.L.sph_tuning, self.h)

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


class Batch:
    """Independent cases stepped together (one sph_batch)."""

    def __init__(self):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.sph_batch_create(C.byref(h))
        if rc:
            raise SPHError(rc, "sph_batch_create failed")
        self.h = h

    def step(self, cases, counts, dts, times=None):
        """One step of every case in ``cases`` (Case objects).  Raises
        SPHError on failure; rc SPH_E_NOCONV (-5) is raised after the step
        completed."""
        k = len(cases)
        hs = (C.c_void_p * k)(*[c.h for c in cases])
        cnt = np.ascontiguousarray(counts, np.int32)
        dt = np.ascontiguousarray(dts, np.float64)
        tm = cnt * dt if times is None else np.ascontiguousarray(times, np.float64)
        rc = self.L.sph_batch_step(self.h, hs, k, abi.ptr(cnt), abi.ptr(dt), abi.ptr(tm))
        if rc:
            raise SPHError(rc, f"sph_batch_step: {self.L.sph_batch_last_error(self.h).decode()}")

    def sync(self):
        rc = self.L.sph_batch_synchronize(self.h)
        if rc:
            raise SPHError(rc, f"sph_batch_synchronize: {self.L.sph_batch_last_error(self.h).decode()}")

    def options(self, graphs=True, autotune=True):
        self.L.sph_batch_set_options(self.h, int(graphs), int(autotune))

    def profile(self, on=True):
        self.L.sph_batch_profile_enable(self.h, 1 if on else 0)

    def profile_report(self):
        buf = (abi.KernelTime * 128)()
        return _profile_rows(buf, self.L.sph_batch_profile_report(self.h, buf, 128))

    def tuning(self):
        return _tuning_rows(self.L.sph_batch_tuning, self.h)

    def close(self):
        if getattr(self, "h", None):
            self.L.sph_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate_bytes(params, n):
    return lib().sph_estimate_bytes(C.byref(params), int(n))


class Server:
    """Continuous-batching request loop (one sph_server).  dry_run_rate > 0
    simulates the GPU at that many particle-steps/s."""

    def __init__(self, budget_bytes, max_cases=256, dry_run_rate=0.0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.sph_server_create(int(budget_bytes), int(max_cases), float(dry_run_rate), C.byref(h))
        if rc:
            raise SPHError(rc, "sph_server_create failed")
        self.h = h
        self._fin = np.zeros(4096, np.int64)

    def _check(self, rc, what):
        if rc:
            raise SPHError(rc, f"{what}: {self.L.sph_server_last_error(self.h).decode()}")

    def submit(self, params, P1, steps, dt, normals=None, t_arrival=-1.0):
        P1 = np.ascontiguousarray(P1)
        assert P1.dtype == abi.PART1
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32)
        rid = C.c_longlong()
        self._check(self.L.sph_server_submit(self.h, C.byref(params), abi.ptr(P1), abi.ptr(nrm), len(P1),
                                             int(steps), float(dt), float(t_arrival), C.byref(rid)), "submit")
        return rid.value

    def tick(self):
        """One co-batched step; returns the ids that finished in it."""
        nf = C.c_int()
        self._check(self.L.sph_server_tick(self.h, abi.ptr(self._fin), len(self._fin), C.byref(nf)), "tick")
        return [int(x) for x in self._fin[:min(nf.value, len(self._fin))]]

    def idle_until(self, t):
        self.L.sph_server_idle_until(self.h, float(t))

    def clock(self):
        return self.L.sph_server_clock(self.h)

    def status(self):
        i = abi.ServerInfo()
        self._check(self.L.sph_server_status(self.h, C.byref(i)), "status")
        return i

    def request(self, rid):
        i = abi.RequestInfo()
        self._check(self.L.sph_server_request(self.h, int(rid), C.byref(i)), "request")
        return i

    def result(self, rid, name, n):
        out = np.empty(n, np.float64)
        self._check(self.L.sph_server_result(self.h, int(rid), abi.F[name], abi.ptr(out), n), "result")
        return out

    def result_part1(self, rid, n):
        out = np.zeros(n, abi.PART1)
        self._check(self.L.sph_server_result_part1(self.h, int(rid), abi.ptr(out), n), "result_part1")
        return out

    def release(self, rid):
        self._check(self.L.sph_server_release(self.h, int(rid)), "release")

    def close(self):
        if getattr(self, "h", None):
            self.L.sph_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
