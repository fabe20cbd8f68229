"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "urban-sph_a100_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)
import abi  # noqa: E402

LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    lib.oracle_params_default.argtypes = [C.POINTER(abi.Params)]
    lib.oracle_create.restype = vp
    lib.oracle_create.argtypes = [C.POINTER(abi.Params), C.c_int]
    lib.oracle_upload.argtypes = [vp, vp, vp, vp, C.c_int]
    lib.oracle_step.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    lib.oracle_apr_refine.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
    lib.oracle_get_field_f64.argtypes = [vp, C.c_int, vp, C.c_int]
    lib.oracle_get_field_i32.argtypes = [vp, C.c_int, vp, C.c_int]
    lib.oracle_nnps.argtypes = [vp, vp, vp, vp, vp]
    lib.oracle_domain.argtypes = [vp, C.POINTER(abi.Domain)]
    lib.oracle_neighbors.argtypes = [vp, C.c_int, vp, vp, C.POINTER(C.c_longlong)]
    lib.oracle_ppe_info.argtypes = [vp, C.POINTER(abi.PPEReport)]
    lib.oracle_ppe_system.argtypes = [vp, ip, C.POINTER(C.c_longlong), vp, vp, vp, vp, vp, vp, vp, vp]
    lib.oracle_stats.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_canon_sum.restype = C.c_double
    lib.oracle_canon_sum.argtypes = [vp, C.c_long]
    for f in ("oracle_tmpA",):
        getattr(lib, f).restype = C.c_double
        getattr(lib, f).argtypes = [C.c_int, C.c_double, C.c_double]
    for f in ("oracle_kernel_w", "oracle_kernel_dw", "oracle_kernel_wp", "oracle_kernel_dwp"):
        getattr(lib, f).restype = C.c_double
        getattr(lib, f).argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    lib.oracle_interp_ux_caseH.restype = C.c_double
    lib.oracle_interp_ux_caseH.argtypes = [C.c_double]
    lib.oracle_phi_m.restype = C.c_double
    lib.oracle_phi_m.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_search.argtypes = [vp]
    lib.oracle_run_pass.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double]
    lib.oracle_get_work_f64.argtypes = [vp, C.c_int, vp, C.c_int]
    lib.oracle_set_work_f64.argtypes = [vp, C.c_int, vp, C.c_int]
    lib.oracle_set_work_i32.argtypes = [vp, C.c_int, vp, C.c_int]
    lib.oracle_set_mutant.argtypes = [C.c_int]
    return lib


lib = _load()


def default_params():
    p = abi.Params()
    lib.oracle_params_default(C.byref(p))
    return p


def set_threads(n):
    return lib.oracle_set_threads(int(n))


# Passes of oracle_run_pass (the oracle's OP_* order).
PASSES = ("grad_corr", "prep", "boundary", "most", "neumann", "advection", "projection", "periodic",
          "freeslip", "ppe", "outlet", "pressure_force", "ibm_force", "ibm_spread", "ib_normals")


def set_mutant(k):
    """Plants mistake k in the oracle (0 = none); returns the previous one."""
    return lib.oracle_set_mutant(int(k))


def canon_sum(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib.oracle_canon_sum(abi.ptr(v), v.size)


class Oracle:
    """One case on the CPU, mirroring the product's context API."""

    def __init__(self, params, n):
        self.params = params
        self.h = lib.oracle_create(C.byref(params), int(n))
        if not self.h:
            raise ValueError("oracle_create failed")
        self.n = 0

    def upload(self, P1, P2=None, normals=None):
        P1 = np.ascontiguousarray(P1)
        nrm = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32)
        rc = lib.oracle_upload(self.h, abi.ptr(P1), abi.ptr(P2), abi.ptr(nrm), len(P1))
        if rc:
            raise RuntimeError(f"oracle_upload rc={rc}")
        self.n = len(P1)

    def step(self, count, dt, time=None):
        if time is None:
            time = count * dt
        return lib.oracle_step(self.h, int(count), float(dt), float(time))

    def search(self):
        """The count == 0 neighbour search alone (then single passes)."""
        rc = lib.oracle_search(self.h)
        if rc:
            raise RuntimeError(f"oracle_search rc={rc}")

    def run_pass(self, name, count=0, dt=1e-4, time=None):
        if time is None:
            time = count * dt
        return lib.oracle_run_pass(self.h, PASSES.index(name), int(count), float(dt), float(time))

    def work(self, name):
        """A field of the working state (SP1 / SP2 / P3), sorted order."""
        out = np.empty(self.n, np.float64)
        rc = lib.oracle_get_work_f64(self.h, abi.F[name], abi.ptr(out), self.n)
        if rc:
            raise RuntimeError(rc)
        return out

    def set_work(self, name, values):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.float64), (self.n,)))
        rc = lib.oracle_set_work_f64(self.h, abi.F[name], abi.ptr(v), self.n)
        if rc:
            raise RuntimeError(rc)

    def set_work_i(self, name, values):
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, np.int32), (self.n,)))
        rc = lib.oracle_set_work_i32(self.h, abi.I[name], abi.ptr(v), self.n)
        if rc:
            raise RuntimeError(rc)

    def apr_refine(self, target=None, max_passes=1):
        """Refinement passes; target: per-particle target level in the
        current (cell-sorted) order, or None to keep the current targets."""
        t = None if target is None else np.ascontiguousarray(target, np.float64)
        n_out = C.c_int(0)
        rc = lib.oracle_apr_refine(self.h, abi.ptr(t), int(max_passes), C.byref(n_out))
        if rc:
            raise RuntimeError(f"oracle_apr_refine rc={rc}")
        self.n = n_out.value
        return self.n

    def field(self, name):
        out = np.empty(self.n, np.float64)
        rc = lib.oracle_get_field_f64(self.h, abi.F[name], abi.ptr(out), self.n)
        if rc:
            raise RuntimeError(rc)
        return out

    def ifield(self, name):
        out = np.empty(self.n, np.int32)
        rc = lib.oracle_get_field_i32(self.h, abi.I[name], abi.ptr(out), self.n)
        if rc:
            raise RuntimeError(rc)
        return out

    def domain(self):
        d = abi.Domain()
        lib.oracle_domain(self.h, C.byref(d))
        return d

    def nnps(self):
        d = self.domain()
        keys = np.empty(self.n, np.int32)
        perm = np.empty(self.n, np.int32)
        gs = np.empty(d.num_cells, np.int32)
        ge = np.empty(d.num_cells, np.int32)
        rc = lib.oracle_nnps(self.h, abi.ptr(keys), abi.ptr(perm), abi.ptr(gs), abi.ptr(ge))
        if rc:
            raise RuntimeError(rc)
        return keys, perm, gs, ge

    def neighbors(self, kind):
        tot = C.c_longlong(0)
        lib.oracle_neighbors(self.h, kind, None, None, C.byref(tot))
        off = np.empty(self.n + 1, np.int32)
        nbr = np.empty(max(tot.value, 1), np.int32)
        lib.oracle_neighbors(self.h, kind, abi.ptr(off), abi.ptr(nbr), C.byref(tot))
        return off, nbr[:tot.value]

    def ppe_report(self):
        r = abi.PPEReport()
        lib.oracle_ppe_info(self.h, C.byref(r))
        return r

    def ppe_system(self):
        nr, nnz = C.c_int(0), C.c_longlong(0)
        lib.oracle_ppe_system(self.h, C.byref(nr), C.byref(nnz), *([None] * 9))
        n, m = nr.value, nnz.value
        rows = np.empty(n, np.int32)
        diag = np.empty(n)
        rhs = np.empty(n)
        rowptr = np.empty(n + 1, np.int32)
        cols = np.empty(max(m, 1), np.int32)
        vals = np.empty(max(m, 1))
        x0 = np.empty(n)
        x = np.empty(n)
        lib.oracle_ppe_system(self.h, C.byref(nr), C.byref(nnz), abi.ptr(rows), abi.ptr(diag),
                              abi.ptr(rhs), abi.ptr(rowptr), abi.ptr(cols), abi.ptr(vals),
                              abi.ptr(x0), abi.ptr(x))
        return dict(rows=rows, diag=diag, rhs=rhs, rowptr=rowptr, cols=cols[:m],
                    vals=vals[:m], x0=x0, x=x)

    def stats(self):
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        lib.oracle_stats(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def close(self):
        if self.h:
            lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kernel_w(ktype, kappa, h, r):
    return lib.oracle_kernel_w(ktype, kappa, h, r)


def kernel_dw(ktype, kappa, h, r):
    return lib.oracle_kernel_dw(ktype, kappa, h, r)


def kernel_wp(ktype, kappa, h, r):
    return lib.oracle_kernel_wp(ktype, kappa, h, r)


def kernel_dwp(ktype, kappa, h, r):
    return lib.oracle_kernel_dwp(ktype, kappa, h, r)
