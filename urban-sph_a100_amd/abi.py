"""ctypes mirror of include/urbansph.h (records, parameters, enums).

The record dtypes reproduce the reference's AoS layout
(class_Cuda_Particle_Array.cuh:11-112 with Real=float, Variable_Type.cuh:2-4)
so numpy arrays of them can be handed to the C ABI unchanged.
"""
import ctypes as C

import numpy as np

_F = "<f4"
_I = "<i4"

PART1_FIELDS = (
    [("i_type", _I), ("buffer_type", _I), ("MOST_buffer", _I), ("p_type", _I)]
    + [(n, _F) for n in (
        "x", "y", "z", "ux", "uy", "uz", "m", "h", "temp", "pres", "pres0", "rho",
        "flt_s", "enthalpy", "concn", "concentration", "k_turb", "e_turb", "eli",
        "vol", "m_ref", "h_ref", "vis_t", "fbx", "fby", "fbz", "fmx", "fmy", "hmz",
        "float1", "float2", "float3")]
    + [("ncell", _I), ("source", _I)]
)
PART2_FIELDS = [(n, _F) for n in (
    "rho_ref", "SR", "x0", "y0", "z0", "ux0", "uy0", "uz0", "rho0", "drho0",
    "temp0", "concn0", "enthalpy0")]
PART3_FIELDS = (
    [(n, _F) for n in (
        "drho", "dconcn", "denthalpy", "ftotalx", "ftotaly", "ftotalz", "ftotal",
        "fpx", "fpy", "fpz", "fbody", "dtemp", "dk_turb", "de_trub", "cc",
        "nx", "ny", "nz")]
    + [("Cm", _F, (3, 3))]
    + [(n, _F) for n in (
        "inv_cm_xx", "inv_cm_yy", "inv_cm_zz", "inv_cm_xy", "inv_cm_yz",
        "inv_cm_zx", "Gxx", "Gxy", "Gyx", "Gyy", "L11", "L12", "L21", "L22")]
)
PART1 = np.dtype(PART1_FIELDS)
PART2 = np.dtype(PART2_FIELDS)
PART3 = np.dtype(PART3_FIELDS)
assert PART1.itemsize == 152 and PART2.itemsize == 52 and PART3.itemsize == 164

# Option values (include/urbansph.h)
GAUSSIAN, QUINTIC, QUARTIC, WENDLAND2, WENDLAND4, WENDLAND6 = range(6)
EULER, PRE_COR = 0, 1
FV_NONE, FV_CLEARY, FV_MONAGHAN = 0, 1, 2
LAMINAR, SSM, DDF = 0, 4, 5
PPE_REF_JACOBI, PPE_BICGSTAB = 0, 1
FP64, MIXED = 0, 1
MOST_REFERENCE, MOST_PRESCRIBED_L = 0, 1

E_ARG, E_HIP, E_NOMEM, E_STATE, E_NOCONV, E_NONFINITE, E_UNSUPPORTED = (
    -1, -2, -3, -4, -5, -6, -7)

F64_FIELDS = [
    "X", "Y", "Z", "UX", "UY", "UZ", "M", "H", "TEMP", "PRES", "PRES0", "RHO",
    "FLT_S", "CONCN", "CONCENTRATION", "K_TURB", "E_TURB", "VIS_T", "FBX", "FBY",
    "FBZ", "FMX", "FMY", "HMZ", "FLOAT1", "FLOAT2", "FLOAT3", "VOL", "ENTHALPY",
    "RHO_REF", "SR", "X0", "Y0", "Z0", "UX0", "UY0", "UZ0", "TEMP0", "CONCN0",
    "FTOTALX", "FTOTALY", "FTOTALZ", "FTOTAL", "FPX", "FPY", "FPZ", "DTEMP",
    "DCONCN", "INV_CM_XX", "INV_CM_YY", "INV_CM_ZZ", "INV_CM_XY", "INV_CM_YZ",
    "INV_CM_ZX", "NX", "NY", "NZ",
]
F = {name: i for i, name in enumerate(F64_FIELDS)}
I32_FIELDS = ["I_TYPE", "BUFFER_TYPE", "MOST_BUFFER", "P_TYPE", "NCELL", "SOURCE"]
I = {name: i for i, name in enumerate(I32_FIELDS)}

_INT_PARAMS = (
    "dim", "kernel_type", "time_type", "freq_filt", "fp_solve", "fv_solve",
    "fg_solve", "con_solve", "boussinesq_solve", "kgc_solve",
    "turbulence_model", "concn_solve", "noslip_bc", "artificial_viscosity")


class Params(C.Structure):
    """sph_params (include/urbansph.h), member for member."""
    _fields_ = (
        [(n, C.c_int) for n in _INT_PARAMS]
        + [(n, C.c_double) for n in (
            "kappa", "u_limit", "mu_fluid", "mu_wall", "gravity", "T_ini",
            "T_con", "T_start_t", "T_con_t", "tempb", "Cs_sgs", "Ck_sgs",
            "Ce_sgs", "rho_air", "av_alpha", "av_beta", "av_c", "solid_k",
            "air_k", "solid_cp", "air_cp", "solid_rho", "diff_coef", "L2", "L3",
            "L_left", "L_right", "outlet_copy_dx", "copy_tol",
            "neumann_ghost_r")]
        + [("most_mode", C.c_int)]
        + [(n, C.c_double) for n in (
            "most_z0", "most_z0h", "most_L", "most_ghost_dz", "most_ghost_r",
            "ibm_ut_factor")]
        + [("ppe_solver", C.c_int), ("precision", C.c_int),
           ("ppe_tol", C.c_double), ("ppe_maxit", C.c_int),
           ("ppe_inner_tol", C.c_double), ("ppe_max_refine", C.c_int), ("ppe_index16", C.c_int),
           ("ppe_reorder", C.c_int), ("apr_h_change", C.c_int),
           ("ale_positions", C.c_int), ("ppe_precond", C.c_int), ("ib_normals", C.c_int)]
        + [(n, C.c_double) for n in (
            "ib_face_xlo", "ib_face_xhi", "ib_face_ylo", "ib_face_yhi", "ib_face_zlo", "ib_face_zhi",
            "ib_face_tol")]
        + [("mg_sort_all", C.c_int), ("tile_cells", C.c_int), ("tile_particles", C.c_int),
           ("tile_halo_max", C.c_int), ("compact_records", C.c_int),
           ("ppe_stencil", C.c_int)]
    )

    def update(self, **kw):
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(f"unknown sph_params member {k!r}")
            setattr(self, k, v)
        return self

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Domain(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "x_min", "x_max", "y_min", "y_max", "z_min", "z_max", "dcell")] + [
        (n, C.c_int) for n in ("NI", "NJ", "NK", "num_cells")]


class PPEReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refinements", C.c_int),
                ("unknowns", C.c_int), ("nnz", C.c_longlong),
                ("rel_residual", C.c_double), ("converged", C.c_int),
                ("ordering", C.c_int), ("slots", C.c_longlong), ("index16", C.c_int), ("tail_rows", C.c_int),
                ("precond", C.c_int), ("mg_levels", C.c_int), ("mg_coarsest", C.c_int),
                ("mg_status", C.c_int), ("mg_sorted_rows", C.c_longlong), ("mg_long_rows", C.c_int),
                ("heavy_lists", C.c_int), ("tile_chunks", C.c_int), ("tile_raw_chunks", C.c_int),
                ("tile_halo_largest", C.c_int), ("tile_halo_total", C.c_longlong),
                ("stencil_slices", C.c_int), ("stencil_nnz", C.c_longlong)]


ORDER_CELL, ORDER_LATTICE = 0, 1
PRECOND_JACOBI, PRECOND_MULTIGRID = 0, 1
IBN_REFERENCE, IBN_PLANES, IBN_UPLOADED = 0, 1, 2
(MG_OK, MG_UNSUITED, MG_TERMS, MG_COLUMNS, MG_SPAN, MG_PIVOT, MG_STALLED, MG_LEVELS,
 MG_STALLED_SOLVE) = range(9)


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_longlong),
                ("total_ms", C.c_double), ("alg_bytes", C.c_double)]


class ServerInfo(C.Structure):
    _fields_ = [("queued", C.c_int), ("running", C.c_int), ("finished", C.c_int),
                ("reserved_bytes", C.c_longlong), ("budget_bytes", C.c_longlong),
                ("ticks", C.c_longlong), ("particle_steps", C.c_longlong), ("clock", C.c_double),
                ("admit_s", C.c_double), ("step_s", C.c_double), ("retire_s", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("rc", C.c_int), ("n", C.c_int), ("steps", C.c_int),
                ("steps_done", C.c_int), ("noconv_steps", C.c_int), ("ppe_iterations", C.c_longlong),
                ("est_bytes", C.c_longlong), ("t_arrival", C.c_double), ("t_admit", C.c_double),
                ("t_finish", C.c_double)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

MAX_PLOT = 64
TABLE_CAP, MAX_TABLES = 512, 32


class PropTables(C.Structure):
    """sph_prop_tables: material property tables, one per p_type."""
    _fields_ = [("ntables", C.c_int), ("ndata", C.c_int), ("index", C.c_int * MAX_TABLES),
                ("size", C.c_int * MAX_TABLES)] + [(n, C.c_float * TABLE_CAP) for n in ("T", "h", "k", "cp", "mu")]


class RunConfig(C.Structure):
    """sph_run_config: run controls and solv.txt settings outside sph_params."""
    _fields_ = [(n, C.c_int) for n in (
        "solver_type", "prop_table", "atmos_stab", "terrain_type", "structure", "open_boundary",
        "switch_ptype", "apr_solve")] + [("block_config", C.c_int * 3)] + [(n, C.c_int) for n in (
        "nb_cell_type", "timestep_update", "freq_output", "freq_apr", "freq_bapr", "freq_temp",
        "freq_h_reset", "min_iteration", "max_iteration")] + [(n, C.c_double) for n in (
        "buffersize", "p_ref", "dp_th", "drho_th", "p_relaxation", "dt_structure", "dt", "time",
        "time_end")] + [("num_plot", C.c_int), ("plot", (C.c_char * 32) * MAX_PLOT), ("unknown", C.c_char * 256)]

    def as_dict(self):
        d = {}
        for n, _ in self._fields_:
            v = getattr(self, n)
            if n == "block_config":
                v = list(v)
            elif n == "plot":
                v = [v[i].value.decode() for i in range(self.num_plot)]
            elif n == "unknown":
                v = v.decode().split()
            d[n] = v
        return d


def ptr(a):
    """Raw data pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)
