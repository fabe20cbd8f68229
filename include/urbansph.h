/*
 * urbansph.h -- C ABI of the MI355X-native ISPH projection step.
 *
 * This is the drop-in boundary for the reference's one-step entry point
 *   SOPHIA_single_ISPH(...)                      ISPH_Calc.cuh:4-233
 * and for the buffer ownership that its caller, the time loop, holds
 *   ISPH_Calc(...)  allocations / upload / loop  ISPH_Calc.cuh:238-577
 * The reference has no FFI; its host C++ calls the step directly with raw
 * device pointers, host globals (vii/vif, function_init.cuh:1-163) and
 * __constant__ mirrors.  Here the caller owns one opaque context per case
 * (or per batch of cases) and passes plain pointers and sizes.  No torch or
 * HIP types appear in any signature.
 *
 * All entry points return 0 on success and a negative SPH_E_* code on
 * failure; sph_last_error() then holds a message.  The reference checks no
 * errors at all (it never calls cudaGetLastError on the path).
 *
 * Threading: one host thread per context at a time.  Every call is ordered
 * on the context's own HIP stream; calls that return host data synchronise
 * that stream.
 */
#ifndef URBANSPH_H
#define URBANSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Particle records.  Field order and widths are those of the reference's AoS
 * structs with Real=float and int_t=uint_t=int (Variable_Type.cuh:2-4):
 *   part1  class_Cuda_Particle_Array.cuh:11-67   152 bytes
 *   part2  class_Cuda_Particle_Array.cuh:69-85    52 bytes
 *   part3  class_Cuda_Particle_Array.cuh:87-112  164 bytes
 * A reference caller can pass its part1/part2/part3 arrays unchanged.
 * ------------------------------------------------------------------------- */
typedef struct sph_part1 {
    int   i_type, buffer_type, MOST_buffer, p_type;
    float x, y, z;
    float ux, uy, uz;
    float m, h, temp, pres, pres0, rho, flt_s, enthalpy, concn, concentration;
    float k_turb, e_turb;
    float eli, vol;
    float m_ref, h_ref;
    float vis_t;
    float fbx, fby, fbz;
    float fmx, fmy, hmz;
    float float1, float2, float3;
    int   ncell, source;
} sph_part1;

typedef struct sph_part2 {
    float rho_ref, SR;
    float x0, y0, z0;
    float ux0, uy0, uz0;
    float rho0, drho0, temp0, concn0, enthalpy0;
} sph_part2;

typedef struct sph_part3 {
    float drho, dconcn, denthalpy;
    float ftotalx, ftotaly, ftotalz, ftotal;
    float fpx, fpy, fpz, fbody;
    float dtemp, dk_turb, de_trub, cc;
    float nx, ny, nz;
    float Cm[3][3];
    float inv_cm_xx, inv_cm_yy, inv_cm_zz;
    float inv_cm_xy, inv_cm_yz, inv_cm_zx;
    float Gxx, Gxy, Gyx, Gyy;
    float L11, L12, L21, L22;
} sph_part3;

/* ---------------------------------------------------------------------------
 * Option values (Parameters.cuh:26-31,33-34,58-61,64-69,73-76)
 * ------------------------------------------------------------------------- */
enum { SPH_GAUSSIAN = 0, SPH_QUINTIC = 1, SPH_QUARTIC = 2,
       SPH_WENDLAND2 = 3, SPH_WENDLAND4 = 4, SPH_WENDLAND6 = 5 };
enum { SPH_EULER = 0, SPH_PRE_COR = 1 };
enum { SPH_FV_NONE = 0, SPH_FV_CLEARY = 1, SPH_FV_MONAGHAN = 2 };
enum { SPH_LAMINAR = 0, SPH_SSM = 4, SPH_DDF = 5 };
/* PPE solver: REF_JACOBI = the reference's single sweep of KERNEL_PPE3D
 * (functions_PPE.cuh:2-175, ISPH_Calc.cuh:137) made deterministic (every row
 * reads the pressures of the previous iterate); BICGSTAB = Jacobi-
 * preconditioned BiCGSTAB on the same operator to a relative residual. */
enum { SPH_PPE_REF_JACOBI = 0, SPH_PPE_BICGSTAB = 1 };
/* Precision: FP64 = every field and sum in double.  MIXED = interaction
 * sums and the inner PPE solve in float with double accumulation; Krylov
 * dots, residual norms and an outer iterative-refinement loop in double. */
enum { SPH_FP64 = 0, SPH_MIXED = 1 };
/* MOST wall model (functions_MOST.cuh:175-322): REFERENCE = bulk-Richardson
 * Newton iteration on the Obukhov length; PRESCRIBED_L = the length is given. */
enum { SPH_MOST_REFERENCE = 0, SPH_MOST_PRESCRIBED_L = 1 };

/* ---------------------------------------------------------------------------
 * Solver parameters.  The integer/real knobs are the vii/vif entries the step
 * reads (function_init.cuh:1-81, parsed at :166-463); the remaining members
 * are compile-time constants of the reference made runtime (Parameters.cuh,
 * functions_PROP.cuh:121-131, functions_NNPS.cuh:4-14).  sph_params_default()
 * fills every member with the reference's value.
 * ------------------------------------------------------------------------- */
typedef struct sph_params {
    /* vii */
    int    dim;                 /* only 3 is supported                       */
    int    kernel_type;         /* SPH_GAUSSIAN .. SPH_WENDLAND6              */
    int    time_type;           /* SPH_EULER / SPH_PRE_COR                    */
    int    freq_filt;           /* Shepard filter period, must be > 0         */
    int    fp_solve;            /* pressure force                             */
    int    fv_solve;            /* SPH_FV_*                                   */
    int    fg_solve;            /* gravity                                    */
    int    con_solve;           /* temperature equation                       */
    int    boussinesq_solve;
    int    kgc_solve;           /* only 0 (NO): the reference never builds
                                   Cm, so KGC would zero the gradients        */
    int    turbulence_model;    /* SPH_LAMINAR / SPH_SSM / SPH_DDF            */
    int    concn_solve;         /* concentration diffusion                    */
    int    noslip_bc;
    int    artificial_viscosity;/* 1 = always on, as functions_PPE.cuh:268    */
    /* vif */
    double kappa;               /* support radius factor k in k*h             */
    double u_limit;             /* velocity limiter, functions_TIME_ISPH:182  */
    /* constants made runtime */
    double mu_fluid;            /* viscosity2(): p_type 1 and others          */
    double mu_wall;             /* viscosity2(): p_type 0 and 1000            */
    double gravity;             /* Gravitational_CONST                        */
    double T_ini, T_con, T_start_t, T_con_t, tempb;
    double Cs_sgs, Ck_sgs, Ce_sgs;
    double rho_air;             /* IBM penalty density (Rho_air)              */
    double av_alpha, av_beta, av_c;   /* artificial viscosity 0.05,0.05,30   */
    double solid_k, air_k, solid_cp, air_cp, solid_rho;  /* con_solve only   */
    double diff_coef;           /* diffusion_coefficient() (concn_solve)      */
    /* open-boundary geometry (functions_NNPS.cuh:4-14, functions_BC.cuh:732) */
    double L2, L3, L_left, L_right;
    double outlet_copy_dx;      /* source plane at L3 - outlet_copy_dx        */
    double copy_tol;            /* exact-position match radius (0.001 m)      */
    double neumann_ghost_r;     /* Neumann inlet ghost match radius (0.1 m)   */
    /* MOST */
    int    most_mode;
    double most_z0, most_z0h, most_L;
    double most_ghost_dz, most_ghost_r;   /* heat-flux deposit: z+5 m, 0.1 m  */
    /* IBM */
    double ibm_ut_factor;       /* target tangential speed = f * u_caseH(z)   */
    /* PPE solve */
    int    ppe_solver;          /* SPH_PPE_*                                  */
    int    precision;           /* SPH_FP64 / SPH_MIXED                       */
    double ppe_tol;             /* relative residual ||b-Ax||/||b||           */
    int    ppe_maxit;           /* max BiCGSTAB iterations per solve          */
    double ppe_inner_tol;       /* MIXED: floor of each inner solve's relative
                                   tolerance (each aims at ppe_tol absolute)  */
    int    ppe_max_refine;      /* MIXED: max outer refinement steps          */
    int    ppe_index16;         /* 16-bit PPE column indices where they fit
                                   (1, default) or 32-bit (0); the results are
                                   identical, only the bytes moved differ      */
    int    ppe_reorder;         /* MIXED: number the PPE unknowns in lattice
                                   (z, y, x) order for contiguous SpMV gathers
                                   (1, default) or keep the particles' cell
                                   order (0).  A case with a row of more than
                                   160 entries (coarse particles next to
                                   refined ones) keeps cell order.            */
    int    apr_h_change;        /* h_change (vii[36]): refined children get
                                   h/2 (1, default) or keep h (0)             */
    int    ale_positions;       /* 1: the corrector moves particles,
                                   x += (u0 + u)/2 dt eli for p_type > 0
                                   (functions_TIME_ISPH.cuh:193-195, computed
                                   but never stored by the reference), and the
                                   neighbour search runs every step
                                   (Lagrangian / ALE); 0 (default): positions
                                   fixed, search at count == 0 only          */
    int    ppe_precond;         /* MIXED: preconditioner of the inner float
                                   BiCGSTAB.  SPH_PRECOND_JACOBI (default):
                                   the reference's diagonal scaling.
                                   SPH_PRECOND_MULTIGRID: one V-cycle of an
                                   aggregation multigrid (2x2x2 lattice blocks
                                   per level, Galerkin coarse operators, damped
                                   Jacobi smoothing, dense inverse of the
                                   coarsest level), built once per set of PPE
                                   values; the outer FP64 refinement and the
                                   tolerance are unchanged.  Used for a case
                                   whose PPE has more than 1024 unknowns and
                                   no row of more than 256 entries, alone or
                                   in a batch (each case of a batch takes its
                                   own); elsewhere (small or refined cases
                                   with long rows) and after a multigrid
                                   inner solve that stalls, the solve uses
                                   SPH_PRECOND_JACOBI.  The report says which
                                   ran.  FP64 (BiCGSTAB) uses the same float
                                   V-cycle, built from a float copy of its
                                   operator, on float copies of p and s; its
                                   Krylov vectors, operator and tolerance stay
                                   double.                                    */
    int    ib_normals;          /* where the IB markers' (p_type >= 1000)
                                   normals come from, SPH_IBN_*.
                                   SPH_IBN_REFERENCE (default): every
                                   count == 0 step sets them from the face
                                   planes below (KERNEL_clc_IB_normal_vector3D,
                                   functions_PREP.cuh:923-1015), and steps with
                                   count > 0 run with zero normals, as the
                                   reference's per-step part3 reset leaves them
                                   (ISPH_Calc.cuh:61).  SPH_IBN_PLANES: the
                                   same plane normals, kept at later steps.
                                   SPH_IBN_UPLOADED: the normals passed to
                                   sph_upload (zero where none), kept.         */
    double ib_face_xlo, ib_face_xhi, ib_face_ylo, ib_face_yhi, ib_face_zlo, ib_face_zhi;
                                /* marker face planes of the plane rule: a
                                   marker within ib_face_tol of x = xlo gets
                                   n_x = -1, of x = xhi n_x = +1 (y, z alike),
                                   and the sum is normalised in float.
                                   Defaults 0.764, 0.836, -0.036, 0.036, 0,
                                   0.152 (functions_PREP.cuh:958-963)         */
    double ib_face_tol;         /* 1e-5 m                                     */
    int    mg_sort_all;         /* check of the multigrid set-up: 1 forms
                                   every coarse row with the sorting Galerkin
                                   kernels (term order, double sums) instead
                                   of the fixed-point hash; 0 (default) only
                                   rows of more than 256 distinct columns     */
    int    tile_cells;          /* MIXED passes over float-exact positions
                                   (prep, advection, pressure force, PPE
                                   assembly): blocks of cells whose particles
                                   one workgroup runs with the neighbours'
                                   records staged in LDS, x | y << 8 | z << 16
                                   cells (1..16 each); 0 gathers each record
                                   from memory.  Results are the same bitwise.
                                   Default 0: on the MI355X the per-particle
                                   gathers of the compact records run faster
                                   (the staged halo limits residency to one or
                                   two workgroups a CU, and its slot reads
                                   conflict in the LDS banks)                 */
    int    tile_particles;      /* particles of one workgroup of those passes
                                   (64, 128, 192 or 256; a block of more is
                                   split evenly; 0 means 128).  Default 128   */
    int    tile_halo_max;       /* most distinct neighbours a workgroup stages
                                   (1..2048); a workgroup with more gathers
                                   from memory; 0 means 1536).  Default 1536  */
    int    compact_records;     /* MIXED with float positions: 1 (default)
                                   gathers 32/64-byte float records of the
                                   neighbours; 0 the 128-byte double records
                                   and field arrays.  The same results
                                   bitwise; 0 is for checking that           */
    int    ppe_stencil;         /* MIXED, lattice row order, fixed particles:
                                   1 (default) stores a 64-row slice whose
                                   rows' column offsets (col - row) have a
                                   union of at most 5/4 of its widest row
                                   over that union, without columns (an
                                   offset a row lacks holds the value 0); the
                                   products read values only.  After the
                                   first assembly of a search the slice drops
                                   the offsets whose values are zero in all
                                   its rows (MIXED stores coefficients below
                                   1e-30 as 0).  The same results bitwise; 0
                                   stores every column                       */
} sph_params;

/* Sources of the IB marker normals (ib_normals). */
enum { SPH_IBN_REFERENCE = 0, SPH_IBN_PLANES = 1, SPH_IBN_UPLOADED = 2 };

/* Preconditioners of the BiCGSTAB solve (ppe_precond). */
enum { SPH_PRECOND_JACOBI = 0, SPH_PRECOND_MULTIGRID = 1 };

/* ---------------------------------------------------------------------------
 * Errors
 * ------------------------------------------------------------------------- */
enum {
    SPH_OK = 0,
    SPH_E_ARG = -1,          /* invalid argument / configuration            */
    SPH_E_HIP = -2,          /* HIP runtime or kernel launch failure        */
    SPH_E_NOMEM = -3,        /* device allocation failed / budget exceeded  */
    SPH_E_STATE = -4,        /* call out of order (e.g. step before upload) */
    SPH_E_NOCONV = -5,       /* PPE did not reach ppe_tol within ppe_maxit  */
    SPH_E_NONFINITE = -6,    /* NaN/Inf detected in the state               */
    SPH_E_UNSUPPORTED = -7   /* option the build does not implement         */
};

/* ---------------------------------------------------------------------------
 * Field identifiers for sph_get_field_f64 / sph_get_field_i32: state in the
 * device (cell-sorted) order.  P_* = persistent state (P1/P2 after the step),
 * W_* = step working copies (SP1/SP2), Q_* = per-step scratch (P3).
 * ------------------------------------------------------------------------- */
enum {
    SPH_F_X = 0, SPH_F_Y, SPH_F_Z, SPH_F_UX, SPH_F_UY, SPH_F_UZ,
    SPH_F_M, SPH_F_H, SPH_F_TEMP, SPH_F_PRES, SPH_F_PRES0, SPH_F_RHO,
    SPH_F_FLT_S, SPH_F_CONCN, SPH_F_CONCENTRATION, SPH_F_K_TURB,
    SPH_F_E_TURB, SPH_F_VIS_T, SPH_F_FBX, SPH_F_FBY, SPH_F_FBZ,
    SPH_F_FMX, SPH_F_FMY, SPH_F_HMZ, SPH_F_FLOAT1, SPH_F_FLOAT2,
    SPH_F_FLOAT3, SPH_F_VOL, SPH_F_ENTHALPY,
    /* part2 */
    SPH_F_RHO_REF, SPH_F_SR, SPH_F_X0, SPH_F_Y0, SPH_F_Z0,
    SPH_F_UX0, SPH_F_UY0, SPH_F_UZ0, SPH_F_TEMP0, SPH_F_CONCN0,
    /* part3 (scratch of the last step) */
    SPH_F_FTOTALX, SPH_F_FTOTALY, SPH_F_FTOTALZ, SPH_F_FTOTAL,
    SPH_F_FPX, SPH_F_FPY, SPH_F_FPZ, SPH_F_DTEMP, SPH_F_DCONCN,
    SPH_F_INV_CM_XX, SPH_F_INV_CM_YY, SPH_F_INV_CM_ZZ,
    SPH_F_INV_CM_XY, SPH_F_INV_CM_YZ, SPH_F_INV_CM_ZX,
    SPH_F_NX, SPH_F_NY, SPH_F_NZ,
    SPH_F_COUNT_F64
};
enum {
    SPH_I_I_TYPE = 0, SPH_I_BUFFER_TYPE, SPH_I_MOST_BUFFER, SPH_I_P_TYPE,
    SPH_I_NCELL, SPH_I_SOURCE, SPH_I_COUNT
};

/* Domain and cell grid as the reference sizes it (ISPH.cuh:75-101,
 * function_init.cuh:678-708, functions_NNPS.cuh:534-546).  All float. */
typedef struct sph_domain {
    float x_min, x_max, y_min, y_max, z_min, z_max;
    float dcell;
    int   NI, NJ, NK;
    int   num_cells;            /* NI*NJ*NK + 1 (last = dummy key)           */
} sph_domain;

/* PPE report of the last step. */
typedef struct sph_ppe_report {
    int    iterations;          /* BiCGSTAB iterations (inner, summed)       */
    int    refinements;         /* MIXED: outer refinement steps             */
    int    unknowns;            /* rows of the solve                         */
    long long nnz;              /* stored off-diagonal entries               */
    double rel_residual;        /* final ||b-Ax||/||b|| (true, FP64)         */
    int    converged;
    int    ordering;            /* row numbering: SPH_ORDER_*                 */
    long long slots;            /* stored SELL slots, padding included        */
    int    index16;             /* 16-bit column indices in use               */
    int    tail_rows;           /* MIXED: rows of more than 256 entries, whose
                                   entries past the 256th are stored
                                   contiguously and summed by a wave each     */
    int    precond;             /* preconditioner of the last inner solve:
                                   SPH_PRECOND_*                              */
    int    mg_levels;           /* multigrid levels incl. the PPE itself and
                                   the coarsest (0 without multigrid)         */
    int    mg_coarsest;         /* unknowns of the coarsest level             */
    int    mg_status;           /* SPH_MG_*: why a multigrid that was asked
                                   for is not in use (SPH_MG_OK otherwise)   */
    long long mg_sorted_rows;   /* coarse rows the sorting Galerkin kernels
                                   formed (rows of more than 256 distinct
                                   columns, or all with mg_sort_all)         */
    int    mg_long_rows;        /* coarse rows of more than 128 entries,
                                   summed by a wave each in every pass       */
    int    heavy_lists;         /* particles of the last search with more than
                                   512 neighbour-list entries (the record
                                   passes run them last)                     */
    int    tile_chunks;         /* workgroups of the LDS-staged passes (0: the
                                   step gathered per particle)               */
    int    tile_raw_chunks;     /* of those, workgroups whose halo exceeded
                                   tile_halo_max (memory gathers)            */
    int    tile_halo_largest;   /* largest staged halo (records)             */
    long long tile_halo_total;  /* staged records of all workgroups          */
    int    stencil_slices;      /* 64-row slices stored over one set of
                                   column offsets (ppe_stencil)              */
    long long stencil_nnz;      /* off-diagonal entries in those slices     */
} sph_ppe_report;

/* Why the multigrid preconditioner was not used (sph_ppe_report.mg_status). */
enum {
    SPH_MG_OK = 0,
    SPH_MG_UNSUITED = 1,        /* at most 1024 unknowns, long rows with 32-bit
                                   columns, or a lattice beyond 32-bit keys  */
    SPH_MG_TERMS = 2,           /* a coarse row gathers more than 8192 terms
                                   or 64 members                              */
    SPH_MG_COLUMNS = 3,         /* reserved: a coarse row's columns are no
                                   longer bounded apart from its terms       */
    SPH_MG_SPAN = 4,            /* a coarse row's columns lie too far apart   */
    SPH_MG_PIVOT = 5,           /* zero or non-finite pivot of the coarsest
                                   level                                      */
    SPH_MG_STALLED = 6,         /* a level kept more than 90 % of the unknowns
                                   of the one before                          */
    SPH_MG_LEVELS = 7,          /* more than 12 levels                        */
    SPH_MG_STALLED_SOLVE = 8    /* an inner solve stalled with it             */
};
/* PPE row numbering: the particles' cell order or lattice (z, y, x) order
 * (see ppe_reorder). */
enum { SPH_ORDER_CELL = 0, SPH_ORDER_LATTICE = 1 };

typedef struct sph_ctx sph_ctx;

/* Fills every member with the reference's value (see sph_params comments). */
void sph_params_default(sph_params* p);

/* Creates a context for one case with room for n_part2 particles
 * (num_part2, ISPH.cuh:111-128).  Validates the parameters: dim==3,
 * freq_filt>0, kgc_solve==0, kappa>0, supported option values. */
int sph_create(const sph_params* prm, int n_part2, sph_ctx** out);

/* Host AoS -> device.  P2 may be NULL (the reference uploads zeros,
 * ISPH_Calc.cuh:351).  ib_normals (3 floats per particle, may be NULL) are
 * the outward normals of IBM markers (p_type >= 1000); they persist for the
 * life of the case.  Sizes the domain and cell grid from the positions and
 * resets the step state: the next sph_step must pass count == 0, which runs
 * the neighbour search (ISPH_Calc.cuh:23-52). */
int sph_upload(sph_ctx* c, const sph_part1* P1, const sph_part2* P2,
               const float* ib_normals, int n);

/* One projection step == SOPHIA_single_ISPH (ISPH_Calc.cuh:4-233) for the
 * given step counter, time-step and time.  Asynchronous apart from the
 * BiCGSTAB convergence polls: returns after enqueueing the last pass.
 * Returns SPH_E_NOCONV, after completing the step, when the PPE solve did
 * not reach ppe_tol within ppe_maxit, and SPH_E_NONFINITE, after completing
 * it, when the PPE system held NaN or Inf (the solve then stops at once).
 * The reference checks neither. */
int sph_step(sph_ctx* c, int count, double dt, double time);

/* Options of the context's step engine.  use_graphs: replay each BiCGSTAB
 * iteration as one captured HIP graph (default 1).  autotune: run the
 * neighbour passes at 64/128/256/512 threads per workgroup on the first
 * four steps of each size class and keep the fastest (default 1).  Neither
 * changes any result. */
int sph_set_engine_options(sph_ctx* c, int use_graphs, int autotune);

/* Workgroup sizes the autotuner picked: one row of 1 + SPH_TUNE_PASSES ints
 * per size class (floor(log2(particles in the step)), then the size of each
 * tunable pass: grad_corr, prep, boundary3d, neumann, advection, freeslip,
 * pressure_force, ibm_force, ibm_spread, ppe_assemble; 0 while that class is
 * still being timed; neumann always runs one 64-thread wave per boundary
 * particle and reads 64).  Returns the number of rows written (<= max_rows). */
#define SPH_TUNE_PASSES 10
int sph_tuning(const sph_ctx* c, int* out, int max_rows);

/* Device -> host AoS, in the device (cell-sorted) order.  Any pointer may be
 * NULL.  Doubles are rounded to float. */
int sph_download(sph_ctx* c, sph_part1* P1, sph_part2* P2, sph_part3* P3, int n);

/* One field in full precision, device order.  n must equal the particle
 * count. */
int sph_get_field_f64(sph_ctx* c, int field, double* out, int n);
int sph_get_field_i32(sph_ctx* c, int field, int* out, int n);

/* Neighbour-search products of the last count==0 step: sorted cell keys and
 * the sort permutation (n each), and cell start/end (num_cells each; empty
 * cells have start -1 and end 0).  Any pointer may be NULL. */
int sph_nnps_debug(sph_ctx* c, int* keys, int* perm, int* cell_start,
                   int* cell_end);

/* Neighbour lists of the count==0 search, CSR in device order: for particle
 * i, entries [offsets[i], offsets[i+1]) are ascending neighbour indices.
 * kind 0 = the dcell-query window list (r < kappa*max(hi,hj), self
 * included); kind 1 = the binning-window list (0 < r < kappa*hi).
 * Call with nbr == NULL to get the entry count in *total. */
int sph_neighbors(sph_ctx* c, int kind, int* offsets, int* nbr, long long* total);

int sph_domain_info(const sph_ctx* c, sph_domain* d);
int sph_ppe_info(const sph_ctx* c, sph_ppe_report* r);

/* Max |u|, max %drho and max |f| over fluid particles (kernel_copy_max +
 * 3x DeviceReduce::Max, functions_OUTPUT.cuh:19-33, ISPH_Calc.cuh:504-514). */
int sph_stats(sph_ctx* c, float* umax, float* drho_pct_max, float* fmax);

/* sph_stats without any device allocation, memset or atomic: the same three
 * maxima, bit for bit.  No counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Its working memory is one
 * library-owned array of about 768 KB per GPU, held until the process ends
 * and shared by every call on that GPU: calls that overlap on one GPU race.
 * Cases of more than 16,777,216 particles get SPH_E_ARG (use sph_stats). */
int sph_stats_noalloc(sph_ctx* c, float* umax, float* drho_pct_max, float* fmax);

/* Blocks until all enqueued work of the context is done; reports deferred
 * errors (non-convergence, non-finite state). */
int sph_synchronize(sph_ctx* c);

/* Per-kernel timing with HIP events on the context's stream.  enable=1
 * starts recording (adds an event pair around every launch); the report
 * lists, per kernel name, launches, total milliseconds and algorithmic
 * bytes moved.  Returns the number of kernels written (<= max). */
int sph_profile_enable(sph_ctx* c, int enable);
typedef struct sph_kernel_time {
    char   name[48];
    long long launches;
    double total_ms;
    double alg_bytes;           /* sum over launches of algorithmic bytes    */
} sph_kernel_time;
int sph_profile_report(sph_ctx* c, sph_kernel_time* out, int max);

/* Adaptive particle refinement, the splitting half of Particle_refinement_cell
 * (functions_APR_etc.cuh:354-375; KERNEL_APR_condition3D_cell and
 * KERNEL_cell_APR_3D, functions_APR.cuh:270-500).  target_level (n values,
 * device order, or NULL to keep the current ones, initially 0) is the
 * reference's M_num; children inherit it.  Each pass splits, in every cell,
 * the first particle whose target exceeds its level log8(m_ref/m) by >= 0.75
 * into 8 children (m/8, vol/8, h/2, offsets +-cbrt(m/rho)/4), appended after
 * the current particles; the parent becomes a dummy (i_type 3).  The search
 * runs again after every pass; passes stop early when none splits.  *n_out
 * receives the new particle count (dummies included).  Needs a search (a
 * count == 0 step) beforehand and room in the context (n_part2). */
int sph_apr_refine(sph_ctx* c, const double* target_level, int max_passes, int* n_out);

/* Device memory held by the context, bytes. */
long long sph_device_bytes(const sph_ctx* c);

const char* sph_last_error(const sph_ctx* c);
void sph_destroy(sph_ctx* c);

/* ---- Batches: independent cases stepped as one launch sequence ----------
 * No counterpart in the reference, which runs one case per GPU thread
 * (ISPH.cuh:159-162).  A batch steps contexts made by sph_create together:
 * every pass is one launch over the particles of all listed cases, the
 * Krylov solves of all cases share their launches, and each case keeps its
 * own scalars, tolerance and convergence.  A case's results are bitwise the
 * same as when it is stepped alone with sph_step.  A context belongs to no
 * batch: any subset of contexts may be stepped together, one batch step at
 * a time per context. */
typedef struct sph_batch sph_batch;
int sph_batch_create(sph_batch** out);

/* One step of each listed case with its own count, dt and time.  The cases
 * must share precision and ppe_solver.  Returns 0, the first error, or
 * SPH_E_NOCONV / SPH_E_NONFINITE after completing the step when a case's
 * PPE did not converge or met NaN/Inf (sph_ppe_info of each case says
 * which). */
int sph_batch_step(sph_batch* b, sph_ctx* const* cases, int ncases, const int* count,
                   const double* dt, const double* time);
int sph_batch_synchronize(sph_batch* b);
int sph_batch_set_options(sph_batch* b, int use_graphs, int autotune);
int sph_batch_profile_enable(sph_batch* b, int enable);
int sph_batch_profile_report(sph_batch* b, sph_kernel_time* out, int max);
int sph_batch_tuning(const sph_batch* b, int* out, int max_rows);
const char* sph_batch_last_error(const sph_batch* b);
void sph_batch_destroy(sph_batch* b);

/* ---- Serving: continuous batching of scenario requests ------------------
 * No counterpart in the reference (its ISPH_Calc runs one deck to the end).
 * Requests (a particle deck, a step count, dt) are queued on submit,
 * admitted in arrival order while their estimated device memory fits the
 * HBM budget and fewer than max_cases run, kept resident, and advanced one
 * co-batched step per tick (sph_batch_step over every running case).  A
 * finished request frees its memory in the same tick, so queued requests
 * join at the next one.  Every case computes bitwise what it computes alone.
 * All requests of one server share precision and ppe_solver.
 * dry_run_rate > 0: no GPU is used; each tick advances a simulated clock by
 * (particles of the running cases) / dry_run_rate seconds, for testing and
 * capacity planning of the admission policy. */
enum { SPH_REQ_QUEUED = 0, SPH_REQ_RUNNING = 1, SPH_REQ_DONE = 2, SPH_REQ_FAILED = 3 };
typedef struct sph_server sph_server;
typedef struct sph_server_info {
    int queued, running, finished;
    long long reserved_bytes;       /* device memory held for running cases  */
    long long budget_bytes;
    long long ticks;                /* co-batched steps so far               */
    long long particle_steps;       /* sum over ticks of running particles   */
    double clock;                   /* seconds since creation (or simulated) */
    double admit_s, step_s, retire_s;  /* host seconds spent admitting (create +
                                       upload), in batch steps, retiring (result
                                       copies + free)                           */
} sph_server_info;
typedef struct sph_request_info {
    int state;                      /* SPH_REQ_*                             */
    int rc;                         /* 0 or the error that failed it         */
    int n, steps, steps_done, noconv_steps;
    long long ppe_iterations;       /* summed over its steps                 */
    long long est_bytes;            /* admission estimate                    */
    double t_arrival, t_admit, t_finish;   /* server clock; -1 = not yet     */
} sph_request_info;

/* Upper bound of the device memory of one case, used for admission; the
 * measured footprint replaces it after the case's first step. */
long long sph_estimate_bytes(const sph_params* prm, int n);
int sph_server_create(long long hbm_budget_bytes, int max_cases, double dry_run_rate, sph_server** out);
/* Copies the deck; t_arrival < 0 means now (server clock). */
int sph_server_submit(sph_server* s, const sph_params* prm, const sph_part1* P1, const float* ib_normals,
                      int n, int steps, double dt, double t_arrival, long long* id_out);
/* Admit what fits, run one co-batched step, retire finished requests (their
 * ids go to finished[], up to max_finished; *n_finished counts all). */
int sph_server_tick(sph_server* s, long long* finished, int max_finished, int* n_finished);
/* Dry run only: move the simulated clock forward to t (idle time). */
int sph_server_idle_until(sph_server* s, double t);
double sph_server_clock(const sph_server* s);
int sph_server_status(const sph_server* s, sph_server_info* info);
int sph_server_request(const sph_server* s, long long id, sph_request_info* info);
/* Final FP64 field of a finished request in input order: SPH_F_UX, _UY, _UZ
 * or _PRES. */
int sph_server_result(const sph_server* s, long long id, int field, double* out, int n);
/* Final state of a finished request as reference records, input order. */
int sph_server_result_part1(const sph_server* s, long long id, sph_part1* out, int n);
int sph_server_release(sph_server* s, long long id);
const char* sph_server_last_error(const sph_server* s);
void sph_server_destroy(sph_server* s);

/* ---- File formats (host only, no device work) ---------------------------
 * The reference's solver settings file (solv.txt), particle deck and VTK
 * output.  None of these touch a context; they fill or read the records
 * above. */

/* Run controls and the settings of solv.txt that are not step parameters.
 * Codes follow the reference's globals (Parameters.cuh, function_init.cuh). */
#define SPH_MAX_PLOT 64
typedef struct sph_run_config {
    int    solver_type;         /* 0 WCSPH (the reference's default), 1 ISPH  */
    int    prop_table;          /* property_table(YES/NO):                    */
    int    atmos_stab;          /* 0 Unstable, 1 Neutral, 2 Stable            */
    int    terrain_type;        /* 0 Flat .. 4 Ocean                          */
    int    structure;           /* structure_solv                             */
    int    open_boundary;       /* 0 NO, 1 OPEN, 2 PERIODIC                   */
    int    switch_ptype;
    int    apr_solve;           /* 0 NO, 1 Cell, 2 Switch, 3 Block            */
    int    block_config[3];     /* NBI, NBJ, NBK                              */
    int    nb_cell_type;        /* 0 3X3, 1 5X5                               */
    int    timestep_update;     /* flag_timestep_update                       */
    int    freq_output;         /* plot-output                                */
    int    freq_apr, freq_bapr, freq_temp, freq_h_reset;
    int    min_iteration, max_iteration;
    double buffersize, p_ref, dp_th, drho_th, p_relaxation, dt_structure;
    double dt, time, time_end;  /* time-step, start-time, end-time            */
    int    num_plot;            /* plot-variables: count and names            */
    char   plot[SPH_MAX_PLOT][32];
    char   unknown[256];        /* keywords whose value matched no option,
                                   space separated (the reference keeps the
                                   previous value for these, as here)         */
} sph_run_config;

/* Reads solv.txt: replaces read_solv_input (function_init.cuh:166-463).
 * Tokens are whitespace separated; each keyword is followed by the
 * reference's number of filler tokens and then its value.  Step settings
 * go into *prm (members not named by the file keep their values, except
 * the solve switches that the reference clears first: fp, fv, fg, con,
 * boussinesq, kgc, turbulence, concn); everything else goes into *run,
 * which is reset first.  Values are stored as the reference codes them,
 * including options sph_create rejects (e.g. viscous-force Violeau = 3).
 * Returns SPH_E_ARG if the file cannot be read. */
int sph_read_solv(const char* path, sph_params* prm, sph_run_config* run);

/* Reads a particle deck: replaces read_input (function_init.cuh:550-676)
 * and gpu_count_particle_numbers2 (function_init.cuh:483-498).  The first
 * line lists integer column codes separated by tabs (1 x, 2 y, 3 z, 4 ux,
 * 5 uy, 6 uz, 7 m, 8 p_type, 9 h, 10 temp, 11 pres, 12 rho, 15 concn,
 * 26 h_ref, 27 m_ref; every other code is read and skipped); then one
 * whitespace-separated value per column and particle.  Each particle read
 * gets i_type 1 and zero elsewhere; an incomplete last record is dropped.
 * With P1 == NULL only counts (*n_out).  SPH_E_NOMEM if more than cap
 * particles; SPH_E_ARG if the file or its header cannot be read. */
int sph_read_deck(const char* path, sph_part1* P1, int cap, int* n_out);

/* Writes legacy VTK (version 3.0, BINARY, POLYDATA, big-endian): replaces
 * save_vtk_bin_single_flag (functions_OUTPUT.cuh:121-616).  Particles with
 * i_type 1 or 2 are written: POINTS (float x, y, z), then POINT_DATA with
 * one FIELD array for every table entry whose leading characters match a
 * label, as the reference's strncmp table does (e.g. "velocity" -> velocity
 * x3, "smoothing" -> h, "rho_ref" -> rho and rho_ref).  The FIELD count is
 * the number of arrays written.  P2 and P3 may be NULL when no label needs
 * them (rho_ref; normal, dtemp), else SPH_E_ARG. */
int sph_write_vtk(const char* path, const sph_part1* P1, const sph_part2* P2, const sph_part3* P3, int n,
                  const char* const* labels, int nlabels);

/* Material property tables (temperature T, enthalpy h, conductivity k,
 * specific heat cp, viscosity mu), one table per p_type, stored back to
 * back: table i holds entries index[i] .. index[i] + size[i] - 1.  The
 * reference's capacity is 50 entries and 10 tables (Parameters.cuh:66,
 * SOPHIA_gpu.cu:56-57); these are larger. */
#define SPH_TABLE_CAP 512
#define SPH_MAX_TABLES 32
typedef struct sph_prop_tables {
    int   ntables, ndata;
    int   index[SPH_MAX_TABLES], size[SPH_MAX_TABLES];
    float T[SPH_TABLE_CAP], h[SPH_TABLE_CAP], k[SPH_TABLE_CAP], cp[SPH_TABLE_CAP], mu[SPH_TABLE_CAP];
} sph_prop_tables;

/* Reads a property-table file: replaces read_table (function_init.cuh:724-970).
 * A block "#p_type N" ... "#end" holds lines starting //T, //h, //k, //cp,
 * //mu, each followed by one line of comma-separated values.  A block's
 * length is that of the last array read in it; index/size follow the
 * reference (tables numbered 0 .. ntables-1 in file order).  SPH_E_NOMEM
 * beyond the capacities; SPH_E_ARG for an unreadable file or a block
 * without #end. */
int sph_read_tables(const char* path, sph_prop_tables* t);

/* Table lookup: replaces interp2 (functions_PROP.cuh:8-40).  Piecewise-
 * linear in float through (x[i], y[i]), extended linearly past both ends;
 * NaN for size < 2 or a NaN argument. */
float sph_table_interp(const float* x, const float* y, int size, float v);

#ifdef __cplusplus
}
#endif
#endif /* URBANSPH_H */
