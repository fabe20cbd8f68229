// C ABI of the MI355X ISPH projection step (include/urbansph.h).
//
// The context owns every device buffer of one case, the way the reference's
// per-GPU thread ISPH_Calc owns them (ISPH_Calc.cuh:327-344,553-571), and
// sph_step enqueues the step sequence of SOPHIA_single_ISPH
// (ISPH_Calc.cuh:4-233) on the context's stream.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "flags.h"
#include "sph_internal.h"

namespace sph {

// Device blocks are recycled within a context: a freed block goes to the
// context's cache and serves a later request of its size up to 1/8 more.
// The steps that search every step (moving particles) or after refinement
// would otherwise pay a hipMalloc/hipFree pair, and hipFree's device-wide
// wait, per buffer per step.
int dmalloc(sph_ctx* c, void** p, size_t bytes) {
    *p = nullptr;
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    auto it = c->cache.lower_bound(bytes);
    if (it != c->cache.end() && it->first <= bytes + bytes / 8) {
        *p = it->second;
        c->live[*p] = it->first;
        c->cache.erase(it);
        return 0;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !c->cache.empty()) {
        (void)hipGetLastError();
        dcache_release(c);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, SPH_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    c->live[*p] = bytes;
    c->bytes += (long long)bytes;
    return 0;
}
void dfree(sph_ctx* c, void* p) {
    if (!p) return;
    auto it = c->live.find(p);
    if (it == c->live.end()) {
        (void)hipFree(p);
        return;
    }
    c->cache.emplace(it->second, p);
    c->live.erase(it);
}
// Returns the cached blocks to HIP (hipFree waits for the device first).
void dcache_release(sph_ctx* c) {
    for (auto& b : c->cache) {
        (void)hipFree(b.second);
        c->bytes -= (long long)b.first;
    }
    c->cache.clear();
}

// AoS (reference layout) -> SoA (same order)
__global__ void __launch_bounds__(BLOCK) k_unpack(int n, const sph_part1* A, const sph_part2* B, const float* nrm,
                                                  sph_ctx* dummy, PView P, P2View P2, QView Q, double* eli,
                                                  double* m_ref, double* h_ref, int* it, int* bt, int* mb, int* pt,
                                                  int* nc, int* src) {
    (void)dummy;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sph_part1& a = A[i];
    it[i] = a.i_type; bt[i] = a.buffer_type; mb[i] = a.MOST_buffer; pt[i] = a.p_type; nc[i] = a.ncell; src[i] = a.source;
    P.x[i] = a.x; P.y[i] = a.y; P.z[i] = a.z; P.ux[i] = a.ux; P.uy[i] = a.uy; P.uz[i] = a.uz;
    P.m[i] = a.m; P.h[i] = a.h; P.temp[i] = a.temp; P.pres[i] = a.pres; P.pres0[i] = a.pres0; P.rho[i] = a.rho;
    P.flt_s[i] = a.flt_s; P.enthalpy[i] = a.enthalpy; P.concn[i] = a.concn; P.concentration[i] = a.concentration;
    P.k_turb[i] = a.k_turb; P.e_turb[i] = a.e_turb; P.vol[i] = a.vol; P.vis_t[i] = a.vis_t;
    P.fbx[i] = a.fbx; P.fby[i] = a.fby; P.fbz[i] = a.fbz; P.fmx[i] = a.fmx; P.fmy[i] = a.fmy; P.hmz[i] = a.hmz;
    P.float1[i] = a.float1; P.float2[i] = a.float2; P.float3[i] = a.float3;
    eli[i] = a.eli; m_ref[i] = a.m_ref; h_ref[i] = a.h_ref;
    if (B) {
        const sph_part2& b = B[i];
        P2.rho_ref[i] = b.rho_ref; P2.SR[i] = b.SR; P2.x0[i] = b.x0; P2.y0[i] = b.y0; P2.z0[i] = b.z0;
        P2.ux0[i] = b.ux0; P2.uy0[i] = b.uy0; P2.uz0[i] = b.uz0; P2.temp0[i] = b.temp0; P2.concn0[i] = b.concn0;
    } else {
        P2.rho_ref[i] = 0; P2.SR[i] = 0; P2.x0[i] = 0; P2.y0[i] = 0; P2.z0[i] = 0;
        P2.ux0[i] = 0; P2.uy0[i] = 0; P2.uz0[i] = 0; P2.temp0[i] = 0; P2.concn0[i] = 0;
    }
    if (nrm) { Q.nx[i] = nrm[3 * i]; Q.ny[i] = nrm[3 * i + 1]; Q.nz[i] = nrm[3 * i + 2]; }
    else { Q.nx[i] = 0; Q.ny[i] = 0; Q.nz[i] = 0; }
}

// SoA -> AoS for download (rounded to float)
__global__ void __launch_bounds__(BLOCK) k_pack(int n, PView P, P2View S2, QView Q, const double* eli,
                                                const double* m_ref, const double* h_ref, sph_part1* A, sph_part2* B,
                                                sph_part3* C) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (A) {
        sph_part1 a;
        a.i_type = P.i_type[i]; a.buffer_type = P.buffer_type[i]; a.MOST_buffer = P.MOST_buffer[i]; a.p_type = P.p_type[i];
        a.x = (float)P.x[i]; a.y = (float)P.y[i]; a.z = (float)P.z[i];
        a.ux = (float)P.ux[i]; a.uy = (float)P.uy[i]; a.uz = (float)P.uz[i];
        a.m = (float)P.m[i]; a.h = (float)P.h[i]; a.temp = (float)P.temp[i]; a.pres = (float)P.pres[i];
        a.pres0 = (float)P.pres0[i]; a.rho = (float)P.rho[i]; a.flt_s = (float)P.flt_s[i];
        a.enthalpy = (float)P.enthalpy[i]; a.concn = (float)P.concn[i]; a.concentration = (float)P.concentration[i];
        a.k_turb = (float)P.k_turb[i]; a.e_turb = (float)P.e_turb[i]; a.eli = (float)eli[i]; a.vol = (float)P.vol[i];
        a.m_ref = (float)m_ref[i]; a.h_ref = (float)h_ref[i]; a.vis_t = (float)P.vis_t[i];
        a.fbx = (float)P.fbx[i]; a.fby = (float)P.fby[i]; a.fbz = (float)P.fbz[i];
        a.fmx = (float)P.fmx[i]; a.fmy = (float)P.fmy[i]; a.hmz = (float)P.hmz[i];
        a.float1 = (float)P.float1[i]; a.float2 = (float)P.float2[i]; a.float3 = (float)P.float3[i];
        a.ncell = P.ncell[i]; a.source = P.source[i];
        A[i] = a;
    }
    if (B) {
        sph_part2 b{};
        b.rho_ref = (float)S2.rho_ref[i]; b.SR = (float)S2.SR[i];
        b.x0 = (float)S2.x0[i]; b.y0 = (float)S2.y0[i]; b.z0 = (float)S2.z0[i];
        b.ux0 = (float)S2.ux0[i]; b.uy0 = (float)S2.uy0[i]; b.uz0 = (float)S2.uz0[i];
        b.temp0 = (float)S2.temp0[i]; b.concn0 = (float)S2.concn0[i];
        B[i] = b;
    }
    if (C) {
        sph_part3 q{};
        q.ftotalx = (float)Q.ftotalx[i]; q.ftotaly = (float)Q.ftotaly[i]; q.ftotalz = (float)Q.ftotalz[i];
        q.ftotal = (float)Q.ftotal[i]; q.fpx = (float)Q.fpx[i]; q.fpy = (float)Q.fpy[i]; q.fpz = (float)Q.fpz[i];
        q.dtemp = (float)Q.dtemp[i]; q.dconcn = (float)Q.dconcn[i];
        q.nx = (float)Q.nx[i]; q.ny = (float)Q.ny[i]; q.nz = (float)Q.nz[i];
        q.inv_cm_xx = (float)Q.inv_xx[i]; q.inv_cm_yy = (float)Q.inv_yy[i]; q.inv_cm_zz = (float)Q.inv_zz[i];
        q.inv_cm_xy = (float)Q.inv_xy[i]; q.inv_cm_yz = (float)Q.inv_yz[i]; q.inv_cm_zx = (float)Q.inv_zx[i];
        C[i] = q;
    }
}

}  // namespace sph

using namespace sph;

// --------------------------------------------------------------- views ----
PView sph_ctx::pview() const {
    PView v;
    v.i_type = pi[SPH_I_I_TYPE]; v.buffer_type = pi[SPH_I_BUFFER_TYPE]; v.MOST_buffer = pi[SPH_I_MOST_BUFFER];
    v.p_type = pi[SPH_I_P_TYPE]; v.ncell = pi[SPH_I_NCELL]; v.source = pi[SPH_I_SOURCE];
    v.x = pf[SPH_F_X]; v.y = pf[SPH_F_Y]; v.z = pf[SPH_F_Z]; v.ux = pf[SPH_F_UX]; v.uy = pf[SPH_F_UY]; v.uz = pf[SPH_F_UZ];
    v.m = pf[SPH_F_M]; v.h = pf[SPH_F_H]; v.temp = pf[SPH_F_TEMP]; v.pres = pf[SPH_F_PRES]; v.pres0 = pf[SPH_F_PRES0];
    v.rho = pf[SPH_F_RHO]; v.flt_s = pf[SPH_F_FLT_S]; v.enthalpy = pf[SPH_F_ENTHALPY]; v.concn = pf[SPH_F_CONCN];
    v.concentration = pf[SPH_F_CONCENTRATION]; v.k_turb = pf[SPH_F_K_TURB]; v.e_turb = pf[SPH_F_E_TURB];
    v.eli = pf[SPH_F_COUNT_F64 + 0]; v.vol = pf[SPH_F_VOL]; v.m_ref = pf[SPH_F_COUNT_F64 + 1];
    v.h_ref = pf[SPH_F_COUNT_F64 + 2]; v.vis_t = pf[SPH_F_VIS_T];
    v.fbx = pf[SPH_F_FBX]; v.fby = pf[SPH_F_FBY]; v.fbz = pf[SPH_F_FBZ];
    v.fmx = pf[SPH_F_FMX]; v.fmy = pf[SPH_F_FMY]; v.hmz = pf[SPH_F_HMZ];
    v.float1 = pf[SPH_F_FLOAT1]; v.float2 = pf[SPH_F_FLOAT2]; v.float3 = pf[SPH_F_FLOAT3];
    return v;
}
PView sph_ctx::sview() const {
    PView v = pview();
    v.ux = sm[MF_UX]; v.uy = sm[MF_UY]; v.uz = sm[MF_UZ]; v.pres = sm[MF_PRES]; v.temp = sm[MF_TEMP];
    v.vis_t = sm[MF_VIS_T]; v.k_turb = sm[MF_K_TURB]; v.e_turb = sm[MF_E_TURB]; v.flt_s = sm[MF_FLT_S];
    v.fmx = sm[MF_FMX]; v.fmy = sm[MF_FMY]; v.hmz = sm[MF_HMZ];
    v.float1 = sm[MF_FLOAT1]; v.float2 = sm[MF_FLOAT2]; v.float3 = sm[MF_FLOAT3];
    v.fbx = sm[MF_FBX]; v.fby = sm[MF_FBY]; v.fbz = sm[MF_FBZ];
    v.concentration = sm[MF_CONCENTRATION]; v.concn = sm[MF_CONCN]; v.enthalpy = sm[MF_ENTHALPY];
    return v;
}
static P2View mk2(double* const* a) {
    return P2View{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9]};
}
P2View sph_ctx::p2view() const { return mk2(p2); }
P2View sph_ctx::s2view() const { return mk2(s2); }
QView sph_ctx::qview() const {
    QView v;
    v.ftotalx = q[0]; v.ftotaly = q[1]; v.ftotalz = q[2]; v.ftotal = q[3];
    v.fpx = q[4]; v.fpy = q[5]; v.fpz = q[6]; v.dtemp = q[7]; v.dconcn = q[8];
    v.inv_xx = q[9]; v.inv_yy = q[10]; v.inv_zz = q[11]; v.inv_xy = q[12]; v.inv_yz = q[13]; v.inv_zx = q[14];
    v.nx = q[15]; v.ny = q[16]; v.nz = q[17];
    return v;
}
Grid sph_ctx::grid() const { return Grid{dom, gstr, gend, cstart, keys}; }
NList sph_ctx::nlist() const { return NList{nl_slice_off, nl_cnt, nl_ent}; }

// ----------------------------------------------------------- C ABI ----
extern "C" {

void sph_params_default(sph_params* p) {
    // Reference defaults: Parameters.cuh, functions_PROP.cuh:121-131,
    // functions_NNPS.cuh:4-14, functions_PPE.cuh:268-271, functions_BC.cuh:732.
    std::memset(p, 0, sizeof(*p));
    p->dim = 3; p->kernel_type = SPH_WENDLAND2; p->time_type = SPH_PRE_COR; p->freq_filt = 1;
    p->fp_solve = 1; p->fv_solve = SPH_FV_CLEARY; p->noslip_bc = 1; p->artificial_viscosity = 1;
    p->turbulence_model = SPH_LAMINAR;
    p->kappa = 2.0; p->u_limit = 1000.0;
    p->mu_fluid = 1.8e-5; p->mu_wall = 1.0e-2; p->gravity = 9.810;
    p->T_ini = 300.0; p->T_con = 300.0; p->T_start_t = 0.0; p->T_con_t = 10.0; p->tempb = 300.0;
    p->Cs_sgs = 0.15; p->Ck_sgs = 0.1; p->Ce_sgs = 0.93; p->rho_air = 1.225;
    p->av_alpha = 0.05; p->av_beta = 0.05; p->av_c = 30.0;
    p->solid_k = 1.0; p->air_k = 0.025; p->solid_cp = 900.0; p->air_cp = 1005.0; p->solid_rho = 2300.0;
    p->diff_coef = 2.0;
    p->L2 = 0.028; p->L3 = 13.304; p->L_left = -0.412; p->L_right = 0.412;
    p->outlet_copy_dx = 0.008; p->copy_tol = 0.001; p->neumann_ghost_r = 0.1;
    p->most_mode = SPH_MOST_REFERENCE; p->most_z0 = 1.0; p->most_z0h = 0.1; p->most_L = 1e6;
    p->most_ghost_dz = 5.0; p->most_ghost_r = 0.1;
    p->ibm_ut_factor = 0.2;
    p->ppe_solver = SPH_PPE_REF_JACOBI; p->precision = SPH_FP64; p->ppe_tol = 1e-8; p->ppe_maxit = 5000;
    p->ppe_inner_tol = 1e-5; p->ppe_max_refine = 20; p->ppe_index16 = 1; p->ppe_reorder = 1; p->apr_h_change = 1; p->ale_positions = 0;
    p->ppe_precond = SPH_PRECOND_JACOBI;
    // IB normals: the face planes of functions_PREP.cuh:958-963
    p->ib_normals = SPH_IBN_REFERENCE;
    p->ib_face_xlo = 0.764; p->ib_face_xhi = 0.836; p->ib_face_ylo = -0.036; p->ib_face_yhi = 0.036;
    p->ib_face_zlo = 0.0; p->ib_face_zhi = 0.152; p->ib_face_tol = 1e-5;
    // LDS-staged record passes: off (measured slower than the per-particle gathers)
    p->tile_cells = 0; p->tile_particles = 128; p->tile_halo_max = 1536; p->compact_records = 1; p->ppe_stencil = 1;
}

static int validate(const sph_params* p, std::string& why) {
    if (p->dim != 3) { why = "dim must be 3"; return SPH_E_UNSUPPORTED; }
    if (p->kernel_type < SPH_GAUSSIAN || p->kernel_type > SPH_WENDLAND6) { why = "kernel_type"; return SPH_E_ARG; }
    if (p->freq_filt <= 0) { why = "freq_filt must be > 0 (functions_PREP.cuh:395)"; return SPH_E_ARG; }
    if (p->kgc_solve != 0) { why = "kgc_solve != NO is not supported (its Cm matrix is never built)"; return SPH_E_UNSUPPORTED; }
    if (!(p->kappa > 0)) { why = "kappa must be > 0"; return SPH_E_ARG; }
    if (p->time_type != SPH_EULER && p->time_type != SPH_PRE_COR) { why = "time_type"; return SPH_E_ARG; }
    if (p->fv_solve < 0 || p->fv_solve > 2) { why = "fv_solve"; return SPH_E_UNSUPPORTED; }
    if (p->turbulence_model != SPH_LAMINAR && p->turbulence_model != SPH_SSM && p->turbulence_model != SPH_DDF) {
        why = "turbulence_model"; return SPH_E_UNSUPPORTED;
    }
    if (p->ppe_solver != SPH_PPE_REF_JACOBI && p->ppe_solver != SPH_PPE_BICGSTAB) { why = "ppe_solver"; return SPH_E_ARG; }
    if (p->precision != SPH_FP64 && p->precision != SPH_MIXED) { why = "precision"; return SPH_E_ARG; }
    if (p->ppe_solver == SPH_PPE_BICGSTAB && (!(p->ppe_tol > 0) || p->ppe_maxit <= 0)) { why = "ppe_tol/ppe_maxit"; return SPH_E_ARG; }
    if (p->precision == SPH_MIXED && (!(p->ppe_inner_tol > 0) || p->ppe_max_refine < 0)) { why = "ppe_inner_tol"; return SPH_E_ARG; }
    if (p->most_mode != SPH_MOST_REFERENCE && p->most_mode != SPH_MOST_PRESCRIBED_L) { why = "most_mode"; return SPH_E_ARG; }
    if (p->ppe_precond != SPH_PRECOND_JACOBI && p->ppe_precond != SPH_PRECOND_MULTIGRID) { why = "ppe_precond"; return SPH_E_ARG; }
    if (p->ib_normals < SPH_IBN_REFERENCE || p->ib_normals > SPH_IBN_UPLOADED) { why = "ib_normals"; return SPH_E_ARG; }
    if (p->tile_cells != 0) {
        for (int a = 0; a < 3; a++) {
            const int v = (p->tile_cells >> (8 * a)) & 0xff;
            if (v < 1 || v > 16) { why = "tile_cells: 1..16 cells an axis"; return SPH_E_ARG; }
        }
        if (p->tile_cells >> 24) { why = "tile_cells"; return SPH_E_ARG; }
    }
    if (p->tile_particles != 0 && (p->tile_particles < 64 || p->tile_particles > 256 || p->tile_particles % 64)) {
        why = "tile_particles: 64, 128, 192 or 256"; return SPH_E_ARG;
    }
    if (p->tile_halo_max < 0 || p->tile_halo_max > 2048) { why = "tile_halo_max: 1..2048"; return SPH_E_ARG; }
    if (p->ppe_stencil != 0 && p->ppe_stencil != 1) { why = "ppe_stencil: 0 or 1"; return SPH_E_ARG; }
    return 0;
}

int sph_create(const sph_params* prm, int n_part2, sph_ctx** out) {
    if (!out) return SPH_E_ARG;
    *out = nullptr;
    if (!prm || n_part2 <= 0 || n_part2 > MAX_PARTICLES) return SPH_E_ARG;
    std::string why;
    int rc = validate(prm, why);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPH_E_HIP;
    sph_ctx* c = new sph_ctx();
    c->prm = *prm;
    c->cap = n_part2;
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return SPH_E_HIP; }
    c->c_pi15 = std::pow(3.141592653, 1.5);
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) { delete c; return SPH_E_HIP; }
    c->stream = c->own_stream;
    c->solo = new Eng();
    c->profp = &c->solo->prof;
    if (eng_init(*c->solo, c->own_stream) != 0) { sph_destroy(c); return SPH_E_HIP; }
    const size_t N = (size_t)n_part2;
    auto A = [&](double** p) { return dalloc(c, p, N); };
    int e = 0;
    static const int pfields[] = {SPH_F_X, SPH_F_Y, SPH_F_Z, SPH_F_UX, SPH_F_UY, SPH_F_UZ, SPH_F_M, SPH_F_H,
                                  SPH_F_TEMP, SPH_F_PRES, SPH_F_PRES0, SPH_F_RHO, SPH_F_FLT_S, SPH_F_CONCN,
                                  SPH_F_CONCENTRATION, SPH_F_K_TURB, SPH_F_E_TURB, SPH_F_VIS_T, SPH_F_FBX,
                                  SPH_F_FBY, SPH_F_FBZ, SPH_F_FMX, SPH_F_FMY, SPH_F_HMZ, SPH_F_FLOAT1,
                                  SPH_F_FLOAT2, SPH_F_FLOAT3, SPH_F_VOL, SPH_F_ENTHALPY};
    for (int f : pfields) e = e ? e : A(&c->pf[f]);
    for (int f = 0; f < 3; f++) e = e ? e : A(&c->pf[SPH_F_COUNT_F64 + f]);
    for (int f = 0; f < SPH_I_COUNT; f++) e = e ? e : dalloc(c, &c->pi[f], N);
    for (int f = 0; f < MF_COUNT; f++) e = e ? e : A(&c->sm[f]);
    for (int f = 0; f < 10; f++) e = e ? e : A(&c->p2[f]);
    for (int f = 0; f < 10; f++) e = e ? e : A(&c->s2[f]);
    for (int f = 0; f < 18; f++) e = e ? e : A(&c->q[f]);
    for (int f = 0; f < 5; f++) e = e ? e : A(&c->snap[f]);
    e = e ? e : dalloc(c, &c->stat_buf, 16);
    if (e) { sph_destroy(c); return e; }
    *out = c;
    return 0;
}

int sph_upload(sph_ctx* c, const sph_part1* P1, const sph_part2* P2, const float* ib_normals, int n) {
    if (!c || !P1) return SPH_E_ARG;
    if (n <= 0 || n > c->cap) return fail(c, SPH_E_ARG, "n out of range");
    for (int i = 0; i < n; i++) {
        if (!(P1[i].h > 0.0f) || !(P1[i].h_ref > 0.0f))
            return fail(c, SPH_E_ARG, "h and h_ref must be > 0 (functions_NNPS.cuh:953)");
        if (!std::isfinite(P1[i].x) || !std::isfinite(P1[i].y) || !std::isfinite(P1[i].z))
            return fail(c, SPH_E_ARG, "non-finite position");
    }
    c->n = n;
    c->pos_f32 = true;
    dfree(c, c->porder);   // the pass order of the last search; the next search rebuilds it
    c->porder = nullptr;
    // find_minmax (skips the last particle), function_init.cuh:678-708; cell
    // sizing ISPH.cuh:89-97, functions_NNPS.cuh:534-546.  float arithmetic.
    float mnx = P1[0].x, mxx = P1[0].x, mny = P1[0].y, mxy = P1[0].y, mnz = P1[0].z, mxz = P1[0].z;
    for (int i = 0; i < n - 1; i++) {
        float tx = P1[i].x, ty = P1[i].y, tz = P1[i].z;
        if (tx < mnx) mnx = tx;
        if (tx > mxx) mxx = tx;
        if (ty < mny) mny = ty;
        if (ty > mxy) mxy = ty;
        if (tz < mnz) mnz = tz;
        if (tz > mxz) mxz = tz;
    }
    sph_domain d{};
    d.x_min = mnx; d.x_max = mxx; d.y_min = mny; d.y_max = mxy; d.z_min = mnz; d.z_max = mxz;
    volatile float crf = 1.1f;
    volatile float kap = (float)c->prm.kappa;
    volatile float h0 = P1[0].h_ref;
    volatile float t1 = crf * kap;
    volatile float t2 = t1 * h0;
    d.dcell = t2 / 2.0f;
    volatile float ex = (d.x_max - d.x_min) / d.dcell, ey = (d.y_max - d.y_min) / d.dcell, ez = (d.z_max - d.z_min) / d.dcell;
    d.NI = (int)ex + 1;
    d.NJ = (int)ey + 1;
    d.NK = (int)ez + 1;
    long long cells = (long long)d.NI * d.NJ * d.NK + 1;
    if (cells > (1LL << 30)) return fail(c, SPH_E_ARG, "cell grid too large");
    if (c->gstr && cells != (long long)c->dom.num_cells) {
        dfree(c, c->gstr); dfree(c, c->gend); dfree(c, c->cstart);
        c->gstr = c->gend = c->cstart = nullptr;
    }
    d.num_cells = (int)cells;
    c->dom = d;
    // staging copy + unpack
    void* a1 = nullptr; void* a2 = nullptr; void* an = nullptr;
    SPH_TRY(dmalloc(c, &a1, sizeof(sph_part1) * (size_t)n));
    SPH_HIP(c, hipMemcpyAsync(a1, P1, sizeof(sph_part1) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    if (P2) {
        SPH_TRY(dmalloc(c, &a2, sizeof(sph_part2) * (size_t)n));
        SPH_HIP(c, hipMemcpyAsync(a2, P2, sizeof(sph_part2) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    }
    if (ib_normals) {
        SPH_TRY(dmalloc(c, &an, sizeof(float) * 3 * (size_t)n));
        SPH_HIP(c, hipMemcpyAsync(an, ib_normals, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    }
    hipLaunchKernelGGL(k_unpack, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const sph_part1*)a1,
                       (const sph_part2*)a2, (const float*)an, (sph_ctx*)nullptr, c->pview(), c->p2view(), c->qview(),
                       c->pf[SPH_F_COUNT_F64], c->pf[SPH_F_COUNT_F64 + 1], c->pf[SPH_F_COUNT_F64 + 2],
                       c->pi[SPH_I_I_TYPE], c->pi[SPH_I_BUFFER_TYPE], c->pi[SPH_I_MOST_BUFFER], c->pi[SPH_I_P_TYPE],
                       c->pi[SPH_I_NCELL], c->pi[SPH_I_SOURCE]);
    SPH_TRY(launch_check(c, "unpack"));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, a1); dfree(c, a2); dfree(c, an);
    ppe_free(c);
    c->uploaded = true;
    c->searched = false;
    c->generation++;
    return 0;
}

int sph_step(sph_ctx* c, int count, double dt, double time) {
    if (!c) return SPH_E_ARG;
    sph_ctx* one[1] = {c};
    int rc = engine_step(*c->solo, one, 1, &count, &dt, &time);
    if (rc && !c->solo->err.empty()) c->err = c->solo->err;
    return rc;
}

int sph_synchronize(sph_ctx* c) {
    if (!c) return SPH_E_ARG;
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    int rc = prof_collect(*c->solo);
    if (rc) c->err = c->solo->err;
    return rc;
}

int sph_set_engine_options(sph_ctx* c, int use_graphs, int autotune) {
    if (!c) return SPH_E_ARG;
    c->solo->use_graphs = use_graphs != 0;
    c->solo->autotune = autotune != 0;
    return 0;
}

static int tuning_rows(const Eng& e, int* out, int max_rows) {
    if (!out || max_rows <= 0) return 0;
    int k = 0;
    for (auto& t : e.tune) {
        if (k >= max_rows) break;
        int* row = out + (size_t)k * (1 + TK_COUNT);
        row[0] = t.first;
        for (int q = 0; q < TK_COUNT; q++) row[1 + q] = t.second.phase >= TUNE_CANDIDATES ? t.second.block[q] : 0;
        k++;
    }
    return k;
}

int sph_tuning(const sph_ctx* c, int* out, int max_rows) {
    if (!c) return SPH_E_ARG;
    return tuning_rows(*c->solo, out, max_rows);
}

int sph_download(sph_ctx* c, sph_part1* P1, sph_part2* P2, sph_part3* P3, int n) {
    if (!c) return SPH_E_ARG;
    if (!c->uploaded || n != c->n) return fail(c, SPH_E_ARG, "n must equal the uploaded particle count");
    void *a = nullptr, *b = nullptr, *q = nullptr;
    if (P1) SPH_TRY(dmalloc(c, &a, sizeof(sph_part1) * (size_t)n));
    if (P2) SPH_TRY(dmalloc(c, &b, sizeof(sph_part2) * (size_t)n));
    if (P3) SPH_TRY(dmalloc(c, &q, sizeof(sph_part3) * (size_t)n));
    hipLaunchKernelGGL(k_pack, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, c->pview(), c->s2view(), c->qview(),
                       (const double*)c->pf[SPH_F_COUNT_F64], (const double*)c->pf[SPH_F_COUNT_F64 + 1],
                       (const double*)c->pf[SPH_F_COUNT_F64 + 2], (sph_part1*)a, (sph_part2*)b, (sph_part3*)q);
    SPH_TRY(launch_check(c, "pack"));
    if (P1) SPH_HIP(c, hipMemcpyAsync(P1, a, sizeof(sph_part1) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (P2) SPH_HIP(c, hipMemcpyAsync(P2, b, sizeof(sph_part2) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (P3) SPH_HIP(c, hipMemcpyAsync(P3, q, sizeof(sph_part3) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, a); dfree(c, b); dfree(c, q);
    return 0;
}

int sph_get_field_f64(sph_ctx* c, int field, double* out, int n) {
    if (!c || !out) return SPH_E_ARG;
    if (!c->uploaded || n != c->n) return fail(c, SPH_E_ARG, "n must equal the uploaded particle count");
    const double* src = nullptr;
    if (field >= 0 && field < SPH_F_RHO_REF) src = c->pf[field];
    else if (field >= SPH_F_RHO_REF && field < SPH_F_FTOTALX) src = c->s2[field - SPH_F_RHO_REF];
    else if (field >= SPH_F_FTOTALX && field < SPH_F_COUNT_F64) src = c->q[field - SPH_F_FTOTALX];
    if (!src) return fail(c, SPH_E_ARG, "unknown field");
    SPH_HIP(c, hipMemcpyAsync(out, src, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sph_get_field_i32(sph_ctx* c, int field, int* out, int n) {
    if (!c || !out) return SPH_E_ARG;
    if (!c->uploaded || n != c->n) return fail(c, SPH_E_ARG, "n must equal the uploaded particle count");
    if (field < 0 || field >= SPH_I_COUNT) return fail(c, SPH_E_ARG, "unknown field");
    SPH_HIP(c, hipMemcpyAsync(out, c->pi[field], sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sph_nnps_debug(sph_ctx* c, int* keys, int* perm, int* cell_start, int* cell_end) {
    if (!c) return SPH_E_ARG;
    if (!c->searched) return fail(c, SPH_E_STATE, "no neighbour search yet");
    const size_t n = (size_t)c->n, nc = (size_t)c->dom.num_cells;
    if (keys) SPH_HIP(c, hipMemcpyAsync(keys, c->keys, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    if (perm) SPH_HIP(c, hipMemcpyAsync(perm, c->perm, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    if (cell_start) SPH_HIP(c, hipMemcpyAsync(cell_start, c->gstr, sizeof(int) * nc, hipMemcpyDeviceToHost, c->stream));
    if (cell_end) SPH_HIP(c, hipMemcpyAsync(cell_end, c->gend, sizeof(int) * nc, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sph_neighbors(sph_ctx* c, int kind, int* offsets, int* nbr, long long* total) {
    if (!c) return SPH_E_ARG;
    if (!c->searched) return fail(c, SPH_E_STATE, "no neighbour search yet");
    if (kind != 0 && kind != 1) return fail(c, SPH_E_ARG, "kind must be 0 or 1");
    const int n = c->n;
    int nslices = (n + SLICE - 1) / SLICE;
    std::vector<int> cnt(n);
    std::vector<long long> off(nslices + 1);
    std::vector<unsigned> ent((size_t)c->nl_total + 1);
    SPH_HIP(c, hipMemcpyAsync(cnt.data(), c->nl_cnt, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipMemcpyAsync(off.data(), c->nl_slice_off, sizeof(long long) * nslices, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipMemcpyAsync(ent.data(), c->nl_ent, sizeof(unsigned) * (size_t)c->nl_total, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    const unsigned want = kind == 0 ? F_DP : F_B;
    long long tot = 0;
    for (int i = 0; i < n; i++) {
        if (offsets) offsets[i] = (int)tot;
        const unsigned* row = ent.data() + off[i >> 6] + (i & 63);
        for (int k = 0; k < cnt[i]; k++) {
            unsigned e = row[(size_t)k * SLICE];
            if (e & want) {
                if (nbr) nbr[tot] = (int)(e & NB_IDX);
                tot++;
            }
        }
    }
    if (offsets) offsets[n] = (int)tot;
    if (total) *total = tot;
    return 0;
}

int sph_domain_info(const sph_ctx* c, sph_domain* d) {
    if (!c || !d) return SPH_E_ARG;
    if (!c->uploaded) return SPH_E_STATE;
    *d = c->dom;
    return 0;
}

int sph_ppe_info(const sph_ctx* c, sph_ppe_report* r) {
    if (!c || !r) return SPH_E_ARG;
    *r = c->rep;
    return 0;
}

int sph_stats(sph_ctx* c, float* umax, float* drho_pct_max, float* fmax_) {
    if (!c || !umax || !drho_pct_max || !fmax_) return SPH_E_ARG;
    if (!c->searched) return fail(c, SPH_E_STATE, "no step yet");
    return stats_run(c, umax, drho_pct_max, fmax_);
}

int sph_stats_noalloc(sph_ctx* c, float* umax, float* drho_pct_max, float* fmax_) {
    if (!c || !umax || !drho_pct_max || !fmax_) return SPH_E_ARG;
    if (!c->searched) return fail(c, SPH_E_STATE, "no step yet");
    return stats_run_noalloc(c, umax, drho_pct_max, fmax_);
}

static int profile_enable(Eng& e, int enable) {
    if (!enable) SPH_TRY(prof_collect(e));
    e.prof.on = enable != 0;
    if (enable) e.prof.recs.clear();
    return 0;
}
static int profile_report(Eng& e, sph_kernel_time* out, int max) {
    SPH_HIP(&e, hipStreamSynchronize(e.st));
    SPH_TRY(prof_collect(e));
    int k = 0;
    for (auto& r : e.prof.recs) {
        if (k >= max || !out) break;
        sph_kernel_time& t = out[k++];
        std::memset(&t, 0, sizeof(t));
        std::snprintf(t.name, sizeof(t.name), "%s", r.name.c_str());
        t.launches = r.launches;
        t.total_ms = r.ms;
        t.alg_bytes = r.bytes;
    }
    return k;
}

int sph_profile_enable(sph_ctx* c, int enable) {
    if (!c) return SPH_E_ARG;
    return profile_enable(*c->solo, enable);
}

int sph_profile_report(sph_ctx* c, sph_kernel_time* out, int max) {
    if (!c) return SPH_E_ARG;
    return profile_report(*c->solo, out, max);
}

int sph_apr_refine(sph_ctx* c, const double* target_level, int max_passes, int* n_out) {
    if (!c || max_passes < 0) return SPH_E_ARG;
    if (!c->uploaded || !c->searched)
        return fail(c, SPH_E_STATE, "refinement needs a neighbour search (a count == 0 step) first");
    c->err.clear();
    return apr_refine(c, target_level, max_passes, n_out);
}

long long sph_device_bytes(const sph_ctx* c) { return c ? c->bytes : 0; }

// ------------------------------------------------------------ batches ----
struct sph_batch {
    Eng e;
};

int sph_batch_create(sph_batch** out) {
    if (!out) return SPH_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SPH_E_HIP;
    sph_batch* b = new sph_batch();
    if (eng_init(b->e, nullptr) != 0) { eng_free(b->e); delete b; return SPH_E_HIP; }
    *out = b;
    return 0;
}

int sph_batch_step(sph_batch* b, sph_ctx* const* cases, int ncases, const int* count, const double* dt,
                   const double* time) {
    if (!b) return SPH_E_ARG;
    return engine_step(b->e, cases, ncases, count, dt, time);
}

int sph_batch_synchronize(sph_batch* b) {
    if (!b) return SPH_E_ARG;
    SPH_HIP(&b->e, hipStreamSynchronize(b->e.st));
    return prof_collect(b->e);
}

int sph_batch_set_options(sph_batch* b, int use_graphs, int autotune) {
    if (!b) return SPH_E_ARG;
    b->e.use_graphs = use_graphs != 0;
    b->e.autotune = autotune != 0;
    return 0;
}

int sph_batch_profile_enable(sph_batch* b, int enable) {
    if (!b) return SPH_E_ARG;
    return profile_enable(b->e, enable);
}

int sph_batch_profile_report(sph_batch* b, sph_kernel_time* out, int max) {
    if (!b) return SPH_E_ARG;
    return profile_report(b->e, out, max);
}

int sph_batch_tuning(const sph_batch* b, int* out, int max_rows) {
    if (!b) return SPH_E_ARG;
    return tuning_rows(b->e, out, max_rows);
}

const char* sph_batch_last_error(const sph_batch* b) { return b ? b->e.err.c_str() : "null batch"; }

void sph_batch_destroy(sph_batch* b) {
    if (!b) return;
    eng_free(b->e);
    delete b;
}

const char* sph_last_error(const sph_ctx* c) { return c ? c->err.c_str() : "null context"; }

void sph_destroy(sph_ctx* c) {
    if (!c) return;
    if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    if (c->solo) { eng_free(*c->solo); delete c->solo; c->solo = nullptr; }
    ppe_free(c);
    for (auto p : c->pf) dfree(c, p);
    for (auto p : c->pi) dfree(c, p);
    for (auto p : c->sm) dfree(c, p);
    for (auto p : c->p2) dfree(c, p);
    for (auto p : c->s2) dfree(c, p);
    for (auto p : c->q) dfree(c, p);
    for (auto p : c->snap) dfree(c, p);
    tile_release(c);
    void* rest[] = {c->keys, c->perm, c->keys_tmp, c->perm_tmp, c->gstr, c->gend, c->cstart, c->nm_list, c->nl_slice_off, c->nl_wprev, c->porder, c->nl_cnt,
                    c->nl_ent, c->stat_buf, c->apr_target, c->xnew[0], c->xnew[1], c->xnew[2], c->rec, c->crec, c->crec_adv, c->ibm_near, c->nmp_off, c->nmp_ent, c->nmp_cent, c->nm_set, c->nm_rec};
    for (auto p : rest) dfree(c, p);
    dcache_release(c);
    for (auto& b : c->live) (void)hipFree(b.first);   // none expected: every buffer is freed above
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

}  // extern "C"
