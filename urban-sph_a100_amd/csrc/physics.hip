// Per-step physics passes of the ISPH projection step on MI355X.
//
// One thread per particle over structure-of-arrays state.  Neighbour passes
// walk the per-particle neighbour list built once by the count==0 search
// (the reference rescans the 5^3 cell window every pass; in its Eulerian mode
// positions never move, ISPH_Calc.cuh:23-63).  Each list entry carries
// precomputed FP64 window/radius flags, so a pass selects exactly the
// reference's neighbour set for that kernel in both precision modes.
// Passes whose search is centred on a ghost node (Neumann inlet mirror,
// periodic copy, outlet copy, MOST deposit) scan the cell window directly.
//
// R is the pair-math type (double: FP64 mode, float: mixed mode).  Sums are
// always accumulated in double.  Differences of positions and velocities are
// taken in double and then rounded to R, so the float path loses no
// significance to cancellation.
#include <mutex>
#include <string>

#include "flags.h"
#include "sph_internal.h"
#include "sph_math.h"

namespace sph {

// This thread's particle: case descriptor `cd` and index `i` in the case;
// threads past the case's end return.
#define SEG_PARTICLE(sg)                          \
    int first_;                                   \
    const int cid_ = seg_case((sg), first_);      \
    const CaseDesc& cd = (sg).cd[cid_];           \
    const int i = first_ + (int)threadIdx.x;      \
    if (i >= cd.n) return;
// prep, advection and the pressure force: the particles of short lists, a
// thread each in the pass order; in MIXED the last cd.n_heavy of that order
// (the heavy lists) are left to a wave each (SEG_HEAVY)
#define SEG_PARTICLE_LIGHT(sg)                            \
    int first_;                                           \
    const int cid_ = seg_case((sg), first_);              \
    const CaseDesc& cd = (sg).cd[cid_];                   \
    const int t_ = first_ + (int)threadIdx.x;             \
    if (t_ >= cd.n - cd.n_heavy) return;                  \
    const int i = cd.porder ? cd.porder[t_] : t_;
#define SEG_HEAVY(sg)                                         \
    int first_;                                               \
    const int cid_ = seg_case((sg), first_);                  \
    const CaseDesc& cd = (sg).cd[cid_];                       \
    const int u_ = (first_ + (int)threadIdx.x) >> 6;          \
    const int lane_ = threadIdx.x & 63;                       \
    if (u_ >= cd.n_heavy) return;                             \
    const int i = cd.porder[cd.n - cd.n_heavy + u_];
// The same through the case's pass order when it has one (long neighbour
// lists last): for the record-gathering passes, where it measured faster.
#define SEG_PARTICLE_ORDERED(sg)                          \
    int first_;                                           \
    const int cid_ = seg_case((sg), first_);              \
    const CaseDesc& cd = (sg).cd[cid_];                   \
    const int t_ = first_ + (int)threadIdx.x;             \
    if (t_ >= cd.n) return;                               \
    const int i = cd.porder ? cd.porder[t_] : t_;


// Particles of the cells (i0..i1, cj, ck) that lie inside the grid, as one
// range [rg.x, rg.y) of the sorted order; false when the row is outside the
// grid or the clipped x-range is empty.
__device__ __forceinline__ bool row_range(const Grid& g, int i0, int i1, int cj, int ck, int2& rg) {
    const sph_domain& d = g.d;
    if (cj < 0 || cj > d.NJ - 1 || ck < 0 || ck > d.NK - 1) return false;
    i0 = max(i0, 0);
    i1 = min(i1, d.NI - 1);
    if (i0 > i1) return false;
    const int base = d.NI * cj + d.NI * d.NJ * ck;
    rg = make_int2(g.cstart[base + i0], g.cstart[base + i1 + 1]);
    return true;
}

// Bin-cell range that contains every particle within distance r of p
// (binning is monotone in each coordinate; a small margin absorbs the
// double -> float rounding of the bounds).
__device__ __forceinline__ void ball_bins(const sph_domain& d, double px, double py, double pz, double r,
                                          Cell& lo, Cell& hi) {
    const double m = 1e-9 + 1e-6 * r;
    lo = Cell{bin_axis((float)(px - r - m), d.x_min, d.x_max, d.NI), bin_axis((float)(py - r - m), d.y_min, d.y_max, d.NJ),
              bin_axis((float)(pz - r - m), d.z_min, d.z_max, d.NK)};
    hi = Cell{bin_axis((float)(px + r + m), d.x_min, d.x_max, d.NI), bin_axis((float)(py + r + m), d.y_min, d.y_max, d.NJ),
              bin_axis((float)(pz + r + m), d.z_min, d.z_max, d.NK)};
}


// ------------------------------------------------------------------------
// Step start: SP1 <- P1 and SP2 <- P2 for the mutable fields, P3 <- 0
// (ISPH_Calc.cuh:53-63), and KERNEL_set_ncell (functions_NNPS.cuh:943-955).
__global__ void __launch_bounds__(MAX_BLOCK) k_step_begin(Seg sg) {
    SEG_PARTICLE(sg);
    const PView P = cd.P, S = cd.S;
    const P2View P2 = cd.P2, S2 = cd.S2;
    const QView Q = cd.Q;
    if (P.i_type[i] != 3) P.ncell[i] = (int)(1.0 * ceilf(2.0f / (float)(P.h_ref[i] / P.h[i])));
    S.ux[i] = P.ux[i]; S.uy[i] = P.uy[i]; S.uz[i] = P.uz[i]; S.pres[i] = P.pres[i]; S.temp[i] = P.temp[i];
    S.vis_t[i] = P.vis_t[i]; S.k_turb[i] = P.k_turb[i]; S.e_turb[i] = P.e_turb[i]; S.flt_s[i] = P.flt_s[i];
    S.fmx[i] = P.fmx[i]; S.fmy[i] = P.fmy[i]; S.hmz[i] = P.hmz[i];
    S.float1[i] = P.float1[i]; S.float2[i] = P.float2[i]; S.float3[i] = P.float3[i];
    S.fbx[i] = P.fbx[i]; S.fby[i] = P.fby[i]; S.fbz[i] = P.fbz[i];
    S.concentration[i] = P.concentration[i]; S.concn[i] = P.concn[i]; S.enthalpy[i] = P.enthalpy[i];
    S2.rho_ref[i] = P2.rho_ref[i]; S2.SR[i] = P2.SR[i]; S2.x0[i] = P2.x0[i]; S2.y0[i] = P2.y0[i];
    S2.z0[i] = P2.z0[i]; S2.ux0[i] = P2.ux0[i]; S2.uy0[i] = P2.uy0[i]; S2.uz0[i] = P2.uz0[i];
    S2.temp0[i] = P2.temp0[i]; S2.concn0[i] = P2.concn0[i];
    Q.ftotalx[i] = 0.0; Q.ftotaly[i] = 0.0; Q.ftotalz[i] = 0.0; Q.ftotal[i] = 0.0;
    Q.fpx[i] = 0.0; Q.fpy[i] = 0.0; Q.fpz[i] = 0.0; Q.dtemp[i] = 0.0; Q.dconcn[i] = 0.0;
    if (!cd.gc_fresh) {
        Q.inv_xx[i] = 0.0; Q.inv_yy[i] = 0.0; Q.inv_zz[i] = 0.0; Q.inv_xy[i] = 0.0; Q.inv_yz[i] = 0.0; Q.inv_zx[i] = 0.0;
    }
    if (cd.xnew[0]) { cd.xnew[0][i] = P.x[i]; cd.xnew[1][i] = P.y[i]; cd.xnew[2][i] = P.z[i]; }
}

// ALE: the corrector's positions become the state at the end of the step,
// after the boundary passes that still search the step's cells.
__global__ void __launch_bounds__(MAX_BLOCK) k_commit_positions(Seg sg) {
    SEG_PARTICLE(sg);
    if (!cd.xnew[0]) return;
    const PView P = cd.P;
    P.x[i] = cd.xnew[0][i]; P.y[i] = cd.xnew[1][i]; P.z[i] = cd.xnew[2][i];
}

// Snapshot of the fields a copy/mirror pass reads while it writes them.
__global__ void __launch_bounds__(MAX_BLOCK) k_snapshot(Seg sg, int which, int nm_pack) {
    SEG_PARTICLE(sg);
    const PView A = which ? cd.P : cd.S;
    cd.snap[0][i] = A.ux[i]; cd.snap[1][i] = A.uy[i]; cd.snap[2][i] = A.uz[i];
    cd.snap[3][i] = A.pres[i]; cd.snap[4][i] = A.temp[i];
    if (nm_pack && cd.nm_rec && i < cd.nm_nset) {   // the Neumann pass's records, from the same state
        const int j = cd.nm_set[i];
        float4* r = (float4*)(cd.nm_rec + (size_t)i * NMREC_FLOATS);
        r[0] = make_float4((float)A.x[j], (float)A.y[j], (float)A.z[j], (float)A.m[j]);
        r[1] = make_float4((float)A.rho[j], (float)A.ux[j], (float)A.uy[j], (float)A.uz[j]);
        *(double2*)(r + 2) = make_double2(A.pres[j], 0.0);
    }
}

// ------------------------------------------------------------------------
// KERNEL_clc_gradient_correction_3D, functions_KNL.cuh:216-324
template <class R>
__global__ void __launch_bounds__(MAX_BLOCK) k_grad_corr(Seg sg) {
    SEG_PARTICLE(sg);
    if (cd.gc_fresh) return;   // unchanged since the last step of this search
    const KConst kc = cd.kc;
    const PView S = cd.S;
    const QView Q = cd.Q;
    const NList nl = cd.nl;
    int ptype_i = S.p_type[i];
    R hi = (R)S.h[i];
    R tmp_A = calc_tmpA<R>(kc, hi);
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double tmpxx = 0, tmpyy = 0, tmpzz = 0, tmpxy = 0, tmpyz = 0, tmpzx = 0;
    NLRow row = nl_row(nl, i);
    for (int k = 0; k < row.cnt; k++) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_B)) continue;
        int j = e & NB_IDX;
        R dx = (R)(xi - S.x[j]), dy = (R)(yi - S.y[j]), dz = (R)(zi - S.z[j]);
        R tdist = sqrt(dx * dx + dy * dy + dz * dz);
        R tdwij = kernel_dw<R>(kc, tmp_A, hi, tdist);
        R mj = (R)S.m[j], rhoj = (R)S.rho[j];
        R mtd = mj * tdwij;
        R rtd = (R)1.0 / (rhoj * tdist);
        R txx = -mtd * dx * dx; txx *= rtd;
        R txy = -mtd * dy * dx; txy *= rtd;
        R tyy = -mtd * dy * dy; tyy *= rtd;
        R tzx = -mtd * dx * dz; tzx *= rtd;
        R tyz = -mtd * dy * dz; tyz *= rtd;
        R tzz = -mtd * dz * dz; tzz *= rtd;
        R f = (R)(S.p_type[j] == 1);
        tmpxx += (double)(txx * f); tmpxy += (double)(txy * f); tmpyy += (double)(tyy * f);
        tmpzx += (double)(tzx * f); tmpyz += (double)(tyz * f); tmpzz += (double)(tzz * f);
    }
    double tmpcmd = tmpxx * (tmpyy * tmpzz - tmpyz * tmpyz);
    tmpcmd -= tmpxy * (tmpxy * tmpzz - tmpyz * tmpzx);
    tmpcmd += tmpzx * (tmpxy * tmpyz - tmpyy * tmpzx);
    if ((fabs(tmpcmd) > 2.2204460492503131e-16) && (ptype_i == 1)) {
        double rtcmd = 1.0 / tmpcmd;
        Q.inv_xx[i] = (tmpyy * tmpzz - tmpyz * tmpyz) * rtcmd;
        Q.inv_xy[i] = (tmpzx * tmpyz - tmpxy * tmpzz) * rtcmd;
        Q.inv_zx[i] = (tmpxy * tmpyz - tmpzx * tmpyy) * rtcmd;
        Q.inv_yy[i] = (tmpxx * tmpzz - tmpzx * tmpzx) * rtcmd;
        Q.inv_yz[i] = (tmpzx * tmpxy - tmpxx * tmpyz) * rtcmd;
        Q.inv_zz[i] = (tmpxx * tmpyy - tmpxy * tmpxy) * rtcmd;
    } else {
        Q.inv_xx[i] = 1.0; Q.inv_xy[i] = 0.0; Q.inv_zx[i] = 0.0;
        Q.inv_yy[i] = 1.0; Q.inv_yz[i] = 0.0; Q.inv_zz[i] = 1.0;
    }
}

// ------------------------------------------------------------------------
// KERNEL_clc_prep3D, functions_PREP.cuh:287-488
// Neighbour records of every particle from the working state (see
// rec_pos): all pairs, or only those that hold velocities and pressures, or
// pressures.
template <int WHICH>
__global__ void __launch_bounds__(BLOCK) k_pack_records(Seg sg) {
    SEG_PARTICLE(sg);
    const PView S = cd.S;
    if (cd.crec) {   // the gathering passes of this case read the compact records
        if (WHICH == REC_PQ) {
            float4* c = (float4*)(cd.crec + (size_t)i * CREC_FLOATS);
            c[0] = make_float4((float)S.x[i], (float)S.y[i], (float)S.z[i], (float)S.m[i]);
            c[1] = make_float4((float)S.rho[i], (float)(S.pres[i] - S.pres0[i]), (float)S.h[i],
                               __int_as_float(S.p_type[i]));
            return;
        }
        float4* c = (float4*)(cd.crec_adv + (size_t)i * CREC_ADV_FLOATS);
        double2* d = (double2*)(c + 2);
        d[0] = make_double2(S.ux[i], S.uy[i]);
        d[1] = make_double2(S.uz[i], S.pres[i] - S.pres0[i]);
        if (WHICH == REC_VEL_PQ) return;
        const int t = (S.p_type[i] & 0xffff) | (S.i_type[i] << 16) |
                      (cd.ppe.rowof && cd.ppe.rowof[i] >= 0 ? CREC_ROW_BIT : 0);
        if (WHICH == REC_PRE) c[0] = make_float4((float)S.x[i], (float)S.y[i], (float)S.z[i], (float)S.m[i]);
        c[1] = make_float4((float)S.rho[i], (float)S.h[i],
                           (float)(viscosity2(cd.pr, S.p_type[i]) + S.vis_t[i] * S.rho[i]), __int_as_float(t));
        return;
    }
    if (WHICH == REC_PRE) return;
    double2* r = (double2*)(cd.rec + (size_t)i * REC_DOUBLES);
    r[2] = make_double2(S.rho[i], S.pres[i] - S.pres0[i]);
    if (WHICH == REC_PQ) return;
    r[3] = make_double2(S.ux[i], S.uy[i]);
    r[4] = make_double2(S.uz[i], S.h[i]);
    if (WHICH == REC_VEL_PQ) return;
    const long long t = (long long)(unsigned)S.p_type[i] | ((long long)S.i_type[i] << 32);
    r[0] = make_double2(S.x[i], S.y[i]);
    r[1] = make_double2(S.z[i], S.m[i]);
    r[5] = make_double2(S.vis_t[i], __longlong_as_double(t));
    r[6] = make_double2(S.temp[i], S.concn[i]);
}

// W = 64 (MIXED, a heavy list): the calling wave's lanes take entries lane,
// lane + 64, ... and wave_fold adds their sums; lane 0 writes.
template <class R, class Src, int W = 1>
__device__ __forceinline__ void prep_body(const CaseDesc& cd, int i, const NList& nl, int li, const Src src,
                                          int lane = 0) {
    const KConst kc = cd.kc;
    const sph_params pr = cd.pr;
    const PView S = cd.S;
    const P2View S2 = cd.S2;
    const QView Q = cd.Q;
    const int count = cd.count;
    const double dt = cd.dt;
    if (S.i_type[i] > 2) return;
    if (S.p_type[i] >= 1000) return;
    if (S.p_type[i] <= 0) return;
    double ki = fmax(S.k_turb[i], 0.0);
    double nui = S.vis_t[i];
    R gc11 = (R)Q.inv_xx[i], gc12 = (R)Q.inv_xy[i], gc22 = (R)Q.inv_yy[i];
    R gc23 = (R)Q.inv_yz[i], gc31 = (R)Q.inv_zx[i], gc33 = (R)Q.inv_zz[i];
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double uxi = S.ux[i], uyi = S.uy[i], uzi = S.uz[i];
    double hi_d = S.h[i];
    R hi = (R)hi_d;
    R tmp_A = calc_tmpA<R>(kc, hi);
    double tmp_flt = 0.0, tmp_sigma = 0.0;
    double st11 = 0, st12 = 0, st22 = 0, st23 = 0, st31 = 0, st33 = 0;
    const bool filt = (count % pr.freq_filt) == 0;
    const bool les = (pr.fv_solve != 0) && (pr.turbulence_model != SPH_LAMINAR);
    const bool compact = Src::lds || (sizeof(R) == 4 && src.ok());
    const NLRow row = nl_row(nl, li);
    for (int k = W > 1 ? lane : 0; k < row.cnt; k += W) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_HP)) continue;
        int j = e & NB_IDX;
        // the neighbour's position, m, rho and type from the compact record
        // (the same values as the field arrays hold) or the field arrays
        double xj, yj, zj, mj_d, rhoj_d;
        int ptj;
        const float* C = compact ? src.at(j) : nullptr;
        if (compact) {
            const rec_f4 a = Src::f4(C), b = Src::f4(src.piece(C, 1));
            xj = a.x; yj = a.y; zj = a.z; mj_d = a.w; rhoj_d = b.x; ptj = crec_ptype(b.w);
        } else {
            if constexpr (Src::lds) continue;   // staged records are compact
            xj = S.x[j]; yj = S.y[j]; zj = S.z[j]; mj_d = S.m[j]; rhoj_d = S.rho[j]; ptj = S.p_type[j];
        }
        R dx = (R)(xi - xj), dy = (R)(yi - yj), dz = (R)(zi - zj);
        R mj = (R)mj_d, rhoj = (R)rhoj_d;
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) + (R)1e-20;
        R tmp_wij = kernel_w<R>(kc, tmp_A, hi, tdist);
        R tmp_dwij = kernel_dw<R>(kc, tmp_A, hi, tdist);
        R tdwx = tmp_dwij * dx / tdist;
        R tdwy = tmp_dwij * dy / tdist;
        R tdwz = tmp_dwij * dz / tdist;
        if (filt) tmp_flt += (double)(mj / rhoj * tmp_wij);
        if (les) {
            tmp_sigma += (double)(tdist * tdist * tmp_wij * mj / rhoj);
            if (ptj == 1) {
                R tdwxc = gc11 * tdwx + gc12 * tdwy + gc31 * tdwz;
                R tdwyc = gc12 * tdwx + gc22 * tdwy + gc23 * tdwz;
                R tdwzc = gc31 * tdwx + gc23 * tdwy + gc33 * tdwz;
                double uxj, uyj, uzj;
                if (compact) {
                    const rec_d2 u = Src::d2(src.piece(C, 2)), w = Src::d2(src.piece(C, 3));
                    uxj = u.x; uyj = u.y; uzj = w.x;
                } else {
                    uxj = S.ux[j]; uyj = S.uy[j]; uzj = S.uz[j];
                }
                R dux = (R)(uxj - uxi), duy = (R)(uyj - uyi), duz = (R)(uzj - uzi);
                st11 += (double)(mj / rhoj * dux * tdwxc);
                st12 += (double)((R)0.5 * mj / rhoj * (dux * tdwyc + duy * tdwxc));
                st22 += (double)(mj / rhoj * duy * tdwyc);
                st23 += (double)((R)0.5 * mj / rhoj * (duy * tdwzc + duz * tdwyc));
                st31 += (double)((R)0.5 * mj / rhoj * (dux * tdwzc + duz * tdwxc));
                st33 += (double)(mj / rhoj * duz * tdwzc);
            }
        }
    }
    if constexpr (W > 1) {
        wave_fold(tmp_flt); wave_fold(tmp_sigma);
        wave_fold(st11); wave_fold(st12); wave_fold(st22); wave_fold(st23); wave_fold(st31); wave_fold(st33);
        if (lane != 0) return;
    }
    if (les) {
        double tmp_SR = 2 * (st11 * st11 + st22 * st22 + st33 * st33 + 2 * st12 * st12 + 2 * st31 * st31 + 2 * st23 * st23);
        tmp_SR = sqrt(tmp_SR);
        S2.SR[i] = tmp_SR;
        tmp_sigma = 0.3333 * tmp_sigma;
        tmp_sigma = sqrt(tmp_sigma);
        double tmp_cl = tmp_sigma > hi_d ? tmp_sigma : hi_d;
        if (pr.turbulence_model == SPH_SSM) {
            double tvis_t = pr.Cs_sgs * pr.Cs_sgs * tmp_cl * tmp_cl * tmp_SR;
            if (S.buffer_type[i] != 1) S.vis_t[i] = tvis_t;
        } else if (pr.turbulence_model == SPH_DDF) {
            double p = nui * tmp_SR * tmp_SR;
            double b = 0.0;
            double eps = pr.Ce_sgs * ki * sqrt(ki) / tmp_cl;
            double dk = p + b - eps + 0.0;
            ki += dk * dt;
            ki = ki > 1e-4 ? ki : 1e-4;
            double tvis_t = pr.Ck_sgs * tmp_cl * sqrt(ki);
            if (S.buffer_type[i] != 1) { S.vis_t[i] = tvis_t; S.k_turb[i] = ki; S.e_turb[i] = eps; }
        }
    }
    if (count == 0) S2.rho_ref[i] = S.rho[i];
    if (filt) S.flt_s[i] = tmp_flt;
}
template <class R>
__global__ void __launch_bounds__(MAX_BLOCK) k_prep(Seg sg) {
    SEG_PARTICLE_LIGHT(sg);
    prep_body<R>(cd, i, cd.nl, i, GSrc<CREC_ADV_FLOATS>{sizeof(R) == 4 ? &cd.crec_adv : nullptr});
}
__global__ void __launch_bounds__(WAVE_BLOCK) k_prep_heavy(Seg sg) {
    SEG_HEAVY(sg);
    prep_body<float, GSrc<CREC_ADV_FLOATS>, 64>(cd, i, cd.nl, i, GSrc<CREC_ADV_FLOATS>{&cd.crec_adv}, lane_);
}
__global__ void __launch_bounds__(TILE_MAX) k_prep_tile(Seg sg) {
    TILE_PARTICLE(sg, cd.crec_adv, CREC_ADV_FLOATS);
    if (chunk_.y >= 0)
        prep_body<float>(cd, i, cd.tile.nl, lane_, LSrc<CREC_ADV_FLOATS>{(const float*)tile_lds_, chunk_.y,
                                                     cd.tile.halo + chunk_.x,
                                             cd.tile.self[lane_]});
    else prep_body<float>(cd, i, cd.tile.nl, lane_, GSrc<CREC_ADV_FLOATS>{&cd.crec_adv});
}

// ------------------------------------------------------------------------
// KERNEL_boundary3D, functions_BC.cuh:100-190 (wall velocity)
template <class R>
__global__ void __launch_bounds__(MAX_BLOCK) k_boundary3d(Seg sg) {
    SEG_PARTICLE(sg);
    const sph_params pr = cd.pr;
    if (pr.noslip_bc != 1) return;   // ISPH_Calc.cuh:83
    const KConst kc = cd.kc;
    const PView S = cd.S;
    const NList nl = cd.nl;
    if (S.i_type[i] > 2) return;
    int p_type_i = S.p_type[i];
    if (p_type_i > 0) return;
    R hi = (R)S.h[i];
    R tmp_A = calc_tmpA<R>(kc, hi);
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double uxi = S.ux[i], uyi = S.uy[i], uzi = S.uz[i];
    double tmpx = 0.0, tmpy = 0.0, tmpz = 0.0, flt = 1.0e-10;
    NLRow row = nl_row(nl, i);
    for (int k = 0; k < row.cnt; k++) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_HM)) continue;
        int j = e & NB_IDX;
        R dx = (R)(xi - S.x[j]), dy = (R)(yi - S.y[j]), dz = (R)(zi - S.z[j]);
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
        R twij = kernel_w<R>(kc, tmp_A, hi, tdist);
        int p_type_j = S.p_type[j];
        if (pr.noslip_bc == 1) {
            if ((p_type_j > 0) && (p_type_j != 9)) {
                tmpx += (double)((R)S.ux[j] * twij);
                tmpy += (double)((R)S.uy[j] * twij);
                tmpz += (double)((R)S.uz[j] * twij);
                flt += (double)twij;
            }
        }
    }
    if (pr.noslip_bc == 1 && (p_type_i == 0 || p_type_i == 9)) {
        S.ux[i] = 2 * uxi * (p_type_i == 9) - tmpx / flt;
        S.uy[i] = 2 * uyi * (p_type_i == 9) - tmpy / flt;
        S.uz[i] = 2 * uzi * (p_type_i == 9) - tmpz / flt;
    }
}

// The same in MIXED, a 64-lane wave per boundary particle (walls are a few
// per cent of the case): lane sums over entries lane, lane + 64, ..., then
// a fixed butterfly; only the particles the serial kernel writes (p_type 0
// and 9) are visited.
__global__ void __launch_bounds__(WAVE_BLOCK) k_boundary3d_wave(Seg sg) {
    int first_;
    const int cid_ = seg_case(sg, first_);
    const CaseDesc& cd = sg.cd[cid_];
    const int u = (first_ + (int)threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (u >= cd.nm_count) return;
    const sph_params pr = cd.pr;
    if (pr.noslip_bc != 1) return;   // ISPH_Calc.cuh:83
    const PView S = cd.S;
    const int i = cd.nm_list[u];
    if (S.i_type[i] > 2) return;
    const int p_type_i = S.p_type[i];
    if (p_type_i != 0 && p_type_i != 9) return;
    const KConst kc = cd.kc;
    const NList nl = cd.nl;
    float hi = (float)S.h[i];
    float tmp_A = calc_tmpA<float>(kc, hi);
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    const NLRow row = nl_row(nl, i);
    for (int k = lane; k < row.cnt; k += 64) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_HM)) continue;
        int j = e & NB_IDX;
        const int p_type_j = S.p_type[j];
        if (!((p_type_j > 0) && (p_type_j != 9))) continue;
        float dx = (float)(xi - S.x[j]), dy = (float)(yi - S.y[j]), dz = (float)(zi - S.z[j]);
        float tdist = sqrt(dx * dx + dy * dy + dz * dz) - 1e-20f;
        float twij = kernel_w<float>(kc, tmp_A, hi, tdist);
        t[0] += (double)((float)S.ux[j] * twij);
        t[1] += (double)((float)S.uy[j] * twij);
        t[2] += (double)((float)S.uz[j] * twij);
        t[3] += (double)twij;
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 4; q++) t[q] += __shfl_xor(t[q], off, 64);
    if (lane != 0) return;
    const double flt = 1.0e-10 + t[3];
    const double uxi = S.ux[i], uyi = S.uy[i], uzi = S.uz[i];
    S.ux[i] = 2 * uxi * (p_type_i == 9) - t[0] / flt;
    S.uy[i] = 2 * uyi * (p_type_i == 9) - t[1] / flt;
    S.uz[i] = 2 * uzi * (p_type_i == 9) - t[2] / flt;
}

// ------------------------------------------------------------------------
// KERNEL_MOST_boundary3D, functions_MOST.cuh:175-322.  Part 1: the
// surface-layer fluxes of each MOST particle; the neighbour heat deposit is
// stored per particle for the gather in part 2.
__global__ void __launch_bounds__(MAX_BLOCK) k_most_self(Seg sg) {
    SEG_PARTICLE(sg);
    const sph_params pr = cd.pr;
    const PView S = cd.S;
    double* dep = cd.snap[4];   // the snapshots are not live between the copy passes
    const int count = cd.count;
    const double dt = cd.dt;
    if (S.MOST_buffer[i] != 1) { dep[i] = 0.0; return; }
    const double KVK = 0.4;
    double zi = S.z[i];
    double uxi = S.ux[i], uyi = S.uy[i];
    double tempi = S.temp[i], rhoi = S.rho[i], mi = S.m[i];
    double space = S.h[i] / 1.5;
    double z0 = pr.most_z0, z0h = pr.most_z0h;
    double u_star = 0.0, theta_star = 0.0, OB_length;
    double theta_surface = wall_temperature(pr, dt, count);
    double wind_speed = sqrt(uxi * uxi + uyi * uyi);
    wind_speed = fmax(0.001, fmin(wind_speed, 100.0));
    if (pr.most_mode == SPH_MOST_PRESCRIBED_L) {
        OB_length = pr.most_L;
        u_star = KVK * wind_speed / Phi_m(OB_length, zi, z0);
        theta_star = KVK * (tempi - theta_surface) / Phi_h(OB_length, zi, z0h);
    } else {
        double Ri_b = (pr.gravity / tempi) * (tempi - theta_surface) * zi / (wind_speed * wind_speed + 1e-10);
        if (fabs(Ri_b) > 1e-3) {
            OB_length = zi / Ri_b;
            OB_length = fmax(-500.0, fmin(OB_length, 500.0));
            for (int iter = 0; iter < 200; iter++) {
                double ol_m = OB_length, ol_l = 0.999 * ol_m, ol_u = 1.001 * ol_m;
                double pm = Phi_m(ol_m, zi, z0);
                double fL = Ri_b - (zi / ol_m) * Phi_h(ol_m, zi, z0h) / (pm * pm);
                double dfL = ((-(zi / ol_u) * Phi_h(ol_u, zi, z0h) / (Phi_m(ol_u, zi, z0) * Phi_m(ol_u, zi, z0))) -
                              (-(zi / ol_l) * Phi_h(ol_l, zi, z0h) / (Phi_m(ol_l, zi, z0) * Phi_m(ol_l, zi, z0)))) /
                             (ol_u - ol_l);
                OB_length -= 1.0 * fL / dfL;
                if (fabs(OB_length) > 1e-6) {
                    if (fabs((OB_length - ol_m) / OB_length) < 1e-5) break;
                    else if (fabs(OB_length - ol_m) < 1e-5) break;
                }
            }
            u_star = KVK * wind_speed / Phi_m(OB_length, zi, z0);
            theta_star = KVK * (tempi - theta_surface) / Phi_h(OB_length, zi, z0h);
        } else {
            OB_length = 1e6;
            u_star = KVK * wind_speed / log(zi / z0);
            theta_star = KVK * (tempi - theta_surface) / log(zi / z0h);
        }
    }
    double momentum_x = -rhoi * u_star * (uxi * u_star / wind_speed);
    double momentum_y = -rhoi * u_star * (uyi * u_star / wind_speed);
    double shf = -rhoi * pr.air_cp * u_star * theta_star * 100.0;
    dep[i] = 0.5 * shf / (space * rhoi * pr.air_cp);
    S.fmx[i] = 1.0 * momentum_x * space * space / mi;
    S.fmy[i] = 1.0 * momentum_y * space * space / mi;
    S.float1[i] = u_star; S.float2[i] = theta_star; S.float3[i] = OB_length;
}

// Part 2: every particle j gathers the deposits of the MOST particles i whose
// ghost node (x_i, y_i, z_i + dz) lies within ghost_r of j and whose 3x3x3
// ghost window holds j's cell, in ascending i, with i's own deposit applied
// after its neighbour deposits (functions_MOST.cuh:286-316 serialised).
__global__ void __launch_bounds__(MAX_BLOCK) k_most_deposit(Seg sg) {
    SEG_PARTICLE(sg);
    const int j = i;
    const sph_params pr = cd.pr;
    const PView S = cd.S;
    const QView Q = cd.Q;
    const Grid g = cd.g;
    const double* dep = cd.snap[4];
    const int* keys = cd.keys;
    const sph_domain& d = g.d;
    const bool self_most = S.MOST_buffer[j] == 1;
    double xj = S.x[j], yj = S.y[j], zj = S.z[j];
    double acc = 0.0, last = 0.0;
    bool any = false, self_done = !self_most;
    int key_j = keys[j];
    int cj_i = key_j % d.NI, cj_j = (key_j / d.NI) % d.NJ, cj_k = key_j / (d.NI * d.NJ);
    double r = pr.most_ghost_r;
    double px = xj, py = yj, pz = zj - pr.most_ghost_dz;
    double mg = 1e-6 + 1e-6 * r;
    bool box_ok = !(pz + r + mg < (double)d.z_min || pz - r - mg > (double)d.z_max ||
                    px + r + mg < (double)d.x_min || px - r - mg > (double)d.x_max ||
                    py + r + mg < (double)d.y_min || py - r - mg > (double)d.y_max);
    if (box_ok && key_j < d.num_cells - 1) {
        int i0 = bin_axis((float)(px - r - mg), d.x_min, d.x_max, d.NI), i1 = bin_axis((float)(px + r + mg), d.x_min, d.x_max, d.NI);
        int j0 = bin_axis((float)(py - r - mg), d.y_min, d.y_max, d.NJ), j1 = bin_axis((float)(py + r + mg), d.y_min, d.y_max, d.NJ);
        int k0 = bin_axis((float)(pz - r - mg), d.z_min, d.z_max, d.NK), k1 = bin_axis((float)(pz + r + mg), d.z_min, d.z_max, d.NK);
        for (int ck = k0; ck <= k1; ck++)
            for (int cj = j0; cj <= j1; cj++)
                for (int ci = i0; ci <= i1; ci++) {
                    int cell = ci + d.NI * cj + d.NI * d.NJ * ck;
                    int s = g.gstr[cell];
                    if (s == -1) continue;
                    int e = g.gend[cell];
                    for (int i = s; i < e; i++) {
                        if (!self_done && i > j) { acc += dep[j]; last = dep[j]; any = true; self_done = true; }
                        if (S.MOST_buffer[i] != 1) continue;
                        double xg = S.x[i], yg = S.y[i], zg = S.z[i] + pr.most_ghost_dz;
                        Cell c = query_cell(d, xg, yg, zg);
                        bool inwin = abs(cj_i - c.i) <= 1 && abs(cj_j - c.j) <= 1 && abs(cj_k - c.k) <= 1;
                        double tdist = dist3<double>(xg, yg, zg, xj, yj, zj) - 1e-20;
                        if (inwin && tdist < pr.most_ghost_r) { acc += dep[i]; last = dep[i]; any = true; }
                        if (i == j) { acc += dep[j]; last = dep[j]; any = true; self_done = true; }
                    }
                }
    }
    if (!self_done) { acc += dep[j]; last = dep[j]; any = true; }
    if (any) { Q.dtemp[j] += acc; S.hmz[j] = last; }
}

// ------------------------------------------------------------------------
// KERNEL_Neumann_boundary3D, functions_BC.cuh:290-446.  Pressures are read
// from a snapshot (pold) taken before the launch.
// The window of the Neumann scan around cell c (ncell nc), in scan order
// (z-major rows of x-cells, the particles of a row in sorted order), as
// chunks of 64 candidates: f(j, count) with lane l holding candidate l of
// the chunk (j < 0: none) and count the chunk's candidates.  Every lane
// makes the same calls.
template <class F>
__device__ __forceinline__ void neumann_window(const Grid& g, const Cell& c, int nc, int lane, F&& f) {
    const int side = 2 * nc + 1, nrows = side * side;
    if (nrows <= 64) {
        // Lane k holds x-row k of the window (z-major, then y): its particle
        // range and the exclusive prefix of the lengths, so the window's
        // candidates are numbered v = 0 .. total-1 in scan order.
        int rs = 0, rl = 0;
        if (lane < nrows) {
            int2 rg;
            if (row_range(g, c.i - nc, c.i + nc, c.j + lane % side - nc, c.k + lane / side - nc, rg)) {
                rs = rg.x;
                rl = rg.y - rg.x;
            }
        }
        int incl = rl;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        const int excl = incl - rl, total = __shfl(incl, 63, 64);
        for (int v0 = 0; v0 < total; v0 += 64) {
            const int v = v0 + lane, vv = min(v, total - 1);
            // the row holding vv: the last row whose prefix is <= vv (never
            // an empty row); every lane takes part in every shuffle
            int lo = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const int cand = lo + step;
                const int pe = __shfl(excl, min(cand, 63), 64);
                if (cand < nrows && pe <= vv) lo = cand;
            }
            const int start = __shfl(rs, lo, 64), pre = __shfl(excl, lo, 64);
            f(v < total ? start + (vv - pre) : -1, min(64, total - v0));
        }
    } else {
        for (int z = -nc; z <= nc; z++)
            for (int y = -nc; y <= nc; y++) {
                int2 rg;
                if (!row_range(g, c.i - nc, c.i + nc, c.j + y, c.k + z, rg)) continue;
                for (int b = rg.x; b < rg.y; b += 64) f(b + lane < rg.y ? b + lane : -1, min(64, rg.y - b));
            }
    }
}

// Neumann pairs of one boundary particle per 64-thread block: the
// candidates of its window that carry weight in the Neumann sums (fluid,
// non-buffer, inside the 2h search range, with the pass's own precision R
// for the distance), in scan order.  COUNT: their number; FILL: the list.
template <class R, bool FILL>
__global__ void __launch_bounds__(64) k_nmp_build(PView A, Grid g, KConst kc, double L2, const int* nm_list,
                                                  int nm_count, int* cnt, const long long* off, int* ent) {
    const int u = blockIdx.x, lane = threadIdx.x;
    if (u >= nm_count) return;
    const int i = nm_list[u];
    const R hi = (R)(2.0 * A.h[i]);
    const R search_range = (R)kc.kappa * hi;
    const double xi = A.x[i], yi = A.y[i], zi = A.z[i];
    const int bt = A.buffer_type[i];
    const double xg = (bt == 1 || bt == -1) ? 2.0 * L2 - xi : xi;
    const Cell c = query_cell(g.d, xg, yi, zi);
    const long long at = FILL ? off[u] : 0;
    int n = 0;
    neumann_window(g, c, A.ncell[i], lane, [&](int j, int count) {
        (void)count;
        bool hit = false;
        if (j >= 0 && A.p_type[j] == 1 && A.buffer_type[j] == 0) {
            R dx = (R)(xg - A.x[j]), dy = (R)(yi - A.y[j]), dz = (R)(zi - A.z[j]);
            R tdist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
            hit = tdist < search_range;
        }
        const unsigned long long m = __ballot(hit);
        if (FILL && hit) ent[at + n + __popcll(m & ((1ULL << lane) - 1ULL))] = j;
        n += __popcll(m);
    });
    if (!FILL && lane == 0) cnt[u] = n;
}
__global__ void __launch_bounds__(BLOCK) k_nmp_widen(int n, const int* a, long long* b) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) b[r] = a[r];
}

template <class R>
__global__ void __launch_bounds__(64) k_neumann(Seg sg, int which, int only_count0) {
    // one 64-lane wave per boundary particle of the case's compact list
    int first_;
    const int cid_ = seg_case(sg, first_);
    const CaseDesc& cd = sg.cd[cid_];
    const int u = first_ >> 6, lane = threadIdx.x;
    if (u >= cd.nm_count) return;
    if (only_count0 && cd.count != 0) return;   // the extra pre-advection call, ISPH_Calc.cuh:88-90
    const KConst kc = cd.kc;
    const sph_params pr = cd.pr;
    const PView A = which ? cd.P : cd.S;
    const Grid g = cd.g;
    const double* pold = cd.snap[3];
    const int i = cd.nm_list[u];
    const sph_domain& d = g.d;
    R hi = (R)(2.0 * A.h[i]);
    R tmp_A = calc_tmpA<R>(kc, hi);
    R search_range = (R)kc.kappa * hi;
    double xi = A.x[i], yi = A.y[i], zi = A.z[i];
    double pres_g = 0;
    int bt = A.buffer_type[i];
    double xg = (bt == 1 || bt == -1) ? 2.0 * pr.L2 - xi : xi;
    double yg = yi, zg = zi;
    Cell c = query_cell(d, xg, yg, zg);
    // the ghost pressure and the gradient only enter the inlet formula
    const bool inlet = (bt == 1 || bt == -1);
    // The window's cells of one x-row are consecutive in cell order, so
    // their particles form one range of the sorted order, visited in the
    // same order as cell by cell.  Lanes test consecutive candidates; the
    // last match in that order wins, as in the serial scan.
    for (int z = -1; z <= 1 && inlet; z++)
        for (int y = -1; y <= 1; y++) {
            int2 rg;
            if (!row_range(g, c.i - 1, c.i + 1, c.j + y, c.k + z, rg)) continue;
            for (int b = rg.x; b < rg.y; b += 64) {
                const int j = b + lane;
                bool hit = false;
                double pj = 0.0;
                if (j < rg.y) {
                    R dx = (R)(xg - A.x[j]), dy = (R)(yg - A.y[j]), dz = (R)(zg - A.z[j]);
                    R dist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
                    hit = dist < (R)pr.neumann_ghost_r;
                    pj = pold[j];
                }
                const unsigned long long m = __ballot(hit);
                if (m) pres_g = __shfl(pj, 63 - __clzll(m), 64);
            }
        }
    const int nt = inlet ? 8 : 5;
    if (sizeof(R) == 4 && cd.nmp_off) {
        // MIXED: each lane adds the terms of its share of the listed pairs
        // (pairs lane, lane + 64, ... in list order), then a fixed butterfly
        // adds the 64 lane sums; lane 0's total is used, so the sums are the
        // same on every call, alone or in a batch
        double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const long long e0 = cd.nmp_off[u], e1 = cd.nmp_off[u + 1];
        const float* const nrec = cd.nm_rec;   // the same values, packed (float-exact positions)
        for (long long e = e0 + lane; e < e1; e += 64) {
            double xj, yj, zj, presj;
            R mj, rhoj, uxj, uyj, uzj;
            if (nrec) {
                const float* C = nrec + (size_t)cd.nmp_cent[e] * NMREC_FLOATS;
                const rec_f4 a = ld_f4(C), b = ld_f4(C + 4);
                xj = a.x; yj = a.y; zj = a.z; mj = (R)a.w; rhoj = (R)b.x; uxj = (R)b.y; uyj = (R)b.z; uzj = (R)b.w;
                presj = ld_d2(C + 8).x;
            } else {
                const int j = cd.nmp_ent[e];
                xj = A.x[j]; yj = A.y[j]; zj = A.z[j]; mj = (R)A.m[j]; rhoj = (R)A.rho[j];
                uxj = (R)A.ux[j]; uyj = (R)A.uy[j]; uzj = (R)A.uz[j];
                presj = pold[j];
            }
            R dx = (R)(xg - xj), dy = (R)(yg - yj), dz = (R)(zg - zj);
            R tdist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
            if (!(tdist < search_range)) continue;
            R twij = kernel_w<R>(kc, tmp_A, hi, tdist);
            s[0] += (double)(mj / rhoj * (R)presj * twij);
            s[1] += (double)(mj / rhoj * uxj * twij);
            s[2] += (double)(mj / rhoj * uyj * twij);
            s[3] += (double)(mj / rhoj * uzj * twij);
            s[4] += (double)(mj / rhoj * twij);
            if (inlet) {
                R tdwij = kernel_dw<R>(kc, tmp_A, hi, tdist);
                R tdwx = dx / tdist * tdwij, tdwy = dy / tdist * tdwij, tdwz = dz / tdist * tdwij;
                R dp = (R)(presj - pres_g);
                s[5] += (double)(mj / rhoj * dp * tdwx);
                s[6] += (double)(mj / rhoj * dp * tdwy);
                s[7] += (double)(mj / rhoj * dp * tdwz);
            }
        }
        for (int off = 32; off > 0; off >>= 1)
            for (int q = 0; q < 8; q++)
                if (q < nt) s[q] += __shfl_xor(s[q], off, 64);
        if (lane != 0) return;
        const double tflt = s[4];
        if (tflt < 1e-6) {
            A.pres[i] = 0.0;
        } else {
            A.pres[i] = s[0] / tflt;
            if (A.p_type[i] == -1) { A.ux[i] = s[1] / tflt; A.uy[i] = s[2] / tflt; A.uz[i] = s[3] / tflt; }
            if (inlet) A.pres[i] = pres_g + (xi - xg) * s[5] + (yi - yg) * s[6] + (zi - zg) * s[7];
        }
        return;
    }
    // FP64: each lane forms the terms of one candidate exactly as the serial
    // scan does; lanes 0..7 then add term k of every contributing candidate
    // in candidate order, so every sum is bit for bit the serial one.
    __shared__ double term[8][64];
    __shared__ int use[64];
    double acc = 0.0;   // lane k < 8 holds sum k
    const bool listed = cd.nmp_off != nullptr;
    // one chunk of candidates v0 .. v0 + 63 of the window's flattened order:
    // lane l holds candidate v0 + l (particle j, or j < 0 for none)
    auto chunk = [&](int j, int count) {
        int ok = 0;
        // only fluid, non-buffer neighbours carry weight (the reference
        // multiplies the others by 0; skipping those exact zeros leaves every
        // sum bit for bit the same); the pair list holds only those
        if (j >= 0 && (listed || (A.p_type[j] == 1 && A.buffer_type[j] == 0))) {
            R dx = (R)(xg - A.x[j]), dy = (R)(yg - A.y[j]), dz = (R)(zg - A.z[j]);
            R tdist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
            if (tdist < search_range) {
                ok = 1;
                R mj = (R)A.m[j], rhoj = (R)A.rho[j];
                double presj = pold[j];
                R twij = kernel_w<R>(kc, tmp_A, hi, tdist);
                term[0][lane] = (double)(mj / rhoj * (R)presj * twij);
                term[1][lane] = (double)(mj / rhoj * (R)A.ux[j] * twij);
                term[2][lane] = (double)(mj / rhoj * (R)A.uy[j] * twij);
                term[3][lane] = (double)(mj / rhoj * (R)A.uz[j] * twij);
                term[4][lane] = (double)(mj / rhoj * twij);
                if (inlet) {
                    R tdwij = kernel_dw<R>(kc, tmp_A, hi, tdist);
                    R tdwx = dx / tdist * tdwij, tdwy = dy / tdist * tdwij, tdwz = dz / tdist * tdwij;
                    R dp = (R)(presj - pres_g);
                    term[5][lane] = (double)(mj / rhoj * dp * tdwx);
                    term[6][lane] = (double)(mj / rhoj * dp * tdwy);
                    term[7][lane] = (double)(mj / rhoj * dp * tdwz);
                }
            }
        }
        use[lane] = ok;
        __syncthreads();   // one wave per block: a wave barrier
        if (lane < nt)
            for (int q = 0; q < count; q++)
                if (use[q]) acc += term[lane][q];
        __syncthreads();
    };
    if (cd.nmp_off) {
        // the pairs of this search (the window's weighted candidates, scan order)
        const long long e0 = cd.nmp_off[u], e1 = cd.nmp_off[u + 1];
        for (long long b = e0; b < e1; b += 64) {
            const long long e = b + lane;
            chunk(e < e1 ? cd.nmp_ent[e] : -1, (int)min((long long)64, e1 - b));
        }
    } else {
        neumann_window(g, c, A.ncell[i], lane, chunk);
    }
    const double tpres = __shfl(acc, 0, 64), tux = __shfl(acc, 1, 64), tuy = __shfl(acc, 2, 64);
    const double tuz = __shfl(acc, 3, 64), tflt_s = __shfl(acc, 4, 64);
    const double grad_px = __shfl(acc, 5, 64), grad_py = __shfl(acc, 6, 64), grad_pz = __shfl(acc, 7, 64);
    if (lane != 0) return;
    if (tflt_s < 1e-6) {
        A.pres[i] = 0.0;
    } else {
        A.pres[i] = tpres / tflt_s;
        if (A.p_type[i] == -1) { A.ux[i] = tux / tflt_s; A.uy[i] = tuy / tflt_s; A.uz[i] = tuz / tflt_s; }
        if (bt == 1 || bt == -1) A.pres[i] = pres_g + (xi - xg) * grad_px + (yi - yg) * grad_py + (zi - zg) * grad_pz;
    }
}

// ------------------------------------------------------------------------
// KERNEL_advection_force3D, functions_PPE.cuh:191-480
// SCAL: some case of the launch transports temperature or concentration or
// uses the Boussinesq buoyancy; without them (the canyon) those terms and
// their registers are compiled out (130 -> fewer VGPRs, more waves per SIMD)
template <class R, bool SCAL, class Src, int W = 1>
__device__ __forceinline__ void advection_body(const CaseDesc& cd, int i, const NList& nl, int li, const Src src,
                                               int lane = 0) {
    const KConst kc = cd.kc;
    const sph_params pr = cd.pr;
    const bool con_solve = SCAL && pr.con_solve, concn_solve = SCAL && pr.concn_solve;
    const bool boussinesq_solve = SCAL && pr.boussinesq_solve;
    const PView S = cd.S;
    const QView Q = cd.Q;
    if (S.i_type[i] > 2) return;
    int ptypei = S.p_type[i];
    if (ptypei == 1000) return;
    if (ptypei <= 0) return;
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double uxi = S.ux[i], uyi = S.uy[i], uzi = S.uz[i];
    double tempi = SCAL ? S.temp[i] : 0.0;
    R hi = (R)S.h[i], mi = (R)S.m[i], rhoi = (R)S.rho[i];
    double hi_d = S.h[i];
    double eli = S.eli[i];
    R tmp_A = calc_tmpA<R>(kc, hi);
    R kci = 0, cpi = 1, concni = 0, diffi = 0;
    if (con_solve) {
        kci = (R)(pr.solid_k * (ptypei == 0 || ptypei == 1000) + pr.air_k * (ptypei == 1 || ptypei == -1));
        cpi = (R)(pr.solid_cp * (ptypei == 0 || ptypei == 1000) + pr.air_cp * (ptypei == 1 || ptypei == -1));
    }
    if (concn_solve) { concni = (R)S.concn[i]; diffi = (R)pr.diff_coef; }
    R visi = (R)(viscosity2(pr, ptypei) + S.vis_t[i] * S.rho[i]);
    double betai = 1.0 / pr.T_ini;
    double tmpx = 0, tmpy = 0, tmpz = 0, tmp_Rc = 0, tmp_Rd = 0;
    double eulerx = 0, eulery = 0, eulerz = 0, eulert = 0, tmp_temp = 0, tmp_flt = 0;
    const R Alpha = (R)pr.av_alpha, Beta = (R)pr.av_beta, soundspeed = (R)pr.av_c;
    const R ome = (R)(1 - eli);
    const bool compact = Src::lds || (sizeof(R) == 4 && src.ok());
    const NLRow row = nl_row(nl, li);
    for (int k = W > 1 ? lane : 0; k < row.cnt; k += W) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_DP)) continue;
        int j = e & NB_IDX;
        double hj_d, visj_d, xj, yj, zj, mj_d, uxj, uyj, uzj, rhoj_d;
        int ptypej;
        if (compact) {
            const float* C = src.at(j);
            const rec_f4 b = Src::f4(src.piece(C, 1));
            ptypej = crec_ptype(b.w);
            if (ptypej >= 1000) continue;
            const rec_f4 a = Src::f4(C);
            const rec_d2 u = Src::d2(src.piece(C, 2)), w = Src::d2(src.piece(C, 3));
            xj = a.x; yj = a.y; zj = a.z; mj_d = a.w;
            rhoj_d = b.x; hj_d = b.y; visj_d = b.z;
            uxj = u.x; uyj = u.y; uzj = w.x;
        } else {
            if constexpr (Src::lds) continue;   // staged records are compact
            double vis_tj, pqj_;
            int itj_;
            rec_aux(cd.rec, j, vis_tj, ptypej, itj_);
            if (ptypej >= 1000) continue;
            rec_pos(cd.rec, j, xj, yj, zj, mj_d, rhoj_d, pqj_);
            rec_vel(cd.rec, j, uxj, uyj, uzj, hj_d);
            visj_d = viscosity2(pr, ptypej) + vis_tj * rhoj_d;
        }
        R dx = (R)(xi - xj), dy = (R)(yi - yj), dz = (R)(zi - zj);
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) + (R)1e-20;
        R twij = kernel_w<R>(kc, tmp_A, hi, tdist);
        R tdwij = kernel_dw<R>(kc, tmp_A, hi, tdist);
        R tdwx = tdwij * dx / tdist;
        R tdwy = tdwij * dy / tdist;
        R tdwz = tdwij * dz / tdist;
        R mj = (R)mj_d, rhoj = (R)rhoj_d;
        R dux = (R)(uxi - uxj), duy = (R)(uyi - uyj), duz = (R)(uzi - uzj);   // u_i - u_j
        double tempj = 0.0, concnj_d = 0.0;
        if (con_solve || concn_solve || boussinesq_solve) {
            if (compact) { tempj = S.temp[src.gidx(j)]; concnj_d = S.concn[src.gidx(j)]; }
            else rec_scalar(cd.rec, j, tempj, concnj_d);
        }
        if (pr.fv_solve == SPH_FV_CLEARY) {
            R visj = (R)visj_d;
            R C_v = dx * tdwx + dy * tdwy + dz * tdwz;
            C_v *= (visi * visj) / (visi + visj);
            C_v *= (R)4 * (mj / (rhoi * rhoj));
            C_v /= tdist;
            C_v /= tdist;
            tmpx += (double)(C_v * dux); tmpy += (double)(C_v * duy); tmpz += (double)(C_v * duz);
        } else if (pr.fv_solve == SPH_FV_MONAGHAN) {
            R uij_xij = dux * dx + duy * dy + duz * dz;
            R visj = (R)visj_d;
            R phi_ij = uij_xij;
            phi_ij /= (tdist * tdist + (R)0.01 * hi * hi);
            R P_ij = (R)(-2.0) * (R)(3 + 2.0) * phi_ij * mj / rhoj;
            R C_v = (R)2.0 * (visi * visj) / (visi + visj - (R)1e-20);
            R nui = C_v / rhoi;
            P_ij *= nui;
            if (ptypei <= 0 || ptypei >= 1000) P_ij = 0;
            if (ptypej <= 0 || ptypei >= 1000) P_ij = 0;
            tmpx += (double)(-(P_ij)*tdwx); tmpy += (double)(-(P_ij)*tdwy); tmpz += (double)(-(P_ij)*tdwz);
        }
        if (eli < 1.0) {
            R Rux = (R)uxi, Ruy = (R)uyi, Ruz = (R)uzi;
            R vx = -dux, vy = -duy, vz = -duz;   // u_j - u_i
            eulerx += (double)((Rux * vx * tdwx + Ruy * vx * tdwy + Ruz * vx * tdwz) * mj / rhoj * ome);
            eulery += (double)((Rux * vy * tdwx + Ruy * vy * tdwy + Ruz * vy * tdwz) * mj / rhoj * ome);
            eulerz += (double)((Rux * vz * tdwx + Ruy * vz * tdwy + Ruz * vz * tdwz) * mj / rhoj * ome);
            if (con_solve)
                eulert -= (double)((R)2.0 * mj / (rhoi + rhoj) * (R)(tempj - tempi) * (dux * tdwx + duy * tdwy + duz * tdwz) * ome);
            if (concn_solve) {
                R dc = (R)(concnj_d - (double)concni);
                tmp_Rd -= (double)((Rux * dc * tdwx + Ruy * dc * tdwy + Ruz * dc * tdwz) * mj / rhoj * ome);
            }
        }
        if (con_solve) {
            R kcj = (R)(pr.solid_k * (ptypej == 0 || ptypej == 1000) + pr.air_k * (ptypej == 1 || ptypej == -1));
            R rhoj_temp = (ptypej == 1 || ptypej == -1) ? rhoj : (R)pr.solid_rho;
            R mj_temp = mj * rhoj_temp / rhoj;
            R sum_con_H = (R)4.0 * mj_temp * kcj * kci * (R)(tempi - tempj) * tdwij;
            sum_con_H /= (tdist + (R)1e-10) * rhoi * rhoj_temp * (kci + kcj);
            tmp_Rc += (double)sum_con_H;
        }
        if (concn_solve) {
            R concnj = (R)concnj_d;
            R diffj = (R)pr.diff_coef;
            R tmprd = mi * (diffi * rhoi + diffj * rhoj);
            tmprd *= (dx * tdwx + dy * tdwy + dz * tdwz);
            tmprd /= (rhoi * rhoj * (tdist * tdist + (R)0.01 * hi * hi));
            R tmprd1 = tmprd * (concni - concnj) * (R)(ptypei == ptypej);
            tmp_Rd += (double)tmprd1;
        }
        if (boussinesq_solve) {
            tmp_temp += (double)((mj / rhoj) * (R)tempj * twij);
            tmp_flt += (double)((mj / rhoj) * twij);
        }
        if (pr.artificial_viscosity) {
            R uij_xij = dux * dx + duy * dy + duz * dz;
            if (uij_xij < 0) {
                R h_ij = (R)((hi_d + hj_d) * 0.5);
                R phi_ij = h_ij * uij_xij;
                phi_ij /= (tdist * tdist + (R)0.01 * h_ij * h_ij);
                R P_ij = mi * phi_ij * (-Alpha * soundspeed + Beta * phi_ij);
                P_ij /= (rhoi + rhoj);
                P_ij *= (R)0.5;
                tmpx += (double)(-(P_ij)*tdwx); tmpy += (double)(-(P_ij)*tdwy); tmpz += (double)(-(P_ij)*tdwz);
            }
        }
    }
    if constexpr (W > 1) {
        wave_fold(tmpx); wave_fold(tmpy); wave_fold(tmpz); wave_fold(tmp_Rc); wave_fold(tmp_Rd);
        wave_fold(eulerx); wave_fold(eulery); wave_fold(eulerz); wave_fold(eulert);
        wave_fold(tmp_temp); wave_fold(tmp_flt);
        if (lane != 0) return;
    }
    if (pr.fg_solve) tmpz += -pr.gravity;
    if ((boussinesq_solve) && (ptypei > 0) && (ptypei < 1000))
        tmpz = betai * pr.gravity * (tmp_temp / tmp_flt - pr.tempb);
    Q.ftotalx[i] = tmpx - eulerx;
    Q.ftotaly[i] = tmpy - eulery;
    Q.ftotalz[i] = tmpz - eulerz;
    if (con_solve) Q.dtemp[i] += tmp_Rc / (double)cpi - eulert;
    Q.ftotal[i] = sqrt(tmpx * tmpx + tmpy * tmpy + tmpz * tmpz);
    if (concn_solve) Q.dconcn[i] = tmp_Rd;
}
template <class R, bool SCAL>
__global__ void __launch_bounds__(MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(SCAL ? 1 : 4)))
k_advection(Seg sg) {
    SEG_PARTICLE_LIGHT(sg);
    advection_body<R, SCAL>(cd, i, cd.nl, i, GSrc<CREC_ADV_FLOATS>{sizeof(R) == 4 ? &cd.crec_adv : nullptr});
}
template <bool SCAL>
__global__ void __launch_bounds__(WAVE_BLOCK) k_advection_heavy(Seg sg) {
    SEG_HEAVY(sg);
    advection_body<float, SCAL, GSrc<CREC_ADV_FLOATS>, 64>(cd, i, cd.nl, i, GSrc<CREC_ADV_FLOATS>{&cd.crec_adv}, lane_);
}
template <bool SCAL>
__global__ void __launch_bounds__(TILE_MAX) k_advection_tile(Seg sg) {
    TILE_PARTICLE(sg, cd.crec_adv, CREC_ADV_FLOATS);
    if (chunk_.y >= 0)
        advection_body<float, SCAL>(cd, i, cd.tile.nl, lane_, LSrc<CREC_ADV_FLOATS>{(const float*)tile_lds_, chunk_.y,
                                                     cd.tile.halo + chunk_.x,
                                                         cd.tile.self[lane_]});
    else advection_body<float, SCAL>(cd, i, cd.tile.nl, lane_, GSrc<CREC_ADV_FLOATS>{&cd.crec_adv});
}

// ------------------------------------------------------------------------
// KERNEL_clc_projection, functions_TIME_ISPH.cuh:2-112 (predictor)
__global__ void __launch_bounds__(MAX_BLOCK) k_projection(Seg sg) {
    SEG_PARTICLE(sg);
    const sph_params pr = cd.pr;
    if (pr.time_type != SPH_PRE_COR) return;   // ISPH_Calc.cuh:105
    const PView S = cd.S;
    const P2View S2 = cd.S2;
    const QView Q = cd.Q;
    const double dt = cd.dt, time = cd.time;
    if (S.i_type[i] >= 3) return;
    if (S.p_type[i] >= 1000) return;
    int buffer_type = S.buffer_type[i], MOST = S.MOST_buffer[i], p_typei = S.p_type[i];
    if (p_typei == 9) {
        S2.x0[i] = S.x[i]; S2.y0[i] = S.y[i]; S2.z0[i] = S.z[i];
        S.ux[i] = 1.0; S.uy[i] = 0; S.uz[i] = 0;
        S2.ux0[i] = 1.0; S2.uy0[i] = 0; S2.uz0[i] = 0;
    } else {
        double tux0 = S.ux[i], tuy0 = S.uy[i], tuz0 = S.uz[i];
        double tuxp, tuyp, tuzp;
        if (p_typei > 0) {
            double b = (double)(buffer_type == 0 || MOST == 1);
            double tdux = Q.ftotalx[i] * b + S.fmx[i];
            double tduy = Q.ftotaly[i] * b + S.fmy[i];
            double tduz = Q.ftotalz[i] * b;
            tuxp = tux0 + tdux * (dt);
            tuyp = tuy0 + tduy * (dt);
            tuzp = tuz0 + tduz * (dt);
        } else {
            tuxp = tux0; tuyp = tuy0; tuzp = tuz0;
        }
        S.ux[i] = tuxp; S.uy[i] = tuyp; S.uz[i] = tuzp;
        S2.x0[i] = S.x[i]; S2.y0[i] = S.y[i]; S2.z0[i] = S.z[i];
        S2.ux0[i] = tux0; S2.uy0[i] = tuy0; S2.uz0[i] = tuz0;
    }
    if (time == 0) S2.rho_ref[i] = S.rho[i];
    if (pr.con_solve == 1) S2.temp0[i] = S.temp[i];
    if (pr.concn_solve == 1) {
        double tconcn = S.concn[i];
        S2.concn0[i] = tconcn;
        S.concn[i] = tconcn + Q.dconcn[i] * (dt);
    }
}

// ------------------------------------------------------------------------
// KERNEL_periodic_boundary3D, functions_BC.cuh:1015-1154 (reads the snapshot
// arrays; last match inside the first cell that has one).
__global__ void __launch_bounds__(MAX_BLOCK) k_periodic(Seg sg, int which) {
    SEG_PARTICLE(sg);
    const sph_params pr = cd.pr;
    const PView A = which ? cd.P : cd.S;
    const Grid g = cd.g;
    const double *oux = cd.snap[0], *ouy = cd.snap[1], *ouz = cd.snap[2], *opres = cd.snap[3], *otemp = cd.snap[4];
    int bt = A.buffer_type[i];
    if (bt == 0) return;
    if (A.p_type[i] >= 1000) return;
    double xi = A.x[i], yi = A.y[i], zi = A.z[i];
    double xg, yg, zg;
    if (bt == 3) { xg = xi; yg = yi + pr.L_right - pr.L_left; zg = zi; }
    else if (bt == 4) { xg = xi; yg = yi - pr.L_right + pr.L_left; zg = zi; }
    else return;
    const sph_domain& d = g.d;
    int nc = A.ncell[i];
    Cell c = query_cell(d, xg, yg, zg);
    int found = -1;
    // only cells inside the bin range of the copy_tol ball can hold a match
    // (the first such cell in window order is unchanged)
    Cell lo, hi;
    ball_bins(d, xg, yg, zg, pr.copy_tol, lo, hi);
    for (int z = max(-nc, lo.k - c.k); z <= min(nc, hi.k - c.k) && found < 0; z++)
        for (int y = max(-nc, lo.j - c.j); y <= min(nc, hi.j - c.j) && found < 0; y++)
            for (int x = max(-nc, lo.i - c.i); x <= min(nc, hi.i - c.i) && found < 0; x++) {
                if (!cell_ok(d, c.i + x, c.j + y, c.k + z)) continue;
                int cell = (c.i + x) + d.NI * (c.j + y) + d.NI * d.NJ * (c.k + z);
                int s = g.gstr[cell];
                if (s == -1) continue;
                int e = g.gend[cell];
                for (int j = s; j < e; j++) {
                    double tdist = dist3<double>(xg, yg, zg, A.x[j], A.y[j], A.z[j]) - 1e-20;
                    if (tdist < pr.copy_tol) found = j;
                }
            }
    if (found >= 0) {
        A.ux[i] = oux[found]; A.uy[i] = ouy[found]; A.uz[i] = ouz[found];
        A.pres[i] = opres[found]; A.temp[i] = otemp[found];
    } else {
        A.ux[i] = 0; A.uy[i] = 0; A.uz[i] = 0; A.pres[i] = 0; A.temp[i] = 0;
    }
}

// ------------------------------------------------------------------------
// KERNEL_freeslip_boundary3D, functions_BC.cuh:1156-1271
template <class R>
__global__ void __launch_bounds__(MAX_BLOCK) k_freeslip(Seg sg, int which) {
    SEG_PARTICLE(sg);
    const KConst kc = cd.kc;
    const PView A = which ? cd.P : cd.S;
    const NList nl = cd.nl;
    if (A.p_type[i] != -1) return;
    R hi = (R)A.h[i];
    R tmp_A = calc_tmpA<R>(kc, hi);
    double xg = A.x[i], yg = A.y[i], zg = A.z[i];
    double tux = 0, tuy = 0, tuz = 0, tpres = 0, ttemp = 0, tflt_s = 0;
    NLRow row = nl_row(nl, i);
    for (int k = 0; k < row.cnt; k++) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_HM)) continue;
        int j = e & NB_IDX;
        // fluid neighbours only (the reference weights the others by an
        // exact 0, which leaves the sums unchanged)
        if (A.p_type[j] != 1) continue;
        R dx = (R)(xg - A.x[j]), dy = (R)(yg - A.y[j]), dz = (R)(zg - A.z[j]);
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
        R mj = (R)A.m[j], rhoj = (R)A.rho[j];
        R twij = kernel_w<R>(kc, tmp_A, hi, tdist);
        tux += (double)((R)A.ux[j] * (mj / rhoj) * twij);
        tuy += (double)((R)A.uy[j] * (mj / rhoj) * twij);
        tuz += (double)((R)A.uz[j] * (mj / rhoj) * twij);
        tpres += (double)((R)A.pres[j] * (mj / rhoj) * twij);
        ttemp += (double)((R)A.temp[j] * (mj / rhoj) * twij);
        tflt_s += (double)((mj / rhoj) * twij);
    }
    if (tflt_s > 1e-5) {
        A.ux[i] = tux / tflt_s; A.uy[i] = tuy / tflt_s; A.uz[i] = tuz / tflt_s;
        A.pres[i] = tpres / tflt_s; A.temp[i] = ttemp / tflt_s;
    }
}

// The same sums in MIXED with a 64-lane wave per boundary particle (the
// lid's particles are a few per cent of the case, one thread each left
// the device nearly idle): lane l adds entries l, l + 64, ... of the list,
// then a fixed butterfly adds the lane sums, so a case gives the same bits
// alone and in a batch.
__global__ void __launch_bounds__(WAVE_BLOCK) k_freeslip_wave(Seg sg, int which) {
    int first_;
    const int cid_ = seg_case(sg, first_);
    const CaseDesc& cd = sg.cd[cid_];
    const int u = (first_ + (int)threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (u >= cd.nm_count) return;
    const PView A = which ? cd.P : cd.S;
    const int i = cd.nm_list[u];
    if (A.p_type[i] != -1) return;
    const KConst kc = cd.kc;
    const NList nl = cd.nl;
    float hi = (float)A.h[i];
    float tmp_A = calc_tmpA<float>(kc, hi);
    double xg = A.x[i], yg = A.y[i], zg = A.z[i];
    double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const NLRow row = nl_row(nl, i);
    for (int k = lane; k < row.cnt; k += 64) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_HM)) continue;
        int j = e & NB_IDX;
        if (A.p_type[j] != 1) continue;
        float dx = (float)(xg - A.x[j]), dy = (float)(yg - A.y[j]), dz = (float)(zg - A.z[j]);
        float tdist = sqrt(dx * dx + dy * dy + dz * dz) - 1e-20f;
        float mj = (float)A.m[j], rhoj = (float)A.rho[j];
        float twij = kernel_w<float>(kc, tmp_A, hi, tdist);
        t[0] += (double)((float)A.ux[j] * (mj / rhoj) * twij);
        t[1] += (double)((float)A.uy[j] * (mj / rhoj) * twij);
        t[2] += (double)((float)A.uz[j] * (mj / rhoj) * twij);
        t[3] += (double)((float)A.pres[j] * (mj / rhoj) * twij);
        t[4] += (double)((float)A.temp[j] * (mj / rhoj) * twij);
        t[5] += (double)((mj / rhoj) * twij);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 6; q++) t[q] += __shfl_xor(t[q], off, 64);
    if (lane != 0) return;
    if (t[5] > 1e-5) {
        A.ux[i] = t[0] / t[5]; A.uy[i] = t[1] / t[5]; A.uz[i] = t[2] / t[5];
        A.pres[i] = t[3] / t[5]; A.temp[i] = t[4] / t[5];
    }
}

// ------------------------------------------------------------------------
// KERNEL_donothing_buffer3D, functions_BC.cuh:673-761 (snapshot reads)
__global__ void __launch_bounds__(MAX_BLOCK) k_outlet(Seg sg, int which) {
    SEG_PARTICLE(sg);
    const sph_params pr = cd.pr;
    const PView A = which ? cd.P : cd.S;
    const Grid g = cd.g;
    const double *oux = cd.snap[0], *ouy = cd.snap[1], *ouz = cd.snap[2], *opres = cd.snap[3], *otemp = cd.snap[4];
    if (A.buffer_type[i] != 2) return;
    const sph_domain& d = g.d;
    double xi = A.x[i], yi = A.y[i], zi = A.z[i];
    Cell c = query_cell(d, xi, yi, zi);
    int nc = 2 * A.ncell[i];
    double xg = pr.L3 - pr.outlet_copy_dx;
    int found = -1;
    // A match lies within copy_tol of (xg, yi, zi); binning is monotone, so
    // only window cells inside the bin range of that ball can hold one.
    // Skipping the others leaves the last match in window order unchanged.
    Cell lo, hi;
    ball_bins(d, xg, yi, zi, pr.copy_tol, lo, hi);
    for (int z = max(-nc, lo.k - c.k); z <= min(nc, hi.k - c.k); z++)
        for (int y = max(-nc, lo.j - c.j); y <= min(nc, hi.j - c.j); y++)
            for (int x = max(-nc, lo.i - c.i); x <= min(nc, hi.i - c.i); x++) {
                if (!cell_ok(d, c.i + x, c.j + y, c.k + z)) continue;
                int cell = (c.i + x) + d.NI * (c.j + y) + d.NI * d.NJ * (c.k + z);
                int s = g.gstr[cell];
                if (s == -1) continue;
                int e = g.gend[cell];
                for (int j = s; j < e; j++) {
                    double tdist = dist3<double>(xg, yi, zi, A.x[j], A.y[j], A.z[j]) - 1e-20;
                    if (A.p_type[j] < 1000 && tdist < pr.copy_tol) found = j;
                }
            }
    if (found >= 0) {
        A.pres[i] = opres[found]; A.ux[i] = oux[found]; A.uy[i] = ouy[found]; A.uz[i] = ouz[found];
        A.temp[i] = otemp[found];
    } else {
        A.pres[i] = 0; A.ux[i] = 0; A.uy[i] = 0; A.uz[i] = 0; A.temp[i] = 0;
    }
}

// ------------------------------------------------------------------------
// KERNEL_pressureforce3D, functions_PPE.cuh:482-647.  row: i's neighbour
// entries; src: where their records are (compact records in MIXED).
template <class R, class Src, int W = 1>
__device__ __forceinline__ void pressure_force_body(const CaseDesc& cd, int i, const NList& nl, int li,
                                                    const Src src, int lane = 0) {
    const KConst kc = cd.kc;
    const sph_params pr = cd.pr;
    const PView S = cd.S;
    const QView Q = cd.Q;
    if (S.i_type[i] > 2) return;
    if (S.p_type[i] >= 1000) return;
    if (S.p_type[i] <= 0) return;
    if (S.buffer_type[i] != 0) return;
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double hi_d = S.h[i];
    double pi_d = S.pres[i] - S.pres0[i];
    R pi = (R)pi_d, rhoi = (R)S.rho[i];
    double tmpx = 0, tmpy = 0, tmpz = 0;
    // the pair's kernel normalisation and derivative factor depend only on
    // h_ij: kept while consecutive pairs share it
    R h_c = (R)-1, A_c = (R)0, f_c = (R)0;
    const NLRow row = nl_row(nl, li);
    if (sizeof(R) == 4 && src.ok()) {
        // the same sums from the compact float records
        for (int k = W > 1 ? lane : 0; k < row.cnt; k += W) {
            unsigned e = NB_AT(row, k);
            if (!(e & F_DM)) continue;
            const float* C = src.at(e & NB_IDX);
            const rec_f4 b = Src::f4(src.piece(C, 1));
            if (__float_as_int(b.w) >= 1000) continue;
            const rec_f4 a = Src::f4(C);
            R dx = (R)(xi - (double)a.x), dy = (R)(yi - (double)a.y), dz = (R)(zi - (double)a.z);
            R tdist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
            R hij = (R)(0.5 * (hi_d + (double)b.z));
            if (hij != h_c) { h_c = hij; A_c = calc_tmpA<R>(kc, hij); f_c = dw_factor<R>(kc, hij); }
            R tdwij = kernel_dw_f<R>(kc, A_c, hij, f_c, tdist);
            R tdwx = tdwij * dx / tdist;
            R tdwy = tdwij * dy / tdist;
            R tdwz = tdwij * dz / tdist;
            R mj = (R)a.w, rhoj = (R)b.x, pj = (R)b.y;
            if (pr.fp_solve) {
                const R Pb = (R)0.0;
                R C_p = (R)(-2.0) * mj / (rhoi * rhoj) * ((pi + Pb) * rhoj + (pj + Pb) * rhoi) / (rhoi + rhoj);
                tmpx += (double)(C_p * tdwx); tmpy += (double)(C_p * tdwy); tmpz += (double)(C_p * tdwz);
            }
        }
        if constexpr (W > 1) {
            wave_fold(tmpx); wave_fold(tmpy); wave_fold(tmpz);
            if (lane != 0) return;
        }
        Q.fpx[i] = tmpx; Q.fpy[i] = tmpy; Q.fpz[i] = tmpz;
        return;
    }
    if constexpr (Src::lds) return;   // staged records are compact
    for (int k = W > 1 ? lane : 0; k < row.cnt; k += W) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_DM)) continue;
        int j = e & NB_IDX;
        if (S.p_type[j] >= 1000) continue;
        double hj, pqj, xj, yj, zj, mj_d, rhoj_d;
        rec_pos(cd.rec, j, xj, yj, zj, mj_d, rhoj_d, pqj);
        hj = ld_d2(rec_at(cd.rec, j, 4)).y;
        R dx = (R)(xi - xj), dy = (R)(yi - yj), dz = (R)(zi - zj);
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) - (R)1e-20;
        R hij = (R)(0.5 * (hi_d + hj));
        if (hij != h_c) { h_c = hij; A_c = calc_tmpA<R>(kc, hij); f_c = dw_factor<R>(kc, hij); }
        R tdwij = kernel_dw_f<R>(kc, A_c, hij, f_c, tdist);
        R tdwx = tdwij * dx / tdist;
        R tdwy = tdwij * dy / tdist;
        R tdwz = tdwij * dz / tdist;
        R mj = (R)mj_d, rhoj = (R)rhoj_d;
        R pj = (R)pqj;
        if (pr.fp_solve) {
            const R Pb = (R)0.0;
            R C_p = (R)(-2.0) * mj / (rhoi * rhoj) * ((pi + Pb) * rhoj + (pj + Pb) * rhoi) / (rhoi + rhoj);
            tmpx += (double)(C_p * tdwx); tmpy += (double)(C_p * tdwy); tmpz += (double)(C_p * tdwz);
        }
    }
    if constexpr (W > 1) {
        wave_fold(tmpx); wave_fold(tmpy); wave_fold(tmpz);
        if (lane != 0) return;
    }
    Q.fpx[i] = tmpx; Q.fpy[i] = tmpy; Q.fpz[i] = tmpz;
}
template <class R>
__global__ void __launch_bounds__(MAX_BLOCK) k_pressure_force(Seg sg) {
    SEG_PARTICLE_LIGHT(sg);
    pressure_force_body<R>(cd, i, cd.nl, i, GSrc<CREC_FLOATS>{sizeof(R) == 4 ? &cd.crec : nullptr});
}
__global__ void __launch_bounds__(WAVE_BLOCK) k_pressure_force_heavy(Seg sg) {
    SEG_HEAVY(sg);
    pressure_force_body<float, GSrc<CREC_FLOATS>, 64>(cd, i, cd.nl, i, GSrc<CREC_FLOATS>{&cd.crec}, lane_);
}
__global__ void __launch_bounds__(TILE_MAX) k_pressure_force_tile(Seg sg) {
    TILE_PARTICLE(sg, cd.crec, CREC_FLOATS);
    if (chunk_.y >= 0)
        pressure_force_body<float>(cd, i, cd.tile.nl, lane_, LSrc<CREC_FLOATS>{(const float*)tile_lds_, chunk_.y,
                                                     cd.tile.halo + chunk_.x,
                                                      cd.tile.self[lane_]});
    else pressure_force_body<float>(cd, i, cd.tile.nl, lane_, GSrc<CREC_FLOATS>{&cd.crec});
}

// ------------------------------------------------------------------------
// IBM_force_interpolation3D, functions_IBM.cuh:342-490
template <class R>
__global__ void __launch_bounds__(MAX_BLOCK) k_ibm_force(Seg sg) {
    SEG_PARTICLE(sg);
    const KConst kc = cd.kc;
    const sph_params pr = cd.pr;
    const PView S = cd.S;
    const QView Q = cd.Q;
    const NList nl = cd.nl;
    const double dt = cd.dt;
    if (S.p_type[i] < 1000) return;
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    R tmp_h = (R)(0.5 * S.h[i]);
    R tmp_A = calc_tmpA<R>(kc, tmp_h);
    double tmpux = 0, tmpuy = 0, tmpuz = 0, filt = 0;
    NLRow row = nl_row(nl, i);
    for (int k = 0; k < row.cnt; k++) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_BH)) continue;
        int j = e & NB_IDX;
        if (S.i_type[j] == 4 || S.p_type[j] != 1) continue;
        R dx = (R)(xi - S.x[j]), dy = (R)(yi - S.y[j]), dz = (R)(zi - S.z[j]);
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) + (R)1e-20;
        R twij = kernel_w<R>(kc, tmp_A, tmp_h, tdist);
        R tmprho = (R)S.m[j] / (R)S.rho[j];
        tmpux += (double)(((R)S.ux[j]) * tmprho * twij);
        tmpuy += (double)(((R)S.uy[j]) * tmprho * twij);
        tmpuz += (double)(((R)S.uz[j]) * tmprho * twij);
        filt += (double)(tmprho * twij);
    }
    S.concentration[i] = filt;
    double ufx = tmpux / (filt + 1e-10), ufy = tmpuy / (filt + 1e-10), ufz = tmpuz / (filt + 1e-10);
    double nx = Q.nx[i], ny = Q.ny[i], nz = Q.nz[i];
    double nmag = sqrt(nx * nx + ny * ny + nz * nz) + 1e-20;
    nx = nx / nmag; ny = ny / nmag; nz = nz / nmag;
    double un = ufx * nx + ufy * ny + ufz * nz;
    double utx = ufx - un * nx, uty = ufy - un * ny, utz = ufz - un * nz;
    double Ubtx = pr.ibm_ut_factor * interp_ux_caseH(zi, 1.0);
    double Ubty = 0.0, Ubtz = 0.0;
    double beta_t = 1.0;
    double corrx = (-un) * nx + beta_t * (Ubtx - utx);
    double corry = (-un) * ny + beta_t * (Ubty - uty);
    double corrz = (-un) * nz + beta_t * (Ubtz - utz);
    S.fbx[i] = pr.rho_air * corrx / dt;
    S.fby[i] = pr.rho_air * corry / dt;
    S.fbz[i] = pr.rho_air * corrz / dt;
}

// The marker force in MIXED, a 64-lane wave per marker (markers are in the
// boundary list): lane sums over entries lane, lane + 64, ..., then a fixed
// butterfly; lane 0 forms the force from the sums as the serial kernel does.
__global__ void __launch_bounds__(WAVE_BLOCK) k_ibm_force_wave(Seg sg) {
    int first_;
    const int cid_ = seg_case(sg, first_);
    const CaseDesc& cd = sg.cd[cid_];
    const int u = (first_ + (int)threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (u >= cd.nm_count) return;
    const PView S = cd.S;
    const int i = cd.nm_list[u];
    if (S.p_type[i] < 1000) return;
    const KConst kc = cd.kc;
    const sph_params pr = cd.pr;
    const QView Q = cd.Q;
    const NList nl = cd.nl;
    const double dt = cd.dt;
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    float tmp_h = (float)(0.5 * S.h[i]);
    float tmp_A = calc_tmpA<float>(kc, tmp_h);
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    const NLRow row = nl_row(nl, i);
    for (int k = lane; k < row.cnt; k += 64) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_BH)) continue;
        int j = e & NB_IDX;
        if (S.i_type[j] == 4 || S.p_type[j] != 1) continue;
        float dx = (float)(xi - S.x[j]), dy = (float)(yi - S.y[j]), dz = (float)(zi - S.z[j]);
        float tdist = sqrt(dx * dx + dy * dy + dz * dz) + 1e-20f;
        float twij = kernel_w<float>(kc, tmp_A, tmp_h, tdist);
        float tmprho = (float)S.m[j] / (float)S.rho[j];
        t[0] += (double)(((float)S.ux[j]) * tmprho * twij);
        t[1] += (double)(((float)S.uy[j]) * tmprho * twij);
        t[2] += (double)(((float)S.uz[j]) * tmprho * twij);
        t[3] += (double)(tmprho * twij);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 4; q++) t[q] += __shfl_xor(t[q], off, 64);
    if (lane != 0) return;
    const double filt = t[3];
    S.concentration[i] = filt;
    double ufx = t[0] / (filt + 1e-10), ufy = t[1] / (filt + 1e-10), ufz = t[2] / (filt + 1e-10);
    double nx = Q.nx[i], ny = Q.ny[i], nz = Q.nz[i];
    double nmag = sqrt(nx * nx + ny * ny + nz * nz) + 1e-20;
    nx = nx / nmag; ny = ny / nmag; nz = nz / nmag;
    double un = ufx * nx + ufy * ny + ufz * nz;
    double utx = ufx - un * nx, uty = ufy - un * ny, utz = ufz - un * nz;
    double Ubtx = pr.ibm_ut_factor * interp_ux_caseH(zi, 1.0);
    double Ubty = 0.0, Ubtz = 0.0;
    double beta_t = 1.0;
    double corrx = (-un) * nx + beta_t * (Ubtx - utx);
    double corry = (-un) * ny + beta_t * (Ubty - uty);
    double corrz = (-un) * nz + beta_t * (Ubtz - utz);
    S.fbx[i] = pr.rho_air * corrx / dt;
    S.fby[i] = pr.rho_air * corry / dt;
    S.fbz[i] = pr.rho_air * corrz / dt;
}

// IBM_spreading_interpolation3D, functions_IBM.cuh:492-613
template <class R>
__global__ void __launch_bounds__(MAX_BLOCK) k_ibm_spread(Seg sg) {
    SEG_PARTICLE(sg);
    const KConst kc = cd.kc;
    const PView S = cd.S;
    const NList nl = cd.nl;
    if (S.p_type[i] != 1) return;
    if (S.buffer_type[i] == 2) return;
    if (cd.ibm_near && !cd.ibm_near[i]) {   // no marker in the list: what the sums below give
        S.fbx[i] = 0.0; S.fby[i] = 0.0; S.fbz[i] = 0.0; S.concentration[i] = 0.0;
        return;
    }
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    R tmp_h = (R)(0.5 * S.h[i]);
    R tmp_A = calc_tmpA<R>(kc, tmp_h);
    double filt = 0, fb_x = 0, fb_y = 0, fb_z = 0;
    NLRow row = nl_row(nl, i);
    for (int k = 0; k < row.cnt; k++) {
        unsigned e = NB_AT(row, k);
        if (!(e & F_BH)) continue;
        int j = e & NB_IDX;
        if (S.i_type[j] == 4 || S.p_type[j] != 1000) continue;
        R dx = (R)(xi - S.x[j]), dy = (R)(yi - S.y[j]), dz = (R)(zi - S.z[j]);
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) + (R)1e-20;
        R twij = kernel_w<R>(kc, tmp_A, tmp_h, tdist);
        R tmprho = (R)S.m[j] / (R)S.rho[j];
        fb_x += (double)((R)S.fbx[j] * twij * tmprho);
        fb_y += (double)((R)S.fby[j] * twij * tmprho);
        fb_z += (double)((R)S.fbz[j] * twij * tmprho);
        filt += (double)(tmprho * twij);
    }
    if (filt < 1e-12) {
        S.fbx[i] = 0.0; S.fby[i] = 0.0; S.fbz[i] = 0.0; S.concentration[i] = 0.0;
        return;
    }
    S.fbx[i] = fb_x / (filt + 1e-10);
    S.fby[i] = fb_y / (filt + 1e-10);
    S.fbz[i] = fb_z / (filt + 1e-10);
    S.concentration[i] = filt;
}

// ------------------------------------------------------------------------
// KERNEL_time_update_projection, functions_TIME_ISPH.cuh:115-296 (corrector)
// and IBM_corrector, functions_IBM.cuh:615-725 (only its copies are live).
__global__ void __launch_bounds__(MAX_BLOCK) k_time_update(Seg sg) {
    SEG_PARTICLE(sg);
    const sph_params pr = cd.pr;
    const PView S = cd.S, P = cd.P;
    const P2View S2 = cd.S2, P2 = cd.P2;
    const QView Q = cd.Q;
    const double dt = cd.dt;
    const int count = cd.count;
    if (S.p_type[i] >= 1000) {   // IBM_corrector (no i_type filter there)
        P.fbx[i] = S.fbx[i]; P.fby[i] = S.fby[i]; P.fbz[i] = S.fbz[i];
        P.concentration[i] = S.concentration[i]; P.vol[i] = S.vol[i]; P.pres[i] = S.pres[i];
        return;
    }
    if (S.i_type[i] >= 3) return;
    int buffer_type = S.buffer_type[i], p_type_i = S.p_type[i];
    if (p_type_i == 9) {
        P.ux[i] = S.ux[i]; P.uy[i] = S.uy[i]; P.uz[i] = S.uz[i];
    } else {
        double tux0, tuy0, tuz0, dux_dt, duy_dt, duz_dt;
        if (p_type_i > 0) {
            tux0 = S.ux[i]; tuy0 = S.uy[i]; tuz0 = S.uz[i];
            if (buffer_type == 1 || buffer_type == -1) { tux0 = S2.ux0[i]; tuy0 = S2.uy0[i]; tuz0 = S2.uz0[i]; }
            double b = (double)(buffer_type == 0);
            dux_dt = Q.fpx[i] * b + S.fbx[i] / S.rho[i] + S.fmx[i];
            duy_dt = Q.fpy[i] * b + S.fby[i] / S.rho[i] + S.fmy[i];
            duz_dt = Q.fpz[i] * b + S.fbz[i] / S.rho[i];
        } else {
            tux0 = S2.ux0[i]; tuy0 = S2.uy0[i]; tuz0 = S2.uz0[i];
            dux_dt = duy_dt = duz_dt = 0.0;
        }
        double uxc = tux0 + dux_dt * (dt), uyc = tuy0 + duy_dt * (dt), uzc = tuz0 + duz_dt * (dt);
        if ((uxc * uxc + uyc * uyc + uzc * uzc) >= pr.u_limit * pr.u_limit) { uxc = tux0; uyc = tuy0; uzc = tuz0; }
        P.ux[i] = uxc; P.uy[i] = uyc; P.uz[i] = uzc;
        if (cd.xnew[0]) {
            // functions_TIME_ISPH.cuh:193-195 (computed there, written here)
            const double mv = (double)(p_type_i > 0) * S.eli[i];
            cd.xnew[0][i] = S.x[i] + (S2.ux0[i] + uxc) / 2.0 * (dt) * mv;
            cd.xnew[1][i] = S.y[i] + (S2.uy0[i] + uyc) / 2.0 * (dt) * mv;
            cd.xnew[2][i] = S.z[i] + (S2.uz0[i] + uzc) / 2.0 * (dt) * mv;
        }
    }
    if (pr.con_solve == 1) P.temp[i] = S2.temp0[i] + Q.dtemp[i] * dt * (buffer_type == 0);
    else { P.enthalpy[i] = S.enthalpy[i]; P.temp[i] = S.temp[i]; }
    if (pr.concn_solve == 1) {
        double cc = S2.concn0[i] + Q.dconcn[i] * dt;
        P.concn[i] = cc < 0.0 ? 0.0 : cc;
    } else P.concn[i] = S.concn[i];
    P.pres[i] = S.pres[i]; P.flt_s[i] = S.flt_s[i];
    P.k_turb[i] = S.k_turb[i]; P.e_turb[i] = S.e_turb[i];
    P2.rho_ref[i] = S2.rho_ref[i]; P.vis_t[i] = S.vis_t[i];
    P.fbx[i] = S.fbx[i]; P.fby[i] = S.fby[i]; P.fbz[i] = S.fbz[i];
    P.fmx[i] = S.fmx[i]; P.fmy[i] = S.fmy[i]; P.hmz[i] = S.hmz[i];
    P.float1[i] = S.float1[i]; P.float2[i] = S.float2[i]; P.float3[i] = S.float3[i];
    P.concentration[i] = S.concentration[i];
    if (P.source[i] == 1) P.concn[i] += 0.5 * dt;
    if (P.p_type[i] == 0) P.temp[i] = wall_temperature(pr, dt, count);
}

// ------------------------------------------------------------------------
// kernel_copy_max + 3 x DeviceReduce::Max, functions_OUTPUT.cuh:19-33
__device__ __forceinline__ unsigned f2ord(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__global__ void __launch_bounds__(BLOCK) k_stats(int n, PView P, P2View S2, QView Q, unsigned* out) {
    __shared__ unsigned s[3][BLOCK];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float a = 0, b = 0, c = 0;
    if (i < n && !(P.p_type[i] == 0 || P.i_type[i] >= 2)) {
        a = (float)sqrt(P.ux[i] * P.ux[i] + P.uy[i] * P.uy[i] + P.uz[i] * P.uz[i]);
        b = (float)((P.rho[i] - S2.rho_ref[i]) / S2.rho_ref[i] * 100.0);
        c = (float)Q.ftotal[i];
    }
    unsigned ua = i < n ? f2ord(a) : 0u, ub = i < n ? f2ord(b) : 0u, uc = i < n ? f2ord(c) : 0u;
    s[0][threadIdx.x] = ua; s[1][threadIdx.x] = ub; s[2][threadIdx.x] = uc;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] = max(s[q][threadIdx.x], s[q][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) atomicMax(&out[q], s[q][0]);
}
__global__ void k_stats_final(const unsigned* in, float* out) {
    if (threadIdx.x < 3) out[threadIdx.x] = ord2f(in[threadIdx.x]);
}

// ========================================================================
// Host orchestration: every launch covers all cases of the engine's step.
// Per-case switches (noslip_bc, time_type, the count==0 Neumann call) are
// tested inside the kernels, so one launch sequence serves any mix.
// ========================================================================

// launch over the particles of every case; tunable passes (kid >= 0) run at
// the workgroup size the autotuner picked for this batch size
// a boundary pass with a 64-lane wave per boundary particle of each case
#define WAVEL(name, kern, ...)                                                                        \
    do {                                                                                            \
        Shape sh_;                                                                                  \
        SPH_TRY(shape(e, 3, WAVE_BLOCK, &sh_));                                                     \
        if (sh_.grid > 0) {                                                                         \
            SPH_TRY(prof_begin(e, name, 0.0));                                                      \
            hipLaunchKernelGGL(kern, dim3(sh_.grid), dim3(WAVE_BLOCK), 0, e.st, Seg{e.d_desc, sh_.tab}, ##__VA_ARGS__); \
            SPH_TRY(launch_check(&e, name));                                                        \
            SPH_TRY(prof_end(e));                                                                   \
        }                                                                                           \
    } while (0)
// the heavy-list share of a MIXED record pass: a wave per such particle
#define HEAVYL(name, kern, ...)                                                                       \
    do {                                                                                            \
        Shape sh_;                                                                                  \
        SPH_TRY(shape(e, 5, WAVE_BLOCK, &sh_));                                                     \
        if (sh_.grid > 0) {                                                                         \
            SPH_TRY(prof_begin(e, name, 0.0));                                                      \
            hipLaunchKernelGGL(kern, dim3(sh_.grid), dim3(WAVE_BLOCK), 0, e.st, Seg{e.d_desc, sh_.tab}, ##__VA_ARGS__); \
            SPH_TRY(launch_check(&e, name));                                                        \
            SPH_TRY(prof_end(e));                                                                   \
        }                                                                                           \
    } while (0)
// the cell-tile variant of a record pass: a workgroup per chunk, the halo
// records (recf floats each) in dynamic LDS
#define TILEL(name, kern, recf)                                                                      \
    do {                                                                                            \
        Shape sh_;                                                                                  \
        SPH_TRY(shape(e, 4, e.tile_lanes, &sh_));                                                   \
        if (sh_.grid > 0) {                                                                         \
            const size_t lds_ = tile_lds_bytes(e, (recf));                                          \
            SPH_TRY(tile_lds_allow(e, (const void*)(kern), lds_));                                  \
            SPH_TRY(prof_begin(e, name, 0.0));                                                      \
            hipLaunchKernelGGL(kern, dim3(sh_.grid), dim3(e.tile_lanes), lds_, e.st, Seg{e.d_desc, sh_.tab}); \
            SPH_TRY(launch_check(&e, name));                                                        \
            SPH_TRY(prof_end(e));                                                                   \
        }                                                                                           \
    } while (0)
#define SEGL(name, kid, kern, ...)                                                                   \
    do {                                                                                            \
        const int blk_ = (kid) >= 0 ? tuned_block(e, (kid)) : BLOCK;                                \
        Shape sh_;                                                                                  \
        SPH_TRY(shape(e, 0, blk_, &sh_));                                                           \
        if (sh_.grid > 0) {                                                                         \
            SPH_TRY(prof_begin(e, name, 0.0));                                                      \
            if ((kid) >= 0) SPH_TRY(tune_begin(e, (kid), blk_));                                    \
            hipLaunchKernelGGL(kern, dim3(sh_.grid), dim3(blk_), 0, e.st, Seg{e.d_desc, sh_.tab}, ##__VA_ARGS__); \
            SPH_TRY(launch_check(&e, name));                                                        \
            if ((kid) >= 0) SPH_TRY(tune_end(e));                                                   \
            SPH_TRY(prof_end(e));                                                                   \
        }                                                                                           \
    } while (0)

static bool mixed_of(const Eng& e) { return e.cs[0]->prm.precision == SPH_MIXED; }

// which: 0 = step working copy (SP1), 1 = persistent state (P1)
// Neumann pairs of the case's boundary particles for its current search,
// in the precision of its steps.
// 1 for a particle whose spreading list (the F_BH entries) holds an IB
// marker (p_type 1000): a superset of the particles the spreading sums
// reach, whatever the types of i and the markers' i_type become in the step
__global__ void __launch_bounds__(BLOCK) k_ibm_near(int n, PView S, NList nl, unsigned char* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned char hit = 0;
    NLRow row = nl_row(nl, i);
    for (int k = 0; k < row.cnt && !hit; k++) {
        const unsigned e = NB_AT(row, k);
        if (e & F_BH) hit = S.p_type[e & NB_IDX] == 1000;
    }
    out[i] = hit;
}
int ibm_near_build(sph_ctx* c) {
    if (c->ibm_near_gen == c->generation && c->ibm_near) return 0;
    if (!c->ibm_near) SPH_TRY(dalloc(c, &c->ibm_near, (size_t)c->cap));
    if (c->n > 0) {
        hipLaunchKernelGGL(k_ibm_near, dim3(grid1(c->n)), dim3(BLOCK), 0, c->stream, c->n, c->sview(), c->nlist(),
                           c->ibm_near);
        SPH_TRY(launch_check(c, "ibm_near"));
    }
    c->ibm_near_gen = c->generation;
    return 0;
}

// The distinct particles of the Neumann pairs (a thin layer of fluid along
// walls, lid and buildings) and the pairs as indices into that set.
__global__ void __launch_bounds__(BLOCK) k_zero_i32(int n, int* a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}
__global__ void __launch_bounds__(BLOCK) k_nms_mark(long long m, const int* ent, int* mark) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m) mark[ent[e]] = 1;
}
__global__ void __launch_bounds__(BLOCK) k_nms_list(int n, const int* mark, const int* idx, int* set) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && mark[i]) set[idx[i]] = i;
}
__global__ void __launch_bounds__(BLOCK) k_nms_remap(long long m, const int* ent, const int* idx, int* cent) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m) cent[e] = idx[ent[e]];
}
static int neumann_set(sph_ctx* c, long long total) {
    const int n = c->n;
    int *mark = nullptr, *idx = nullptr;
    SPH_TRY(dalloc(c, &mark, n));
    SPH_TRY(dalloc(c, &idx, n));
    hipLaunchKernelGGL(k_zero_i32, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, mark);
    hipLaunchKernelGGL(k_nms_mark, dim3(grid1(total)), dim3(BLOCK), 0, c->stream, total, (const int*)c->nmp_ent, mark);
    SPH_TRY(launch_check(c, "neumann_set_mark"));
    int nset = 0;
    SPH_TRY(exclusive_scan_i32(c, mark, idx, n, &nset));
    SPH_TRY(dalloc(c, &c->nm_set, (size_t)std::max(1, nset)));
    SPH_TRY(dalloc(c, &c->nmp_cent, (size_t)total));
    SPH_TRY(dalloc(c, &c->nm_rec, (size_t)std::max(1, nset) * NMREC_FLOATS));
    hipLaunchKernelGGL(k_nms_list, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)mark, (const int*)idx,
                       c->nm_set);
    hipLaunchKernelGGL(k_nms_remap, dim3(grid1(total)), dim3(BLOCK), 0, c->stream, total, (const int*)c->nmp_ent,
                       (const int*)idx, c->nmp_cent);
    SPH_TRY(launch_check(c, "neumann_set"));
    dfree(c, mark);
    dfree(c, idx);
    c->nm_nset = nset;
    return 0;
}

int neumann_pairs(sph_ctx* c) {
    dfree(c, c->nmp_off);
    dfree(c, c->nmp_ent);
    dfree(c, c->nmp_cent);
    dfree(c, c->nm_set);
    dfree(c, c->nm_rec);
    c->nmp_off = nullptr;
    c->nmp_ent = nullptr;
    c->nmp_cent = nullptr;
    c->nm_set = nullptr;
    c->nm_rec = nullptr;
    c->nm_nset = 0;
    const int m = c->nm_count;
    if (m > 0 && c->nm_list) {
        const bool mixed = c->prm.precision == SPH_MIXED;
        const KConst kc{c->prm.kernel_type, c->prm.kappa, c->c_pi15};
        int* cnt = nullptr;
        long long* wide = nullptr;
        long long total = 0;
        SPH_TRY(dalloc(c, &cnt, m));
        SPH_TRY(dalloc(c, &wide, m));
        SPH_TRY(dalloc(c, &c->nmp_off, (size_t)m + 1));
        if (mixed)
            hipLaunchKernelGGL((k_nmp_build<float, false>), dim3(m), dim3(64), 0, c->stream, c->pview(), c->grid(), kc,
                               c->prm.L2, (const int*)c->nm_list, m, cnt, (const long long*)nullptr, (int*)nullptr);
        else
            hipLaunchKernelGGL((k_nmp_build<double, false>), dim3(m), dim3(64), 0, c->stream, c->pview(), c->grid(),
                               kc, c->prm.L2, (const int*)c->nm_list, m, cnt, (const long long*)nullptr, (int*)nullptr);
        SPH_TRY(launch_check(c, "neumann_pairs_count"));
        hipLaunchKernelGGL(k_nmp_widen, dim3(grid1(m)), dim3(BLOCK), 0, c->stream, m, (const int*)cnt, wide);
        SPH_TRY(launch_check(c, "neumann_pairs_widen"));
        SPH_TRY(exclusive_scan_i64(c, wide, c->nmp_off, m, &total));
        SPH_HIP(c, hipMemcpyAsync(c->nmp_off + m, &total, sizeof(long long), hipMemcpyHostToDevice, c->stream));
        SPH_TRY(dalloc(c, &c->nmp_ent, (size_t)std::max(1LL, total)));
        if (mixed)
            hipLaunchKernelGGL((k_nmp_build<float, true>), dim3(m), dim3(64), 0, c->stream, c->pview(), c->grid(), kc,
                               c->prm.L2, (const int*)c->nm_list, m, (int*)nullptr, (const long long*)c->nmp_off,
                               c->nmp_ent);
        else
            hipLaunchKernelGGL((k_nmp_build<double, true>), dim3(m), dim3(64), 0, c->stream, c->pview(), c->grid(),
                               kc, c->prm.L2, (const int*)c->nm_list, m, (int*)nullptr, (const long long*)c->nmp_off,
                               c->nmp_ent);
        SPH_TRY(launch_check(c, "neumann_pairs_fill"));
        SPH_HIP(c, hipStreamSynchronize(c->stream));   // total lives on this frame
        dfree(c, cnt);
        dfree(c, wide);
        if (mixed && c->pos_f32 && !c->prm.ale_positions && total > 0) SPH_TRY(neumann_set(c, total));
    }
    c->nmp_gen = c->generation;
    c->nmp_for = c->nm_list;
    c->nmp_prec = c->prm.precision;
    return 0;
}

static int run_neumann(Eng& e, int which, int only_count0) {
    SEGL("snapshot", -1, k_snapshot, which, 1);
    // one 64-thread block (one wave) per boundary particle of each case
    Shape sh;
    SPH_TRY(shape(e, 3, 64, &sh));
    if (sh.grid > 0) {
        SPH_TRY(prof_begin(e, "neumann", 0.0));
        if (mixed_of(e))
            hipLaunchKernelGGL(k_neumann<float>, dim3(sh.grid), dim3(64), 0, e.st, Seg{e.d_desc, sh.tab}, which,
                               only_count0);
        else
            hipLaunchKernelGGL(k_neumann<double>, dim3(sh.grid), dim3(64), 0, e.st, Seg{e.d_desc, sh.tab}, which,
                               only_count0);
        SPH_TRY(launch_check(&e, "neumann"));
        SPH_TRY(prof_end(e));
    }
    return 0;
}
static int run_outlet(Eng& e, int which) {
    SEGL("snapshot", -1, k_snapshot, which, 0);
    SEGL("outlet", -1, k_outlet, which);
    return 0;
}
static int run_periodic(Eng& e, int which) {
    SEGL("snapshot", -1, k_snapshot, which, 0);
    SEGL("periodic", -1, k_periodic, which);
    return 0;
}
static int run_freeslip(Eng& e, int which) {
    if (mixed_of(e)) WAVEL("freeslip", k_freeslip_wave, which);   // the lid's particles among the boundary ones
    else SEGL("freeslip", TK_FREESLIP, k_freeslip<double>, which);
    return 0;
}

// Neighbour records of every case of the step, from its working state as
// it is now; run before each pass that gathers from them.
int pack_records(Eng& e, int which) {
    if (which == REC_PRE) {
        bool any = false;
        for (size_t k = 0; k < e.cs.size(); k++) any |= e.h_desc[k].crec_adv != nullptr;
        if (any) SEGL("pack_records", -1, k_pack_records<REC_PRE>);
        return 0;
    }
    if (which == REC_PQ) SEGL("pack_records", -1, k_pack_records<REC_PQ>);
    else if (which == REC_VEL_PQ) SEGL("pack_records", -1, k_pack_records<REC_VEL_PQ>);
    else SEGL("pack_records", -1, k_pack_records<REC_ALL>);
    return 0;
}

// KERNEL_clc_IB_normal_vector3D, functions_PREP.cuh:923-1015 (count == 0,
// ISPH_Calc.cuh:70-74): markers within ib_face_tol of a face plane get -1 / +1
// on that axis (the upper plane wins a tie, as it is tested last), normalised
// in float with IEEE round-to-nearest square root and division.  In
// SPH_IBN_REFERENCE mode a step with count > 0 writes zero normals, the
// values the reference's per-step part3 reset leaves (ISPH_Calc.cuh:61).
__global__ void __launch_bounds__(BLOCK) k_ib_normals(Seg sg) {
    SEG_PARTICLE(sg);
    const sph_params& pr = cd.pr;
    if (pr.ib_normals == SPH_IBN_UPLOADED) return;
    if (cd.count > 0 && pr.ib_normals == SPH_IBN_PLANES) return;
    const PView S = cd.S;
    if (S.p_type[i] < 1000) return;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    if (cd.count == 0) {
        const double xi = S.x[i], yi = S.y[i], zi = S.z[i], tol = pr.ib_face_tol;
        if (fabs(xi - pr.ib_face_xlo) < tol) nx = -1.0f;
        if (fabs(xi - pr.ib_face_xhi) < tol) nx = 1.0f;
        if (fabs(yi - pr.ib_face_ylo) < tol) ny = -1.0f;
        if (fabs(yi - pr.ib_face_yhi) < tol) ny = 1.0f;
        if (fabs(zi - pr.ib_face_zlo) < tol) nz = -1.0f;
        if (fabs(zi - pr.ib_face_zhi) < tol) nz = 1.0f;
    }
    const float mag = __fsqrt_rn(nx * nx + ny * ny + nz * nz);
    const QView Q = cd.Q;
    if ((double)mag > 1e-9) {
        Q.nx[i] = (double)__fdiv_rn(nx, mag);
        Q.ny[i] = (double)__fdiv_rn(ny, mag);
        Q.nz[i] = (double)__fdiv_rn(nz, mag);
    } else {
        Q.nx[i] = 0.0; Q.ny[i] = 0.0; Q.nz[i] = 0.0;
    }
}

// The host-side choices of step_physics_front, one bit each.
long long step_physics_front_key(const Eng& e) {
    long long key = mixed_of(e) ? 1 : 0;
    bool count0 = false, gc = false, nrm = false, scal = false, compact = false;
    for (size_t k = 0; k < e.cs.size(); k++) {
        const sph_ctx* c = e.cs[k];
        count0 |= e.h_desc[k].count == 0;
        gc |= !e.h_desc[k].gc_fresh;
        const int m = c->prm.ib_normals;
        nrm |= m == SPH_IBN_REFERENCE || (m == SPH_IBN_PLANES && e.h_desc[k].count == 0);
        scal |= c->prm.con_solve || c->prm.concn_solve || c->prm.boussinesq_solve;
        compact |= e.h_desc[k].crec_adv != nullptr;
    }
    return key | (count0 ? 2 : 0) | (gc ? 4 : 0) | (nrm ? 8 : 0) | (scal ? 16 : 0) | (compact ? 32 : 0) |
           (e.tile_lanes ? 64 : 0);
}
long long step_physics_back_key(const Eng& e) {
    bool ale = false;
    for (const sph_ctx* c : e.cs) ale |= c->prm.ale_positions != 0;
    return (mixed_of(e) ? 1 : 0) | (ale ? 2 : 0) | (e.tile_lanes ? 4 : 0);
}

// Everything of SOPHIA_single_ISPH before the pressure solve
// (ISPH_Calc.cuh:20-132).  The count==0 neighbour searches have already run.
int step_physics_front(Eng& e) {
    const bool mixed = mixed_of(e);
    bool any_count0 = false;
    for (size_t k = 0; k < e.cs.size(); k++) any_count0 |= e.h_desc[k].count == 0;
    SEGL("step_begin", -1, k_step_begin);
    bool any_nrm = false;   // a case whose marker normals this step sets
    for (size_t k = 0; k < e.cs.size(); k++) {
        const int m = e.cs[k]->prm.ib_normals;
        any_nrm |= m == SPH_IBN_REFERENCE || (m == SPH_IBN_PLANES && e.h_desc[k].count == 0);
    }
    if (any_nrm) SEGL("ib_normals", -1, k_ib_normals);
    bool any_gc = false;
    for (size_t k = 0; k < e.cs.size(); k++) any_gc |= !e.h_desc[k].gc_fresh;
    if (any_gc) {
        if (mixed) SEGL("grad_corr", TK_GRAD_CORR, k_grad_corr<float>);
        else SEGL("grad_corr", TK_GRAD_CORR, k_grad_corr<double>);
    }
    SPH_TRY(pack_records(e, REC_PRE));   // compact records for the strain pass
    if (mixed && e.tile_lanes) TILEL("prep", k_prep_tile, CREC_ADV_FLOATS);
    else if (mixed) {
        SEGL("prep", TK_PREP, k_prep<float>);
        HEAVYL("prep_heavy", k_prep_heavy);
    } else SEGL("prep", TK_PREP, k_prep<double>);
    if (mixed) WAVEL("boundary3d", k_boundary3d_wave);
    else SEGL("boundary3d", TK_BOUNDARY, k_boundary3d<double>);
    SEGL("most_self", -1, k_most_self);
    SEGL("most_deposit", -1, k_most_deposit);
    if (any_count0) SPH_TRY(run_neumann(e, 0, 1));
    SPH_TRY(pack_records(e, REC_ALL));
    bool scal = false;   // scalar transport or buoyancy in some case of the step
    for (sph_ctx* c : e.cs) scal |= c->prm.con_solve || c->prm.concn_solve || c->prm.boussinesq_solve;
    if (mixed) {
        if (e.tile_lanes) {
            if (scal) TILEL("advection", k_advection_tile<true>, CREC_ADV_FLOATS);
            else TILEL("advection", k_advection_tile<false>, CREC_ADV_FLOATS);
        } else if (scal) {
            SEGL("advection", TK_ADVECTION, (k_advection<float, true>));
            HEAVYL("advection_heavy", k_advection_heavy<true>);
        } else {
            SEGL("advection", TK_ADVECTION, (k_advection<float, false>));
            HEAVYL("advection_heavy", k_advection_heavy<false>);
        }
    } else {
        if (scal) SEGL("advection", TK_ADVECTION, (k_advection<double, true>));
        else SEGL("advection", TK_ADVECTION, (k_advection<double, false>));
    }
    SEGL("projection", -1, k_projection);
    SPH_TRY(run_periodic(e, 0));
    SPH_TRY(run_freeslip(e, 0));
    return 0;
}

// Everything after the pressure solve (ISPH_Calc.cuh:143-214).
int step_physics_back(Eng& e) {
    const bool mixed = mixed_of(e);
    SPH_TRY(run_neumann(e, 0, 0));
    SPH_TRY(run_outlet(e, 0));
    SPH_TRY(pack_records(e, REC_PQ));   // pressures after the solve and the boundary passes
    if (mixed && e.tile_lanes) TILEL("pressure_force", k_pressure_force_tile, CREC_FLOATS);
    else if (mixed) {
        SEGL("pressure_force", TK_PRESSURE, k_pressure_force<float>);
        HEAVYL("pressure_force_heavy", k_pressure_force_heavy);
    }
    else SEGL("pressure_force", TK_PRESSURE, k_pressure_force<double>);
    if (mixed) WAVEL("ibm_force", k_ibm_force_wave);
    else SEGL("ibm_force", TK_IBM_FORCE, k_ibm_force<double>);
    if (mixed) SEGL("ibm_spread", TK_IBM_SPREAD, k_ibm_spread<float>);
    else SEGL("ibm_spread", TK_IBM_SPREAD, k_ibm_spread<double>);
    SEGL("time_update", -1, k_time_update);
    SPH_TRY(run_neumann(e, 1, 0));
    SPH_TRY(run_outlet(e, 1));
    SPH_TRY(run_periodic(e, 1));
    SPH_TRY(run_freeslip(e, 1));
    bool ale = false;
    for (sph_ctx* c : e.cs) ale |= c->prm.ale_positions != 0;
    if (ale) SEGL("commit_positions", -1, k_commit_positions);
    return 0;
}

int stats_run(sph_ctx* c, float* umax, float* drho, float* fmax_) {
    const int n = c->n;
    unsigned* acc = (unsigned*)c->stat_buf;
    SPH_HIP(c, hipMemsetAsync(acc, 0, 3 * sizeof(unsigned), c->stream));
    hipLaunchKernelGGL(k_stats, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, c->pview(), c->s2view(), c->qview(), acc);
    SPH_TRY(launch_check(c, "stats"));
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, c->stream, (const unsigned*)acc, c->stat_buf + 4);
    SPH_TRY(launch_check(c, "stats_final"));
    float h[3];
    SPH_HIP(c, hipMemcpyAsync(h, c->stat_buf + 4, 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    *umax = h[0]; *drho = h[1]; *fmax_ = h[2];
    return 0;
}

// ---- statistics without allocation (sph_stats_noalloc) -------------------
// Per-workgroup maxima go to a file-scope device array instead of the
// context's buffer, and a one-workgroup fold reduces them, so the call needs
// no device allocation, no memset and no atomics.  The array is this
// library's own on each GPU: every call on a GPU shares it, so calls that
// overlap on one GPU race.  Particles up to NA_STATS_MAX_BLOCKS * BLOCK.
constexpr int NA_STATS_MAX_BLOCKS = 1 << 16;
constexpr int NA_STATS_OUT = 3 * NA_STATS_MAX_BLOCKS;   // first word of the 3 float results
constexpr long long NA_STATS_MAX_N = (long long)NA_STATS_MAX_BLOCKS * BLOCK;
static __device__ __align__(256) unsigned scratch[NA_STATS_OUT + 64];

// The maxima of k_stats for this workgroup's particles, stored at part[3b..3b+2].
__global__ void __launch_bounds__(BLOCK) k_stats_part(int n, PView P, P2View S2, QView Q, unsigned* part) {
    __shared__ unsigned s[3][BLOCK];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    float a = 0, b = 0, c = 0;
    if (i < n && !(P.p_type[i] == 0 || P.i_type[i] >= 2)) {
        a = (float)sqrt(P.ux[i] * P.ux[i] + P.uy[i] * P.uy[i] + P.uz[i] * P.uz[i]);
        b = (float)((P.rho[i] - S2.rho_ref[i]) / S2.rho_ref[i] * 100.0);
        c = (float)Q.ftotal[i];
    }
    s[0][threadIdx.x] = i < n ? f2ord(a) : 0u;
    s[1][threadIdx.x] = i < n ? f2ord(b) : 0u;
    s[2][threadIdx.x] = i < n ? f2ord(c) : 0u;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] = max(s[q][threadIdx.x], s[q][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}
// One workgroup: the maxima over the nb partials (starting from 0, as the
// zeroed accumulator of k_stats does), as floats into out[0..2].
__global__ void __launch_bounds__(BLOCK) k_stats_fold(int nb, const unsigned* part, float* out) {
    __shared__ unsigned s[3][BLOCK];
    unsigned a = 0u, b = 0u, c = 0u;
    for (int k = threadIdx.x; k < nb; k += BLOCK) {
        a = max(a, part[3 * k]);
        b = max(b, part[3 * k + 1]);
        c = max(c, part[3 * k + 2]);
    }
    s[0][threadIdx.x] = a; s[1][threadIdx.x] = b; s[2][threadIdx.x] = c;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] = max(s[q][threadIdx.x], s[q][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = ord2f(s[threadIdx.x][0]);
}

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (the host-side name of the array is no device address).
static std::mutex na_mu;
static void* na_addr[64] = {};

static int na_scratch(sph_ctx* c, unsigned** out) {
    const int dev = c->device;
    if (dev < 0 || dev >= 64) return fail(c, SPH_E_ARG, "sph_stats_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_addr[dev]) {
        int cur = -1;
        SPH_HIP(c, hipGetDevice(&cur));
        if (cur != dev) SPH_HIP(c, hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t er = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) SPH_HIP(c, hipSetDevice(cur));
        if (er != hipSuccess || !p)
            return fail(c, SPH_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(er));
        na_addr[dev] = p;
    }
    *out = (unsigned*)na_addr[dev];
    return 0;
}

int stats_run_noalloc(sph_ctx* c, float* umax, float* drho, float* fmax_) {
    const int n = c->n;
    if ((long long)n > NA_STATS_MAX_N)
        return fail(c, SPH_E_ARG, "sph_stats_noalloc: " + std::to_string(n) + " particles exceed its bound of " +
                                      std::to_string(NA_STATS_MAX_N) + " (use sph_stats)");
    unsigned* sc = nullptr;
    SPH_TRY(na_scratch(c, &sc));
    const unsigned nb = grid1(n);
    hipLaunchKernelGGL(k_stats_part, dim3(nb), dim3(BLOCK), 0, c->stream, n, c->pview(), c->s2view(), c->qview(), sc);
    SPH_TRY(launch_check(c, "stats_part"));
    float* res = (float*)(sc + NA_STATS_OUT);
    hipLaunchKernelGGL(k_stats_fold, dim3(1), dim3(BLOCK), 0, c->stream, (int)nb, (const unsigned*)sc, res);
    SPH_TRY(launch_check(c, "stats_fold"));
    float h[3];
    SPH_HIP(c, hipMemcpyAsync(h, res, 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    *umax = h[0]; *drho = h[1]; *fmax_ = h[2];
    return 0;
}

}  // namespace sph
