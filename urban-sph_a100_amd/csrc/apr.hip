// Adaptive particle refinement (APR): cell-based 1 -> 8 splitting.
//
//   KERNEL_APR_condition3D_cell, splitting branch  functions_APR.cuh:270-330  -> k_apr_pick
//   Cell_index (prefix sum of the cell flags)     functions_APR_etc.cuh:366   -> exclusive scan
//   KERNEL_cell_APR_3D, splitting branch          functions_APR.cuh:420-500  -> k_apr_split
//   Particle_refinement_cell                      functions_APR_etc.cuh:354-375 -> sph_apr_refine
//
// A particle's refinement level is log8(m_ref / m).  In every occupied cell
// the first particle (in cell order) whose target level exceeds its level
// by at least 0.75 is split into 8 children at +-a along each axis,
// a = cbrt(m / rho) / 4, rounded to float positions as the reference's
// float records hold them, with m/8, vol/8 and h/2 (when apr_h_change),
// every other field copied.  Children of the k-th splitting cell (ascending cell
// index) go to n + 8k..n + 8k + 7 in the reference's order; the parent
// becomes a dummy (i_type 3).  One split per cell per pass; the neighbour
// search runs again after each pass.
//
// The reference keeps the target level (M_num) and the split flag
// (apr_cond) in its particle record, but both are commented out of the
// compiled record (class_Cuda_Particle_Array.cuh:37-43); here the target
// level is a per-particle array of the context, inherited by the children.
#include "flags.h"
#include "sph_internal.h"
#include "sph_math.h"

namespace sph {

// Cube root by Newton's method on the mantissa.  The same IEEE operations
// run on every device, so children's positions agree bit for bit with any
// other evaluation of this routine.
__device__ __forceinline__ double apr_cbrt(double v) {
    int e = 0;
    double m = frexp(v, &e);          // v = m 2^e, m in [0.5, 1)
    const int r = ((e % 3) + 3) % 3;
    m = ldexp(m, r);                  // m in [0.5, 4), e - r divisible by 3
    e -= r;
    double y = 1.0;
#pragma unroll
    for (int it = 0; it < 8; it++) y = y - (y * y * y - m) / (3.0 * y * y);
    return ldexp(y, e / 3);
}

__global__ void __launch_bounds__(BLOCK) k_apr_pick(int ncells, const int* gstr, const int* gend, PView P,
                                                    const double* target, int* flag, int* pick) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncells) return;
    flag[cell] = 0;
    pick[cell] = -1;
    const int s = gstr[cell];
    if (s == -1) return;
    const int e = gend[cell];
    const double l8 = log(8.0);
    for (int j = s; j < e; j++) {
        const double level = log(P.m_ref[j] / P.m[j]) / l8;
        if (target[j] - level >= 0.75) {
            // the split loop passes over a non-fluid particle whose target
            // exceeds 2, and the condition marked only this one
            // (functions_APR.cuh:423)
            if (!(P.p_type[j] != 1 && target[j] > 2.0)) {
                flag[cell] = 1;
                pick[cell] = j;
            }
            return;
        }
    }
}

constexpr int APR_MAX_F64 = 64;
struct AprFields {
    double* f[APR_MAX_F64];
    int* iv[SPH_I_COUNT];
    int nf;
    double *x, *y, *z, *m, *vol, *h, *rho;
};

__global__ void __launch_bounds__(BLOCK) k_apr_split(int ncells, const int* flag, const int* rank, const int* pick,
                                                     int n0, AprFields F, int h_change) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncells || !flag[cell]) return;
    const int j = pick[cell];
    const int base = n0 + 8 * rank[cell];
    const double xm = F.x[j], ym = F.y[j], zm = F.z[j];
    const double a = 0.25 * apr_cbrt(F.m[j] / F.rho[j]);
    for (int k = 0; k < 8; k++) {
        const int c = base + k;
        for (int f = 0; f < F.nf; f++) F.f[f][c] = F.f[f][j];
        for (int f = 0; f < SPH_I_COUNT; f++) F.iv[f][c] = F.iv[f][j];
        // child order of functions_APR.cuh:440-478: x +,+,+,+,-,-,-,-;
        // y +,+,-,-,+,+,-,-; z +,-,+,-,+,-,+,-.  The reference stores the
        // children's positions in its float records (functions_APR.cuh:
        // 544-590): rounded to float here too, so a refined case's positions
        // stay float values
        F.x[c] = (double)(float)(k < 4 ? xm + a : xm - a);
        F.y[c] = (double)(float)((k & 2) ? ym - a : ym + a);
        F.z[c] = (double)(float)((k & 1) ? zm - a : zm + a);
        F.m[c] = 0.125 * F.m[j];
        F.vol[c] = 0.125 * F.vol[j];
        if (h_change) F.h[c] = 0.5 * F.h[j];
    }
    F.iv[SPH_I_I_TYPE][j] = 3;
}

int apr_refine(sph_ctx* c, const double* target, int max_passes, int* n_out) {
    const int ncells = c->dom.num_cells - 1;
    if (!c->apr_target) {
        SPH_TRY(dalloc(c, &c->apr_target, c->cap));
        SPH_HIP(c, hipMemsetAsync(c->apr_target, 0, sizeof(double) * (size_t)c->cap, c->stream));
    }
    if (target)
        SPH_HIP(c, hipMemcpyAsync(c->apr_target, target, sizeof(double) * (size_t)c->n, hipMemcpyHostToDevice, c->stream));
    int *flag = nullptr, *rank = nullptr, *pick = nullptr;
    SPH_TRY(dalloc(c, &flag, ncells));
    SPH_TRY(dalloc(c, &rank, ncells));
    SPH_TRY(dalloc(c, &pick, ncells));
    int rc = 0;
    bool split_any = false;
    for (int pass = 0; pass < max_passes && rc == 0; pass++) {
        hipLaunchKernelGGL(k_apr_pick, dim3(grid1(ncells)), dim3(BLOCK), 0, c->stream, ncells, (const int*)c->gstr,
                           (const int*)c->gend, c->pview(), (const double*)c->apr_target, flag, pick);
        if ((rc = launch_check(c, "apr_pick"))) break;
        int nsplit = 0;
        if ((rc = exclusive_scan_i32(c, flag, rank, ncells, &nsplit))) break;
        if (nsplit == 0) break;
        if ((long long)c->n + 8LL * nsplit > c->cap) {
            rc = fail(c, SPH_E_NOMEM, "APR needs room for " + std::to_string(c->n + 8LL * nsplit) +
                                          " particles; create the context with a larger n_part2");
            break;
        }
        AprFields F{};
        for (int f = 0; f < SPH_F_COUNT_F64 + 3; f++)
            if (c->pf[f]) F.f[F.nf++] = c->pf[f];
        for (int f = 0; f < 10; f++) F.f[F.nf++] = c->p2[f];
        for (int f = 15; f < 18; f++) F.f[F.nf++] = c->q[f];
        F.f[F.nf++] = c->apr_target;
        for (int f = 0; f < SPH_I_COUNT; f++) F.iv[f] = c->pi[f];
        F.x = c->pf[SPH_F_X]; F.y = c->pf[SPH_F_Y]; F.z = c->pf[SPH_F_Z]; F.m = c->pf[SPH_F_M];
        F.vol = c->pf[SPH_F_VOL]; F.h = c->pf[SPH_F_H]; F.rho = c->pf[SPH_F_RHO];
        hipLaunchKernelGGL(k_apr_split, dim3(grid1(ncells)), dim3(BLOCK), 0, c->stream, ncells, (const int*)flag,
                           (const int*)rank, (const int*)pick, c->n, F, c->prm.apr_h_change);
        if ((rc = launch_check(c, "apr_split"))) break;
        c->n += 8 * nsplit;   // the children's positions are float values like their parents'
        dfree(c, c->porder);   // the pass order covers the old count; the next list rebuilds it
        c->porder = nullptr;
        ppe_free(c);
        if ((rc = set_ncell_run(c))) break;
        // the next pass needs only the cells; the neighbour list is built
        // once after the last pass, on the cells and order of its search
        rc = nnps_run(c, false);
        split_any = true;
    }
    if (rc == 0 && split_any) rc = nnps_list(c);
    dfree(c, flag);
    dfree(c, rank);
    dfree(c, pick);
    if (n_out) *n_out = c->n;
    return rc;
}

}  // namespace sph
