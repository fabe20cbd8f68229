// Pressure Poisson equation of the projection step on MI355X.
//
// Operator: the reference's KERNEL_PPE3D (functions_PPE.cuh:2-175) sweep
//     p_i = p0_i + (b_i + sum_j c_ij (p_j - p0_j)) / sum_j c_ij
// is the Jacobi iteration of   A_i q_i - sum_{j in U} c_ij q_j = b_i + kn_i
// with q = p - p0, A_i = sum_j c_ij over every PPE pair, U the particles the
// sweep updates (fluid, i_type<=2) and kn_i = sum over the other pairs, whose
// pressures the boundary kernels hold fixed during the solve.
//
//   SPH_PPE_REF_JACOBI : one sweep, every row reading the previous iterate
//                         (removes the read/write race of functions_PPE.cuh:93
//                         vs :166).
//   SPH_PPE_BICGSTAB   : Jacobi-preconditioned BiCGSTAB to ppe_tol.
//
// Storage: sliced ELLPACK with 64-row slices (one wave64 per slice, entries
// of a slice column-major, so a wave's loads of one slot are contiguous).
// The column structure is built once per neighbour search; values, diagonal
// and right-hand side are assembled every step from the neighbour list.
//
// Reductions: every dot product is the canonical tree of 256-element chunks
// (one 256-thread block per chunk, pairwise offsets 128..1, chunk partials
// reduced the same way), so results are bit-reproducible and independent of
// which other cases run.  Scalars live on the device; kernels after
// convergence return immediately, so the host polls only every few
// iterations and the iterate sequence matches a sequential implementation.
//
// MIXED precision: matrix values and Krylov vectors of an inner BiCGSTAB are
// float, every SpMV row sum and every dot is accumulated in double; an outer
// iterative-refinement loop computes the residual with double vectors and
// double accumulation and adds the float correction to a double solution.
#include "flags.h"
#include "sph_internal.h"
#include "sph_math.h"

namespace sph {

enum Scal {
    SC_RHO, SC_RHO_OLD, SC_ALPHA, SC_OMEGA, SC_BETA, SC_BNORM, SC_RNORM, SC_SNORM, SC_TOL,
    SC_DONE, SC_CONV, SC_HALF, SC_IT, SC_ITERS, SC_MAXIT, SC_ZEROX, SC_COUNT = 32
};

struct PPEState {
    int nr = 0;                 // unknowns
    int nslices = 0;
    long long nnz_slots = 0;    // SELL storage incl. padding
    long long nnz = 0;          // real entries
    int* rowof = nullptr;       // particle -> row or -1
    int* rows = nullptr;        // row -> particle
    int* cnt = nullptr;         // entries per row
    long long* sl_off = nullptr;
    int* cols = nullptr;
    double* vals64 = nullptr;
    float* vals32 = nullptr;
    double* diag = nullptr;
    double* rhs = nullptr;
    double* x = nullptr;        // solution q (double)
    double* pnew = nullptr;     // Jacobi output
    // Krylov work vectors (double set and float set)
    double *r, *rhat, *p, *phat, *v, *s, *shat, *t;
    float *fr, *frhat, *fp, *fphat, *fv, *fs, *fshat, *ft, *fx, *frhs;
    double* partA = nullptr;
    double* partB = nullptr;
    long long nb = 0;
    int last_iters = 16;
};

// ------------------------------------------------------------------------
// Canonical reductions.  A chunk of 256 values is summed by the pairwise
// tree t[i] += t[i+off], off = 128, 64, ..., 1.  One wave64 evaluates it in
// registers: lane l holds a[l], a[l+64], a[l+128], a[l+192]; the two upper
// levels are (a[l]+a[l+128]) + (a[l+64]+a[l+192]); the last six are
// shuffles with offsets 32..1.  The pairing is exactly the tree's, so the
// result is bit-identical to any other evaluation of the same tree.
__device__ __forceinline__ double wave_tree(double a0, double a1, double a2, double a3) {
    double v = (a0 + a2) + (a1 + a3);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
// one value per thread of a 256-thread block -> canonical sum on thread 0
__device__ __forceinline__ double block_canon(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    double r = 0.0;
    if (t < 64) r = wave_tree(sh[t], sh[t + 64], sh[t + 128], sh[t + 192]);
    __syncthreads();
    return r;
}
// two sums at once (waves 0 and 1)
__device__ __forceinline__ void block_canon2(double a, double b, double* sh, double* ra, double* rb) {
    const int t = threadIdx.x;
    sh[t] = a;
    sh[BLOCK + t] = b;
    __syncthreads();
    if (t < 64) *ra = wave_tree(sh[t], sh[t + 64], sh[t + 128], sh[t + 192]);
    else if (t < 128) {
        const int l = t - 64;
        *rb = wave_tree(sh[BLOCK + l], sh[BLOCK + l + 64], sh[BLOCK + l + 128], sh[BLOCK + l + 192]);
    }
    __syncthreads();
}
__device__ __forceinline__ double chunk_from_global(const double* part, long long base, long long nb, int lane) {
    double a[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        long long idx = base + lane + 64 * q;
        a[q] = idx < nb ? part[idx] : 0.0;
    }
    return wave_tree(a[0], a[1], a[2], a[3]);
}
// One 256-thread block reduces nb partials canonically (nb <= 65536); the
// result is valid on thread 0.
__device__ double canon_partials(const double* part, long long nb, double* lvl) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (nb <= 256) {
        double r = 0.0;
        if (w == 0) r = chunk_from_global(part, 0, nb, lane);
        return r;
    }
    int nc = (int)((nb + 255) / 256);
    for (int c = w; c < nc; c += (int)(blockDim.x >> 6)) {
        double v = chunk_from_global(part, (long long)c * 256, nb, lane);
        if (lane == 0) lvl[c] = v;
    }
    __syncthreads();
    double r = 0.0;
    if (w == 0) {
        double a[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int idx = lane + 64 * q;
            a[q] = idx < nc ? lvl[idx] : 0.0;
        }
        r = wave_tree(a[0], a[1], a[2], a[3]);
    }
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------------
// structure
__global__ void __launch_bounds__(BLOCK) k_ppe_flag(int n, PView S, int* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (S.i_type[i] <= 2 && S.p_type[i] < 1000 && S.p_type[i] > 0) ? 1 : 0;
}
__global__ void __launch_bounds__(BLOCK) k_ppe_rows(int n, const int* flag, const int* scan, int* rowof, int* rows) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) { rowof[i] = scan[i]; rows[scan[i]] = i; }
    else rowof[i] = -1;
}
template <bool FILL>
__global__ void __launch_bounds__(BLOCK) k_ppe_struct(int nr, const int* rows, const int* rowof, PView S, NList nl,
                                                      int* cnt, const long long* sl_off, int* cols) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    int i = rows[r];
    const unsigned* row = nl.ent + nl.slice_off[i >> 6] + (i & 63);
    int m = nl.cnt[i];
    int slot = 0;
    int* out = FILL ? cols + sl_off[r >> 6] + (r & 63) : nullptr;
    for (int k = 0; k < m; k++) {
        unsigned e = row[(size_t)k * SLICE];
        if (!(e & F_DP)) continue;
        int j = e & NB_IDX;
        if (S.p_type[j] >= 1000 || S.i_type[j] == 4) continue;
        if (j == i) continue;
        int rj = rowof[j];
        if (rj < 0) continue;
        if (FILL) out[(size_t)slot * SLICE] = rj;
        slot++;
    }
    if (!FILL) cnt[r] = slot;
}
__global__ void __launch_bounds__(BLOCK) k_row_slice_width(int nr, int nslices, const int* cnt, long long* w) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int m = 0;
    for (int l = 0; l < SLICE; l++) {
        int r = s * SLICE + l;
        if (r < nr) m = max(m, cnt[r]);
    }
    w[s] = (long long)m * SLICE;
}

// ------------------------------------------------------------------------
// per-step assembly, functions_PPE.cuh:2-160
template <class R, class V>
__global__ void __launch_bounds__(BLOCK) k_ppe_assemble(int nr, const int* rows, const int* rowof, KConst kc, PView S,
                                                        QView Q, NList nl, const long long* sl_off, V* vals,
                                                        double* diag, double* rhs, double* x0, double* pnew,
                                                        double dt, int jacobi) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    int i = rows[r];
    R hi = (R)(1.0 * S.h[i]);
    R tmp_A = calc_tmpA<R>(kc, hi);
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double uxi = S.ux[i], uyi = S.uy[i], uzi = S.uz[i];
    R rhoi = (R)S.rho[i];
    R gc11 = (R)Q.inv_xx[i], gc12 = (R)Q.inv_xy[i], gc22 = (R)Q.inv_yy[i];
    R gc23 = (R)Q.inv_yz[i], gc31 = (R)Q.inv_zx[i], gc33 = (R)Q.inv_zz[i];
    double bix = 0, biy = 0, biz = 0, Aij = 0, AijPij = 0, kn = 0;
    const unsigned* row = nl.ent + nl.slice_off[i >> 6] + (i & 63);
    int m = nl.cnt[i];
    V* out = vals + sl_off[r >> 6] + (r & 63);
    int slot = 0;
    for (int k = 0; k < m; k++) {
        unsigned e = row[(size_t)k * SLICE];
        if (!(e & F_DP)) continue;
        int j = e & NB_IDX;
        if (S.p_type[j] >= 1000 || S.i_type[j] == 4) continue;
        R dx = (R)(xi - S.x[j]), dy = (R)(yi - S.y[j]), dz = (R)(zi - S.z[j]);
        R rhoj = (R)S.rho[j], mj = (R)S.m[j];
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) + (R)1e-20;
        R tdwij = kernel_dwp<R>(kc, tmp_A, hi, tdist);
        R tdwx = tdwij * dx / tdist;
        R tdwy = tdwij * dy / tdist;
        R tdwz = tdwij * dz / tdist;
        R tdwxc = gc11 * tdwx + gc12 * tdwy + gc31 * tdwz;
        R tdwyc = gc12 * tdwx + gc22 * tdwy + gc23 * tdwz;
        R tdwzc = gc31 * tdwx + gc23 * tdwy + gc33 * tdwz;
        R hm = (R)2.0 * rhoi * rhoj / (rhoi + rhoj);
        bix += (double)(mj / rhoj * (R)(S.ux[j] - uxi) * tdwxc * hm);
        biy += (double)(mj / rhoj * (R)(S.uy[j] - uyi) * tdwyc * hm);
        biz += (double)(mj / rhoj * (R)(S.uz[j] - uzi) * tdwzc * hm);
        R cc = (R)2.0 * mj / rhoj * (dx * tdwxc + dy * tdwyc + dz * tdwzc) / tdist / tdist;
        double ccd = (double)cc;
        Aij += ccd;
        double qj = S.pres[j] - S.pres0[j];
        if (jacobi) AijPij += ccd * qj;
        if (j != i) {
            if (rowof[j] >= 0) {
                if (!jacobi) out[(size_t)slot * SLICE] = (V)cc;
                slot++;
            } else {
                kn += ccd * qj;
            }
        }
    }
    double bi = 1.0 * (bix + biy + biz + 0.0) / dt;
    double q0 = S.pres[i] - S.pres0[i];
    if (jacobi) {
        pnew[r] = S.pres0[i] + (bi + AijPij) / Aij;
    } else {
        diag[r] = Aij;
        rhs[r] = bi + kn;
        x0[r] = q0;
    }
}

__global__ void __launch_bounds__(BLOCK) k_ppe_scatter(int nr, const int* rows, const double* pres0, const double* src,
                                                       int add_p0, double* pres) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    int i = rows[r];
    pres[i] = add_p0 ? pres0[i] + src[r] : src[r];
}

// ------------------------------------------------------------------------
// Krylov kernels.  X = vector type, V = matrix value type.  Every kernel is
// gated on the device flag SC_DONE.
#define GATE(sc) if ((sc)[SC_DONE] != 0.0) return

template <class V, class X>
__device__ __forceinline__ double spmv_row(int r, int nr, const long long* sl_off, const int* cnt, const int* cols,
                                           const V* vals, const double* diag, const X* x) {
    const long long base = sl_off[r >> 6] + (r & 63);
    const int m = cnt[r];
    const int* cp = cols + base;
    const V* vp = vals + base;
    double s = 0.0;
    int k = 0;
    // matrix entries stream through once: nontemporal loads keep the
    // caches for the x gathers; four independent loads in flight per lane
    for (; k + 4 <= m; k += 4) {
        const int c0 = __builtin_nontemporal_load(cp + (size_t)(k + 0) * SLICE);
        const int c1 = __builtin_nontemporal_load(cp + (size_t)(k + 1) * SLICE);
        const int c2 = __builtin_nontemporal_load(cp + (size_t)(k + 2) * SLICE);
        const int c3 = __builtin_nontemporal_load(cp + (size_t)(k + 3) * SLICE);
        const V v0 = __builtin_nontemporal_load(vp + (size_t)(k + 0) * SLICE);
        const V v1 = __builtin_nontemporal_load(vp + (size_t)(k + 1) * SLICE);
        const V v2 = __builtin_nontemporal_load(vp + (size_t)(k + 2) * SLICE);
        const V v3 = __builtin_nontemporal_load(vp + (size_t)(k + 3) * SLICE);
        const X x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        s += (double)v0 * (double)x0;
        s += (double)v1 * (double)x1;
        s += (double)v2 * (double)x2;
        s += (double)v3 * (double)x3;
    }
    for (; k < m; k++) {
        const int c0 = __builtin_nontemporal_load(cp + (size_t)k * SLICE);
        const V v0 = __builtin_nontemporal_load(vp + (size_t)k * SLICE);
        s += (double)v0 * (double)x[c0];
    }
    return diag[r] * (double)x[r] - s;
}

// r = rhs - A x ; rhat = r ; partials rhs.rhs (A) and r.r (B)
template <class V, class X>
__global__ void __launch_bounds__(BLOCK) k_kr_init(int nr, const long long* sl_off, const int* cnt, const int* cols,
                                                   const V* vals, const double* diag, const X* rhs, const X* x,
                                                   X* r, X* rhat, double* pa, double* pb) {
    __shared__ double sh[2 * BLOCK];
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (row < nr) {
        double y = spmv_row<V, X>(row, nr, sl_off, cnt, cols, vals, diag, x);
        X rr = (X)((double)rhs[row] - y);
        r[row] = rr;
        rhat[row] = rr;
        a = (double)rhs[row] * (double)rhs[row];
        b = (double)rr * (double)rr;
    }
    double ta = 0.0, tb = 0.0;
    block_canon2(a, b, sh, &ta, &tb);
    if (threadIdx.x == 0) pa[blockIdx.x] = ta;
    if (threadIdx.x == 64) pb[blockIdx.x] = tb;
}

// p = r (it==1) or r + beta (p - omega v); phat = p / diag
template <class X>
__global__ void __launch_bounds__(BLOCK) k_kr_p(int nr, const double* sc, const X* r, const X* v, X* p, X* phat,
                                                const double* diag) {
    GATE(sc);
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nr) return;
    X pv;
    if (sc[SC_IT] == 1.0) pv = r[row];
    else pv = (X)((double)r[row] + sc[SC_BETA] * ((double)p[row] - sc[SC_OMEGA] * (double)v[row]));
    p[row] = pv;
    phat[row] = (X)((double)pv / diag[row]);
}

// y = A x ; partial d.y (A)   [and y.y (B) when two dots are needed]
template <class V, class X, bool TWO>
__global__ void __launch_bounds__(BLOCK) k_kr_spmv(int nr, const double* sc, const long long* sl_off, const int* cnt,
                                                   const int* cols, const V* vals, const double* diag, const X* x,
                                                   X* y, const X* d, double* pa, double* pb) {
    GATE(sc);
    __shared__ double sh[2 * BLOCK];
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (row < nr) {
        X yy = (X)spmv_row<V, X>(row, nr, sl_off, cnt, cols, vals, diag, x);
        y[row] = yy;
        if (TWO) {
            a = (double)yy * (double)yy;        // t.t
            b = (double)yy * (double)d[row];    // t.s
        } else {
            a = (double)d[row] * (double)yy;    // rhat.v
        }
    }
    if (TWO) {
        double ta = 0.0, tb = 0.0;
        block_canon2(a, b, sh, &ta, &tb);
        if (threadIdx.x == 0) pa[blockIdx.x] = ta;
        if (threadIdx.x == 64) pb[blockIdx.x] = tb;
    } else {
        double ta = block_canon(a, sh);
        if (threadIdx.x == 0) pa[blockIdx.x] = ta;
    }
}

// s = r - alpha v ; shat = s / diag ; partial s.s
template <class X>
__global__ void __launch_bounds__(BLOCK) k_kr_s(int nr, const double* sc, const X* r, const X* v, X* s, X* shat,
                                                const double* diag, double* pa) {
    GATE(sc);
    __shared__ double sh[BLOCK];
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (row < nr) {
        X sv = (X)((double)r[row] - sc[SC_ALPHA] * (double)v[row]);
        s[row] = sv;
        shat[row] = (X)((double)sv / diag[row]);
        a = (double)sv * (double)sv;
    }
    double ta = block_canon(a, sh);
    if (threadIdx.x == 0) pa[blockIdx.x] = ta;
}

// x = x + alpha phat + omega shat ; r = s - omega t ; partials r.r, rhat.r
// (after a half-step exit only x = x + alpha phat)
template <class X, class XS>
__global__ void __launch_bounds__(BLOCK) k_kr_xr(int nr, const double* sc, XS* x, const X* phat, const X* shat,
                                                 const X* s, const X* t, X* r, const X* rhat, double* pa, double* pb) {
    const bool half = sc[SC_HALF] == 1.0;
    if (sc[SC_DONE] != 0.0 && !half) return;
    __shared__ double sh[2 * BLOCK];
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (half) {
        if (row < nr) x[row] = (XS)((double)x[row] + sc[SC_ALPHA] * (double)phat[row]);
        return;
    }
    double a = 0.0, b = 0.0;
    if (row < nr) {
        x[row] = (XS)((double)x[row] + sc[SC_ALPHA] * (double)phat[row] + sc[SC_OMEGA] * (double)shat[row]);
        X rv = (X)((double)s[row] - sc[SC_OMEGA] * (double)t[row]);
        r[row] = rv;
        a = (double)rv * (double)rv;
        b = (double)rhat[row] * (double)rv;
    }
    double ta = 0.0, tb = 0.0;
    block_canon2(a, b, sh, &ta, &tb);
    if (threadIdx.x == 0) pa[blockIdx.x] = ta;
    if (threadIdx.x == 64) pb[blockIdx.x] = tb;
}

enum FinOp { FIN_INIT, FIN_ALPHA, FIN_S, FIN_OMEGA, FIN_R, FIN_NORM };
constexpr int FIN_BLOCK = 1024;
__global__ void __launch_bounds__(FIN_BLOCK) k_kr_fin(int op, long long nb, const double* pa, const double* pb,
                                                  double* sc) {
    __shared__ double lvl[256];
    __shared__ double resB;
    const bool half = sc[SC_HALF] == 1.0;
    if (op == FIN_R) {
        if (sc[SC_DONE] != 0.0 && !half) return;
    } else if (op != FIN_INIT && op != FIN_NORM) {
        if (sc[SC_DONE] != 0.0) return;
    }
    if (op == FIN_R && half) {
        if (threadIdx.x == 0) sc[SC_HALF] = 2.0;
        return;
    }
    double A = canon_partials(pa, nb, lvl);
    double B = 0.0;
    if (op == FIN_INIT || op == FIN_OMEGA || op == FIN_R || op == FIN_NORM) {
        B = canon_partials(pb, nb, lvl);
        if (threadIdx.x == 0) resB = B;
    }
    if (threadIdx.x != 0) return;
    const double tol = sc[SC_TOL];
    switch (op) {
    case FIN_INIT: {
        double bnorm = sqrt(A), rnorm = sqrt(B);
        sc[SC_BNORM] = bnorm; sc[SC_RNORM] = rnorm;
        sc[SC_RHO] = B; sc[SC_RHO_OLD] = 1.0; sc[SC_ALPHA] = 1.0; sc[SC_OMEGA] = 1.0; sc[SC_BETA] = 0.0;
        sc[SC_IT] = 1.0; sc[SC_HALF] = 0.0; sc[SC_ZEROX] = 0.0;
        sc[SC_DONE] = 0.0; sc[SC_CONV] = 0.0; sc[SC_ITERS] = 0.0;
        if (bnorm == 0.0) { sc[SC_DONE] = 1.0; sc[SC_CONV] = 1.0; sc[SC_ZEROX] = 1.0; sc[SC_RNORM] = 0.0; }
        else if (rnorm <= tol * bnorm) { sc[SC_DONE] = 1.0; sc[SC_CONV] = 1.0; }
        else if (B == 0.0) { sc[SC_DONE] = 1.0; sc[SC_ITERS] = 1.0; }
        break;
    }
    case FIN_ALPHA: sc[SC_ALPHA] = sc[SC_RHO] / A; break;
    case FIN_S: {
        double snorm = sqrt(A);
        sc[SC_SNORM] = snorm;
        if (snorm <= tol * sc[SC_BNORM]) {
            sc[SC_DONE] = 1.0; sc[SC_HALF] = 1.0; sc[SC_CONV] = 1.0;
            sc[SC_ITERS] = sc[SC_IT]; sc[SC_RNORM] = snorm;
        }
        break;
    }
    case FIN_OMEGA: sc[SC_OMEGA] = B / A; break;
    case FIN_R: {
        double rnorm = sqrt(A);
        sc[SC_RNORM] = rnorm;
        if (rnorm <= tol * sc[SC_BNORM]) { sc[SC_DONE] = 1.0; sc[SC_CONV] = 1.0; sc[SC_ITERS] = sc[SC_IT]; }
        else if (sc[SC_OMEGA] == 0.0) { sc[SC_DONE] = 1.0; sc[SC_ITERS] = sc[SC_IT]; }
        else {
            double rho_new = B;
            sc[SC_RHO_OLD] = sc[SC_RHO];
            sc[SC_RHO] = rho_new;
            sc[SC_IT] = sc[SC_IT] + 1.0;
            if (sc[SC_IT] > sc[SC_MAXIT]) { sc[SC_DONE] = 1.0; sc[SC_ITERS] = sc[SC_MAXIT]; }
            else if (rho_new == 0.0) { sc[SC_DONE] = 1.0; sc[SC_ITERS] = sc[SC_IT]; }
            else sc[SC_BETA] = (sc[SC_RHO] / sc[SC_RHO_OLD]) * (sc[SC_ALPHA] / sc[SC_OMEGA]);
        }
        break;
    }
    case FIN_NORM: sc[SC_RNORM] = sqrt(A); sc[SC_BNORM] = sqrt(B); break;
    }
}

// outer residual of the mixed solve: r = rhs - A x (double x, float values,
// double accumulation); partials r.r (A) and rhs.rhs (B); float copy of r
__global__ void __launch_bounds__(BLOCK) k_outer_res(int nr, const long long* sl_off, const int* cnt, const int* cols,
                                                     const float* vals, const double* diag, const double* rhs,
                                                     const double* x, float* r32, double* pa, double* pb) {
    __shared__ double sh[2 * BLOCK];
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (row < nr) {
        double y = spmv_row<float, double>(row, nr, sl_off, cnt, cols, vals, diag, x);
        double rr = rhs[row] - y;
        r32[row] = (float)rr;
        a = rr * rr;
        b = rhs[row] * rhs[row];
    }
    double ta = 0.0, tb = 0.0;
    block_canon2(a, b, sh, &ta, &tb);
    if (threadIdx.x == 0) pa[blockIdx.x] = ta;
    if (threadIdx.x == 64) pb[blockIdx.x] = tb;
}
__global__ void __launch_bounds__(BLOCK) k_add_correction(int nr, double* x, const float* e) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < nr) x[row] = x[row] + (double)e[row];
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_zero(long long n, T* a) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (T)0;
}

// ========================================================================
// host side
// ========================================================================
static KConst kconst(const sph_ctx* c) { return KConst{c->prm.kernel_type, c->prm.kappa, c->c_pi15}; }

void ppe_free(sph_ctx* c) {
    PPEState* P = c->ppe;
    if (!P) return;
    void* ptrs[] = {P->rowof, P->rows, P->cnt, P->sl_off, P->cols, P->vals64, P->vals32, P->diag, P->rhs, P->x,
                    P->pnew, P->r, P->rhat, P->p, P->phat, P->v, P->s, P->shat, P->t, P->fr, P->frhat, P->fp,
                    P->fphat, P->fv, P->fs, P->fshat, P->ft, P->fx, P->frhs, P->partA, P->partB};
    for (void* q : ptrs) dfree(c, q);
    delete P;
    c->ppe = nullptr;
}

static int ppe_build(sph_ctx* c) {
    ppe_free(c);
    PPEState* P = new PPEState();
    c->ppe = P;
    const int n = c->n;
    PView S = c->sview();
    int* flag = nullptr;
    int* scan = nullptr;
    SPH_TRY(dalloc(c, &flag, n));
    SPH_TRY(dalloc(c, &scan, n));
    hipLaunchKernelGGL(k_ppe_flag, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, S, flag);
    SPH_TRY(launch_check(c, "ppe_flag"));
    int nr = 0;
    SPH_TRY(exclusive_scan_i32(c, flag, scan, n, &nr));
    P->nr = nr;
    SPH_TRY(dalloc(c, &P->rowof, n));
    SPH_TRY(dalloc(c, &P->rows, nr + 1));
    hipLaunchKernelGGL(k_ppe_rows, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)flag, (const int*)scan,
                       P->rowof, P->rows);
    SPH_TRY(launch_check(c, "ppe_rows"));
    dfree(c, flag);
    dfree(c, scan);
    if (nr == 0) return 0;
    P->nslices = (nr + SLICE - 1) / SLICE;
    SPH_TRY(dalloc(c, &P->cnt, nr));
    SPH_TRY(dalloc(c, &P->sl_off, P->nslices + 1));
    hipLaunchKernelGGL(k_ppe_struct<false>, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows,
                       (const int*)P->rowof, S, c->nlist(), P->cnt, (const long long*)nullptr, (int*)nullptr);
    SPH_TRY(launch_check(c, "ppe_count"));
    long long* w = nullptr;
    SPH_TRY(dalloc(c, &w, P->nslices + 1));
    hipLaunchKernelGGL(k_row_slice_width, dim3(grid1(P->nslices)), dim3(BLOCK), 0, c->stream, nr, P->nslices,
                       (const int*)P->cnt, w);
    SPH_TRY(launch_check(c, "ppe_width"));
    SPH_TRY(exclusive_scan_i64(c, w, P->sl_off, P->nslices, &P->nnz_slots));
    dfree(c, w);
    SPH_TRY(dalloc(c, &P->cols, P->nnz_slots + 1));
    hipLaunchKernelGGL(k_zero<int>, dim3(grid1(P->nnz_slots + 1)), dim3(BLOCK), 0, c->stream, P->nnz_slots + 1, P->cols);
    hipLaunchKernelGGL(k_ppe_struct<true>, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows,
                       (const int*)P->rowof, S, c->nlist(), P->cnt, (const long long*)P->sl_off, P->cols);
    SPH_TRY(launch_check(c, "ppe_cols"));
    // real nnz (for reports)
    std::vector<int> hc(nr);
    SPH_HIP(c, hipMemcpyAsync(hc.data(), P->cnt, sizeof(int) * nr, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    P->nnz = 0;
    for (int v : hc) P->nnz += v;
    const bool mixed = c->prm.precision == SPH_MIXED;
    if (mixed) SPH_TRY(dalloc(c, &P->vals32, P->nnz_slots + 1));
    else SPH_TRY(dalloc(c, &P->vals64, P->nnz_slots + 1));
    SPH_TRY(dalloc(c, &P->diag, nr));
    SPH_TRY(dalloc(c, &P->rhs, nr));
    SPH_TRY(dalloc(c, &P->x, nr));
    SPH_TRY(dalloc(c, &P->pnew, nr));
    P->nb = (nr + BLOCK - 1) / BLOCK;
    if (P->nb > 65536) return fail(c, SPH_E_UNSUPPORTED, "PPE larger than 16.7M unknowns");
    SPH_TRY(dalloc(c, &P->partA, P->nb));
    SPH_TRY(dalloc(c, &P->partB, P->nb));
    if (c->prm.ppe_solver == SPH_PPE_BICGSTAB) {
        if (mixed) {
            float** fv[] = {&P->fr, &P->frhat, &P->fp, &P->fphat, &P->fv, &P->fs, &P->fshat, &P->ft, &P->fx, &P->frhs};
            for (auto q : fv) SPH_TRY(dalloc(c, q, nr));
        } else {
            double** dv[] = {&P->r, &P->rhat, &P->p, &P->phat, &P->v, &P->s, &P->shat, &P->t};
            for (auto q : dv) SPH_TRY(dalloc(c, q, nr));
        }
    }
    return 0;
}

static int read_scalars(sph_ctx* c) {
    SPH_HIP(c, hipMemcpyAsync(c->hscal, c->dscal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

#define KF(op)                                                                                     \
    do {                                                                                           \
        SPH_TRY(prof_begin(c, "kr_fin", 0.0));                                                     \
        hipLaunchKernelGGL(k_kr_fin, dim3(1), dim3(FIN_BLOCK), 0, c->stream, (int)(op), P->nb,      \
                           (const double*)P->partA, (const double*)P->partB, sc);                  \
        SPH_TRY(launch_check(c, "kr_fin"));                                                        \
        SPH_TRY(prof_end(c));                                                                      \
    } while (0)
#define KL(name, bytes, grid, kern, ...)                                              \
    do {                                                                             \
        SPH_TRY(prof_begin(c, name, (double)(bytes)));                               \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, c->stream, __VA_ARGS__); \
        SPH_TRY(launch_check(c, name));                                              \
        SPH_TRY(prof_end(c));                                                        \
    } while (0)

// BiCGSTAB on A q = rhs (X vectors, V values), solution in xs (XS type).
// Returns iterations; sets converged / residual in the host scalar mirror.
template <class V, class X, class XS>
static int bicgstab(sph_ctx* c, const V* vals, const X* rhs, XS* xs, X* xinit, X* r, X* rhat, X* p, X* phat, X* v,
                    X* s, X* shat, X* t, double tol, int maxit) {
    PPEState* P = c->ppe;
    const int nr = P->nr;
    const unsigned g = grid1(nr);
    double* sc = c->dscal;
    const double nnzb = (double)P->nnz * (4.0 + sizeof(V)) + (double)nr * (8.0 + 2.0 * sizeof(X));
    // scalars: tol, maxit
    c->hscal[SC_TOL] = tol;
    c->hscal[SC_MAXIT] = (double)maxit;
    SPH_HIP(c, hipMemcpyAsync(sc + SC_TOL, c->hscal + SC_TOL, sizeof(double), hipMemcpyHostToDevice, c->stream));
    SPH_HIP(c, hipMemcpyAsync(sc + SC_MAXIT, c->hscal + SC_MAXIT, sizeof(double), hipMemcpyHostToDevice, c->stream));
    KL("kr_init", nnzb, g, (k_kr_init<V, X>), nr, (const long long*)P->sl_off, (const int*)P->cnt,
       (const int*)P->cols, vals, (const double*)P->diag, rhs, (const X*)xinit, r, rhat, P->partA, P->partB);
    KF(FIN_INIT);
    SPH_TRY(read_scalars(c));
    if (c->hscal[SC_DONE] != 0.0) {
        if (c->hscal[SC_ZEROX] != 0.0)
            hipLaunchKernelGGL(k_zero<XS>, dim3(g), dim3(BLOCK), 0, c->stream, (long long)nr, xs);
        return 0;
    }
    int launched = 0;
    int batch = std::max(2, std::min(32, P->last_iters / 3));
    while (launched < maxit) {
        for (int b = 0; b < batch && launched < maxit; b++, launched++) {
            KL("kr_p", 0, g, k_kr_p<X>, nr, (const double*)sc, (const X*)r, (const X*)v, p, phat,
               (const double*)P->diag);
            KL("spmv", nnzb, g, (k_kr_spmv<V, X, false>), nr, (const double*)sc, (const long long*)P->sl_off,
               (const int*)P->cnt, (const int*)P->cols, vals, (const double*)P->diag, (const X*)phat, v,
               (const X*)rhat, P->partA, P->partB);
            KF(FIN_ALPHA);
            KL("kr_s", 0, g, k_kr_s<X>, nr, (const double*)sc, (const X*)r, (const X*)v, s, shat,
               (const double*)P->diag, P->partA);
            KF(FIN_S);
            KL("spmv", nnzb, g, (k_kr_spmv<V, X, true>), nr, (const double*)sc, (const long long*)P->sl_off,
               (const int*)P->cnt, (const int*)P->cols, vals, (const double*)P->diag, (const X*)shat, t,
               (const X*)s, P->partA, P->partB);
            KF(FIN_OMEGA);
            KL("kr_xr", 0, g, (k_kr_xr<X, XS>), nr, (const double*)sc, xs, (const X*)phat, (const X*)shat,
               (const X*)s, (const X*)t, r, (const X*)rhat, P->partA, P->partB);
            KF(FIN_R);
        }
        SPH_TRY(read_scalars(c));
        if (c->hscal[SC_DONE] != 0.0) break;
        batch = std::min(32, batch * 2);
    }
    if (c->hscal[SC_DONE] == 0.0) {   // launched maxit iterations without the flag
        c->hscal[SC_ITERS] = maxit;
    }
    P->last_iters = std::max(4, (int)c->hscal[SC_ITERS]);
    return 0;
}

int ppe_solve(sph_ctx* c, double dt) {
    if (!c->ppe) SPH_TRY(ppe_build(c));
    PPEState* P = c->ppe;
    const int nr = P->nr;
    c->rep = sph_ppe_report{};
    c->rep.unknowns = nr;
    c->rep.nnz = P->nnz;
    if (nr == 0) { c->rep.converged = 1; return 0; }
    const sph_params& pr = c->prm;
    const bool mixed = pr.precision == SPH_MIXED;
    const bool jacobi = pr.ppe_solver == SPH_PPE_REF_JACOBI;
    PView S = c->sview();
    QView Q = c->qview();
    NList nl = c->nlist();
    KConst kc = kconst(c);
    const unsigned g = grid1(nr);
    const double asm_bytes = (double)nr * (15.0 * 8) + (double)P->nnz * (4.0 + (mixed ? 4.0 : 8.0));
    if (mixed)
        KL("ppe_assemble", asm_bytes, g, (k_ppe_assemble<float, float>), nr, (const int*)P->rows, (const int*)P->rowof,
           kc, S, Q, nl, (const long long*)P->sl_off, P->vals32, P->diag, P->rhs, P->x, P->pnew, dt, jacobi ? 1 : 0);
    else
        KL("ppe_assemble", asm_bytes, g, (k_ppe_assemble<double, double>), nr, (const int*)P->rows,
           (const int*)P->rowof, kc, S, Q, nl, (const long long*)P->sl_off, P->vals64, P->diag, P->rhs, P->x,
           P->pnew, dt, jacobi ? 1 : 0);
    if (jacobi) {
        KL("ppe_scatter", 0, g, k_ppe_scatter, nr, (const int*)P->rows, (const double*)S.pres0,
           (const double*)P->pnew, 0, S.pres);
        c->rep.iterations = 1;
        c->rep.converged = 1;
        c->rep.rel_residual = -1.0;
        return 0;
    }
    int rc = 0;
    if (!mixed) {
        SPH_TRY((bicgstab<double, double, double>(c, P->vals64, P->rhs, P->x, P->x, P->r, P->rhat, P->p, P->phat,
                                                   P->v, P->s, P->shat, P->t, pr.ppe_tol, pr.ppe_maxit)));
        c->rep.iterations = (int)c->hscal[SC_ITERS];
        c->rep.converged = c->hscal[SC_CONV] != 0.0;
        c->rep.rel_residual = c->hscal[SC_BNORM] > 0 ? c->hscal[SC_RNORM] / c->hscal[SC_BNORM] : 0.0;
    } else {
        double* sc = c->dscal;
        int total_it = 0, refine = 0;
        bool conv = false;
        double rel = 0.0;
        for (refine = 0; refine <= pr.ppe_max_refine; refine++) {
            KL("outer_res", (double)P->nnz * 8.0 + (double)nr * 28.0, g, k_outer_res, nr, (const long long*)P->sl_off,
               (const int*)P->cnt, (const int*)P->cols, (const float*)P->vals32, (const double*)P->diag,
               (const double*)P->rhs, (const double*)P->x, P->frhs, P->partA, P->partB);
            KF(FIN_NORM);
            SPH_TRY(read_scalars(c));
            double rn = c->hscal[SC_RNORM], bn = c->hscal[SC_BNORM];
            rel = bn > 0 ? rn / bn : 0.0;
            if (rn <= pr.ppe_tol * bn) { conv = true; break; }
            if (refine == pr.ppe_max_refine) break;
            hipLaunchKernelGGL(k_zero<float>, dim3(g), dim3(BLOCK), 0, c->stream, (long long)nr, P->fx);
            // aim the correction at the final absolute tolerance, but not
            // below what a float Krylov solve reliably attains
            double inner = std::max(pr.ppe_inner_tol, 0.5 * pr.ppe_tol * bn / rn);
            SPH_TRY((bicgstab<float, float, float>(c, P->vals32, P->frhs, P->fx, P->fx, P->fr, P->frhat, P->fp,
                                                    P->fphat, P->fv, P->fs, P->fshat, P->ft, inner, pr.ppe_maxit)));
            total_it += (int)c->hscal[SC_ITERS];
            KL("add_corr", 0, g, k_add_correction, nr, P->x, (const float*)P->fx);
        }
        c->rep.iterations = total_it;
        c->rep.refinements = refine;
        c->rep.converged = conv;
        c->rep.rel_residual = rel;
    }
    KL("ppe_scatter", 0, g, k_ppe_scatter, nr, (const int*)P->rows, (const double*)S.pres0, (const double*)P->x, 1,
       S.pres);
    if (!c->rep.converged) rc = fail(c, SPH_E_NOCONV, "PPE did not converge within ppe_maxit");
    return rc;
}

int reduce_finalize(sph_ctx* c, const double* part, long long nb, double* dst) {
    (void)c; (void)part; (void)nb; (void)dst;
    return 0;
}

}  // namespace sph
