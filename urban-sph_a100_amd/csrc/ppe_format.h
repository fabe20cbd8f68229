// Sliced-ELLPACK storage of the PPE operator and its row product, shared
// by the Krylov kernels and the multigrid preconditioner.
#pragma once
#include "sph_internal.h"

namespace sph {

// V-cycle weights of the multigrid preconditioner: damped Jacobi smoothing
// x = w b / diag, and the over-correction of unsmoothed aggregation
constexpr double MG_OMEGA = 0.9;
constexpr double MG_SCALE = 1.5;

// 16-bit column entries: 10 offset bits, up to 64 window groups per slice
constexpr int EC_OFF_BITS = 10;
constexpr int EC_OFF_MASK = (1 << EC_OFF_BITS) - 1;
constexpr int EC_MAX_GROUPS = 64;
constexpr int EC_MAX_WORDS = (65536 * 256 >> 10) >> 5;   // window bitmap words for 16.7M unknowns

// Matrix layout: sliced ELLPACK over 64-row slices, one wave64 per slice,
// entries column-major within a slice: entry k of row r lives at
//     sl_off[r / 64] + k * 64 + r % 64,
// so the wave's loads of one slot are one contiguous block.
__host__ __device__ __forceinline__ long long prow(const long long* sl_off, int r) {
    return sl_off[r >> 6] + (r & 63);
}
__host__ __device__ __forceinline__ long long pslot(int k) { return (long long)k * SLICE; }

// Off-diagonal part of one row, sum_k vals[k] * x[col[k]], accumulated in
// double in entry order.  Matrix entries stream through once: nontemporal
// loads keep the caches for the x gathers; four independent loads in flight
// per lane.
template <class V, class X>
__device__ __forceinline__ double offdiag_i32(const int* cp, const V* vp, int m, const X* x) {
    double s = 0.0;
    int k = 0;
    for (; k + 4 <= m; k += 4) {
        const int c0 = __builtin_nontemporal_load(cp + pslot(k + 0));
        const int c1 = __builtin_nontemporal_load(cp + pslot(k + 1));
        const int c2 = __builtin_nontemporal_load(cp + pslot(k + 2));
        const int c3 = __builtin_nontemporal_load(cp + pslot(k + 3));
        const V v0 = __builtin_nontemporal_load(vp + pslot(k + 0));
        const V v1 = __builtin_nontemporal_load(vp + pslot(k + 1));
        const V v2 = __builtin_nontemporal_load(vp + pslot(k + 2));
        const V v3 = __builtin_nontemporal_load(vp + pslot(k + 3));
        const X x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
        s += (double)v0 * (double)x0;
        s += (double)v1 * (double)x1;
        s += (double)v2 * (double)x2;
        s += (double)v3 * (double)x3;
    }
    for (; k < m; k++) s += (double)__builtin_nontemporal_load(vp + pslot(k)) * (double)x[__builtin_nontemporal_load(cp + pslot(k))];
    return s;
}
// The same over 16-bit columns: entry e = (group << EC_OFF_BITS) | offset,
// column = wb[group] + offset, wb = the slice's group bases (in LDS).
__device__ __forceinline__ int ecol(const int* wb, int e) { return wb[e >> EC_OFF_BITS] + (e & EC_OFF_MASK); }
template <class V, class X>
__device__ __forceinline__ double offdiag_u16(const unsigned short* ep, const V* vp, int m, const int* wb,
                                              const X* x) {
    double s = 0.0;
    int k = 0;
    for (; k + 4 <= m; k += 4) {
        const int e0 = __builtin_nontemporal_load(ep + pslot(k + 0));
        const int e1 = __builtin_nontemporal_load(ep + pslot(k + 1));
        const int e2 = __builtin_nontemporal_load(ep + pslot(k + 2));
        const int e3 = __builtin_nontemporal_load(ep + pslot(k + 3));
        const V v0 = __builtin_nontemporal_load(vp + pslot(k + 0));
        const V v1 = __builtin_nontemporal_load(vp + pslot(k + 1));
        const V v2 = __builtin_nontemporal_load(vp + pslot(k + 2));
        const V v3 = __builtin_nontemporal_load(vp + pslot(k + 3));
        const X x0 = x[ecol(wb, e0)], x1 = x[ecol(wb, e1)], x2 = x[ecol(wb, e2)], x3 = x[ecol(wb, e3)];
        s += (double)v0 * (double)x0;
        s += (double)v1 * (double)x1;
        s += (double)v2 * (double)x2;
        s += (double)v3 * (double)x3;
    }
    for (; k < m; k++)
        s += (double)__builtin_nontemporal_load(vp + pslot(k)) * (double)x[ecol(wb, __builtin_nontemporal_load(ep + pslot(k)))];
    return s;
}

// The same over a stencil slice: column row + off[k], off the slice's
// ascending offsets, so a slot's x gathers are one contiguous run of the
// wave's rows; the stored zeros add exact zeros.  The offsets are the same
// for the whole wave and constant during the launch: read through the
// constant address space they become scalar loads.  Values and x through the
// global address space.
template <class T> using gmem = const __attribute__((address_space(1))) T*;
template <class T> using cmem = const __attribute__((address_space(4))) T*;
template <class V, class X>
__device__ __forceinline__ double offdiag_stencil(const V* vals, int m, const int* offs, int r, const X* xv) {
    const gmem<V> vp = (gmem<V>)vals;
    const gmem<X> x = (gmem<X>)(xv + r);
    const cmem<int> off = (cmem<int>)offs;
    double s = 0.0;
    int k = 0;
    for (; k + 4 <= m; k += 4) {
        const int o0 = off[k + 0], o1 = off[k + 1], o2 = off[k + 2], o3 = off[k + 3];
        const V v0 = __builtin_nontemporal_load(vp + pslot(k + 0));
        const V v1 = __builtin_nontemporal_load(vp + pslot(k + 1));
        const V v2 = __builtin_nontemporal_load(vp + pslot(k + 2));
        const V v3 = __builtin_nontemporal_load(vp + pslot(k + 3));
        const X x0 = x[o0], x1 = x[o1], x2 = x[o2], x3 = x[o3];
        s += (double)v0 * (double)x0;
        s += (double)v1 * (double)x1;
        s += (double)v2 * (double)x2;
        s += (double)v3 * (double)x3;
    }
    for (; k < m; k++) s += (double)__builtin_nontemporal_load(vp + pslot(k)) * (double)x[off[k]];
    return s;
}
// Row r's slice: the first of its stencil offsets in pp.poff, or -1.  Every
// active lane of the calling wave holds a row of the same slice (rows are
// handed out in whole slices), so the value is read as wave-uniform.
__device__ __forceinline__ int stencil_of(const PPEView& pp, int r) {
    return pp.spat ? __builtin_amdgcn_readfirstlane(pp.spat[r >> 6]) : -1;
}

// Bitonic sort of a wave's Q * 64 keys held in registers, key t = lane +
// 64 q ascending: partners in other lanes through cross-lane moves, partners
// in the same lane (distance 64 and more) by register.
template <class K, int Q>
__device__ __forceinline__ void bitonic_wave(K (&v)[Q], int lane) {
    constexpr int W = Q * SLICE;
#pragma unroll
    for (int k = 2; k <= W; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= SLICE) {
                const int qj = j / SLICE;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    if ((q & qj) == 0) {
                        const bool up = ((lane + q * SLICE) & k) == 0;
                        const K a = v[q], b = v[q | qj];
                        const bool sw = (a > b) == up;
                        v[q] = sw ? b : a;
                        v[q | qj] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const K o = __shfl_xor(v[q], j, SLICE);
                    const bool up = ((lane + q * SLICE) & k) == 0;
                    const bool take_min = ((lane & j) == 0) == up;
                    v[q] = take_min ? (o < v[q] ? o : v[q]) : (o > v[q] ? o : v[q]);
                }
            }
        }
    }
}
// Group bases of the block's slices into LDS (every thread of the block
// takes part; call before the row guard).
__device__ __forceinline__ void load_group_bases(const PPEView& pp, int first, int* wbase) {
    if (pp.ecols) wbase[threadIdx.x] = pp.gbase[first + (int)threadIdx.x];
    __syncthreads();
}

// y_r = diag_r x_r - offdiag_r(x); dg = the diagonal of this operator
template <class V, class X, class DG>
__device__ __forceinline__ double spmv_row(const PPEView& pp, int r, const V* vals, const DG* dg, const X* x,
                                           const int* wbase) {
    const long long base = prow(pp.sl_off, r);
    const int st = stencil_of(pp, r);
    double s;
    if (st >= 0) {
        // a stencil slice's rows all hold its offsets: a uniform count
        const int m = __builtin_amdgcn_readfirstlane(pp.cnt[r]);
        s = offdiag_stencil<V, X>(vals + base, m, pp.poff + st, r, x);
    } else {
        const int m = pp.cnt[r];
        if (pp.ecols) s = offdiag_u16<V, X>(pp.ecols + base, vals + base, m, wbase + (threadIdx.x & ~63), x);
        else s = offdiag_i32<V, X>(pp.cols + base, vals + base, m, x);
    }
    if (pp.ltail) {
        const int u = pp.ltail[r];
        if (u >= 0) s += pp.lsum[u];
    }
    return (double)dg[r] * (double)x[r] - s;
}

}  // namespace sph
