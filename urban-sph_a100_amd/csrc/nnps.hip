// Neighbour search of the count==0 step on MI355X (ISPH_Calc.cuh:23-52):
//   KERNEL_index_particle_to_cell  functions_NNPS.cuh:879-907  -> k_hash
//   cub::DeviceRadixSort::SortPairs ISPH_Calc.cuh:33           -> radix_sort
//   KERNEL_reorder (start/end)     functions_NNPS.cuh:813-838  -> k_cell_bounds
//   KERNEL_reorder (gather)        functions_NNPS.cuh:840-843  -> permute
// plus the per-particle neighbour list that the per-step passes walk
// (the reference rescans the cell window in every pass).
//
// The sort is a stable least-significant-digit radix sort over only the
// ceil(log2(num_cells+1)) key bits, 8 bits per pass.  Each 256-thread tile
// of 2048 keys ranks its keys with wave64 ballots (a match-any over the
// digit bits) and LDS per-wave digit counters, so equal digits keep input
// order and the result equals a stable sort.
#include "flags.h"
#include <climits>

#include "sph_internal.h"
#include "sph_math.h"

namespace sph {

// ------------------------------------------------------------------------
// KERNEL_set_ncell, functions_NNPS.cuh:943-955
__global__ void __launch_bounds__(BLOCK) k_set_ncell(int n, const int* i_type, const double* h, const double* h_ref,
                                                     int* ncell) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i_type[i] == 3) return;
    ncell[i] = (int)(1.0 * ceilf(2.0f / (float)(h_ref[i] / h[i])));
}

// KERNEL_index_particle_to_cell, functions_NNPS.cuh:879-907 (float binning)
__global__ void __launch_bounds__(BLOCK) k_hash(int n, sph_domain d, const int* i_type, const double* x,
                                                const double* y, const double* z, int* keys, int* vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int key;
    if (i_type[i] == 3) key = d.num_cells;
    else {
        Cell c = bin_cell(d, x[i], y[i], z[i]);
        key = c.i + d.NI * c.j + d.NI * d.NJ * c.k;
    }
    keys[i] = key;
    vals[i] = i;
}

// ------------------------------------------------------------------------
// radix sort
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = BLOCK * RS_ITEMS;

__global__ void __launch_bounds__(BLOCK) k_rs_hist(const int* keys, int n, int shift, int ntiles, int* hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    long long base = (long long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        long long idx = base + k * BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(BLOCK) k_rs_scatter(const int* kin, const int* vin, int* kout, int* vout, int n,
                                                      int shift, int ntiles, const int* offs) {
    __shared__ int cnt[256];
    __shared__ int wc[BLOCK / 64][256];
    __shared__ int boff[256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    cnt[tid] = 0;
    boff[tid] = offs[tid * ntiles + blockIdx.x];
#pragma unroll
    for (int q = 0; q < BLOCK / 64; q++) wc[q][tid] = 0;
    __syncthreads();
    long long base = (long long)blockIdx.x * RS_TILE;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = 0; k < RS_ITEMS; k++) {
        long long idx = base + k * BLOCK + tid;
        bool valid = idx < n;
        int key = valid ? kin[idx] : 0;
        int val = valid ? vin[idx] : 0;
        int dg = valid ? ((key >> shift) & 255) : 256;
        unsigned long long peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 9; b++) {
            int bit = (dg >> b) & 1;
            unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        int wrank = __popcll(peers & lt);
        int wcount = __popcll(peers);
        int leader = __ffsll((long long)peers) - 1;
        if (valid && lane == leader) wc[w][dg] = wcount;
        __syncthreads();
        if (valid) {
            int pre = cnt[dg];
            for (int q = 0; q < w; q++) pre += wc[q][dg];
            int pos = boff[dg] + pre + wrank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int q = 0; q < BLOCK / 64; q++) { add += wc[q][tid]; wc[q][tid] = 0; }
        cnt[tid] += add;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// exclusive scans (device-wide), 2048 elements per 256-thread block
template <class T>
__global__ void __launch_bounds__(BLOCK) k_scan_block(const T* in, T* out, long long n, T* sums) {
    __shared__ T sh[BLOCK];
    const int t = threadIdx.x;
    long long base = (long long)blockIdx.x * RS_TILE + (long long)t * RS_ITEMS;
    T v[RS_ITEMS];
    T s = 0;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        long long idx = base + k;
        v[k] = idx < n ? in[idx] : (T)0;
        s += v[k];
    }
    sh[t] = s;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1) {
        T a = t >= off ? sh[t - off] : (T)0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    T run = sh[t] - s;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        long long idx = base + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
    if (t == BLOCK - 1) sums[blockIdx.x] = sh[t];
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_scan_add(T* out, long long n, const T* offs) {
    long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) out[idx] += offs[idx / RS_TILE];
}

template <class T>
static int exclusive_scan(sph_ctx* c, const T* in, T* out, long long n, T* total_out) {
    long long nb = (n + RS_TILE - 1) / RS_TILE;
    if (nb < 1) nb = 1;
    T* sums = nullptr;
    T* sums_scan = nullptr;
    SPH_TRY(dalloc(c, &sums, nb + 1));
    hipLaunchKernelGGL(k_scan_block<T>, dim3((unsigned)nb), dim3(BLOCK), 0, c->stream, in, out, n, sums);
    SPH_TRY(launch_check(c, "scan_block"));
    T total = 0;
    // the total goes to the host (a wait for the stream) only when asked for
    if (nb > 1) {
        SPH_TRY(dalloc(c, &sums_scan, nb + 1));
        SPH_TRY(exclusive_scan<T>(c, sums, sums_scan, nb, total_out ? &total : nullptr));
        hipLaunchKernelGGL(k_scan_add<T>, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, out, n, (const T*)sums_scan);
        SPH_TRY(launch_check(c, "scan_add"));
        dfree(c, sums_scan);
    } else if (total_out) {
        SPH_HIP(c, hipMemcpyAsync(&total, sums, sizeof(T), hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(c, hipStreamSynchronize(c->stream));
    }
    dfree(c, sums);
    if (total_out) *total_out = total;
    return 0;
}
int exclusive_scan_i32(sph_ctx* c, const int* in, int* out, long long n, int* total) {
    return exclusive_scan<int>(c, in, out, n, total);
}
int exclusive_scan_i64(sph_ctx* c, const long long* in, long long* out, long long n, long long* total) {
    return exclusive_scan<long long>(c, in, out, n, total);
}

// Stable LSD radix sort of (key, value) pairs, keys in [0, max_key]; the
// sorted pairs end in *keys / *vals (the buffers may be swapped with the
// temporaries).
int radix_sort_pairs(sph_ctx* c, int** keys, int** vals, int** ktmp, int** vtmp, int n, int max_key) {
    int bits = 1;
    while ((1LL << bits) <= (long long)max_key) bits++;
    int passes = (bits + 7) / 8;
    int ntiles = (n + RS_TILE - 1) / RS_TILE;
    int hist_n = 256 * ntiles;
    int* hist = nullptr;
    int* offs = nullptr;
    SPH_TRY(dalloc(c, &hist, hist_n));
    SPH_TRY(dalloc(c, &offs, hist_n));
    int *kin = *keys, *vin = *vals, *kout = *ktmp, *vout = *vtmp;
    for (int p = 0; p < passes; p++) {
        int shift = 8 * p;
        SPH_TRY(prof_begin(c, "radix_hist", 4.0 * n));
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(BLOCK), 0, c->stream, (const int*)kin, n, shift, ntiles, hist);
        SPH_TRY(launch_check(c, "radix_hist"));
        SPH_TRY(prof_end(c));
        SPH_TRY(exclusive_scan_i32(c, hist, offs, hist_n, nullptr));
        SPH_TRY(prof_begin(c, "radix_scatter", 16.0 * n));
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(BLOCK), 0, c->stream, (const int*)kin, (const int*)vin,
                           kout, vout, n, shift, ntiles, (const int*)offs);
        SPH_TRY(launch_check(c, "radix_scatter"));
        SPH_TRY(prof_end(c));
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != *keys) {   // odd pass count: result is in the tmp buffers
        std::swap(*keys, *ktmp);
        std::swap(*vals, *vtmp);
    }
    dfree(c, hist);
    dfree(c, offs);
    return 0;
}
static int radix_sort(sph_ctx* c, int max_key) {
    return radix_sort_pairs(c, &c->keys, &c->perm, &c->keys_tmp, &c->perm_tmp, c->n, max_key);
}

// ------------------------------------------------------------------------
// cell start / end, KERNEL_reorder functions_NNPS.cuh:825-838
__global__ void __launch_bounds__(BLOCK) k_fill_i32(int* a, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void __launch_bounds__(BLOCK) k_cell_bounds(int n, const int* keys, int* gstr, int* gend) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int h = keys[i];
    if (i == 0 || h != keys[i - 1]) {
        gstr[h] = i;
        if (i > 0) gend[keys[i - 1]] = i;
    }
    if (i == n - 1) gend[h] = i + 1;
}

// gather one array through the permutation: out[i] = in[perm[i]]
template <class T>
__global__ void __launch_bounds__(BLOCK) k_gather(int n, const int* perm, const T* in, T* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}

// First particle of the sorted order whose key is >= cell, for every cell
// and one past the last: the particles of consecutive cells a..b are
// [cstart[a], cstart[b + 1]), in the order the cells hold them.
__global__ void __launch_bounds__(BLOCK) k_cell_start(int ncs, const int* keys, int n, int* cstart) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncs) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < cell) lo = mid + 1;
        else hi = mid;
    }
    cstart[cell] = lo;
}

// ------------------------------------------------------------------------
// Neighbour list: the union of the dcell-query window and the binning
// window of particle i (both +-ncell), scanned in increasing cell index so
// entries come out in ascending j, each tagged with the flags of flags.h.
// One 64-lane wave per particle (particle = global thread index / 64).  The
// window's x-rows are numbered z-major then y (lane k describes row k of a
// batch of 64), their particle ranges are flattened into one candidate
// order v, and each round of 64 candidates is compacted in order with a
// ballot, so the list is the one a serial scan in cell order produces.
template <bool FILL>
__device__ __forceinline__ int nl_wave(int i, const Grid& g, const double* x, const double* y, const double* z,
                                       const double* h, const int* ncell, double kappa, unsigned* out, int cap) {
    const int lane = threadIdx.x & 63;
    const sph_domain& d = g.d;
    const double xi = x[i], yi = y[i], zi = z[i], hi = h[i];
    const int nc = ncell[i];
    const Cell q = query_cell(d, xi, yi, zi);
    const Cell b = bin_cell(d, xi, yi, zi);
    const int jlo = max(min(q.j, b.j) - nc, 0), jhi = min(max(q.j, b.j) + nc, d.NJ - 1);
    const int klo = max(min(q.k, b.k) - nc, 0), khi = min(max(q.k, b.k) + nc, d.NK - 1);
    const int nj = max(jhi - jlo + 1, 0), nk = max(khi - klo + 1, 0), nrows = nj * nk;
    const double khi_r = kappa * hi, khalf = kappa * (0.5 * hi);
    int cnt = 0;
    for (int r0 = 0; r0 < nrows; r0 += 64) {
        // lane k: row r0 + k of the union window
        int rs = 0, rl = 0, base = 0, flags = 0;
        const int rk = r0 + lane;
        if (rk < nrows) {
            const int ck = klo + rk / nj, cj = jlo + rk % nj;
            const bool rowD = abs(cj - q.j) <= nc && abs(ck - q.k) <= nc;
            const bool rowB = abs(cj - b.j) <= nc && abs(ck - b.k) <= nc;
            int i0 = INT_MAX, i1 = INT_MIN;
            if (rowD) { i0 = min(i0, q.i - nc); i1 = max(i1, q.i + nc); }
            if (rowB) { i0 = min(i0, b.i - nc); i1 = max(i1, b.i + nc); }
            i0 = max(i0, 0);
            i1 = min(i1, d.NI - 1);
            base = d.NI * cj + d.NI * d.NJ * ck;
            flags = (rowD ? 1 : 0) | (rowB ? 2 : 0);
            if ((rowD || rowB) && i0 <= i1) {
                rs = g.cstart[base + i0];
                rl = g.cstart[base + i1 + 1] - rs;
            }
        }
        int incl = rl;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        const int excl = incl - rl, total = __shfl(incl, 63, 64);
        const int nr_b = min(64, nrows - r0);
        for (int v0 = 0; v0 < total; v0 += 64) {
            const int v = v0 + lane, vv = min(v, total - 1);
            int lo = 0;   // the row holding vv: the last row whose prefix is <= vv
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const int cand = lo + step;
                const int pe = __shfl(excl, min(cand, 63), 64);
                if (cand < nr_b && pe <= vv) lo = cand;
            }
            const int start = __shfl(rs, lo, 64), pre = __shfl(excl, lo, 64);
            const int rbase = __shfl(base, lo, 64), rflags = __shfl(flags, lo, 64);
            unsigned f = 0;
            int j = -1;
            if (v < total) {
                j = start + (vv - pre);
                const int ci = g.keys[j] - rbase;
                const bool inD = (rflags & 1) && abs(ci - q.i) <= nc;
                const bool inB = (rflags & 2) && abs(ci - b.i) <= nc;
                if (inD || inB) {
                    double r = dist3<double>(xi, yi, zi, x[j], y[j], z[j]);
                    double kmax = kappa * fmax(hi, h[j]);
                    double tp = r + 1e-20, tm = r - 1e-20;
                    if (inD) {
                        if (tp < kmax) f |= F_DP;
                        if (tm < kmax) f |= F_DM;
                        if (tp < khi_r) f |= F_HP;
                        if (tm < khi_r) f |= F_HM;
                    }
                    if (inB) {
                        if (r > 0 && r < khi_r) f |= F_B;
                        if (tp < khalf) f |= F_BH;
                    }
                }
            }
            const unsigned long long m = __ballot(f != 0);
            if (FILL && f) {
                const int below = __popcll(m & ((1ULL << lane) - 1ULL));
                if (cnt + below < cap) out[(size_t)(cnt + below) * SLICE] = (unsigned)j | f;
            }
            cnt += __popcll(m);
        }
    }
    return cnt;
}
__global__ void __launch_bounds__(BLOCK) k_nl_count(int n, Grid g, const double* x, const double* y, const double* z,
                                                    const double* h, const int* ncell, double kappa, int* cnt) {
    // xcd_block: a list slice's 64 particles on one XCD, so their
    // interleaved entries fill whole lines in one L2
    const int i = (int)(((long long)xcd_block() * blockDim.x + threadIdx.x) >> 6);
    if (i >= n) return;   // whole waves
    const int m = nl_wave<false>(i, g, x, y, z, h, ncell, kappa, nullptr, 0);
    if ((threadIdx.x & 63) == 0) cnt[i] = m;
}
// fill pass; with cnt (a fill without a count pass) into the room of each
// slice, writing the counts and over = 1 when a list does not fit
__global__ void __launch_bounds__(BLOCK) k_nl_fill(int n, Grid g, const double* x, const double* y, const double* z,
                                                   const double* h, const int* ncell, double kappa,
                                                   const long long* slice_off, long long total, unsigned* ent, int* cnt,
                                                   int* over) {
    // xcd_block: a list slice's 64 particles on one XCD, so their
    // interleaved entries fill whole lines in one L2
    const int i = (int)(((long long)xcd_block() * blockDim.x + threadIdx.x) >> 6);
    if (i >= n) return;   // whole waves
    const int s = i >> 6;
    const long long end = s + 1 < (n + SLICE - 1) / SLICE ? slice_off[s + 1] : total;
    const int cap = cnt ? (int)((end - slice_off[s]) / SLICE) : INT_MAX;
    const int m = nl_wave<true>(i, g, x, y, z, h, ncell, kappa, ent + slice_off[s] + (i & 63), cap);
    if (cnt && (threadIdx.x & 63) == 0) {
        cnt[i] = m;
        if (m > cap) over[0] = 1;
    }
}
// room of each slice for a fill without a count pass
__global__ void __launch_bounds__(BLOCK) k_slice_spare(int nslices, const long long* prev, long long spare,
                                                       long long* w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nslices) w[s] = prev[s] + spare * SLICE;
}
__global__ void __launch_bounds__(BLOCK) k_slice_width(int n, int nslices, const int* cnt, long long* w) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int m = 0;
    for (int l = 0; l < SLICE; l++) {
        int i = s * SLICE + l;
        if (i < n) m = max(m, cnt[i]);
    }
    w[s] = (long long)m * SLICE;
}

// ------------------------------------------------------------------------
static int permute_state(sph_ctx* c) {
    const int n = c->n;
    double* tmp = nullptr;
    int* itmp = nullptr;
    // each state array is swapped for its gathered copy, so the copies keep
    // the context's capacity: refinement appends particles past n
    SPH_TRY(dalloc(c, &tmp, c->cap));
    SPH_TRY(dalloc(c, &itmp, c->cap));
    auto gd = [&](double*& a) -> int {
        if (!a) return 0;
        hipLaunchKernelGGL(k_gather<double>, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)c->perm,
                           (const double*)a, tmp);
        SPH_TRY(launch_check(c, "permute"));
        std::swap(a, tmp);
        return 0;
    };
    auto gi = [&](int*& a) -> int {
        hipLaunchKernelGGL(k_gather<int>, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)c->perm,
                           (const int*)a, itmp);
        SPH_TRY(launch_check(c, "permute"));
        std::swap(a, itmp);
        return 0;
    };
    SPH_TRY(prof_begin(c, "permute", 0));
    for (int f = 0; f < SPH_F_COUNT_F64 + 3; f++) SPH_TRY(gd(c->pf[f]));
    for (int f = 0; f < SPH_I_COUNT; f++) SPH_TRY(gi(c->pi[f]));
    for (int f = 0; f < 10; f++) SPH_TRY(gd(c->p2[f]));
    for (int f = 15; f < 18; f++) SPH_TRY(gd(c->q[f]));
    SPH_TRY(gd(c->apr_target));
    SPH_TRY(prof_end(c));
    dfree(c, tmp);
    dfree(c, itmp);
    return 0;
}

// Order of the record-gathering particle passes (thread -> particle):
// particles with lists of at most NL_HEAVY entries first, in their order,
// then the longer ones (coarse particles beside refined zones), so that
// those do not hold waves of short lists for their length.  Null when no
// list is that long.  A pass computes each particle on its own, so the
// order changes no result.
constexpr int NL_HEAVY = 512;
__global__ void __launch_bounds__(BLOCK) k_heavy_flag(int n, const int* cnt, int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = cnt[i] > NL_HEAVY ? 1 : 0;
}
__global__ void __launch_bounds__(BLOCK) k_porder(int n, int nlight, const int* flag, const int* scan, int* order) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    order[flag[i] ? nlight + scan[i] : i - scan[i]] = i;
}
static int build_porder(sph_ctx* c) {
    const int n = c->n;
    dfree(c, c->porder);
    c->porder = nullptr;
    c->n_heavy = 0;
    if (n == 0) return 0;
    int *flag = nullptr, *scan = nullptr;
    SPH_TRY(dalloc(c, &flag, n));
    SPH_TRY(dalloc(c, &scan, n));
    hipLaunchKernelGGL(k_heavy_flag, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)c->nl_cnt, flag);
    SPH_TRY(launch_check(c, "porder"));
    int nheavy = 0;
    SPH_TRY(exclusive_scan_i32(c, flag, scan, n, &nheavy));
    c->n_heavy = nheavy;
    if (nheavy > 0) {
        SPH_TRY(dalloc(c, &c->porder, n));
        hipLaunchKernelGGL(k_porder, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, n - nheavy, (const int*)flag,
                           (const int*)scan, c->porder);
        SPH_TRY(launch_check(c, "porder"));
    }
    dfree(c, flag);
    dfree(c, scan);
    return 0;
}

// The list of a search: one wave per particle.  When the previous search
// had as many slices (the particles keep their cell order, and move little
// between searches), the fill runs at once into each slice's previous
// width plus NL_SPARE entries; if a list does not fit, the search counts
// first and fills slices of exact widths.  The entries are the same either
// way.
constexpr int NL_SPARE = 8;
static int build_neighbor_list(sph_ctx* c) {
    const int n = c->n;
    int nslices = (n + SLICE - 1) / SLICE;
    dfree(c, c->nl_cnt); c->nl_cnt = nullptr;
    dfree(c, c->nl_slice_off); c->nl_slice_off = nullptr;
    dfree(c, c->nl_ent); c->nl_ent = nullptr;
    SPH_TRY(dalloc(c, &c->nl_cnt, n));
    SPH_TRY(dalloc(c, &c->nl_slice_off, nslices + 1));
    Grid g = c->grid();
    const double *x = c->pf[SPH_F_X], *y = c->pf[SPH_F_Y], *z = c->pf[SPH_F_Z], *h = c->pf[SPH_F_H];
    const unsigned nlgrid = (unsigned)(((long long)n * 64 + BLOCK - 1) / BLOCK);
    long long* w = nullptr;
    SPH_TRY(dalloc(c, &w, nslices + 1));
    long long total = 0;
    if (c->nl_wprev && c->nl_wprev_n == nslices) {
        int* over = nullptr;
        SPH_TRY(dalloc(c, &over, 1));
        SPH_HIP(c, hipMemsetAsync(over, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(k_slice_spare, dim3(grid1(nslices)), dim3(BLOCK), 0, c->stream, nslices,
                           (const long long*)c->nl_wprev, (long long)NL_SPARE, w);
        SPH_TRY(launch_check(c, "slice_width"));
        SPH_TRY(exclusive_scan_i64(c, w, c->nl_slice_off, nslices, &total));
        SPH_TRY(dalloc(c, &c->nl_ent, (size_t)total + 1));
        SPH_TRY(prof_begin(c, "nl_fill", 0));
        hipLaunchKernelGGL(k_nl_fill, dim3(nlgrid), dim3(BLOCK), 0, c->stream, n, g, x, y, z, h,
                           (const int*)c->pi[SPH_I_NCELL], c->prm.kappa, (const long long*)c->nl_slice_off, total,
                           c->nl_ent, c->nl_cnt, over);
        SPH_TRY(launch_check(c, "nl_fill"));
        SPH_TRY(prof_end(c));
        int h_over = 0;
        SPH_HIP(c, hipMemcpyAsync(&h_over, over, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(c, hipStreamSynchronize(c->stream));
        dfree(c, over);
        if (!h_over) {
            // exact widths of this search: the next one's guess
            hipLaunchKernelGGL(k_slice_width, dim3(grid1(nslices)), dim3(BLOCK), 0, c->stream, n, nslices,
                               (const int*)c->nl_cnt, c->nl_wprev);
            SPH_TRY(launch_check(c, "slice_width"));
            c->nl_total = total;
            dfree(c, w);
            return build_porder(c);
        }
        dfree(c, c->nl_ent);
        c->nl_ent = nullptr;
    }
    SPH_TRY(prof_begin(c, "nl_count", 0));
    hipLaunchKernelGGL(k_nl_count, dim3(nlgrid), dim3(BLOCK), 0, c->stream, n, g, x, y, z, h,
                       (const int*)c->pi[SPH_I_NCELL], c->prm.kappa, c->nl_cnt);
    SPH_TRY(launch_check(c, "nl_count"));
    SPH_TRY(prof_end(c));
    hipLaunchKernelGGL(k_slice_width, dim3(grid1(nslices)), dim3(BLOCK), 0, c->stream, n, nslices,
                       (const int*)c->nl_cnt, w);
    SPH_TRY(launch_check(c, "slice_width"));
    if (c->nl_wprev_n != nslices) {
        dfree(c, c->nl_wprev);
        c->nl_wprev = nullptr;
        SPH_TRY(dalloc(c, &c->nl_wprev, nslices + 1));
        c->nl_wprev_n = nslices;
    }
    SPH_HIP(c, hipMemcpyAsync(c->nl_wprev, w, sizeof(long long) * nslices, hipMemcpyDeviceToDevice, c->stream));
    SPH_TRY(exclusive_scan_i64(c, w, c->nl_slice_off, nslices, &total));
    dfree(c, w);
    c->nl_total = total;
    SPH_TRY(dalloc(c, &c->nl_ent, (size_t)total + 1));
    SPH_TRY(prof_begin(c, "nl_fill", 0));
    hipLaunchKernelGGL(k_nl_fill, dim3(nlgrid), dim3(BLOCK), 0, c->stream, n, g, x, y, z, h,
                       (const int*)c->pi[SPH_I_NCELL], c->prm.kappa, (const long long*)c->nl_slice_off, total,
                       c->nl_ent, (int*)nullptr, (int*)nullptr);
    SPH_TRY(launch_check(c, "nl_fill"));
    SPH_TRY(prof_end(c));
    return build_porder(c);
}

// The particles the Neumann pass updates (every p_type but fluid), in the
// sorted order of the search.
__global__ void __launch_bounds__(BLOCK) k_nm_flag(int n, const int* p_type, int* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = p_type[i] != 1 ? 1 : 0;
}
__global__ void __launch_bounds__(BLOCK) k_nm_fill(int n, const int* flag, const int* scan, int* list) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) list[scan[i]] = i;
}
static int neumann_list(sph_ctx* c) {
    const int n = c->n;
    int *flag = nullptr, *scan = nullptr, cnt = 0;
    SPH_TRY(dalloc(c, &flag, n));
    SPH_TRY(dalloc(c, &scan, n));
    hipLaunchKernelGGL(k_nm_flag, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)c->pi[SPH_I_P_TYPE], flag);
    SPH_TRY(launch_check(c, "nm_flag"));
    SPH_TRY(exclusive_scan_i32(c, flag, scan, n, &cnt));
    dfree(c, c->nm_list);
    c->nm_list = nullptr;
    SPH_TRY(dalloc(c, &c->nm_list, cnt > 0 ? cnt : 1));
    hipLaunchKernelGGL(k_nm_fill, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)flag, (const int*)scan,
                       c->nm_list);
    SPH_TRY(launch_check(c, "nm_fill"));
    c->nm_count = cnt;
    dfree(c, flag);
    dfree(c, scan);
    return 0;
}

// The neighbour list alone, on the cells and order of the last search.
int nnps_list(sph_ctx* c) { return build_neighbor_list(c); }

// Full count==0 search on the current persistent state (any order): cell
// keys, stable sort, cell bounds, permutation of the state and, when
// build_list, the per-particle neighbour list.
int nnps_run(sph_ctx* c, bool build_list) {
    const int n = c->n;
    const sph_domain& d = c->dom;
    if (!c->keys) {
        // sized for the context's capacity: refinement adds particles
        SPH_TRY(dalloc(c, &c->keys, c->cap));
        SPH_TRY(dalloc(c, &c->perm, c->cap));
        SPH_TRY(dalloc(c, &c->keys_tmp, c->cap));
        SPH_TRY(dalloc(c, &c->perm_tmp, c->cap));
    }
    if (!c->gstr) {
        SPH_TRY(dalloc(c, &c->gstr, (size_t)d.num_cells + 1));
        SPH_TRY(dalloc(c, &c->gend, (size_t)d.num_cells + 1));
        SPH_TRY(dalloc(c, &c->cstart, (size_t)d.num_cells + 2));
    }
    SPH_TRY(prof_begin(c, "hash", (3.0 * 8 + 4 + 8) * n));
    hipLaunchKernelGGL(k_hash, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, d, (const int*)c->pi[SPH_I_I_TYPE],
                       (const double*)c->pf[SPH_F_X], (const double*)c->pf[SPH_F_Y], (const double*)c->pf[SPH_F_Z],
                       c->keys, c->perm);
    SPH_TRY(launch_check(c, "hash"));
    SPH_TRY(prof_end(c));
    SPH_TRY(radix_sort(c, d.num_cells));
    SPH_TRY(prof_begin(c, "cell_bounds", 4.0 * n + 8.0 * (d.num_cells + 1)));
    hipLaunchKernelGGL(k_fill_i32, dim3(grid1((long long)d.num_cells + 1)), dim3(BLOCK), 0, c->stream, c->gstr,
                       (long long)d.num_cells + 1, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid1((long long)d.num_cells + 1)), dim3(BLOCK), 0, c->stream, c->gend,
                       (long long)d.num_cells + 1, 0);
    hipLaunchKernelGGL(k_cell_bounds, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)c->keys, c->gstr,
                       c->gend);
    SPH_TRY(launch_check(c, "cell_bounds"));
    hipLaunchKernelGGL(k_cell_start, dim3(grid1((long long)d.num_cells + 2)), dim3(BLOCK), 0, c->stream,
                       d.num_cells + 2, (const int*)c->keys, n, c->cstart);
    SPH_TRY(launch_check(c, "cell_start"));
    SPH_TRY(prof_end(c));
    SPH_TRY(permute_state(c));
    SPH_TRY(neumann_list(c));
    if (build_list) SPH_TRY(build_neighbor_list(c));
    c->searched = true;
    c->generation++;   // state arrays were swapped for their sorted copies
    return 0;
}

int set_ncell_run(sph_ctx* c) {
    const int n = c->n;
    hipLaunchKernelGGL(k_set_ncell, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const int*)c->pi[SPH_I_I_TYPE],
                       (const double*)c->pf[SPH_F_H], (const double*)c->pf[SPH_F_COUNT_F64 + 2],
                       c->pi[SPH_I_NCELL]);
    return launch_check(c, "set_ncell");
}

}  // namespace sph
