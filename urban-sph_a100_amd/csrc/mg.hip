// Aggregation multigrid preconditioner of the MIXED inner BiCGSTAB.
//
// Hierarchy (built once per set of PPE values, on the device):
//   level 0      the PPE operator itself (sliced ELLPACK, float values);
//   level l + 1  one unknown per 2x2x2 block of level-l lattice positions
//                (unsmoothed aggregation), with the Galerkin operator
//                A_{l+1} = P^T A_l P, P the 0/1 aggregate indicator;
//   coarsest     the first level of at most MG_COARSE_MAX unknowns, inverted
//                densely (Gauss-Jordan in double, stored as float).
// Level-0 lattice positions come from the particle positions, each on the
// lattice of its own spacing; a level's block coordinates are the previous
// level's halved.  Every coarse row is formed by one wave that sums its
// columns' terms in 64-bit fixed point in an LDS hash table (integer sums
// do not depend on the order of the lanes) or, for a row of many distinct
// columns, sorts its terms in LDS and adds each column's terms in a fixed
// order; either way the hierarchy is bit-reproducible.
//
// Application (one V-cycle, z = M^-1 b): at every level but the coarsest
//     x = w b / diag;  r = b - A x;  b_next = P^T r
// down to the dense solve, then on the way up
//     x += s P z_next;  z = x + w (b - A x) / diag
// with w = MG_OMEGA (damped Jacobi) and s = MG_SCALE (the over-correction
// that unsmoothed aggregation needs).  All products accumulate in double.
// One launch per pass covers the multigrid cases of a step (block tables of
// each level's rows, as the engine's segment tables do for particles and
// rows); every block returns at once when its case's solve is done, so a
// captured iteration costs nothing after convergence, and a case's V-cycle
// does not depend on the other cases of the step.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "ppe_format.h"
#include "sph_internal.h"

namespace sph {

static_assert(MG_COARSE_MAX <= 1024, "the coarsest level is inverted densely");
// Galerkin rows: the common kernel takes a row of up to 2048 terms (members'
// entries + diagonals) and 512 distinct columns; rows beyond that (coarse
// particles beside refined ones) go to a second kernel with room for 8192
// terms, 64 members and 2048 columns.
constexpr int MG_ENT_CAP = 2048;
constexpr int MG_BIG_CAP = 8192;
constexpr int MG_BIG_MEMBERS = 64;

struct MGLevel {
    int n = 0, nslices = 0;
    long long slots = 0;
    long long nx = 0, ny = 0, nz = 0;   // level 0: extents of the lattice coordinates
    // while building: each row's block coordinates on the lattice of
    // spacing 2^ex q, and ex
    int *cx = nullptr, *cy = nullptr, *cz = nullptr, *ex = nullptr;
    // operator of a coarse level (level 0 is the PPE)
    long long* sl_off = nullptr;
    int* cnt = nullptr;
    int* cols = nullptr;
    float* vals = nullptr;   // -A_IJ, so that y = diag x - sum vals x[cols] as for the PPE
    float* diag = nullptr;
    // link to the next level
    int* agg = nullptr;      // row -> row of the next level
    int* mstart = nullptr;   // next-level row -> first member in mem (n_next + 1 entries)
    int* mem = nullptr;      // this level's rows grouped by aggregate, ascending within each
    // V-cycle vectors (level 0: x is the PPE's mgx0, which the Krylov
    // kernels fill with the pre-smoothing; not owned here)
    float *b = nullptr, *x = nullptr, *y = nullptr, *r = nullptr;
    bool x_shared = false;
    // coarse level: rows of more than MG_CLONG entries, whose products a
    // wave per row forms before each pass (lidx: row -> list index or -1)
    int nlong = 0;
    int *lrows = nullptr, *lidx = nullptr;
    double* lsum = nullptr;
};
constexpr int MG_CLONG = 128;
constexpr int MG_G = 4;   // lanes per row of the coarse residual and post-smoothing passes
static_assert(BLOCK % MG_G == 0 && SLICE % MG_G == 0, "a row's lanes sit in one wave");
struct MGState {
    long long id = 0;        // distinct for every hierarchy built (table signatures)
    int nlev = 0;            // levels, the PPE and the coarsest included
    MGLevel L[MG_MAX_LEVELS];
    float* ainv = nullptr;   // dense inverse of the coarsest operator, row-major
    // level-0 operator of the V-cycle (16-bit PPE columns only): per stored
    // entry one word, the 16-bit column in the low half and the entry over
    // its row's diagonal as a half float in the high half, in the PPE's
    // slot layout; null: the V-cycle reads the PPE's float values
    unsigned* ev = nullptr;
    // its stencil slices: the half values alone, over the slice's offsets
    // whose halves are not all zero (the neighbours at the search radius
    // carry entries of ~1e-39, zero as halves): per slice the kept width
    // (-1: the slice keeps its words in ev), first slot (64 half-words) in eh
    // and first offset in hoff
    int* hw = nullptr;
    long long* hbase = nullptr;
    int* hpat = nullptr;
    unsigned short* eh = nullptr;
    int* hoff = nullptr;
    double* lsum = nullptr;  // level 0: tail sums of the long rows for the next pass
    long long sorted_rows = 0;   // coarse rows the sorting Galerkin kernels formed (all levels)
};

int mg_levels(const MGState* mg) { return mg ? mg->nlev : 0; }
int mg_coarsest(const MGState* mg) { return mg && mg->nlev > 0 ? mg->L[mg->nlev - 1].n : 0; }
long long mg_sorted_rows(const MGState* mg) { return mg ? mg->sorted_rows : 0; }
int mg_long_rows(const MGState* mg) {
    int k = 0;
    if (mg)
        for (int l = 1; l < mg->nlev; l++) k += mg->L[l].nlong;
    return k;
}

// ------------------------------------------------------------------------
// operators as row sources for the Galerkin product
// (a row of more than LONG_ROW entries continues in its tail)
struct FineOp {
    PPEView pp;
    __device__ int tail_len(int m) const {
        if (!pp.ltail) return 0;
        const int u = pp.ltail[m];
        return u < 0 ? 0 : (int)(pp.tail_off[u + 1] - pp.tail_off[u]);
    }
    __device__ int cnt(int m) const { return pp.cnt[m] + tail_len(m); }
    __device__ int col(int m, int k) const {
        if (k >= LONG_ROW && pp.ltail) return pp.tail_cols[pp.tail_off[pp.ltail[m]] + (k - LONG_ROW)];
        const long long e = prow(pp.sl_off, m) + pslot(k);
        return pp.ecols ? ecol(pp.gbase + (m & ~63), pp.ecols[e]) : pp.cols[e];
    }
    __device__ float val(int m, int k) const {
        if (k >= LONG_ROW && pp.ltail) return pp.tail_vals[pp.tail_off[pp.ltail[m]] + (k - LONG_ROW)];
        return pp.vals32[prow(pp.sl_off, m) + pslot(k)];
    }
    __device__ float dg(int m) const { return pp.fdiag[m]; }
};
struct CoarseOp {
    const long long* sl_off;
    const int* cnt_;
    const int* cols;
    const float* vals;
    const float* diag;
    __device__ int cnt(int m) const { return cnt_[m]; }
    __device__ int col(int m, int k) const { return cols[prow(sl_off, m) + pslot(k)]; }
    __device__ float val(int m, int k) const { return vals[prow(sl_off, m) + pslot(k)]; }
    __device__ float dg(int m) const { return diag[m]; }
};
static CoarseOp coarse_op(const MGLevel& L) { return CoarseOp{L.sl_off, L.cnt, L.cols, L.vals, L.diag}; }

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// exclusive prefix sum over the 64 lanes
__device__ __forceinline__ int wave_excl(int v) {
    const int lane = threadIdx.x & 63;
    int s = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(s, off, 64);
        if (lane >= off) s += u;
    }
    return s - v;
}

// ------------------------------------------------------------------------
// hierarchy construction
//
// Aggregates are 2x2x2 blocks of a lattice.  q is the finest particles'
// spacing; a particle of smoothing length h = 2^r h_min starts on the
// lattice of spacing 2^r q (its own spacing), so every resolution zone of a
// refined case coarsens by about 8 per level.  A row on the lattice of
// spacing 2^e q goes to the block of spacing 2^(e+1) q holding it; blocks of
// different spacings are different aggregates.  A block whose rows gather
// more than MG_BIG_CAP terms (or more than MG_BIG_MEMBERS rows) is split
// into runs of its rows, in ascending row order, that do not; the runs keep
// the block's coordinates and meet again in the next level's block.

// lattice position of each PPE row's particle, clamped to the extents (a
// particle outside them joins the boundary blocks)
__device__ __forceinline__ int lattice_index(double v, double v0, double q, long long n) {
    const double k = floor((v - v0) / q + 0.5);
    return (int)fmin(fmax(k, 0.0), (double)(n - 1));
}
__global__ void __launch_bounds__(BLOCK) k_mg_coord0(int nr, const int* rows, PView S, double x0, double y0, double z0,
                                                     double q, long long nx, long long ny, long long nz, int* cx,
                                                     int* cy, int* cz, int* ex, int* rmax) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const int i = rows[r];
    // h_min = 1.5 q; refinement halves h, so h / h_min is a power of two
    const int e = (int)fmin(fmax(rint(log2(S.h[i] / (1.5 * q))), 0.0), 20.0);
    cx[r] = lattice_index(S.x[i], x0, q, nx) >> e;
    cy[r] = lattice_index(S.y[i], y0, q, ny) >> e;
    cz[r] = lattice_index(S.z[i], z0, q, nz) >> e;
    ex[r] = e;
    if (e > 0) atomicMax(rmax, e);
}
// extent of the lattice of spacing 2^e q
__device__ __host__ __forceinline__ long long lattice_extent(long long n0, int e) {
    return (n0 + (1LL << e) - 1) >> e;
}
// first key of the blocks of spacing 2^e q (e >= 1)
__device__ __host__ __forceinline__ long long lattice_key_base(long long nx0, long long ny0, long long nz0, int e) {
    long long b = 0;
    for (int k = 1; k < e; k++) b += lattice_extent(nx0, k) * lattice_extent(ny0, k) * lattice_extent(nz0, k);
    return b;
}
__global__ void __launch_bounds__(BLOCK) k_mg_key(int n, const int* cx, const int* cy, const int* cz, const int* ex,
                                                  long long nx0, long long ny0, long long nz0, int* key, int* val) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int e = ex[r] + 1;
    const long long nx1 = lattice_extent(nx0, e), ny1 = lattice_extent(ny0, e);
    key[r] = (int)(lattice_key_base(nx0, ny0, nz0, e) + ((long long)(cz[r] >> 1) * ny1 + (cy[r] >> 1)) * nx1 +
                   (cx[r] >> 1));
    val[r] = r;
}
__global__ void __launch_bounds__(BLOCK) k_mg_flag(int n, const int* key, int* flag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) flag[t] = (t == 0 || key[t] != key[t - 1]) ? 1 : 0;
}
// the thread of a block's first sorted row walks the block and starts a new
// aggregate where the next row would pass the bounds
template <class OP>
__global__ void __launch_bounds__(BLOCK) k_mg_split(OP op, int n, const int* key, const int* val, int* flag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || !flag[t]) return;
    int terms = 0, members = 0;
    for (int u = t; u < n && (u == t || key[u] == key[t]); u++) {
        const int w = op.cnt(val[u]) + 1;
        if (members > 0 && (terms + w > MG_BIG_CAP || members == MG_BIG_MEMBERS)) {
            flag[u] = 1;
            terms = members = 0;
        }
        terms += w;
        members++;
    }
}
// aggregate of every row, member lists, block coordinates of the aggregates
// (those of any member: all share the block)
__global__ void __launch_bounds__(BLOCK) k_mg_link(int n, const int* key, const int* val, const int* flag,
                                                   const int* scan, const int* cx, const int* cy, const int* cz,
                                                   const int* ex, int* agg, int* mem, int* mstart, int* cx1, int* cy1,
                                                   int* cz1, int* ex1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int I = scan[t] + flag[t] - 1;
    const int m = val[t];
    agg[m] = I;
    mem[t] = m;
    if (flag[t]) {
        mstart[I] = t;
        cx1[I] = cx[m] >> 1;
        cy1[I] = cy[m] >> 1;
        cz1[I] = cz[m] >> 1;
        ex1[I] = ex[m] + 1;
    }
    if (t == n - 1) mstart[I + 1] = n;
}

// One coarse row per 64-thread block (one wave): its terms are the
// members' diagonals and entries mapped to aggregates, numbered in order
// (members ascending, each member's diagonal first, then its entries in
// storage order).  A bitonic sort in LDS of 32-bit keys (column - the row's
// first column) << TB | term number groups each column's terms in term
// order, so each column's sum is formed in that fixed order by one lane:
// the hierarchy is bit-reproducible.  Terms are float values added in
// double.  CAP = MG_ENT_CAP: rows of at most CAP terms, values staged in
// LDS; CAP = MG_BIG_CAP: only the larger rows, values read back from the
// operator.  Output: the off-diagonal count, the columns ascending and
// their values -sum from row_off[I] on (room for the row's terms), the
// diagonal.  Each lane takes the columns whose first sorted term falls in
// its run of sorted positions and sums each of them to its end.  err:
// SPH_MG_TERMS (too many terms or members), SPH_MG_SPAN (columns too far
// apart).
template <class OP, int CAP>
__global__ void __launch_bounds__(SLICE) k_mg_rap(OP op, int n1, const int* agg, const int* mstart, const int* mem,
                                                  const int* redo, int* cnt1, const long long* row_off, int* cols1,
                                                  float* vals1, float* diag1, int* err) {
    constexpr bool BIG = CAP > MG_ENT_CAP;
    constexpr int TB = BIG ? 13 : 11;               // term-number bits
    static_assert(CAP <= (1 << TB), "term numbers fit their key bits");
    __shared__ unsigned sk[CAP];                    // (column - jmin) << TB | term number
    __shared__ float sv[BIG ? 1 : CAP];             // term values by term number
    __shared__ int moff[BIG ? MG_BIG_MEMBERS + 1 : 1];   // first term of each member
    const int I = blockIdx.x;
    const int lane = threadIdx.x;
    if (I >= n1 || !redo[I]) return;
    const int t0 = mstart[I], t1 = mstart[I + 1];
    int nterm = 0;
    for (int t = t0 + lane; t < t1; t += SLICE) nterm += op.cnt(mem[t]) + 1;
    nterm = wave_sum(nterm);
    if (BIG ? nterm <= MG_ENT_CAP : nterm > MG_ENT_CAP) return;   // the other kernel's row
    if (nterm > CAP || (BIG && t1 - t0 > MG_BIG_MEMBERS)) {
        if (lane == 0) err[0] = SPH_MG_TERMS;
        return;
    }
    int np2 = 1;
    while (np2 < nterm) np2 <<= 1;
    int pos = 0, jmin = I, jmax = I;
    for (int t = t0; t < t1; t++) {
        const int m = mem[t];
        const int c = op.cnt(m);
        if (BIG && lane == 0) moff[t - t0] = pos;
        for (int k = lane - 1; k < c; k += SLICE) {
            const int q = pos + k + 1;
            const float v = k < 0 ? op.dg(m) : -op.val(m, k);
            // a stored zero (a stencil slot the row lacks) joins the
            // diagonal's sum, where it adds nothing
            const int J = (k < 0 || v == 0.0f) ? I : agg[op.col(m, k)];
            sk[q] = (unsigned)J;
            jmin = min(jmin, J);
            jmax = max(jmax, J);
            if (!BIG) sv[q] = v;
        }
        pos += c + 1;
    }
    if (BIG && lane == 0) moff[t1 - t0] = pos;
    jmin = -wave_max(-jmin);
    jmax = wave_max(jmax);
    if (jmax - jmin >= (1 << (32 - TB)) - 1) {
        if (lane == 0) err[0] = SPH_MG_SPAN;
        return;
    }
    __syncthreads();
    for (int p = lane; p < np2; p += SLICE)
        sk[p] = p < nterm ? ((sk[p] - (unsigned)jmin) << TB) | (unsigned)p : 0xffffffffu;
    __syncthreads();
    for (int k = 2; k <= np2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < np2; i += SLICE) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned a = sk[i], b = sk[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { sk[i] = b; sk[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    // distinct columns: each lane counts the group starts in its run, and
    // those below the diagonal's column (the diagonal's slot)
    auto col_at = [&](int p) { return (int)(sk[p] >> TB) + jmin; };
    auto starts = [&](int p) { return p == 0 || col_at(p) != col_at(p - 1); };
    const int per = (nterm + SLICE - 1) / SLICE;
    const int p0 = min(nterm, lane * per), p1 = min(nterm, p0 + per);
    int own = 0, below = 0;
    for (int p = p0; p < p1; p++)
        if (starts(p)) { own++; below += col_at(p) < I ? 1 : 0; }
    const int at = wave_excl(own);
    const int total = wave_sum(own);
    const int sI = wave_sum(below);
    if (lane == 0) cnt1[I] = total - 1;
    // a term's value: from LDS, or (large rows) from its member's entry
    auto term_val = [&](int q) -> float {
        if (!BIG) return sv[q];
        int mi = 0;
        while (mi + 1 < t1 - t0 && moff[mi + 1] <= q) mi++;
        const int m = mem[t0 + mi], k = q - moff[mi] - 1;
        return k < 0 ? op.dg(m) : -op.val(m, k);
    };
    const long long base = row_off[I];
    int s = at;
    for (int p = p0; p < p1; p++) {
        if (!starts(p)) continue;
        const int J = col_at(p);
        double acc = 0.0;
        int e = p;
        do {
            acc += (double)term_val((int)(sk[e] & ((1u << TB) - 1u)));
            e++;
        } while (e < nterm && col_at(e) == J);
        if (s == sI) {
            diag1[I] = (float)acc;
        } else {
            const int o = s - (s > sI ? 1 : 0);
            cols1[base + o] = J;
            vals1[base + o] = (float)(-acc);
        }
        s++;
    }
}
// The common form of a coarse row: its terms go into an LDS hash table of
// MG_HASH columns, each summed in 64-bit fixed point (the term times
// `scale`, a power of two that keeps every row's sums below 2^62), so the
// sums do not depend on the order the lanes add them in and the hierarchy is
// bit-reproducible without sorting the terms; then the distinct columns are
// sorted.  A row of more than MG_HASH / 2 distinct columns (the table at most
// half full, so probes stay short; the count does not depend on the order
// of insertion) is flagged in redo, and counted in nredo, for the sorting
// kernels above.
constexpr int MG_HASH = 512;
static_assert(MG_HASH / 2 <= 4 * SLICE, "a row's distinct columns fit the widest register sort");
// key[0 .. total) ascending, by one wave, through registers
template <int Q>
__device__ __forceinline__ void sort_keys_in_place(unsigned long long* key, int total, int lane) {
    unsigned long long v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int t = lane + q * SLICE;
        v[q] = t < total ? key[t] : ~0ull;   // pads sort last
    }
    bitonic_wave<unsigned long long, Q>(v, lane);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int t = lane + q * SLICE;
        if (t < total) key[t] = v[q];
    }
}
template <class OP>
__global__ void __launch_bounds__(SLICE) k_mg_rap_hash(OP op, int n1, const int* agg, const int* mstart,
                                                       const int* mem, double scale, int* redo, int* nredo,
                                                       int* cnt1, const long long* row_off, int* cols1, float* vals1,
                                                       float* diag1, int* err, int force) {
    __shared__ int hk[MG_HASH];                    // column, or -1
    __shared__ long long hs[MG_HASH];              // its fixed-point sum
    __shared__ unsigned long long key[MG_HASH / 2];   // column << 32 | table slot, sorted
    __shared__ int full, ndist;
    const int I = blockIdx.x;
    const int lane = threadIdx.x;
    if (I >= n1) return;
    const int t0 = mstart[I], t1 = mstart[I + 1];
    int nterm = 0;
    for (int t = t0 + lane; t < t1; t += SLICE) nterm += op.cnt(mem[t]) + 1;
    nterm = wave_sum(nterm);
    if (nterm > MG_BIG_CAP) {   // the sums' bound
        if (lane == 0) err[0] = SPH_MG_TERMS;
        return;
    }
    if (force) {   // every row to the sorting kernels (their check)
        if (lane == 0) { redo[I] = 1; atomicAdd(nredo, 1); }
        return;
    }
    for (int h = lane; h < MG_HASH; h += SLICE) { hk[h] = -1; hs[h] = 0; }
    if (lane == 0) { full = 0; ndist = 0; }
    __syncthreads();
    for (int t = t0; t < t1 && !full; t++) {
        const int m = mem[t];
        const int c = op.cnt(m);
        for (int k = lane - 1; k < c && !full; k += SLICE) {
            const double v = k < 0 ? (double)op.dg(m) : -(double)op.val(m, k);
            const int J = (k < 0 || v == 0.0) ? I : agg[op.col(m, k)];   // stored zeros: no column of their own
            const long long f = llrint(v * scale);
            int h = (int)(((unsigned)J * 2654435761u) >> 23) & (MG_HASH - 1);
            int probe = 0;
            for (; probe < MG_HASH; probe++) {
                const int old = atomicCAS(&hk[h], -1, J);
                if (old == -1 && atomicAdd(&ndist, 1) >= MG_HASH / 2) full = 1;
                if (old == -1 || old == J) break;
                h = (h + 1) & (MG_HASH - 1);
            }
            if (probe < MG_HASH) atomicAdd((unsigned long long*)&hs[h], (unsigned long long)f);
            else full = 1;
        }
    }
    __syncthreads();
    if (full) {
        if (lane == 0) { redo[I] = 1; atomicAdd(nredo, 1); }
        return;
    }
    // distinct columns, ascending
    int own = 0;
    for (int h = lane; h < MG_HASH; h += SLICE) own += hk[h] >= 0 ? 1 : 0;
    int w = wave_excl(own);
    const int total = wave_sum(own);
    for (int h = lane; h < MG_HASH; h += SLICE)
        if (hk[h] >= 0) key[w++] = ((unsigned long long)(unsigned)hk[h] << 32) | (unsigned)h;
    __syncthreads();
    // sorted in registers (total <= MG_HASH / 2 = 4 * 64 keys), back to LDS
    if (total <= SLICE) sort_keys_in_place<1>(key, total, lane);
    else if (total <= 2 * SLICE) sort_keys_in_place<2>(key, total, lane);
    else sort_keys_in_place<4>(key, total, lane);
    __syncthreads();
    int below = 0;
    for (int p = lane; p < total; p += SLICE) below += (int)(key[p] >> 32) < I ? 1 : 0;
    const int sI = wave_sum(below);
    if (lane == 0) { cnt1[I] = total - 1; redo[I] = 0; }
    const long long base = row_off[I];
    for (int p = lane; p < total; p += SLICE) {
        const int J = (int)(key[p] >> 32);
        const double v = (double)hs[(int)(key[p] & 0xffffffffu)] / scale;
        if (p == sI) {
            diag1[I] = (float)v;
        } else {
            const int o = p - (p > sI ? 1 : 0);
            cols1[base + o] = J;
            vals1[base + o] = (float)(-v);
        }
    }
}
// largest |entry| of an operator (the diagonal included), as float bits
template <class OP>
__global__ void __launch_bounds__(BLOCK) k_mg_maxabs(OP op, int n, unsigned* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float m = fabsf(op.dg(r));
    const int c = op.cnt(r);
    for (int k = 0; k < c; k++) m = fmaxf(m, fabsf(op.val(r, k)));
    if (m == m) atomicMax(out, __float_as_uint(m));
}

// room for each coarse row's columns: its terms (members' entries + 1)
template <class OP>
__global__ void __launch_bounds__(BLOCK) k_mg_terms(OP op, int n1, const int* mstart, const int* mem, long long* t) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n1) return;
    long long s = 0;
    for (int q = mstart[I]; q < mstart[I + 1]; q++) s += op.cnt(mem[q]) + 1;
    t[I] = s;
}
// rows packed by row_off -> the sliced layout
__global__ void __launch_bounds__(BLOCK) k_mg_pack(int n1, const long long* row_off, const int* cnt, const int* pc,
                                                   const float* pv, const long long* sl_off, int* cols, float* vals) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n1) return;
    const long long src = row_off[r], dst = prow(sl_off, r);
    for (int k = 0; k < cnt[r]; k++) {
        cols[dst + pslot(k)] = pc[src + k];
        vals[dst + pslot(k)] = pv[src + k];
    }
}
__global__ void __launch_bounds__(BLOCK) k_mg_width(int n, int nslices, const int* cnt, long long* w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int m = 0;
    for (int r = s * SLICE; r < min(n, (s + 1) * SLICE); r++) m = max(m, cnt[r]);
    w[s] = (long long)m * SLICE;
}

// coarsest level: dense copy, Gauss-Jordan inversion without pivoting
// (the Galerkin operators of the diagonally dominant PPE keep nonzero
// pivots; a zero or non-finite pivot sets err and the preconditioner is
// not used), float copy of the inverse
__global__ void __launch_bounds__(BLOCK) k_mg_dense_zero(long long n, double* a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}
__global__ void __launch_bounds__(BLOCK) k_mg_dense_fill(CoarseOp op, int n, double* a) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    a[(long long)r * n + r] = (double)op.dg(r);
    const int c = op.cnt(r);
    for (int k = 0; k < c; k++) a[(long long)r * n + op.col(r, k)] = -(double)op.val(r, k);
}
// One launch per pivot k.  rowk / colk hold row k and column k of the
// matrix as the previous step left it; the step writes row k + 1 and
// column k + 1 as it leaves them into rown / coln for the next launch (the
// buffers alternate).  Pivot row: a[k][j] / piv (1 / piv at j = k); other
// rows: a[i][j] - a[i][k] a[k][j] / piv (0 - ... at j = k).
__global__ void __launch_bounds__(BLOCK) k_gj_first(int n, const double* a, double* row0, double* col0) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    row0[j] = a[j];
    col0[j] = a[(long long)j * n];
}
__global__ void __launch_bounds__(BLOCK) k_gj_step(int n, double* a, int k, const double* rowk, const double* colk,
                                                   double* rown, double* coln, int* err) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * n) return;
    const int i = (int)(idx / n), j = (int)(idx % n);
    const double piv = rowk[k];
    if (idx == 0 && (!(fabs(piv) > 0.0) || !(fabs(piv) <= DBL_MAX))) err[0] = SPH_MG_PIVOT;
    const double rj = (j == k ? 1.0 : rowk[j]) / piv;
    const double v = i == k ? rj : (j == k ? 0.0 : a[idx]) - colk[i] * rj;
    a[idx] = v;
    if (i == k + 1) rown[j] = v;
    if (j == k + 1) coln[i] = v;
}
__global__ void __launch_bounds__(BLOCK) k_mg_tofloat(long long n, const double* a, float* b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
// The V-cycle's compact level-0 operator: per stored entry its column word
// | half(value / diag) << 16; stencil slices are packed by k_mg_hpack.
__global__ void __launch_bounds__(BLOCK) k_mg_pack16(PPEView pp, unsigned* ev) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= pp.nr) return;
    if (pp.spat && pp.spat[r >> 6] >= 0) return;
    const long long base = prow(pp.sl_off, r);
    const float d = pp.fdiag[r];
    const int m = pp.cnt[r];
    for (int k = 0; k < m; k++) {
        const long long e = base + pslot(k);
        const unsigned short h = __builtin_bit_cast(unsigned short, (_Float16)(pp.vals32[e] / d));
        ev[e] = (unsigned)pp.ecols[e] | ((unsigned)h << 16);
    }
}
// Stencil slices of the level-0 copy, one 64-thread block per slice: a slot
// is kept when some row of the slice has a nonzero half there (a wave
// ballot, so every lane agrees).  PACK = false: the kept width of each
// slice (hw, -1 for the other slices; wh and wp = the width for the scans);
// PACK = true: the halves of the kept slots at 64 (hbase[s] + k') + lane and
// their offsets at hoff[hpat[s] + k'], k' in slot order.
template <bool PACK>
__global__ void __launch_bounds__(SLICE) k_mg_hpack(PPEView pp, int* hw, long long* wh, int* wp,
                                                     const long long* hbase, const int* hpat, unsigned short* eh,
                                                     int* hoff) {
    const int s = blockIdx.x, l = threadIdx.x, r = s * SLICE + l;
    const int st = pp.spat[s];
    if (st < 0) {
        if (!PACK && l == 0) { hw[s] = -1; wh[s] = 0; wp[s] = 0; }
        return;
    }
    const bool live = r < pp.nr;
    const int n = pp.cnt[s * SLICE];   // every row of a stencil slice holds its offsets
    const long long base = pp.sl_off[s] + l;
    const float d = live ? pp.fdiag[r] : 1.0f;
    int kept = 0;
    for (int k = 0; k < n; k++) {
        const unsigned short h =
            live ? __builtin_bit_cast(unsigned short, (_Float16)(pp.vals32[base + pslot(k)] / d)) : (unsigned short)0;
        // +0 and -0 halves add exact zeros
        if (__ballot((h & 0x7fffu) != 0) == 0) continue;
        if (PACK) {
            eh[pslot(hbase[s] + kept) + l] = h;
            if (l == 0) hoff[hpat[s] + kept] = pp.poff[st + k];
        }
        kept++;
    }
    if (!PACK && l == 0) { hw[s] = kept; wh[s] = kept; wp[s] = kept; }
}

// ------------------------------------------------------------------------
// V-cycle kernels
// A case's level as the V-cycle kernels see it (level 0: the PPE itself,
// read through the case's descriptor).
struct MGLv {
    int n, lc;                    // rows (0: no such level); the case's coarsest level
    int nlong;                    // coarse level: its long rows (lrows, lidx, lsum)
    const int* lrows;
    const int* lidx;
    const long long* sl_off;      // operator of a coarse level
    const int* cnt;
    const int* cols;
    const float* vals;
    const float* diag;
    const int* agg;               // link to the next level
    const int* mstart;
    const int* mem;
    float *b, *x, *y, *r;
    const float* ainv;            // coarsest level: dense inverse
    const unsigned* ev;           // level 0: compact operator, or null
    const int* hw;                // level 0, stencil slices: kept width (or -1), first half-word, first offset
    const long long* hbase;
    const int* hpat;
    const unsigned short* eh;
    const int* hoff;
    double* lsum;                 // level 0: tail sums of the long rows; coarse: long rows' products
};

// sum_k (a_rk / d_r) x[col_k] over the compact level-0 operator, in entry
// order, accumulated in double; four independent loads in flight per lane
__device__ __forceinline__ double mg_row16(const PPEView& pp, const MGLv& L, int r, const float* x,
                                           const int* wb) {
    const unsigned* ev = L.ev;
    const int w = L.hw ? __builtin_amdgcn_readfirstlane(L.hw[r >> 6]) : -1;
    if (w >= 0) {   // a stencil slice: the kept halves only, columns row + offset
        // the slice's first slot and offset, wave-uniform (scalar offset loads)
        const int s6 = r >> 6;
        const int hb = __builtin_amdgcn_readfirstlane((int)L.hbase[s6]);
        const int hq = __builtin_amdgcn_readfirstlane(L.hpat[s6]);
        const gmem<unsigned short> hp = (gmem<unsigned short>)(L.eh + pslot(hb) + (r & 63));
        const cmem<int> off = (cmem<int>)(L.hoff + hq);
        const gmem<float> xr = (gmem<float>)(x + r);
        auto hterm = [&](unsigned short u, int o) {
            return (double)(float)__builtin_bit_cast(_Float16, u) * (double)xr[o];
        };
        double s = 0.0;
        int k = 0;
        for (; k + 4 <= w; k += 4) {
            const unsigned short u0 = __builtin_nontemporal_load(hp + pslot(k + 0));
            const unsigned short u1 = __builtin_nontemporal_load(hp + pslot(k + 1));
            const unsigned short u2 = __builtin_nontemporal_load(hp + pslot(k + 2));
            const unsigned short u3 = __builtin_nontemporal_load(hp + pslot(k + 3));
            s += hterm(u0, off[k + 0]);
            s += hterm(u1, off[k + 1]);
            s += hterm(u2, off[k + 2]);
            s += hterm(u3, off[k + 3]);
        }
        for (; k < w; k++) s += hterm(__builtin_nontemporal_load(hp + pslot(k)), off[k]);
        return s;
    }
    const unsigned* p = ev + prow(pp.sl_off, r);
    const int m = pp.cnt[r];
    auto term = [&](unsigned u) {
        const _Float16 h = __builtin_bit_cast(_Float16, (unsigned short)(u >> 16));
        return (double)(float)h * (double)x[ecol(wb, (int)(u & 0xffffu))];
    };
    double s = 0.0;
    int k = 0;
    for (; k + 4 <= m; k += 4) {
        const unsigned u0 = __builtin_nontemporal_load(p + pslot(k + 0));
        const unsigned u1 = __builtin_nontemporal_load(p + pslot(k + 1));
        const unsigned u2 = __builtin_nontemporal_load(p + pslot(k + 2));
        const unsigned u3 = __builtin_nontemporal_load(p + pslot(k + 3));
        s += term(u0);
        s += term(u1);
        s += term(u2);
        s += term(u3);
    }
    for (; k < m; k++) s += term(__builtin_nontemporal_load(p + pslot(k)));
    return s;
}

// block -> (case, first row of the level); the block returns when the
// case's solve is done
#define MG_HEAD(tab, lvl)                                      \
    const int2 te_ = (tab)[blockIdx.x];                        \
    const CaseDesc& cd = cds[te_.x];                           \
    if (cd.sc[sc_done] != 0.0) return;                         \
    const MGLv& L = lv[te_.x * MG_MAX_LEVELS + (lvl)];         \
    const int row = te_.y + (int)threadIdx.x;

// Prolongation of a level's correction z to the level above, done by the
// thread that forms z[a]: x += s z[a] over the members of aggregate a (each
// member belongs to one aggregate, so no two threads write one x)
__device__ __forceinline__ void mg_prolong_members(const MGLv& F, int a, float z) {
    for (int t = F.mstart[a]; t < F.mstart[a + 1]; t++) {
        const int m = F.mem[t];
        F.x[m] = (float)((double)F.x[m] + MG_SCALE * (double)z);
    }
}
// level 0 (l = 0): r = b - A x
__global__ void __launch_bounds__(BLOCK) k_mg_res(const CaseDesc* cds, const MGLv* lv, const int2* tab, int l,
                                                  int which, int sc_done) {
    __shared__ int wbase[BLOCK];
    MG_HEAD(tab, l);
    {
        const PPEView& pp = cd.ppe;
        load_group_bases(pp, te_.y, wbase);
        if (row >= L.n) return;
        const float* b = which ? pp.fs : pp.fp;
        if (L.ev) {
            const double d = (double)pp.fdiag[row];
            double s = mg_row16(pp, L, row, L.x, wbase + (threadIdx.x & ~63));
            if (pp.ltail && pp.ltail[row] >= 0) s += L.lsum[pp.ltail[row]] / d;
            L.r[row] = (float)((double)b[row] - d * ((double)L.x[row] - s));
        } else {
            L.r[row] = (float)((double)b[row] - spmv_row(pp, row, (const float*)pp.vals32, pp.fdiag, L.x, wbase));
        }
    }
}
// level 0 (l = 0): z = x + w (b - A x) / diag into p-hat or s-hat
__global__ void __launch_bounds__(BLOCK) k_mg_post(const CaseDesc* cds, const MGLv* lv, const int2* tab, int l,
                                                   int which, int sc_done) {
    __shared__ int wbase[BLOCK];
    MG_HEAD(tab, l);
    {
        const PPEView& pp = cd.ppe;
        load_group_bases(pp, te_.y, wbase);
        if (row >= L.n) return;
        const float* b = which ? pp.fs : pp.fp;
        float* out = which ? pp.fshat : pp.fphat;
        const double d = (double)pp.fdiag[row];
        if (L.ev) {
            double s = mg_row16(pp, L, row, L.x, wbase + (threadIdx.x & ~63));
            if (pp.ltail && pp.ltail[row] >= 0) s += L.lsum[pp.ltail[row]] / d;
            out[row] = (float)((double)L.x[row] + MG_OMEGA * ((double)b[row] / d - (double)L.x[row] + s));
        } else {
            const double ax = spmv_row(pp, row, (const float*)pp.vals32, pp.fdiag, L.x, wbase);
            out[row] = (float)((double)L.x[row] + MG_OMEGA * ((double)b[row] - ax) / d);
        }
    }
}
// Coarse levels: MG_G lanes per row.  A coarse row holds a few dozen
// entries, and a level of a few thousand rows gives a thread per row too
// little company to hide its chain of dependent gathers.  Lane g adds
// entries g, g + MG_G, ... in order, then a fixed shuffle tree adds the
// group's sums, so the result does not depend on the launch.  A long row's
// product comes from k_mg_clong.  tab: BLOCK / MG_G rows per block; every
// lane of a group has the same row, so a group returns or stays whole.
__device__ __forceinline__ double coarse_ax_g(const MGLv& L, int row, int g) {
    const int u = L.lidx ? L.lidx[row] : -1;
    double s = 0.0;
    if (u < 0) {
        const long long base = prow(L.sl_off, row);
        const int m = L.cnt[row];
        for (int k = g; k < m; k += MG_G) s += (double)L.vals[base + pslot(k)] * (double)L.x[L.cols[base + pslot(k)]];
    }
#pragma unroll
    for (int off = MG_G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return (double)L.diag[row] * (double)L.x[row] - (u >= 0 ? L.lsum[u] : s);
}
#define MG_HEAD_G(tab, lvl)                                    \
    const int2 te_ = (tab)[blockIdx.x];                        \
    const CaseDesc& cd = cds[te_.x];                           \
    if (cd.sc[sc_done] != 0.0) return;                         \
    const MGLv& L = lv[te_.x * MG_MAX_LEVELS + (lvl)];         \
    const int row = te_.y + (int)threadIdx.x / MG_G;           \
    const int g = (int)threadIdx.x % MG_G;
// r = b - A x on a coarse level
__global__ void __launch_bounds__(BLOCK) k_mg_res_g(const CaseDesc* cds, const MGLv* lv, const int2* tab, int l,
                                                    int sc_done) {
    MG_HEAD_G(tab, l);
    if (row >= L.n) return;
    const double ax = coarse_ax_g(L, row, g);
    if (g == 0) L.r[row] = (float)((double)L.b[row] - ax);
}
// y = x + w (b - A x) / diag on a coarse level, and its prolongation into
// level l - 1 (x += s y over the aggregate's members, a lane each)
__global__ void __launch_bounds__(BLOCK) k_mg_post_g(const CaseDesc* cds, const MGLv* lv, const int2* tab, int l,
                                                     int sc_done) {
    MG_HEAD_G(tab, l);
    if (row >= L.n) return;
    const double ax = coarse_ax_g(L, row, g);
    const float y = (float)((double)L.x[row] + MG_OMEGA * ((double)L.b[row] - ax) / (double)L.diag[row]);
    if (g == 0) L.y[row] = y;
    const MGLv& F = lv[te_.x * MG_MAX_LEVELS + l - 1];
    for (int t = F.mstart[row] + g; t < F.mstart[row + 1]; t += MG_G) {
        const int m = F.mem[t];
        F.x[m] = (float)((double)F.x[m] + MG_SCALE * (double)y);
    }
}
// b of level l + 1: the sum of r over each aggregate's members, in member
// order (tab: rows of level l + 1)
__global__ void __launch_bounds__(BLOCK) k_mg_restrict(const CaseDesc* cds, const MGLv* lv, const int2* tab, int l,
                                                       int sc_done) {
    MG_HEAD(tab, l);
    const MGLv& N = lv[te_.x * MG_MAX_LEVELS + l + 1];
    if (row >= N.n) return;
    double s = 0.0;
    for (int t = L.mstart[row]; t < L.mstart[row + 1]; t++) s += (double)L.r[L.mem[t]];
    const float b = (float)s;
    N.b[row] = b;
    // the next level's pre-smoothing, x = w b / diag, when it has a level below
    if (l + 1 < L.lc) N.x[row] = (float)(MG_OMEGA * (double)b / (double)N.diag[row]);
}
// level 0: the tail sums of the long rows over the current x, one wave per
// long row, lanes over its contiguous tail, fixed shuffle tree (tab: four
// long rows per block)
__global__ void __launch_bounds__(BLOCK) k_mg_long(const CaseDesc* cds, const MGLv* lv, const int2* tab,
                                                   int sc_done) {
    const int2 te_ = tab[blockIdx.x];
    const CaseDesc& cd = cds[te_.x];
    if (cd.sc[sc_done] != 0.0) return;
    const MGLv& L = lv[te_.x * MG_MAX_LEVELS];
    const PPEView& pp = cd.ppe;
    const int u = te_.y + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (u >= pp.nlong) return;
    double s = 0.0;
    for (long long e = pp.tail_off[u] + lane; e < pp.tail_off[u + 1]; e += SLICE)
        s += (double)pp.tail_vals[e] * (double)L.x[pp.tail_cols[e]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) L.lsum[u] = s;
}
// coarse level l: the off-diagonal products of its long rows over x, one
// wave per row, lanes over its entries, fixed shuffle tree (tab: four long
// rows per block)
__global__ void __launch_bounds__(BLOCK) k_mg_clong(const CaseDesc* cds, const MGLv* lv, const int2* tab, int l,
                                                    int sc_done) {
    const int2 te_ = tab[blockIdx.x];
    const CaseDesc& cd = cds[te_.x];
    if (cd.sc[sc_done] != 0.0) return;
    const MGLv& L = lv[te_.x * MG_MAX_LEVELS + l];
    const int u = te_.y + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (u >= L.nlong) return;
    const int row = L.lrows[u];
    const long long base = prow(L.sl_off, row);
    double s = 0.0;
    for (int k = lane; k < L.cnt[row]; k += SLICE)
        s += (double)L.vals[base + pslot(k)] * (double)L.x[L.cols[base + pslot(k)]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) L.lsum[u] = s;
}
__global__ void __launch_bounds__(BLOCK) k_mg_longflag(int n, const int* cnt, int* flag) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) flag[r] = cnt[r] > MG_CLONG ? 1 : 0;
}
__global__ void __launch_bounds__(BLOCK) k_mg_longlist(int n, const int* flag, const int* scan, int* lrows,
                                                       int* lidx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    lidx[r] = flag[r] ? scan[r] : -1;
    if (flag[r]) lrows[scan[r]] = r;
}
// coarsest level: x = Ainv b, one wave per row, lanes over the columns,
// fixed shuffle tree (tab: four rows per block)
__global__ void __launch_bounds__(BLOCK) k_mg_dense(const CaseDesc* cds, const MGLv* lv, const int2* tab,
                                                    int sc_done) {
    const int2 te_ = tab[blockIdx.x];
    const CaseDesc& cd = cds[te_.x];
    if (cd.sc[sc_done] != 0.0) return;
    const int lc = lv[te_.x * MG_MAX_LEVELS].lc;
    const MGLv& C = lv[te_.x * MG_MAX_LEVELS + lc];
    const int row = te_.y + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= C.n) return;
    const int n = C.n;
    const float* a = C.ainv + (long long)row * n;
    double s = 0.0;
    for (int j = lane; j < n; j += SLICE) s += (double)a[j] * (double)C.b[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane != 0) return;
    const float x = (float)s;
    C.x[row] = x;
    if (lc > 0) mg_prolong_members(lv[te_.x * MG_MAX_LEVELS + lc - 1], row, x);   // into the level above
}

// ------------------------------------------------------------------------
// host side
void mg_free(sph_ctx* c, MGState* mg) {
    if (!mg) return;
    for (int l = 0; l < MG_MAX_LEVELS; l++) {
        MGLevel& L = mg->L[l];
        if (L.x_shared) L.x = nullptr;   // the PPE's
        void* ptrs[] = {L.cx, L.cy, L.cz, L.ex, L.sl_off, L.cnt, L.cols, L.vals, L.diag, L.agg, L.mstart, L.mem, L.b, L.x, L.y, L.r,
                        L.lrows, L.lidx, L.lsum};
        for (void* q : ptrs) dfree(c, q);
    }
    dfree(c, mg->ainv);
    dfree(c, mg->ev);
    void* hp[] = {mg->hw, mg->hbase, mg->hpat, mg->eh, mg->hoff};
    for (void* q : hp) dfree(c, q);
    dfree(c, mg->lsum);
    delete mg;
}

// set-up launches, each its own profile record (records do not nest)
#define MGK(name, grid, block, kern, ...)                                                  \
    do {                                                                                   \
        SPH_TRY(prof_begin(c, name, 0.0));                                                 \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, c->stream, ##__VA_ARGS__);    \
        SPH_TRY(launch_check(c, name));                                                    \
        SPH_TRY(prof_end(c));                                                              \
    } while (0)

// Galerkin operator of level l + 1 from level l (src), whose link arrays
// are set.  Returns 0 with *ok = 0 when a row is out of the kernel's bounds.
// Device buffers of one set-up call, freed on every way out of it.
struct ScopedBufs {
    sph_ctx* c;
    std::vector<void*> p;
    template <class T>
    int get(T** out, size_t count) {
        SPH_TRY(dalloc(c, out, count));
        p.push_back(*out);
        return 0;
    }
    ~ScopedBufs() {
        for (void* q : p) dfree(c, q);
    }
};

template <class OP>
static int mg_galerkin(sph_ctx* c, const OP& src, MGLevel& L1, const MGLevel& L0, int* err, int* ok, int* why,
                       int force_sort, long long* nsorted) {
    const int n1 = L1.n;
    ScopedBufs tmp{c, {}};
    // one pass: each row's columns into room for its terms, then packed
    long long *row_off = nullptr, *tcount = nullptr;
    long long room = 0;
    SPH_TRY(tmp.get(&tcount, (size_t)n1 + 1));
    SPH_TRY(tmp.get(&row_off, (size_t)n1 + 1));
    MGK("mg_rap_terms", grid1(n1), BLOCK, k_mg_terms<OP>, src, n1, (const int*)L0.mstart, (const int*)L0.mem, tcount);
    SPH_TRY(exclusive_scan_i64(c, tcount, row_off, n1, &room));
    int* pc = nullptr;
    float* pv = nullptr;
    SPH_TRY(tmp.get(&pc, (size_t)std::max(1LL, room)));
    SPH_TRY(tmp.get(&pv, (size_t)std::max(1LL, room)));
    SPH_TRY(dalloc(c, &L1.cnt, n1));
    SPH_TRY(dalloc(c, &L1.diag, n1));
    // fixed-point scale of the hashed rows: 2^S with |entry| * 2^S <= 2^49,
    // so that MG_BIG_CAP (2^13) terms sum below 2^62
    int* redo = nullptr;
    int* aux = nullptr;   // [0]: largest |entry| (float bits), [1]: rows left to the sorting kernels
    SPH_TRY(tmp.get(&redo, n1));
    SPH_TRY(tmp.get(&aux, 2));
    SPH_HIP(c, hipMemsetAsync(aux, 0, 2 * sizeof(int), c->stream));
    MGK("mg_rap_scale", grid1(L0.n), BLOCK, k_mg_maxabs<OP>, src, L0.n, (unsigned*)aux);
    int h_aux[2] = {0, 0};
    SPH_HIP(c, hipMemcpyAsync(h_aux, aux, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    float amax = 0.0f;
    std::memcpy(&amax, &h_aux[0], sizeof(float));
    if (!(amax > 0.0f) || !std::isfinite(amax)) amax = 1.0f;
    const double scale = std::ldexp(1.0, 49 - (std::ilogb((double)amax) + 1));
    MGK("mg_rap_fill", n1, SLICE, k_mg_rap_hash<OP>, src, n1, (const int*)L0.agg, (const int*)L0.mstart,
        (const int*)L0.mem, scale, redo, aux + 1, L1.cnt, (const long long*)row_off, pc, pv, L1.diag, err, force_sort);
    SPH_HIP(c, hipMemcpyAsync(h_aux + 1, aux + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    *nsorted += h_aux[1];
    if (h_aux[1] > 0) {
        MGK("mg_rap_sort", n1, SLICE, (k_mg_rap<OP, MG_ENT_CAP>), src, n1, (const int*)L0.agg, (const int*)L0.mstart,
            (const int*)L0.mem, (const int*)redo, L1.cnt, (const long long*)row_off, pc, pv, L1.diag, err);
        MGK("mg_rap_sort", n1, SLICE, (k_mg_rap<OP, MG_BIG_CAP>), src, n1, (const int*)L0.agg, (const int*)L0.mstart,
            (const int*)L0.mem, (const int*)redo, L1.cnt, (const long long*)row_off, pc, pv, L1.diag, err);
    }
    int h_err = 0;
    SPH_HIP(c, hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    if (h_err) {
        *ok = 0;
        *why = h_err;
        return 0;
    }
    L1.nslices = (n1 + SLICE - 1) / SLICE;
    long long* w = nullptr;
    SPH_TRY(tmp.get(&w, (size_t)L1.nslices + 1));
    SPH_TRY(dalloc(c, &L1.sl_off, (size_t)L1.nslices + 1));
    MGK("mg_width", grid1(L1.nslices), BLOCK, k_mg_width, n1, L1.nslices, (const int*)L1.cnt, w);
    SPH_TRY(exclusive_scan_i64(c, w, L1.sl_off, L1.nslices, &L1.slots));
    SPH_TRY(dalloc(c, &L1.cols, (size_t)std::max(1LL, L1.slots)));
    SPH_TRY(dalloc(c, &L1.vals, (size_t)std::max(1LL, L1.slots)));
    MGK("mg_rap_pack", grid1(n1), BLOCK, k_mg_pack, n1, (const long long*)row_off, (const int*)L1.cnt,
        (const int*)pc, (const float*)pv, (const long long*)L1.sl_off, L1.cols, L1.vals);
    *ok = 1;
    *why = 0;
    return 0;
}

// Keys of the aggregation of level l (rows on lattices of exponent at most
// emax): the blocks of spacings 2 q .. 2^(emax+1) q.
static long long mg_key_count(const MGLevel& F, int emax) {
    return lattice_key_base(F.nx, F.ny, F.nz, emax + 2);
}

// Aggregates of level l (its coordinates set) -> level l + 1 rows,
// coordinates and link arrays.  F: level 0 (lattice extents); emax: the
// largest lattice exponent of level l's rows.
template <class OP>
static int mg_aggregate(sph_ctx* c, OP op, const MGLevel& F, int emax, MGLevel& L0, MGLevel& L1) {
    const int n = L0.n;
    int *key = nullptr, *val = nullptr, *ktmp = nullptr, *vtmp = nullptr, *flag = nullptr, *scan = nullptr;
    SPH_TRY(dalloc(c, &key, n));
    SPH_TRY(dalloc(c, &val, n));
    SPH_TRY(dalloc(c, &ktmp, n));
    SPH_TRY(dalloc(c, &vtmp, n));
    MGK("mg_key", grid1(n), BLOCK, k_mg_key, n, (const int*)L0.cx, (const int*)L0.cy, (const int*)L0.cz,
        (const int*)L0.ex, F.nx, F.ny, F.nz, key, val);
    SPH_TRY(radix_sort_pairs(c, &key, &val, &ktmp, &vtmp, n, (int)(mg_key_count(F, emax) - 1)));
    dfree(c, ktmp);
    dfree(c, vtmp);
    SPH_TRY(dalloc(c, &flag, n));
    SPH_TRY(dalloc(c, &scan, n));
    MGK("mg_flag", grid1(n), BLOCK, k_mg_flag, n, (const int*)key, flag);
    MGK("mg_split", grid1(n), BLOCK, k_mg_split<OP>, op, n, (const int*)key, (const int*)val, flag);
    int n1 = 0;
    SPH_TRY(exclusive_scan_i32(c, flag, scan, n, &n1));
    L1.n = n1;
    SPH_TRY(dalloc(c, &L0.agg, n));
    SPH_TRY(dalloc(c, &L0.mem, n));
    SPH_TRY(dalloc(c, &L0.mstart, (size_t)n1 + 1));
    SPH_TRY(dalloc(c, &L1.cx, n1));
    SPH_TRY(dalloc(c, &L1.cy, n1));
    SPH_TRY(dalloc(c, &L1.cz, n1));
    SPH_TRY(dalloc(c, &L1.ex, n1));
    MGK("mg_link", grid1(n), BLOCK, k_mg_link, n, (const int*)key, (const int*)val, (const int*)flag,
        (const int*)scan, (const int*)L0.cx, (const int*)L0.cy, (const int*)L0.cz, (const int*)L0.ex, L0.agg, L0.mem,
        L0.mstart, L1.cx, L1.cy, L1.cz, L1.ex);
    dfree(c, key);
    dfree(c, val);
    dfree(c, flag);
    dfree(c, scan);
    return 0;
}

static int mg_vectors(sph_ctx* c, MGLevel& L, bool fine, bool coarsest, float* fine_x) {
    if (!fine) SPH_TRY(dalloc(c, &L.b, L.n));
    if (fine) {
        L.x = fine_x;
        L.x_shared = true;
    } else {
        SPH_TRY(dalloc(c, &L.x, L.n));
    }
    if (!coarsest) SPH_TRY(dalloc(c, &L.r, L.n));
    if (!fine && !coarsest) SPH_TRY(dalloc(c, &L.y, L.n));
    return 0;
}

int mg_build(sph_ctx* c, const PPEView& pp, long long slots, double q, MGState** out, int* ok, int* why) {
    *out = nullptr;
    *ok = 0;
    *why = SPH_MG_UNSUITED;
    // long rows need the compact level-0 copy (its passes add their own tail
    // sums); the finest level's x is the PPE's (the Krylov kernels write the
    // pre-smoothing there)
    if (pp.nr <= MG_COARSE_MAX || (pp.ltail && !pp.ecols) || !pp.vals32 || !pp.fdiag || !pp.mgx0 || !(q > 0)) return 0;
    const sph_domain& d = c->dom;
    const double ex = ((double)d.x_max - d.x_min) / q + 3.0, ey = ((double)d.y_max - d.y_min) / q + 3.0,
                 ez = ((double)d.z_max - d.z_min) / q + 3.0;
    // level-1 keys are 32-bit
    if ((ex / 2 + 1) * (ey / 2 + 1) * (ez / 2 + 1) > 2.0e9) return 0;
    static std::atomic<long long> next_id{0};
    MGState* mg = new MGState();
    mg->id = ++next_id;
    int* err = nullptr;
    auto fail_out = [&](int code) {
        dfree(c, err);
        mg_free(c, mg);
        return code;
    };
#define MGTRY(x)                          \
    do {                                  \
        int rc_ = (x);                    \
        if (rc_) return fail_out(rc_);    \
    } while (0)
    // a HIP error of the build: reported, and the partial hierarchy freed
    auto hipc = [&](hipError_t er, const char* what) {
        return er == hipSuccess ? 0 : fail(c, SPH_E_HIP, std::string(what) + ": " + hipGetErrorString(er));
    };
    MGTRY(dalloc(c, &err, 1));
    int rmax = 0;   // the largest lattice exponent of the particles
    {
        MGLevel& L0 = mg->L[0];
        L0.n = pp.nr;
        L0.nx = (long long)ex;
        L0.ny = (long long)ey;
        L0.nz = (long long)ez;
        MGTRY(dalloc(c, &L0.cx, L0.n));
        MGTRY(dalloc(c, &L0.cy, L0.n));
        MGTRY(dalloc(c, &L0.cz, L0.n));
        MGTRY(dalloc(c, &L0.ex, L0.n));
        MGTRY(hipc(hipMemsetAsync(err, 0, sizeof(int), c->stream), "mg: lattice"));
        const double x0 = (double)d.x_min - q, y0 = (double)d.y_min - q, z0 = (double)d.z_min - q;
        hipLaunchKernelGGL(k_mg_coord0, dim3(grid1(L0.n)), dim3(BLOCK), 0, c->stream, L0.n, pp.rows, c->sview(), x0,
                           y0, z0, q, L0.nx, L0.ny, L0.nz, L0.cx, L0.cy, L0.cz, L0.ex, err);
        MGTRY(launch_check(c, "mg_coord"));
        MGTRY(hipc(hipMemcpyAsync(&rmax, err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "mg: lattice"));
        MGTRY(hipc(hipStreamSynchronize(c->stream), "mg: lattice"));
        // every level's keys are 32-bit
        if (mg_key_count(L0, rmax + MG_MAX_LEVELS) > 2000000000LL) return fail_out(0);
    }
    int nlev = 1;
    bool good = true;
    while (mg->L[nlev - 1].n > MG_COARSE_MAX) {
        if (nlev == MG_MAX_LEVELS) { good = false; *why = SPH_MG_LEVELS; break; }
        MGLevel& L0 = mg->L[nlev - 1];
        MGLevel& L1 = mg->L[nlev];
        MGTRY(hipc(hipMemsetAsync(err, 0, sizeof(int), c->stream), "mg: clear flag"));
        // level l's rows lie on lattices of exponent at most rmax + l
        if (nlev == 1) MGTRY(mg_aggregate(c, FineOp{pp}, mg->L[0], rmax, L0, L1));
        else MGTRY(mg_aggregate(c, coarse_op(L0), mg->L[0], rmax + nlev - 1, L0, L1));
        if (L1.n * 10LL > L0.n * 9LL) { nlev++; good = false; *why = SPH_MG_STALLED; break; }
        int lok = 0;
        int code = 0;
        if (nlev == 1)
            MGTRY(mg_galerkin(c, FineOp{pp}, L1, L0, err, &lok, &code, c->prm.mg_sort_all, &mg->sorted_rows));
        else MGTRY(mg_galerkin(c, coarse_op(L0), L1, L0, err, &lok, &code, c->prm.mg_sort_all, &mg->sorted_rows));
        if (!lok) *why = code;   // SPH_MG_TERMS or SPH_MG_SPAN
        nlev++;
        if (!lok) { good = false; break; }
    }
    mg->nlev = nlev;
    if (!good || nlev < 2) return fail_out(0);
    // dense inverse of the coarsest level
    {
        MGLevel& LC = mg->L[nlev - 1];
        const int n = LC.n;
        const long long nn = (long long)n * n;
        double *a = nullptr, *rowk = nullptr, *colk = nullptr;
        MGTRY(dalloc(c, &a, (size_t)nn));
        MGTRY(dalloc(c, &rowk, 2 * n));
        MGTRY(dalloc(c, &colk, 2 * n));
        MGTRY(hipc(hipMemsetAsync(err, 0, sizeof(int), c->stream), "mg: coarsest level"));
        hipLaunchKernelGGL(k_mg_dense_zero, dim3(grid1(nn)), dim3(BLOCK), 0, c->stream, nn, a);
        MGTRY(launch_check(c, "mg_dense_zero"));
        hipLaunchKernelGGL(k_mg_dense_fill, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, coarse_op(LC), n, a);
        MGTRY(launch_check(c, "mg_dense_fill"));
        MGTRY(prof_begin(c, "mg_gauss_jordan", 0.0));
        hipLaunchKernelGGL(k_gj_first, dim3(grid1(n)), dim3(BLOCK), 0, c->stream, n, (const double*)a, rowk, colk);
        for (int k = 0; k < n; k++) {
            const int cur = (k & 1) * n, nxt = n - cur;
            hipLaunchKernelGGL(k_gj_step, dim3(grid1(nn)), dim3(BLOCK), 0, c->stream, n, a, k,
                               (const double*)rowk + cur, (const double*)colk + cur, rowk + nxt, colk + nxt, err);
        }
        MGTRY(launch_check(c, "mg_gauss_jordan"));
        MGTRY(prof_end(c));
        MGTRY(dalloc(c, &mg->ainv, (size_t)nn));
        hipLaunchKernelGGL(k_mg_tofloat, dim3(grid1(nn)), dim3(BLOCK), 0, c->stream, nn, (const double*)a, mg->ainv);
        MGTRY(launch_check(c, "mg_tofloat"));
        int h_err = 0;
        MGTRY(hipc(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, c->stream), "mg: coarsest level"));
        MGTRY(hipc(hipStreamSynchronize(c->stream), "mg: coarsest level"));
        dfree(c, a);
        dfree(c, rowk);
        dfree(c, colk);
        if (h_err) { *why = SPH_MG_PIVOT; return fail_out(0); }
    }
    if (pp.nlong > 0) MGTRY(dalloc(c, &mg->lsum, pp.nlong));
    if (pp.ecols) {
        MGTRY(dalloc(c, &mg->ev, (size_t)std::max(1LL, slots)));
        hipLaunchKernelGGL(k_mg_pack16, dim3(grid1(pp.nr)), dim3(BLOCK), 0, c->stream, pp, mg->ev);
        MGTRY(launch_check(c, "mg_pack16"));
        if (pp.spat) {   // stencil slices: their kept halves
            const int ns = (pp.nr + SLICE - 1) / SLICE;
            long long* wh = nullptr;
            int* wp = nullptr;
            long long nh = 0;
            int no = 0;
            MGTRY(dalloc(c, &mg->hw, ns));
            MGTRY(dalloc(c, &mg->hbase, ns + 1));
            MGTRY(dalloc(c, &mg->hpat, ns + 1));
            MGTRY(dalloc(c, &wh, ns));
            MGTRY(dalloc(c, &wp, ns));
            hipLaunchKernelGGL(k_mg_hpack<false>, dim3(ns), dim3(SLICE), 0, c->stream, pp, mg->hw, wh, wp,
                               (const long long*)nullptr, (const int*)nullptr, (unsigned short*)nullptr, (int*)nullptr);
            MGTRY(launch_check(c, "mg_hwidth"));
            MGTRY(exclusive_scan_i64(c, wh, mg->hbase, ns, &nh));
            MGTRY(exclusive_scan_i32(c, wp, mg->hpat, ns, &no));
            dfree(c, wh);
            dfree(c, wp);
            if (nh > INT_MAX) { *why = SPH_MG_TERMS; return fail_out(0); }   // slot numbers are read as int
            MGTRY(dalloc(c, &mg->eh, (size_t)std::max(1LL, nh * SLICE)));
            MGTRY(dalloc(c, &mg->hoff, (size_t)std::max(1, no)));
            hipLaunchKernelGGL(k_mg_hpack<true>, dim3(ns), dim3(SLICE), 0, c->stream, pp, (int*)nullptr,
                               (long long*)nullptr, (int*)nullptr, (const long long*)mg->hbase, (const int*)mg->hpat,
                               mg->eh, mg->hoff);
            MGTRY(launch_check(c, "mg_hpack"));
        }
    }
    for (int l = 0; l < nlev; l++) MGTRY(mg_vectors(c, mg->L[l], l == 0, l == nlev - 1, pp.mgx0));
    for (int l = 1; l + 1 < nlev; l++) {
        MGLevel& L = mg->L[l];
        int *flag = nullptr, *scan = nullptr;
        MGTRY(dalloc(c, &flag, L.n));
        MGTRY(dalloc(c, &scan, L.n));
        hipLaunchKernelGGL(k_mg_longflag, dim3(grid1(L.n)), dim3(BLOCK), 0, c->stream, L.n, (const int*)L.cnt, flag);
        MGTRY(launch_check(c, "mg_longflag"));
        int nl = 0;
        MGTRY(exclusive_scan_i32(c, flag, scan, L.n, &nl));
        if (nl > 0) {
            L.nlong = nl;
            MGTRY(dalloc(c, &L.lrows, nl));
            MGTRY(dalloc(c, &L.lidx, L.n));
            MGTRY(dalloc(c, &L.lsum, nl));
            hipLaunchKernelGGL(k_mg_longlist, dim3(grid1(L.n)), dim3(BLOCK), 0, c->stream, L.n, (const int*)flag,
                               (const int*)scan, L.lrows, L.lidx);
            MGTRY(launch_check(c, "mg_longlist"));
        }
        dfree(c, flag);
        dfree(c, scan);
    }
    // coordinates are needed only while building
    for (int l = 0; l < nlev; l++) {
        MGLevel& L = mg->L[l];
        dfree(c, L.cx); dfree(c, L.cy); dfree(c, L.cz); dfree(c, L.ex);
        L.cx = L.cy = L.cz = L.ex = nullptr;
    }
    dfree(c, err);
#undef MGTRY
    *out = mg;
    *ok = 1;
    *why = SPH_MG_OK;
    return 0;
}

// Descriptors and block tables of the step's multigrid cases.
int mg_prepare(Eng& e, const std::vector<const MGState*>& mgs, int sc_done, bool* changed) {
    (void)sc_done;
    const int nc = (int)mgs.size();
    std::vector<long long> sig;
    for (const MGState* mg : mgs) sig.push_back(mg ? mg->id : 0);
    *changed = sig != e.mg_sig;
    if (!*changed) return 0;
    std::vector<MGLv> lv((size_t)nc * MG_MAX_LEVELS);
    std::memset(lv.data(), 0, sizeof(MGLv) * lv.size());
    Eng::MGPlan P;
    std::vector<int2> tab;
    for (int k = 0; k < nc; k++) {
        const MGState* mg = mgs[k];
        if (!mg) continue;
        P.lmax = std::max(P.lmax, mg->nlev);
        for (int l = 0; l < mg->nlev; l++) {
            const MGLevel& S = mg->L[l];
            MGLv& D = lv[(size_t)k * MG_MAX_LEVELS + l];
            D.n = S.n;
            D.lc = mg->nlev - 1;
            D.sl_off = S.sl_off; D.cnt = S.cnt; D.cols = S.cols; D.vals = S.vals; D.diag = S.diag;
            D.agg = S.agg; D.mstart = S.mstart; D.mem = S.mem;
            D.b = S.b; D.x = S.x; D.y = S.y; D.r = S.r;
            D.ainv = l == mg->nlev - 1 ? mg->ainv : nullptr;
            D.ev = l == 0 ? mg->ev : nullptr;
            if (l == 0) { D.hw = mg->hw; D.hbase = mg->hbase; D.hpat = mg->hpat; D.eh = mg->eh; D.hoff = mg->hoff; }
            D.lsum = l == 0 ? mg->lsum : S.lsum;
            D.nlong = l == 0 ? 0 : S.nlong;
            D.lrows = S.lrows;
            D.lidx = S.lidx;
        }
    }
    auto rows_table = [&](int l, bool below, int* off, int* cnt) {
        // the coarse levels' residual and post-smoothing passes run MG_G
        // lanes per row
        const int per = (below && l > 0) ? BLOCK / MG_G : BLOCK;
        *off = (int)tab.size();
        for (int k = 0; k < nc; k++) {
            const MGState* mg = mgs[k];
            if (!mg) continue;
            const int lc = mg->nlev - 1;
            if (below ? !(l < lc) : !(l <= lc)) continue;
            for (int f = 0; f < mg->L[l].n; f += per) tab.push_back(make_int2(k, f));
        }
        *cnt = (int)tab.size() - *off;
    };
    for (int l = 0; l < P.lmax; l++) {
        rows_table(l, true, &P.down_off[l], &P.down_cnt[l]);
        rows_table(l, false, &P.lvl_off[l], &P.lvl_cnt[l]);
    }
    P.dense_off = (int)tab.size();
    for (int k = 0; k < nc; k++) {
        const MGState* mg = mgs[k];
        if (!mg) continue;
        for (int f = 0; f < mg->L[mg->nlev - 1].n; f += BLOCK / SLICE) tab.push_back(make_int2(k, f));
    }
    P.dense_cnt = (int)tab.size() - P.dense_off;
    P.long_off = (int)tab.size();
    for (int k = 0; k < nc; k++) {
        const MGState* mg = mgs[k];
        if (!mg || !mg->lsum) continue;
        const int nlong = e.cs[k]->ppe ? ppe_nlong(e.cs[k]) : 0;
        for (int f = 0; f < nlong; f += BLOCK / SLICE) tab.push_back(make_int2(k, f));
    }
    P.long_cnt = (int)tab.size() - P.long_off;
    for (int l = 0; l < P.lmax; l++) {
        P.clong_off[l] = (int)tab.size();
        for (int k = 0; k < nc; k++) {
            const MGState* mg = mgs[k];
            if (!mg || l == 0 || l + 1 >= mg->nlev) continue;
            for (int f = 0; f < mg->L[l].nlong; f += BLOCK / SLICE) tab.push_back(make_int2(k, f));
        }
        P.clong_cnt[l] = (int)tab.size() - P.clong_off[l];
    }
    // launches in flight may read the old tables
    SPH_HIP(&e, hipStreamSynchronize(e.st));
    const size_t lv_bytes = sizeof(MGLv) * lv.size(), tab_bytes = sizeof(int2) * std::max<size_t>(1, tab.size());
    if (e.mg_lv_cap < lv_bytes) {
        if (e.mg_lv) (void)hipFree(e.mg_lv);
        e.mg_lv = nullptr; e.mg_lv_cap = 0;
        SPH_HIP(&e, hipMalloc(&e.mg_lv, lv_bytes));
        e.mg_lv_cap = lv_bytes;
    }
    if (e.mg_tab_cap < tab_bytes) {
        if (e.mg_tab) (void)hipFree(e.mg_tab);
        e.mg_tab = nullptr; e.mg_tab_cap = 0;
        SPH_HIP(&e, hipMalloc((void**)&e.mg_tab, tab_bytes));
        e.mg_tab_cap = tab_bytes;
    }
    SPH_HIP(&e, hipMemcpy(e.mg_lv, lv.data(), lv_bytes, hipMemcpyHostToDevice));
    if (!tab.empty()) SPH_HIP(&e, hipMemcpy(e.mg_tab, tab.data(), sizeof(int2) * tab.size(), hipMemcpyHostToDevice));
    P.spmv_bytes0 = e.mg_plan.spmv_bytes0;
    e.mg_plan = P;
    e.mg_sig = sig;
    return 0;
}

// One V-cycle for every multigrid case of the step (see mg_prepare).
int mg_apply(Eng& e, int which, int sc_done) {
    const Eng::MGPlan& P = e.mg_plan;
    const CaseDesc* cds = e.d_desc;
    const MGLv* lv = (const MGLv*)e.mg_lv;
    const int2* T = e.mg_tab;
#define MGE(name, bytes, tab_off, cnt, kern, ...)                                                    \
    do {                                                                                             \
        if ((cnt) > 0) {                                                                             \
            SPH_TRY(prof_begin(e, name, (double)(bytes)));                                           \
            hipLaunchKernelGGL(kern, dim3(cnt), dim3(BLOCK), 0, e.st, cds, lv, T + (tab_off), ##__VA_ARGS__); \
            SPH_TRY(launch_check(&e, name));                                                         \
            SPH_TRY(prof_end(e));                                                                    \
        }                                                                                            \
    } while (0)
    // The pointwise passes ride on the passes that form their inputs: the
    // finest level's pre-smoothing on the Krylov kernel that forms p or s
    // (k_kr_p, k_kr_s), a coarse level's on the restriction that forms its
    // b, the prolongation of a level's correction on the pass that forms it
    // (the coarse post-smoothing, or the dense solve)
    for (int l = 0; l + 1 < P.lmax; l++) {
        if (l == 0) MGE("mg_long", 0, P.long_off, P.long_cnt, k_mg_long, sc_done);
        else MGE("mg_clong", 0, P.clong_off[l], P.clong_cnt[l], k_mg_clong, l, sc_done);
        if (l == 0) MGE("mg_res0", P.spmv_bytes0, P.down_off[l], P.down_cnt[l], k_mg_res, l, which, sc_done);
        else MGE("mg_res", 0, P.down_off[l], P.down_cnt[l], k_mg_res_g, l, sc_done);
        MGE("mg_restrict", 0, P.lvl_off[l + 1], P.lvl_cnt[l + 1], k_mg_restrict, l, sc_done);
    }
    MGE("mg_dense", 0, P.dense_off, P.dense_cnt, k_mg_dense, sc_done);
    for (int l = P.lmax - 2; l >= 0; l--) {
        if (l == 0) MGE("mg_long", 0, P.long_off, P.long_cnt, k_mg_long, sc_done);
        else MGE("mg_clong", 0, P.clong_off[l], P.clong_cnt[l], k_mg_clong, l, sc_done);
        if (l == 0) MGE("mg_post0", P.spmv_bytes0, P.down_off[l], P.down_cnt[l], k_mg_post, l, which, sc_done);
        else MGE("mg_post", 0, P.down_off[l], P.down_cnt[l], k_mg_post_g, l, sc_done);
    }
#undef MGE
    return 0;
}

}  // namespace sph
