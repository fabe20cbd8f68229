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
// Batching: every kernel covers the rows of all cases of the step (segment
// table, 256-row blocks that never straddle two cases).  Each case has its
// own scalars, controls (tolerance, iteration cap, skip) and chunk partials,
// and one finalize block per case advances its scalars, so a case's iterate
// sequence does not depend on which other cases share the launches.
//
// Reductions: every dot product is the canonical tree of 256-element chunks
// (one 256-thread block per chunk, pairwise offsets 128..1, chunk partials
// reduced the same way), so results are bit-reproducible.  Scalars live on
// the device; kernels of a case that has converged return immediately, so
// the host polls only every few iterations and the iterate sequence matches
// a sequential implementation.
//
// MIXED precision: matrix values and Krylov vectors of an inner BiCGSTAB are
// float, every SpMV row sum and every dot is accumulated in double; an outer
// iterative-refinement loop computes the residual with double vectors and
// double accumulation and adds the float correction to a double solution.
#include <algorithm>

#include <cfloat>
#include <cmath>

#include "flags.h"
#include "sph_internal.h"
#include "sph_math.h"
#include "ppe_format.h"

namespace sph {

enum Scal {
    SC_RHO, SC_RHO_OLD, SC_ALPHA, SC_OMEGA, SC_BETA, SC_BNORM, SC_RNORM, SC_SNORM, SC_TOL,
    SC_DONE, SC_CONV, SC_HALF, SC_IT, SC_ITERS, SC_MAXIT, SC_ZEROX, SC_COUNT = SCAL_COUNT
};

struct PPEState {
    int nr = 0;                 // unknowns
    int nslices = 0;
    long long nnz_slots = 0;    // SELL storage incl. padding
    long long nnz = 0;          // real entries
    int ordering = SPH_ORDER_CELL;
    int* rowof = nullptr;       // particle -> row or -1
    int* rows = nullptr;        // row -> particle
    int* cnt = nullptr;         // entries per row
    long long* sl_off = nullptr;
    int* cols = nullptr;
    unsigned short* ecols = nullptr;
    int* gbase = nullptr;
    unsigned char* slotmap = nullptr;   // lattice row order: slot of each entry in list order
    // stencil slices (see k_stencil_union): per slice the first of its
    // column offsets in poff, or -1 where the slice stores its columns
    int* spat = nullptr;
    int* poff = nullptr;
    int nstencil = 0;           // stencil slices
    long long nnz_stencil = 0;  // real entries in them (after the squeeze: the nonzero ones)
    bool squeezed = false;      // the stencil slices dropped their zero offsets: values final
    double* vals64 = nullptr;
    float* vals32 = nullptr;
    double* diag = nullptr;
    double* rhs = nullptr;
    double* x = nullptr;        // solution q (double)
    double* pnew = nullptr;     // Jacobi output
    float* fdiag = nullptr;
    // Krylov work vectors (double set and float set)
    double *r = nullptr, *rhat = nullptr, *p = nullptr, *phat = nullptr, *v = nullptr, *s = nullptr,
           *shat = nullptr, *t = nullptr;
    float *fr = nullptr, *frhat = nullptr, *fp = nullptr, *fphat = nullptr, *fv = nullptr, *fs = nullptr,
          *fshat = nullptr, *ft = nullptr, *fx = nullptr, *frhs = nullptr;
    float* mgx0 = nullptr;      // the multigrid V-cycle's finest-level x (its hierarchies use it)
    double* partA = nullptr;
    double* partB = nullptr;
    int nb = 0;
    int* ltail = nullptr;       // MIXED: long rows (more than LONG_ROW entries) and their tails
    int* lrows = nullptr;
    long long* tail_off = nullptr;
    int* tail_cols = nullptr;
    float* tail_vals = nullptr;
    int nlong = 0;
    long long tail_total = 0;
    double* lsum = nullptr;
    // MIXED multigrid preconditioner of the current values: 0 not built,
    // 1 built (mg), -1 not usable for this matrix (Jacobi)
    MGState* mg = nullptr;
    int mg_state = 0;
    int mg_why = SPH_MG_OK;     // SPH_MG_*: why the last build gave no hierarchy
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
// two sums at once (waves 0 and 1): .a valid on thread 0, .b on thread 64
struct Pair { double a, b; };
__device__ __forceinline__ Pair block_canon2(double a, double b, double* sh) {
    const int t = threadIdx.x;
    sh[t] = a;
    sh[BLOCK + t] = b;
    __syncthreads();
    Pair r{0.0, 0.0};
    if (t < 64) r.a = wave_tree(sh[t], sh[t + 64], sh[t + 128], sh[t + 192]);
    else if (t < 128) {
        const int l = t - 64;
        r.b = wave_tree(sh[BLOCK + l], sh[BLOCK + l + 64], sh[BLOCK + l + 128], sh[BLOCK + l + 192]);
    }
    __syncthreads();
    return r;
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
// structure (per case, once per neighbour search)
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
                                                      int* cnt, const long long* sl_off, int* cols, const int* ltail,
                                                      const long long* tail_off, int* tail_cols) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    int i = rows[r];
    const unsigned* row = nl.ent + nl.slice_off[i >> 6] + (i & 63);
    int m = nl.cnt[i];
    int slot = 0;
    int* out = FILL ? cols + prow(sl_off, r) : nullptr;
    for (int k = 0; k < m; k++) {
        unsigned e = row[(size_t)k * SLICE];
        if (!(e & F_DP)) continue;
        int j = e & NB_IDX;
        if (S.p_type[j] >= 1000 || S.i_type[j] == 4) continue;
        if (j == i) continue;
        int rj = rowof[j];
        if (rj < 0) continue;
        if (FILL) {
            if (slot < LONG_ROW || !ltail) out[pslot(slot)] = rj;
            else tail_cols[tail_off[ltail[r]] + (slot - LONG_ROW)] = rj;
        }
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

// Lattice row order of the MIXED matrix.  Rows are renumbered by the
// (z, y, x) lattice position of their particle, so the 64 rows of a slice
// are consecutive lattice sites whose k-th neighbours are consecutive too,
// and the x gathers of a wave fall on a few cache lines.  Only the PPE's
// unknowns are renumbered; particles keep the reference's cell order.
__global__ void __launch_bounds__(BLOCK) k_lattice_key(int nr, const int* rows, PView S, double x0, double y0,
                                                       double z0, double q, long long nx, long long ny,
                                                       long long* key) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const int i = rows[r];
    const long long kx = (long long)floor((S.x[i] - x0) / q + 0.5);
    const long long ky = (long long)floor((S.y[i] - y0) / q + 0.5);
    const long long kz = (long long)floor((S.z[i] - z0) / q + 0.5);
    key[r] = (kz * ny + ky) * nx + kx;
}
// The same key in 32 bits (lattices of fewer than 2^31 sites), paired with
// the row's particle for the device radix sort.
__global__ void __launch_bounds__(BLOCK) k_lattice_key32(int nr, const int* rows, PView S, double x0, double y0,
                                                         double z0, double q, long long nx, long long ny, int* key,
                                                         int* val) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const int i = rows[r];
    const long long kx = (long long)floor((S.x[i] - x0) / q + 0.5);
    const long long ky = (long long)floor((S.y[i] - y0) / q + 0.5);
    const long long kz = (long long)floor((S.z[i] - z0) / q + 0.5);
    key[r] = (int)((kz * ny + ky) * nx + kx);
    val[r] = i;
}
__global__ void __launch_bounds__(BLOCK) k_long_flag(int nr, const int* cnt, int* flag) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr) flag[r] = cnt[r] > LONG_ROW ? 1 : 0;
}
// long-row tables; the row's count becomes its SELL part
__global__ void __launch_bounds__(BLOCK) k_long_index(int nr, int* cnt, const int* flag, const int* scan, int* lrows,
                                                      int* ltail, long long* tlen) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    if (flag[r]) {
        const int u = scan[r];
        lrows[u] = r;
        ltail[r] = u;
        tlen[u] = cnt[r] - LONG_ROW;
        cnt[r] = LONG_ROW;
    } else {
        ltail[r] = -1;
    }
}
__global__ void __launch_bounds__(BLOCK) k_scale_i64(long long n, long long* a, long long f) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] *= f;
}
__global__ void __launch_bounds__(BLOCK) k_widen(int n, const int* a, long long* b) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) b[r] = a[r];
}
__global__ void __launch_bounds__(BLOCK) k_rowof_from_rows(int nr, const int* rows, int* rowof) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr) rowof[rows[r]] = r;
}
// Columns of each row in ascending (lattice) order, and the slot each entry
// takes, indexed by its position in the neighbour list (the order in which
// the assembly visits it).  The fill pass has written every row's columns in
// list order, a thread per row; this pass sorts them in place, one workgroup
// per 64-row slice.  The slice's columns are read slot by slot (whole lines)
// into LDS; then one wave per row sorts the row's (column, position) keys,
// held in registers, by a bitonic network over 64, 128 or 256 keys (the keys
// are distinct, so the order is unique); then the sorted columns and the
// slot map are written slot by slot.  Keys are 32-bit
// (column << 8 | position) while the columns fit 24 bits, else 64-bit.  A
// slice with a row longer than SORT_CAP sets *overflow and writes nothing
// (the case then keeps cell order).
constexpr int SORT_CAP = 160;
static_assert(SORT_CAP <= 4 * 64, "list positions fit 8 bits; a row fits the widest register sort");
// lattice-ordered rows fit the sort, so they never have tails
static_assert(LONG_ROW >= SORT_CAP, "a lattice-ordered row must fit the SELL part");
constexpr int SORT_NARROW_ROWS = 1 << 24;   // rows whose columns fit a 32-bit key
constexpr int SORT_BLOCK = 1024;            // 16 waves; LDS 62 KB a workgroup
template <class K> struct SortKey;
template <> struct SortKey<unsigned> { static constexpr int SHIFT = 8; };
template <> struct SortKey<unsigned long long> { static constexpr int SHIFT = 32; };
// Row l of the staged slice (m <= Q * 64 entries) sorted by (column, list
// position): the slot map (position -> place) and its inverse into LDS.
template <class K, int Q>
__device__ __forceinline__ void sort_row(const int (*col_s)[SLICE + 1], unsigned char (*map_s)[SLICE + 4],
                                         unsigned char (*inv_s)[SLICE + 4], int l, int m, int lane) {
    constexpr int SH = SortKey<K>::SHIFT;
    constexpr K POS = ((K)1 << SH) - 1;
    K v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int t = lane + q * SLICE;
        v[q] = t < m ? ((K)(unsigned)col_s[t][l] << SH) | (K)(unsigned)t : ~(K)0;   // pads sort last
    }
    bitonic_wave<K, Q>(v, lane);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int t = lane + q * SLICE;
        if (t < m) {
            const int pos = (int)(v[q] & POS);
            map_s[pos][l] = (unsigned char)t;
            inv_s[t][l] = (unsigned char)pos;
        }
    }
}
template <class K>
__global__ void __launch_bounds__(SORT_BLOCK) k_ppe_sort_slices(int nr, const int* cnt, const long long* sl_off,
                                                                int* cols, unsigned char* slotmap, int* overflow) {
    constexpr int WAVES = SORT_BLOCK / SLICE;
    // slot-major like the matrix; rows padded so a wave's accesses to one
    // row (one slot per lane) spread over the banks
    __shared__ int col_s[SORT_CAP][SLICE + 1];
    __shared__ unsigned char map_s[SORT_CAP][SLICE + 4];   // list position -> place
    __shared__ unsigned char inv_s[SORT_CAP][SLICE + 4];   // place -> list position
    __shared__ int m_s[SLICE];
    __shared__ int bad;
    const int w = threadIdx.x / SLICE, lane = threadIdx.x % SLICE;
    // consecutive slices on one XCD
    const int sl = (int)xcd_block();
    const long long base = sl_off[sl];
    if (threadIdx.x == 0) bad = 0;
    if (threadIdx.x < SLICE) {
        const int r = sl * SLICE + (int)threadIdx.x;
        m_s[threadIdx.x] = r < nr ? cnt[r] : 0;
    }
    __syncthreads();
    if (threadIdx.x < SLICE && m_s[threadIdx.x] > SORT_CAP) bad = 1;
    for (int t = threadIdx.x; t < SORT_CAP * SLICE; t += SORT_BLOCK) {
        const int k = t / SLICE, l = t % SLICE;
        if (k < m_s[l]) col_s[k][l] = cols[base + pslot(k) + l];
    }
    __syncthreads();
    if (bad) {
        if (threadIdx.x == 0) *overflow = 1;
        return;
    }
    for (int l = w; l < SLICE; l += WAVES) {
        const int m = m_s[l];   // uniform in the wave
        if (m == 0) continue;
        if (m <= SLICE) sort_row<K, 1>(col_s, map_s, inv_s, l, m, lane);
        else if (m <= 2 * SLICE) sort_row<K, 2>(col_s, map_s, inv_s, l, m, lane);
        else sort_row<K, 4>(col_s, map_s, inv_s, l, m, lane);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < SORT_CAP * SLICE; t += SORT_BLOCK) {
        const int k = t / SLICE, l = t % SLICE;
        if (k < m_s[l]) {
            cols[base + pslot(k) + l] = col_s[inv_s[k][l]][l];
            slotmap[base + pslot(k) + l] = map_s[k][l];
        }
    }
}

// 16-bit columns of the PPE matrix, one 64-thread block per 64-row slice.
// The distinct aligned 1024-column windows that the slice's entries touch
// become its groups, in ascending order (collected in an LDS bitmap, so a
// row may list its columns in any order).  Entry k of row r becomes
// (group << 10) | (col & 1023); the slice's group g starts at column
// gbase[64 * slice + g].  More than 64 groups: *overflow = 1 and the case
// keeps 32-bit columns.
__global__ void __launch_bounds__(SLICE) k_ecols_build(int nr, const long long* sl_off, const int* cnt,
                                                       const int* cols, unsigned short* ecols, int* gbase,
                                                       int* overflow) {
    // the windows a slice touches, as a bitmap in LDS; rows may list their
    // columns in any order
    __shared__ unsigned win[EC_MAX_WORDS];
    __shared__ int gk[EC_MAX_GROUPS];
    __shared__ int ngs;
    const int sl = blockIdx.x, l = threadIdx.x, r = sl * SLICE + l;
    const int m = r < nr ? cnt[r] : 0;
    const int* cp = cols + sl_off[sl] + l;
    const int nwords = (((nr - 1) >> EC_OFF_BITS) >> 5) + 1;
    for (int w = l; w < nwords; w += SLICE) win[w] = 0u;
    __syncthreads();
    for (int k = 0; k < m; k++) {
        const int key = cp[pslot(k)] >> EC_OFF_BITS;
        atomicOr(&win[key >> 5], 1u << (key & 31));
    }
    __syncthreads();
    if (l == 0) {
        int ng = 0;
        for (int w = 0; w < nwords; w++) {
            unsigned b = win[w];
            while (b) {
                const int bit = __builtin_ctz(b);
                b &= b - 1;
                if (ng < EC_MAX_GROUPS) gk[ng] = (w << 5) + bit;
                ng++;
            }
        }
        ngs = ng;
    }
    __syncthreads();
    const int ng = ngs;
    if (ng > EC_MAX_GROUPS) {
        if (l == 0) *overflow = 1;
        return;
    }
    gbase[(size_t)sl * SLICE + l] = l < ng ? gk[l] << EC_OFF_BITS : 0;
    unsigned short* ep = ecols + sl_off[sl] + l;
    for (int k2 = 0; k2 < m; k2++) {
        const int c = cp[pslot(k2)];
        const int key = c >> EC_OFF_BITS;
        int lo = 0, hi = ng - 1;   // gk is ascending and holds key
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (gk[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        ep[pslot(k2)] = (unsigned short)((lo << EC_OFF_BITS) | (c & EC_OFF_MASK));
    }
}

// Stencil slices.  With fixed particles on a lattice and rows in lattice
// order, the 64 rows of a slice have nearly the same column offsets
// (col - row): the same lattice stencil, cut by the domain's walls, with
// the neighbours at exactly the search radius kept or dropped by rounding.
// A slice whose rows' offsets have a union of at most 5/4 of its widest row
// (and at most ST_CAP) stores every row over that union, in ascending
// order, an offset the row lacks holding the value 0: its products then
// read no columns, and the wave's x gathers of one slot are the contiguous
// x[row + offset].  The nonzero entries keep their ascending column order,
// so every sum adds the same terms in the same order (plus exact zeros).
constexpr int ST_CAP = SORT_CAP;
constexpr int ST_HASH = 512;
static_assert(ST_CAP <= 255, "stencil slots are numbered in the 8-bit slot map");
// The union of one slice's offsets, ascending, into ptmp[ST_CAP * slice];
// wp[slice] = its size, or 0 when the slice keeps its columns.
__global__ void __launch_bounds__(SLICE) k_stencil_union(int nr, const long long* sl_off, const int* cnt,
                                                         const int* cols, int* wp, int* ptmp) {
    __shared__ int hk[ST_HASH];
    __shared__ int uo[ST_CAP];
    __shared__ int nd, wmax, bad;
    const int s = blockIdx.x, l = threadIdx.x, r = s * SLICE + l;
    for (int h = l; h < ST_HASH; h += SLICE) hk[h] = INT_MIN;
    if (l == 0) { nd = 0; wmax = 0; bad = 0; }
    __syncthreads();
    const int m = r < nr ? cnt[r] : 0;
    atomicMax(&wmax, m);
    const int* cp = cols + sl_off[s] + l;
    for (int k = 0; k < m; k++) {
        const int d = cp[pslot(k)] - r;
        int h = (int)(((unsigned)d * 2654435761u) >> 23) & (ST_HASH - 1);
        for (int probe = 0; probe < ST_HASH; probe++) {
            const int old = atomicCAS(&hk[h], INT_MIN, d);
            if (old == INT_MIN) {
                const int u = atomicAdd(&nd, 1);
                if (u < ST_CAP) uo[u] = d;
                break;
            }
            if (old == d) break;
            h = (h + 1) & (ST_HASH - 1);
        }
    }
    __syncthreads();
    const int n = nd, w = wmax;
    if (n > ST_CAP || n == 0 || 4 * n > 5 * w) {
        if (l == 0) wp[s] = 0;
        return;
    }
    // ascending: each distinct offset's rank
    int* out = ptmp + (size_t)s * ST_CAP;
    int lo = INT_MAX, hi = INT_MIN;
    for (int i = l; i < n; i += SLICE) {
        const int v = uo[i];
        int rank = 0;
        for (int j = 0; j < n; j++) rank += uo[j] < v ? 1 : 0;
        out[rank] = v;
        lo = min(lo, v);
        hi = max(hi, v);
    }
    // every row's columns row + offset stay inside the operator
    __shared__ int slo, shi;
    if (l == 0) { slo = INT_MAX; shi = INT_MIN; }
    __syncthreads();
    atomicMin(&slo, lo);
    atomicMax(&shi, hi);
    __syncthreads();
    if (r < nr && ((long long)r + slo < 0 || (long long)r + shi >= nr)) atomicOr(&bad, 1);
    __syncthreads();
    if (l == 0) wp[s] = bad ? 0 : n;
}
// Entries per row of the new layout: a stencil slice's rows all hold its union.
__global__ void __launch_bounds__(BLOCK) k_stencil_cnt(int nr, const int* cnt, const int* wp, int* cnt2) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const int n = wp[r >> 6];
    cnt2[r] = n > 0 ? n : cnt[r];
}
// The new layout (one 64-thread block per slice): a stencil slice's columns
// row + offset for every slot and its slot map pointed at the offsets' slots;
// other slices copied.  spat[slice] = the first of its offsets in poff, or -1.
__global__ void __launch_bounds__(SLICE) k_stencil_relayout(int nr, const long long* sl_off, const int* cnt,
                                                            const int* cols, const unsigned char* slotmap,
                                                            const int* wp, const int* ptmp,
                                                            const long long* sl_off2, int* cols2,
                                                            unsigned char* slotmap2, int* spat, int* poff) {
    __shared__ int off[ST_CAP];
    const int s = blockIdx.x, l = threadIdx.x, r = s * SLICE + l;
    const int n = wp[s];
    const long long b2 = sl_off2[s];
    if (l == 0) spat[s] = n > 0 ? (int)(b2 / SLICE) : -1;
    for (int k = l; k < n; k += SLICE) {
        off[k] = ptmp[(size_t)s * ST_CAP + k];
        poff[b2 / SLICE + k] = off[k];
    }
    __syncthreads();
    if (r >= nr) return;
    const int m = cnt[r];
    const int* cp = cols + sl_off[s] + l;
    const unsigned char* sm = slotmap + sl_off[s] + l;
    int* cq = cols2 + b2 + l;
    unsigned char* sq = slotmap2 + b2 + l;
    if (n > 0) {
        for (int k = 0; k < n; k++) cq[pslot(k)] = r + off[k];
        for (int t = 0; t < m; t++) {   // t: rank in the neighbour list
            const int d = cp[pslot(sm[pslot(t)])] - r;
            int lo = 0, hi = n - 1;   // off is ascending and holds d
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (off[mid] < d) lo = mid + 1;
                else hi = mid;
            }
            sq[pslot(t)] = (unsigned char)lo;
        }
    } else {
        for (int k = 0; k < m; k++) {
            cq[pslot(k)] = cp[pslot(k)];
            sq[pslot(k)] = sm[pslot(k)];
        }
    }
}
// real entries of the stencil slices, per row (for the byte counts)
__global__ void __launch_bounds__(BLOCK) k_stencil_real(int nr, const int* cnt, const int* wp, long long* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr) out[r] = wp[r >> 6] > 0 ? cnt[r] : 0;
}
// After the first assembly of a search (with fixed particles the values are
// final until the next search): a stencil slice drops the offsets whose
// values are zero in every one of its rows (the search-radius neighbours,
// see stored_value), one 64-thread block per slice; a wave ballot decides,
// so every lane agrees.  MOVE = false: the new width of each slice (w2; the
// old one for slices that store columns) and the stored nonzero values of
// the stencil slices (nzs); MOVE = true: values, columns and offsets into
// the new layout, and spat pointed at the slice's offsets there.
template <bool MOVE>
__global__ void __launch_bounds__(SLICE) k_stencil_squeeze(int nr, const long long* sl_off, const int* cnt,
                                                           int* spat, const int* poff, const float* vals,
                                                           const unsigned short* ecols, const int* cols,
                                                           long long* w2, long long* nzs, const long long* sl_off2,
                                                           float* vals2, unsigned short* ecols2, int* cols2,
                                                           int* poff2, int* cnt2) {
    const int s = blockIdx.x, l = threadIdx.x, r = s * SLICE + l;
    const int st = spat[s];
    const bool live = r < nr;
    const long long b = sl_off[s] + l;
    if (st < 0) {   // columns stored: copied as they are
        if (!MOVE) {
            int m = live ? cnt[r] : 0;
            for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
            if (l == 0) { w2[s] = m; nzs[s] = 0; }
            return;
        }
        if (!live) return;
        const int m = cnt[r];
        const long long b2 = sl_off2[s] + l;
        for (int k = 0; k < m; k++) {
            vals2[b2 + pslot(k)] = vals[b + pslot(k)];
            if (ecols) ecols2[b2 + pslot(k)] = ecols[b + pslot(k)];
            if (cols) cols2[b2 + pslot(k)] = cols[b + pslot(k)];
        }
        cnt2[r] = m;
        return;
    }
    const int n = cnt[s * SLICE];   // every row of a stencil slice holds its offsets
    int kept = 0, nz = 0;
    for (int k = 0; k < n; k++) {
        const float v = live ? vals[b + pslot(k)] : 0.0f;
        const bool keep = __ballot(v != 0.0f) != 0;
        if (MOVE && keep) {
            const long long e2 = sl_off2[s] + pslot(kept) + l;
            if (live) {
                vals2[e2] = v;
                if (ecols) ecols2[e2] = ecols[b + pslot(k)];
                if (cols) cols2[e2] = cols[b + pslot(k)];
            }
            if (l == 0) poff2[sl_off2[s] / SLICE + kept] = poff[st + k];
        }
        nz += v != 0.0f ? 1 : 0;
        kept += keep ? 1 : 0;
    }
    if (!MOVE) {
        for (int o = 32; o > 0; o >>= 1) nz += __shfl_xor(nz, o, 64);
        if (l == 0) { w2[s] = kept; nzs[s] = nz; }
        return;
    }
    // (one wave: every lane read spat[s] above, before this store)
    if (l == 0) spat[s] = (int)(sl_off2[s] / SLICE);
    if (live) cnt2[r] = kept;
}

// This thread's PPE row: case descriptor `cd`, row index `row` in the case
// and the case-local 256-row chunk `chunk` of the block.
#define SEG_ROW(sg)                               \
    int first_;                                   \
    const int cid_ = seg_case((sg), first_);      \
    const CaseDesc& cd = (sg).cd[cid_];           \
    const int row = first_ + (int)threadIdx.x;    \
    const int chunk = first_ / BLOCK;             \
    (void)chunk;

__device__ __forceinline__ const double* vals_of(const PPEView& v, double*) { return v.vals64; }
__device__ __forceinline__ const float* vals_of(const PPEView& v, float*) { return v.vals32; }
__device__ __forceinline__ double* wvals_of(const PPEView& v, double*) { return v.vals64; }
__device__ __forceinline__ float* wvals_of(const PPEView& v, float*) { return v.vals32; }

// Krylov vector set of one precision (x and rhs: the system being solved)
template <class X>
struct KSet { X *r, *rhat, *p, *phat, *v, *s, *shat, *t, *x; const X* rhs; };
// diagonal used with each vector set
__device__ __forceinline__ const double* kdiag(const PPEView& v, double*) { return v.diag; }
__device__ __forceinline__ const float* kdiag(const PPEView& v, float*) { return v.fdiag; }
__device__ __forceinline__ KSet<double> kset(const PPEView& v, double*) {
    return {v.r, v.rhat, v.p, v.phat, v.v, v.s, v.shat, v.t, v.x, v.rhs};
}
__device__ __forceinline__ KSet<float> kset(const PPEView& v, float*) {
    return {v.fr, v.frhat, v.fp, v.fphat, v.fv, v.fs, v.fshat, v.ft, v.fx, v.frhs};
}

// ------------------------------------------------------------------------
// per-step assembly, functions_PPE.cuh:2-160
// The stored PPE coefficient.  Float values (the mixed mode) of magnitude
// below 1e-30 are stored as 0: the neighbours at exactly the search radius,
// where the kernel gradient vanishes, give coefficients of about 1e-39,
// which no float or double sum of this operator can feel (its entries are
// O(1 / dx^2)), and as zeros they can leave the stencil slices.  FP64 stores
// every value as computed.
template <class V, class R>
__device__ __forceinline__ V stored_value(R cc) {
    if constexpr (sizeof(V) == 4) {
        const float v = (float)cc;
        return fabsf(v) < 1e-30f ? 0.0f : v;
    } else {
        return (V)cc;
    }
}
template <class R, class V, class Src>
__device__ __forceinline__ void assemble_body(const CaseDesc& cd, int i, const NList& nl, int li, const Src src,
                                              int jacobi) {
    const PPEView pp = cd.ppe;
    const int r = pp.rowof[i];
    if (r < 0) return;
    const KConst kc = cd.kc;
    const PView S = cd.S;
    const QView Q = cd.Q;
    const double dt = cd.dt;
    const int* rowof = pp.rowof;
    R hi = (R)(1.0 * S.h[i]);
    R tmp_A = calc_tmpA<R>(kc, hi);
    double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    double uxi = S.ux[i], uyi = S.uy[i], uzi = S.uz[i];
    R rhoi = (R)S.rho[i];
    R gc11 = (R)Q.inv_xx[i], gc12 = (R)Q.inv_xy[i], gc22 = (R)Q.inv_yy[i];
    R gc23 = (R)Q.inv_yz[i], gc31 = (R)Q.inv_zx[i], gc33 = (R)Q.inv_zz[i];
    double bix = 0, biy = 0, biz = 0, Aij = 0, AijPij = 0, kn = 0;
    const NLRow nr = nl_row(nl, li);
    const int m = nr.cnt;
    // with fixed particles the values of a search are written by its first
    // step only; later steps compute the right-hand side
    V* out = (jacobi || cd.vals_fresh) ? nullptr : wvals_of(pp, (V*)nullptr) + prow(pp.sl_off, r);
    const bool full = jacobi || out != nullptr;   // the coefficients of every pair are needed
    const unsigned char* smap = pp.slotmap ? pp.slotmap + prow(pp.sl_off, r) : nullptr;
    int slot = 0;
    // read once: the stores of the loop could alias the descriptor
    const bool compact = Src::lds || (sizeof(R) == 4 && src.ok());
    const double* const rec = cd.rec;
    for (int k = 0; k < m; k++) {
        unsigned e = NB_AT(nr, k);
        if (!(e & F_DP)) continue;
        int j = e & NB_IDX;
        double xj, yj, zj, mj_d, rhoj_d, qj, uxj, uyj, uzj;
        bool jrow;   // j is an unknown of the PPE
        if (compact) {   // the compact record (positions float-exact)
            const float* C = src.at(j);
            const rec_f4 b = Src::f4(src.piece(C, 1));
            if (crec_ptype(b.w) >= 1000 || crec_itype(b.w) == 4) continue;
            jrow = crec_has_row(b.w);
            const rec_f4 a = Src::f4(C);
            const rec_d2 u = Src::d2(src.piece(C, 2)), w = Src::d2(src.piece(C, 3));
            xj = a.x; yj = a.y; zj = a.z; mj_d = a.w; rhoj_d = b.x;
            uxj = u.x; uyj = u.y; uzj = w.x; qj = w.y;
        } else {
            if constexpr (Src::lds) continue;   // staged records are compact
            double vtj_, hj_;
            int ptj, itj;
            rec_aux(rec, j, vtj_, ptj, itj);
            if (ptj >= 1000 || itj == 4) continue;
            rec_pos(rec, j, xj, yj, zj, mj_d, rhoj_d, qj);
            rec_vel(rec, j, uxj, uyj, uzj, hj_);
            jrow = rowof[j] >= 0;
        }
        R dx = (R)(xi - xj), dy = (R)(yi - yj), dz = (R)(zi - zj);
        R rhoj = (R)rhoj_d, mj = (R)mj_d;
        R tdist = sqrt(dx * dx + dy * dy + dz * dz) + (R)1e-20;
        R tdwij = kernel_dwp<R>(kc, tmp_A, hi, tdist);
        R tdwx = tdwij * dx / tdist;
        R tdwy = tdwij * dy / tdist;
        R tdwz = tdwij * dz / tdist;
        R tdwxc = gc11 * tdwx + gc12 * tdwy + gc31 * tdwz;
        R tdwyc = gc12 * tdwx + gc22 * tdwy + gc23 * tdwz;
        R tdwzc = gc31 * tdwx + gc23 * tdwy + gc33 * tdwz;
        R hm = (R)2.0 * rhoi * rhoj / (rhoi + rhoj);
        bix += (double)(mj / rhoj * (R)(uxj - uxi) * tdwxc * hm);
        biy += (double)(mj / rhoj * (R)(uyj - uyi) * tdwyc * hm);
        biz += (double)(mj / rhoj * (R)(uzj - uzi) * tdwzc * hm);
        auto coef = [&]() { return (R)2.0 * mj / rhoj * (dx * tdwxc + dy * tdwyc + dz * tdwzc) / tdist / tdist; };
        if (full) {
            R cc = coef();
            double ccd = (double)cc;
            Aij += ccd;
            if (jacobi) AijPij += ccd * qj;
            if (!src.self(j, i)) {
                if (jrow) {
                    if (out) {
                        const V v = stored_value<V>(cc);
                        if (slot < LONG_ROW || !pp.ltail) out[pslot(smap ? smap[pslot(slot)] : slot)] = v;
                        else pp.tail_vals[pp.tail_off[pp.ltail[r]] + (slot - LONG_ROW)] = (float)v;
                    }
                    slot++;
                } else {
                    kn += ccd * qj;
                }
            }
        } else if (!jrow && !src.self(j, i)) {
            // values and diagonal kept from the search's first step: only the
            // known neighbours' share of the right-hand side needs the coefficient
            kn += (double)coef() * qj;
        }
    }
    double bi = 1.0 * (bix + biy + biz + 0.0) / dt;
    double q0 = S.pres[i] - S.pres0[i];
    if (jacobi) {
        pp.pnew[r] = S.pres0[i] + (bi + AijPij) / Aij;
    } else {
        if (out) {
            pp.diag[r] = Aij;
            if (pp.fdiag) pp.fdiag[r] = (float)Aij;
        }
        pp.rhs[r] = bi + kn;
        pp.x[r] = q0;
    }
}
template <class R, class V>
__global__ void __launch_bounds__(MAX_BLOCK) k_ppe_assemble(Seg sg, int jacobi) {
    SEG_ROW(sg);
    if (row >= cd.n) return;
    const int i = cd.porder ? cd.porder[row] : row;   // this thread's particle
    assemble_body<R, V>(cd, i, cd.nl, i, GSrc<CREC_ADV_FLOATS>{sizeof(R) == 4 ? &cd.crec_adv : nullptr}, jacobi);
}
__global__ void __launch_bounds__(TILE_MAX) k_ppe_assemble_tile(Seg sg, int jacobi) {
    TILE_PARTICLE(sg, cd.crec_adv, CREC_ADV_FLOATS);
    if (chunk_.y >= 0)
        assemble_body<float, float>(cd, i, cd.tile.nl, lane_, LSrc<CREC_ADV_FLOATS>{(const float*)tile_lds_, chunk_.y,
                                                     cd.tile.halo + chunk_.x,
                                                        cd.tile.self[lane_]}, jacobi);
    else assemble_body<float, float>(cd, i, cd.tile.nl, lane_, GSrc<CREC_ADV_FLOATS>{&cd.crec_adv}, jacobi);
}

// pressure <- Jacobi sweep output, or p0 + solved increment
__global__ void __launch_bounds__(BLOCK) k_ppe_scatter(Seg sg, int jacobi) {
    SEG_ROW(sg);
    const PPEView pp = cd.ppe;
    if (row >= pp.nr) return;
    const PView S = cd.S;
    const int i = pp.rows[row];
    S.pres[i] = jacobi ? pp.pnew[row] : S.pres0[i] + pp.x[row];
}

// ------------------------------------------------------------------------
// Krylov kernels.  X = vector type, V = matrix value type.  Every kernel of
// a case returns at once when the case's solve is done or skipped.
#define GATE(cd) if ((cd).sc[SC_DONE] != 0.0) return
#define SKIPPED(cd) ((cd).ctl[CTL_SKIP] != 0.0)

// Tail sums of the long rows, one 64-lane wave per row: lane l takes tail
// entries l, l + 64, ... (contiguous, so every load is coalesced) and a
// fixed shuffle tree adds the lanes.  WHICH picks the vector the next
// product multiplies: 0 the Krylov iterate, 1 p-hat, 2 s-hat, 3 the FP64
// solution of the outer residual.
template <class V, class X, int WHICH>
__global__ void __launch_bounds__(BLOCK) k_kr_long(Seg sg) {
    int first_;
    const int cid_ = seg_case(sg, first_);
    const CaseDesc& cd = sg.cd[cid_];
    const int t = first_ + (int)threadIdx.x;
    const PPEView pp = cd.ppe;
    if ((t >> 6) >= pp.nlong) return;   // whole waves
    if (SKIPPED(cd)) return;
    if ((WHICH == 1 || WHICH == 2) && cd.sc[SC_DONE] != 0.0) return;
    const int lane = t & 63, u = t >> 6;
    const KSet<X> K = kset(pp, (X*)nullptr);
    const X* x = WHICH == 0 ? K.x : WHICH == 1 ? K.phat : K.shat;
    const long long e1 = pp.tail_off[u + 1];
    double s = 0.0;
    for (long long e = pp.tail_off[u] + lane; e < e1; e += 64) {
        const int c = pp.tail_cols[e];
        const double xv = WHICH == 3 ? pp.x[c] : (double)x[c];
        s += (double)pp.tail_vals[e] * xv;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) pp.lsum[u] = s;
}

// r = rhs - A x ; rhat = r ; partials rhs.rhs (A) and r.r (B).  ZERO: x is
// zero (the inner solves of the mixed mode start from it), so r = rhs and
// the product is not formed
template <class V, class X, bool ZERO>
__global__ void __launch_bounds__(BLOCK) k_kr_init(Seg sg) {
    __shared__ double sh[2 * BLOCK];
    __shared__ int wbase[BLOCK];
    SEG_ROW(sg);
    if (SKIPPED(cd)) return;
    const PPEView pp = cd.ppe;
    if (!ZERO) load_group_bases(pp, first_, wbase);
    const KSet<X> K = kset(pp, (X*)nullptr);
    double a = 0.0, b = 0.0;
    if (row < pp.nr) {
        double y = ZERO ? 0.0 : spmv_row(pp, row, vals_of(pp, (V*)nullptr), kdiag(pp, (X*)nullptr), K.x, wbase);
        X rr = (X)((double)K.rhs[row] - y);
        K.r[row] = rr;
        K.rhat[row] = rr;
        a = (double)K.rhs[row] * (double)K.rhs[row];
        b = (double)rr * (double)rr;
    }
    const Pair t2 = block_canon2(a, b, sh);
    if (threadIdx.x == 0) pp.partA[chunk] = t2.a;
    if (threadIdx.x == 64) pp.partB[chunk] = t2.b;
}

// x = 0 when the right-hand side vanished (FIN_INIT)
template <class X>
__global__ void __launch_bounds__(BLOCK) k_kr_zerox(Seg sg) {
    SEG_ROW(sg);
    if (SKIPPED(cd) || cd.sc[SC_ZEROX] == 0.0) return;
    const PPEView pp = cd.ppe;
    if (row < pp.nr) kset(pp, (X*)nullptr).x[row] = (X)0;
}

// p = r (it==1) or r + beta (p - omega v); phat = p / diag
template <class X>
__global__ void __launch_bounds__(BLOCK) k_kr_p(Seg sg) {
    SEG_ROW(sg);
    GATE(cd);
    const PPEView pp = cd.ppe;
    if (row >= pp.nr) return;
    const KSet<X> K = kset(pp, (X*)nullptr);
    const double* sc = cd.sc;
    X pv;
    if (sc[SC_IT] == 1.0) pv = K.r[row];
    else pv = (X)((double)K.r[row] + sc[SC_BETA] * ((double)K.p[row] - sc[SC_OMEGA] * (double)K.v[row]));
    K.p[row] = pv;
    if (sizeof(X) == 8 && cd.ctl[CTL_MG] != 0.0) pp.fp[row] = (float)pv;   // the V-cycle's input
    else K.phat[row] = (X)((double)pv / (double)kdiag(pp, (X*)nullptr)[row]);
    // the V-cycle's finest-level pre-smoothing of its input (float p)
    if (pp.mgx0) pp.mgx0[row] = (float)(MG_OMEGA * (double)(float)pv / (double)pp.fdiag[row]);
}

// FP64 solve with the multigrid: p-hat (which = 0) or s-hat (1) from the
// float V-cycle's output
template <class X>
__global__ void __launch_bounds__(BLOCK) k_kr_widen(Seg sg, int which) {
    SEG_ROW(sg);
    GATE(cd);
    if (cd.ctl[CTL_MG] == 0.0) return;
    const PPEView pp = cd.ppe;
    if (row >= pp.nr) return;
    const KSet<X> K = kset(pp, (X*)nullptr);
    if (which) K.shat[row] = (X)pp.fshat[row];
    else K.phat[row] = (X)pp.fphat[row];
}

// TWO = false: v = A phat, partial rhat.v (A)
// TWO = true : t = A shat, partials t.t (A) and t.s (B)
template <class V, class X, bool TWO>
__global__ void __launch_bounds__(BLOCK) k_kr_spmv(Seg sg) {
    __shared__ double sh[2 * BLOCK];
    __shared__ int wbase[BLOCK];
    SEG_ROW(sg);
    GATE(cd);
    const PPEView pp = cd.ppe;
    load_group_bases(pp, first_, wbase);
    const KSet<X> K = kset(pp, (X*)nullptr);
    const X* x = TWO ? K.shat : K.phat;
    X* y = TWO ? K.t : K.v;
    const X* d = TWO ? K.s : K.rhat;
    double a = 0.0, b = 0.0;
    if (row < pp.nr) {
        X yy = (X)spmv_row(pp, row, vals_of(pp, (V*)nullptr), kdiag(pp, (X*)nullptr), x, wbase);
        y[row] = yy;
        if (TWO) {
            a = (double)yy * (double)yy;        // t.t
            b = (double)yy * (double)d[row];    // t.s
        } else {
            a = (double)d[row] * (double)yy;    // rhat.v
        }
    }
    if (TWO) {
        const Pair t2 = block_canon2(a, b, sh);
        if (threadIdx.x == 0) pp.partA[chunk] = t2.a;
        if (threadIdx.x == 64) pp.partB[chunk] = t2.b;
    } else {
        double ta = block_canon(a, sh);
        if (threadIdx.x == 0) pp.partA[chunk] = ta;
    }
}

// s = r - alpha v ; shat = s / diag ; partial s.s
template <class X>
__global__ void __launch_bounds__(BLOCK) k_kr_s(Seg sg) {
    __shared__ double sh[BLOCK];
    SEG_ROW(sg);
    GATE(cd);
    const PPEView pp = cd.ppe;
    const KSet<X> K = kset(pp, (X*)nullptr);
    double a = 0.0;
    if (row < pp.nr) {
        X sv = (X)((double)K.r[row] - cd.sc[SC_ALPHA] * (double)K.v[row]);
        K.s[row] = sv;
        if (sizeof(X) == 8 && cd.ctl[CTL_MG] != 0.0) pp.fs[row] = (float)sv;
        else K.shat[row] = (X)((double)sv / (double)kdiag(pp, (X*)nullptr)[row]);
        // the V-cycle's finest-level pre-smoothing of its input (float s)
        if (pp.mgx0) pp.mgx0[row] = (float)(MG_OMEGA * (double)(float)sv / (double)pp.fdiag[row]);
        a = (double)sv * (double)sv;
    }
    double ta = block_canon(a, sh);
    if (threadIdx.x == 0) pp.partA[chunk] = ta;
}

// x = x + alpha phat + omega shat ; r = s - omega t ; partials r.r, rhat.r
// (after a half-step exit only x = x + alpha phat)
template <class X>
__global__ void __launch_bounds__(BLOCK) k_kr_xr(Seg sg) {
    __shared__ double sh[2 * BLOCK];
    SEG_ROW(sg);
    const double* sc = cd.sc;
    const bool half = sc[SC_HALF] == 1.0;
    if (sc[SC_DONE] != 0.0 && !half) return;
    const PPEView pp = cd.ppe;
    const KSet<X> K = kset(pp, (X*)nullptr);
    if (half) {
        if (row < pp.nr) K.x[row] = (X)((double)K.x[row] + sc[SC_ALPHA] * (double)K.phat[row]);
        return;
    }
    double a = 0.0, b = 0.0;
    if (row < pp.nr) {
        K.x[row] = (X)((double)K.x[row] + sc[SC_ALPHA] * (double)K.phat[row] + sc[SC_OMEGA] * (double)K.shat[row]);
        X rv = (X)((double)K.s[row] - sc[SC_OMEGA] * (double)K.t[row]);
        K.r[row] = rv;
        a = (double)rv * (double)rv;
        b = (double)K.rhat[row] * (double)rv;
    }
    const Pair t2 = block_canon2(a, b, sh);
    if (threadIdx.x == 0) pp.partA[chunk] = t2.a;
    if (threadIdx.x == 64) pp.partB[chunk] = t2.b;
}

// One block per case: reduce the case's chunk partials canonically and
// advance its scalars.
enum FinOp { FIN_INIT, FIN_ALPHA, FIN_S, FIN_OMEGA, FIN_R, FIN_NORM };
constexpr int FIN_BLOCK = 1024;
__global__ void __launch_bounds__(FIN_BLOCK) k_kr_fin(const CaseDesc* cds, int op) {
    __shared__ double lvl[256];
    const CaseDesc& cd = cds[blockIdx.x];
    double* sc = cd.sc;
    const double* ctl = cd.ctl;
    if (ctl[CTL_SKIP] != 0.0) {
        if (op == FIN_INIT && threadIdx.x == 0) {
            sc[SC_DONE] = 1.0; sc[SC_HALF] = 0.0; sc[SC_CONV] = 0.0; sc[SC_ITERS] = 0.0; sc[SC_ZEROX] = 0.0;
        }
        return;
    }
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
    const long long nb = cd.ppe.nb;
    double A = canon_partials(cd.ppe.partA, nb, lvl);
    double B = 0.0;
    if (op == FIN_INIT || op == FIN_OMEGA || op == FIN_R || op == FIN_NORM) B = canon_partials(cd.ppe.partB, nb, lvl);
    if (threadIdx.x != 0) return;
    const double tol = ctl[CTL_TOL];
    switch (op) {
    case FIN_INIT: {
        double bnorm = sqrt(A), rnorm = sqrt(B);
        sc[SC_BNORM] = bnorm; sc[SC_RNORM] = rnorm;
        sc[SC_RHO] = B; sc[SC_RHO_OLD] = 1.0; sc[SC_ALPHA] = 1.0; sc[SC_OMEGA] = 1.0; sc[SC_BETA] = 0.0;
        sc[SC_IT] = 1.0; sc[SC_HALF] = 0.0; sc[SC_ZEROX] = 0.0;
        sc[SC_DONE] = 0.0; sc[SC_CONV] = 0.0; sc[SC_ITERS] = 0.0;
        if (bnorm == 0.0) { sc[SC_DONE] = 1.0; sc[SC_CONV] = 1.0; sc[SC_ZEROX] = 1.0; sc[SC_RNORM] = 0.0; }
        else if (!(bnorm <= DBL_MAX) || !(rnorm <= DBL_MAX)) { sc[SC_DONE] = 1.0; }   // NaN/Inf in the system
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
        else if (sc[SC_OMEGA] == 0.0 || !(rnorm <= DBL_MAX)) { sc[SC_DONE] = 1.0; sc[SC_ITERS] = sc[SC_IT]; }
        else {
            double rho_new = B;
            sc[SC_RHO_OLD] = sc[SC_RHO];
            sc[SC_RHO] = rho_new;
            sc[SC_IT] = sc[SC_IT] + 1.0;
            if (sc[SC_IT] > ctl[CTL_MAXIT]) { sc[SC_DONE] = 1.0; sc[SC_ITERS] = ctl[CTL_MAXIT]; }
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
__global__ void __launch_bounds__(BLOCK) k_outer_res(Seg sg) {
    __shared__ double sh[2 * BLOCK];
    __shared__ int wbase[BLOCK];
    SEG_ROW(sg);
    if (SKIPPED(cd)) return;
    const PPEView pp = cd.ppe;
    load_group_bases(pp, first_, wbase);
    double a = 0.0, b = 0.0;
    if (row < pp.nr) {
        double y = spmv_row(pp, row, (const float*)pp.vals32, (const double*)pp.diag, (const double*)pp.x, wbase);
        double rr = pp.rhs[row] - y;
        pp.frhs[row] = (float)rr;
        a = rr * rr;
        b = pp.rhs[row] * pp.rhs[row];
    }
    const Pair t2 = block_canon2(a, b, sh);
    if (threadIdx.x == 0) pp.partA[chunk] = t2.a;
    if (threadIdx.x == 64) pp.partB[chunk] = t2.b;
}
__global__ void __launch_bounds__(BLOCK) k_add_correction(Seg sg) {
    SEG_ROW(sg);
    if (SKIPPED(cd)) return;
    const PPEView pp = cd.ppe;
    if (row < pp.nr) pp.x[row] = pp.x[row] + (double)pp.fx[row];
}
__global__ void __launch_bounds__(BLOCK) k_zero_fx(Seg sg) {
    SEG_ROW(sg);
    if (SKIPPED(cd)) return;
    const PPEView pp = cd.ppe;
    if (row < pp.nr) pp.fx[row] = 0.0f;
}
template <class T>
__global__ void __launch_bounds__(BLOCK) k_zero(long long n, T* a) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (T)0;
}

// ========================================================================
// host side
// ========================================================================
void ppe_free(sph_ctx* c) {
    PPEState* P = c->ppe;
    if (!P) return;
    void* ptrs[] = {P->rowof, P->rows, P->cnt, P->sl_off, P->cols, P->ecols, P->gbase, P->slotmap, P->spat, P->poff,
                    P->vals64, P->vals32,
                    P->diag, P->fdiag, P->rhs, P->x, P->pnew, P->r, P->rhat, P->p, P->phat, P->v, P->s, P->shat, P->t, P->fr, P->frhat, P->fp,
                    P->fphat, P->fv, P->fs, P->fshat, P->ft, P->fx, P->frhs, P->mgx0, P->partA, P->partB, P->ltail, P->lrows, P->tail_off, P->tail_cols, P->tail_vals, P->lsum};
    for (void* q : ptrs) dfree(c, q);
    mg_free(c, P->mg);
    delete P;
    c->ppe = nullptr;
    c->generation++;
}

// Renumbers the rows by lattice position (see k_lattice_key).  The lattice
// spacing is taken as h / 1.5 of the first row's particle; keys are
// rounded, so jitter below half a spacing leaves the order intact.
static int lattice_spacing(sph_ctx* c, const PPEState* P, double* q) {
    int r0 = 0;
    double h0 = 0.0;
    SPH_HIP(c, hipMemcpyAsync(&r0, P->rows, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    SPH_HIP(c, hipMemcpyAsync(&h0, c->sview().h + r0, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    *q = h0 / 1.5;
    return 0;
}
// The multigrid's lattice spacing: h / 1.5 of the finest row, so that
// refined particles aggregate among themselves first and coarse ones join
// once the blocks reach their spacing.
__global__ void __launch_bounds__(BLOCK) k_min_h(int nr, const int* rows, const double* h, double* out) {
    __shared__ double sh[BLOCK];
    const int r = blockIdx.x * BLOCK + threadIdx.x;
    sh[threadIdx.x] = r < nr ? h[rows[r]] : DBL_MAX;
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}
// the float copy of an FP64 operator's stored values (multigrid set-up)
__global__ void __launch_bounds__(BLOCK) k_vals_narrow(long long n, const double* a, float* b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

static int finest_spacing(sph_ctx* c, const PPEState* P, double* q) {
    const int nb = (int)grid1(P->nr);
    double* part = nullptr;
    SPH_TRY(dalloc(c, &part, nb));
    hipLaunchKernelGGL(k_min_h, dim3(nb), dim3(BLOCK), 0, c->stream, P->nr, (const int*)P->rows,
                       (const double*)c->sview().h, part);
    SPH_TRY(launch_check(c, "min_h"));
    std::vector<double> h(nb);
    SPH_HIP(c, hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, part);
    double m = DBL_MAX;
    for (double v : h) m = std::min(m, v);
    *q = m / 1.5;
    return 0;
}

static int lattice_order(sph_ctx* c, PPEState* P, bool* used) {
    const int nr = P->nr;
    PView S = c->sview();
    double q = 0.0;
    SPH_TRY(lattice_spacing(c, P, &q));
    const sph_domain& d = c->dom;
    const double ex = ((double)d.x_max - d.x_min) / q + 3.0, ey = ((double)d.y_max - d.y_min) / q + 3.0,
                 ez = ((double)d.z_max - d.z_min) / q + 3.0;
    if (!(q > 0) || ex * ey * ez > 4.0e18) { *used = false; return 0; }
    const double x0 = (double)d.x_min - q, y0 = (double)d.y_min - q, z0 = (double)d.z_min - q;
    if (ex * ey * ez < 2.0e9) {
        // device path: 32-bit keys, stable radix sort of (key, particle)
        int *key = nullptr, *val = nullptr, *ktmp = nullptr, *vtmp = nullptr;
        SPH_TRY(dalloc(c, &key, nr));
        SPH_TRY(dalloc(c, &val, nr));
        SPH_TRY(dalloc(c, &ktmp, nr));
        SPH_TRY(dalloc(c, &vtmp, nr));
        const long long nx = (long long)ex, ny = (long long)ey, nz = (long long)ez;
        hipLaunchKernelGGL(k_lattice_key32, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows, S,
                           x0, y0, z0, q, nx, ny, key, val);
        SPH_TRY(launch_check(c, "lattice_key"));
        SPH_TRY(radix_sort_pairs(c, &key, &val, &ktmp, &vtmp, nr, (int)(nx * ny * nz - 1)));
        SPH_HIP(c, hipMemcpyAsync(P->rows, val, sizeof(int) * nr, hipMemcpyDeviceToDevice, c->stream));
        hipLaunchKernelGGL(k_rowof_from_rows, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows,
                           P->rowof);
        SPH_TRY(launch_check(c, "rowof"));
        dfree(c, key);
        dfree(c, val);
        dfree(c, ktmp);
        dfree(c, vtmp);
        return 0;
    }
    long long* key = nullptr;
    SPH_TRY(dalloc(c, &key, nr));
    hipLaunchKernelGGL(k_lattice_key, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows, S, x0, y0,
                       z0, q, (long long)ex, (long long)ey, key);
    SPH_TRY(launch_check(c, "lattice_key"));
    std::vector<long long> hk(nr);
    std::vector<int> hr(nr), order(nr);
    SPH_HIP(c, hipMemcpyAsync(hk.data(), key, sizeof(long long) * nr, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipMemcpyAsync(hr.data(), P->rows, sizeof(int) * nr, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, key);
    for (int r = 0; r < nr; r++) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hk[a] < hk[b]; });
    std::vector<int> nrows(nr);
    for (int r = 0; r < nr; r++) nrows[r] = hr[order[r]];
    SPH_HIP(c, hipMemcpyAsync(P->rows, nrows.data(), sizeof(int) * nr, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_rowof_from_rows, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows, P->rowof);
    SPH_TRY(launch_check(c, "rowof"));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Entry count of every row and the SELL slice offsets for the current row
// numbering.
static int ppe_count_and_slices(sph_ctx* c, PPEState* P, bool tails) {
    const int nr = P->nr;
    SPH_TRY(prof_begin(c, "ppe_count", 0));
    hipLaunchKernelGGL(k_ppe_struct<false>, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows,
                       (const int*)P->rowof, c->sview(), c->nlist(), P->cnt, (const long long*)nullptr, (int*)nullptr,
                       (const int*)nullptr, (const long long*)nullptr, (int*)nullptr);
    SPH_TRY(launch_check(c, "ppe_count"));
    SPH_TRY(prof_end(c));
    if (tails) {
        // rows past LONG_ROW entries keep the rest of their entries in a tail
        int *flag = nullptr, *scan = nullptr, nlong = 0;
        SPH_TRY(dalloc(c, &flag, nr));
        SPH_TRY(dalloc(c, &scan, nr));
        hipLaunchKernelGGL(k_long_flag, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->cnt, flag);
        SPH_TRY(launch_check(c, "long_flag"));
        SPH_TRY(exclusive_scan_i32(c, flag, scan, nr, &nlong));
        if (nlong > 0) {
            long long* tlen = nullptr;
            P->nlong = nlong;
            SPH_TRY(dalloc(c, &P->ltail, nr));
            SPH_TRY(dalloc(c, &P->lrows, nlong));
            SPH_TRY(dalloc(c, &P->tail_off, (size_t)nlong + 1));
            SPH_TRY(dalloc(c, &P->lsum, nlong));
            SPH_TRY(dalloc(c, &tlen, nlong));
            hipLaunchKernelGGL(k_long_index, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, P->cnt, (const int*)flag,
                               (const int*)scan, P->lrows, P->ltail, tlen);
            SPH_TRY(launch_check(c, "long_index"));
            SPH_TRY(exclusive_scan_i64(c, tlen, P->tail_off, nlong, &P->tail_total));
            SPH_HIP(c, hipMemcpyAsync(P->tail_off + nlong, &P->tail_total, sizeof(long long), hipMemcpyHostToDevice,
                                      c->stream));
            SPH_HIP(c, hipStreamSynchronize(c->stream));
            dfree(c, tlen);
            SPH_TRY(dalloc(c, &P->tail_cols, (size_t)P->tail_total));
            SPH_TRY(dalloc(c, &P->tail_vals, (size_t)P->tail_total));
        }
        dfree(c, flag);
        dfree(c, scan);
    }
    long long* w = nullptr;
    SPH_TRY(dalloc(c, &w, P->nslices + 1));
    hipLaunchKernelGGL(k_row_slice_width, dim3(grid1(P->nslices)), dim3(BLOCK), 0, c->stream, nr, P->nslices,
                       (const int*)P->cnt, w);
    SPH_TRY(launch_check(c, "ppe_width"));
    SPH_TRY(exclusive_scan_i64(c, w, P->sl_off, P->nslices, &P->nnz_slots));
    dfree(c, w);
    return 0;
}

// Real entries (for reports, byte counts and the padding test): the sum of
// the row counts.
static int ppe_nnz(sph_ctx* c, PPEState* P) {
    const int nr = P->nr;
    long long *wide = nullptr, *tmp = nullptr, total = 0;
    SPH_TRY(dalloc(c, &wide, nr));
    SPH_TRY(dalloc(c, &tmp, nr));
    hipLaunchKernelGGL(k_widen, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->cnt, wide);
    SPH_TRY(launch_check(c, "widen"));
    SPH_TRY(exclusive_scan_i64(c, wide, tmp, nr, &total));
    dfree(c, wide);
    dfree(c, tmp);
    P->nnz = total + P->tail_total;
    return 0;
}

// Lattice-ordered structure (sorted columns and slot map in place) ->
// stencil slices where they pay (k_stencil_union); the slices' layout, the
// row counts, columns and slot map are replaced.  Runs before the 16-bit
// columns and the values are allocated.
static int stencil_slices(sph_ctx* c, PPEState* P) {
    const int nr = P->nr, ns = P->nslices;
    int *wp = nullptr, *ptmp = nullptr;
    SPH_TRY(dalloc(c, &wp, ns));
    SPH_TRY(dalloc(c, &ptmp, (size_t)ns * ST_CAP));
    hipLaunchKernelGGL(k_stencil_union, dim3(ns), dim3(SLICE), 0, c->stream, nr, (const long long*)P->sl_off,
                       (const int*)P->cnt, (const int*)P->cols, wp, ptmp);
    SPH_TRY(launch_check(c, "stencil_union"));
    long long* real = nullptr;
    long long* tmp = nullptr;
    SPH_TRY(dalloc(c, &real, nr));
    SPH_TRY(dalloc(c, &tmp, nr));
    hipLaunchKernelGGL(k_stencil_real, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->cnt,
                       (const int*)wp, real);
    SPH_TRY(launch_check(c, "stencil_real"));
    long long nnz_st = 0;
    SPH_TRY(exclusive_scan_i64(c, real, tmp, nr, &nnz_st));
    dfree(c, real);
    dfree(c, tmp);
    if (nnz_st == 0) {   // no slice pays: the layout stays
        dfree(c, wp);
        dfree(c, ptmp);
        return 0;
    }
    int* cnt2 = nullptr;
    long long *w = nullptr, *sl_off2 = nullptr, slots2 = 0;
    SPH_TRY(dalloc(c, &cnt2, nr));
    hipLaunchKernelGGL(k_stencil_cnt, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->cnt,
                       (const int*)wp, cnt2);
    SPH_TRY(launch_check(c, "stencil_cnt"));
    SPH_TRY(dalloc(c, &w, ns + 1));
    SPH_TRY(dalloc(c, &sl_off2, ns + 1));
    hipLaunchKernelGGL(k_row_slice_width, dim3(grid1(ns)), dim3(BLOCK), 0, c->stream, nr, ns, (const int*)cnt2, w);
    SPH_TRY(launch_check(c, "stencil_width"));
    SPH_TRY(exclusive_scan_i64(c, w, sl_off2, ns, &slots2));
    dfree(c, w);
    int* cols2 = nullptr;
    unsigned char* slotmap2 = nullptr;
    SPH_TRY(dalloc(c, &cols2, slots2 + 1));
    SPH_TRY(dalloc(c, &slotmap2, slots2 + 1));
    SPH_TRY(dalloc(c, &P->spat, ns));
    SPH_TRY(dalloc(c, &P->poff, slots2 / SLICE + 1));
    hipLaunchKernelGGL(k_zero<int>, dim3(grid1(slots2 + 1)), dim3(BLOCK), 0, c->stream, slots2 + 1, cols2);
    hipLaunchKernelGGL(k_stencil_relayout, dim3(ns), dim3(SLICE), 0, c->stream, nr, (const long long*)P->sl_off,
                       (const int*)P->cnt, (const int*)P->cols, (const unsigned char*)P->slotmap, (const int*)wp,
                       (const int*)ptmp, (const long long*)sl_off2, cols2, slotmap2, P->spat, P->poff);
    SPH_TRY(launch_check(c, "stencil_relayout"));
    std::vector<int> hwp(ns);
    SPH_HIP(c, hipMemcpyAsync(hwp.data(), wp, sizeof(int) * ns, hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    P->nstencil = 0;
    for (int v : hwp) P->nstencil += v > 0 ? 1 : 0;
    P->nnz_stencil = nnz_st;
    dfree(c, wp);
    dfree(c, ptmp);
    dfree(c, P->cnt);
    dfree(c, P->sl_off);
    dfree(c, P->cols);
    dfree(c, P->slotmap);
    P->cnt = cnt2;
    P->sl_off = sl_off2;
    P->cols = cols2;
    P->slotmap = slotmap2;
    P->nnz_slots = slots2;
    return 0;
}

// After the first assembly of a search: the stencil slices drop the offsets
// whose values are zero in all their rows (k_stencil_squeeze).  With fixed
// particles the values are final until the next search, which builds a new
// structure, so the slot map (needed only to write them) goes.
static int stencil_squeeze(sph_ctx* c, PPEState* P) {
    const int nr = P->nr, ns = P->nslices;
    long long *w2 = nullptr, *nzs = nullptr, *tmp = nullptr, *sl_off2 = nullptr, slots2 = 0, nz = 0;
    SPH_TRY(dalloc(c, &w2, ns + 1));
    SPH_TRY(dalloc(c, &nzs, ns));
    SPH_TRY(dalloc(c, &tmp, ns));
    SPH_TRY(dalloc(c, &sl_off2, ns + 1));
    hipLaunchKernelGGL(k_stencil_squeeze<false>, dim3(ns), dim3(SLICE), 0, c->stream, nr, (const long long*)P->sl_off,
                       (const int*)P->cnt, P->spat, (const int*)P->poff, (const float*)P->vals32,
                       (const unsigned short*)P->ecols, (const int*)P->cols, w2, nzs, (const long long*)nullptr,
                       (float*)nullptr, (unsigned short*)nullptr, (int*)nullptr, (int*)nullptr, (int*)nullptr);
    SPH_TRY(launch_check(c, "stencil_widths"));
    SPH_TRY(exclusive_scan_i64(c, nzs, tmp, ns, &nz));
    SPH_TRY(exclusive_scan_i64(c, w2, sl_off2, ns, &slots2));
    slots2 *= SLICE;
    hipLaunchKernelGGL(k_scale_i64, dim3(grid1(ns)), dim3(BLOCK), 0, c->stream, (long long)ns, sl_off2, (long long)SLICE);
    SPH_TRY(launch_check(c, "stencil_offsets"));
    dfree(c, w2);
    dfree(c, nzs);
    dfree(c, tmp);
    float* vals2 = nullptr;
    unsigned short* ecols2 = nullptr;
    int *cols2 = nullptr, *poff2 = nullptr, *cnt2 = nullptr;
    SPH_TRY(dalloc(c, &vals2, slots2 + 1));
    hipLaunchKernelGGL(k_zero<float>, dim3(grid1(slots2 + 1)), dim3(BLOCK), 0, c->stream, slots2 + 1, vals2);
    if (P->ecols) SPH_TRY(dalloc(c, &ecols2, slots2 + 1));
    if (P->cols) SPH_TRY(dalloc(c, &cols2, slots2 + 1));
    SPH_TRY(dalloc(c, &poff2, slots2 / SLICE + 1));
    SPH_TRY(dalloc(c, &cnt2, nr));
    hipLaunchKernelGGL(k_stencil_squeeze<true>, dim3(ns), dim3(SLICE), 0, c->stream, nr, (const long long*)P->sl_off,
                       (const int*)P->cnt, P->spat, (const int*)P->poff, (const float*)P->vals32,
                       (const unsigned short*)P->ecols, (const int*)P->cols, (long long*)nullptr, (long long*)nullptr,
                       (const long long*)sl_off2, vals2, ecols2, cols2, poff2, cnt2);
    SPH_TRY(launch_check(c, "stencil_squeeze"));
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, P->vals32);
    dfree(c, P->ecols);
    dfree(c, P->cols);
    dfree(c, P->poff);
    dfree(c, P->cnt);
    dfree(c, P->sl_off);
    dfree(c, P->slotmap);
    P->vals32 = vals2;
    P->ecols = ecols2;
    P->cols = cols2;
    P->poff = poff2;
    P->cnt = cnt2;
    P->sl_off = sl_off2;
    P->slotmap = nullptr;
    P->nnz_slots = slots2;
    P->nnz = P->nnz - P->nnz_stencil + nz;
    P->nnz_stencil = nz;
    P->squeezed = true;
    // the old layout's blocks would stay in the context's cache, counted in
    // its device memory (the admission estimate bounds that); nothing in
    // flight uses them after the synchronisation above
    dcache_release(c);
    return 0;
}

// Rows and column structure of a case's PPE from its neighbour list.
static int ppe_build_ordered(sph_ctx* c, bool allow_lattice);
int ppe_build(sph_ctx* c) { return ppe_build_ordered(c, true); }

static int ppe_build_ordered(sph_ctx* c, bool allow_lattice) {
    ppe_free(c);
    PPEState* P = new PPEState();
    c->ppe = P;
    c->generation++;
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
    SPH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c, flag);
    dfree(c, scan);
    if (nr == 0) return 0;
    if ((long long)nr > 65536LL * BLOCK) return fail(c, SPH_E_UNSUPPORTED, "PPE larger than 16.7M unknowns");
    const bool mixed = c->prm.precision == SPH_MIXED;
    bool lattice = allow_lattice && mixed && c->prm.ppe_reorder && nr > SLICE;
    if (lattice) SPH_TRY(lattice_order(c, P, &lattice));
    P->nslices = (nr + SLICE - 1) / SLICE;
    SPH_TRY(dalloc(c, &P->cnt, nr));
    SPH_TRY(dalloc(c, &P->sl_off, P->nslices + 1));
    // MIXED keeps the SELL slices at most LONG_ROW wide (FP64 keeps each row
    // whole: its products add the entries in the reference's order)
    SPH_TRY(ppe_count_and_slices(c, P, mixed));
    SPH_TRY(ppe_nnz(c, P));   // real entries (stencil slices add stored zeros)
    SPH_TRY(dalloc(c, &P->cols, P->nnz_slots + 1));
    hipLaunchKernelGGL(k_zero<int>, dim3(grid1(P->nnz_slots + 1)), dim3(BLOCK), 0, c->stream, P->nnz_slots + 1, P->cols);
    if (lattice) {
        int* ovf = nullptr;
        SPH_TRY(dalloc(c, &P->slotmap, P->nnz_slots + 1));
        SPH_TRY(dalloc(c, &ovf, 1));
        hipLaunchKernelGGL(k_zero<int>, dim3(1), dim3(BLOCK), 0, c->stream, 1LL, ovf);
        // the columns in list order, a thread per row, then sorted in place
        SPH_TRY(prof_begin(c, "ppe_cols", 0));
        hipLaunchKernelGGL(k_ppe_struct<true>, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows,
                           (const int*)P->rowof, S, c->nlist(), P->cnt, (const long long*)P->sl_off, P->cols,
                           (const int*)P->ltail, (const long long*)P->tail_off, P->tail_cols);
        SPH_TRY(launch_check(c, "ppe_cols"));
        SPH_TRY(prof_end(c));
        SPH_TRY(prof_begin(c, "ppe_cols_sorted", 0));
        // one workgroup per 64-row slice
        if (nr < SORT_NARROW_ROWS)
            hipLaunchKernelGGL(k_ppe_sort_slices<unsigned>, dim3(P->nslices), dim3(SORT_BLOCK), 0, c->stream, nr,
                               (const int*)P->cnt, (const long long*)P->sl_off, P->cols, P->slotmap, ovf);
        else
            hipLaunchKernelGGL(k_ppe_sort_slices<unsigned long long>, dim3(P->nslices), dim3(SORT_BLOCK), 0,
                               c->stream, nr, (const int*)P->cnt, (const long long*)P->sl_off, P->cols, P->slotmap,
                               ovf);
        SPH_TRY(launch_check(c, "ppe_cols_sorted"));
        SPH_TRY(prof_end(c));
        int hovf = 0;
        SPH_HIP(c, hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(c, hipStreamSynchronize(c->stream));
        dfree(c, ovf);
        if (hovf) {
            // a row with more entries than the in-block sort holds (refined
            // particles next to coarse ones): number this case's rows from
            // cell order instead
            return ppe_build_ordered(c, false);
        }
        P->ordering = SPH_ORDER_LATTICE;
        // fixed particles keep their lattice stencils from step to step
        if (c->prm.ppe_stencil && !c->prm.ale_positions) SPH_TRY(stencil_slices(c, P));
    } else {
        SPH_TRY(prof_begin(c, "ppe_cols", 0));
        hipLaunchKernelGGL(k_ppe_struct<true>, dim3(grid1(nr)), dim3(BLOCK), 0, c->stream, nr, (const int*)P->rows,
                           (const int*)P->rowof, S, c->nlist(), P->cnt, (const long long*)P->sl_off, P->cols,
                           (const int*)P->ltail, (const long long*)P->tail_off, P->tail_cols);
        SPH_TRY(launch_check(c, "ppe_cols"));
        SPH_TRY(prof_end(c));
    }
    P->nb = (nr + BLOCK - 1) / BLOCK;
    if (c->prm.ppe_index16 && (long long)nr <= 65536LL * BLOCK) {   // the window bitmap covers 16.7M columns
        // 16-bit columns where every slice fits 64 windows of 1024 columns
        int* ovf = nullptr;
        SPH_TRY(dalloc(c, &P->ecols, P->nnz_slots + 1));
        SPH_TRY(dalloc(c, &P->gbase, (size_t)P->nb * BLOCK));
        SPH_TRY(dalloc(c, &ovf, 1));
        hipLaunchKernelGGL(k_zero<int>, dim3(1), dim3(BLOCK), 0, c->stream, 1LL, ovf);
        hipLaunchKernelGGL(k_zero<int>, dim3(grid1((long long)P->nb * BLOCK)), dim3(BLOCK), 0, c->stream,
                           (long long)P->nb * BLOCK, P->gbase);
        SPH_TRY(prof_begin(c, "ecols_build", 0));
        hipLaunchKernelGGL(k_ecols_build, dim3(P->nslices), dim3(SLICE), 0, c->stream, nr, (const long long*)P->sl_off,
                           (const int*)P->cnt, (const int*)P->cols, P->ecols, P->gbase, ovf);
        SPH_TRY(launch_check(c, "ecols_build"));
        SPH_TRY(prof_end(c));
        int hovf = 0;
        SPH_HIP(c, hipMemcpyAsync(&hovf, ovf, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(c, hipStreamSynchronize(c->stream));
        dfree(c, ovf);
        if (hovf) {
            dfree(c, P->ecols); P->ecols = nullptr;
            dfree(c, P->gbase); P->gbase = nullptr;
        }
    }
    if (P->ecols) { dfree(c, P->cols); P->cols = nullptr; }   // the 16-bit copy replaces them
    // FP64 with the multigrid: a float copy of the operator builds the
    // hierarchy, and float copies of p and s feed the V-cycle
    const bool fp64_mg = !mixed && c->prm.ppe_precond == SPH_PRECOND_MULTIGRID &&
                         c->prm.ppe_solver == SPH_PPE_BICGSTAB && nr > MG_COARSE_MAX;
    if (mixed || fp64_mg) {
        // the assembly writes the real entries; the stored zeros of stencil
        // slices keep these
        SPH_TRY(dalloc(c, &P->vals32, P->nnz_slots + 1));
        hipLaunchKernelGGL(k_zero<float>, dim3(grid1(P->nnz_slots + 1)), dim3(BLOCK), 0, c->stream, P->nnz_slots + 1,
                           P->vals32);
        SPH_TRY(launch_check(c, "ppe_zero_vals"));
    }
    if (!mixed) SPH_TRY(dalloc(c, &P->vals64, P->nnz_slots + 1));
    SPH_TRY(dalloc(c, &P->diag, nr));
    if (mixed || fp64_mg) SPH_TRY(dalloc(c, &P->fdiag, nr));
    SPH_TRY(dalloc(c, &P->rhs, nr));
    SPH_TRY(dalloc(c, &P->x, nr));
    SPH_TRY(dalloc(c, &P->pnew, nr));
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
            if (fp64_mg) {
                float** fv[] = {&P->fp, &P->fphat, &P->fs, &P->fshat};
                for (auto q : fv) SPH_TRY(dalloc(c, q, nr));
            }
        }
        // the V-cycle's finest-level x, filled with its pre-smoothing by the
        // Krylov kernels that form p and s
        if (c->prm.ppe_precond == SPH_PRECOND_MULTIGRID && nr > MG_COARSE_MAX) SPH_TRY(dalloc(c, &P->mgx0, nr));
    }
    return 0;
}

int ppe_nlong(const sph_ctx* c) { return c->ppe ? c->ppe->nlong : 0; }

PPEView ppe_view(const sph_ctx* c) {
    PPEView v{};
    const PPEState* P = c->ppe;
    if (!P) return v;
    v.nr = P->nr; v.nb = P->nb;
    v.rows = P->rows; v.rowof = P->rowof; v.cnt = P->cnt; v.sl_off = P->sl_off; v.cols = P->cols;
    v.ecols = P->ecols; v.gbase = P->gbase; v.fdiag = P->fdiag; v.slotmap = P->slotmap;
    v.spat = P->spat; v.poff = P->poff;
    v.vals64 = P->vals64; v.vals32 = P->vals32; v.diag = P->diag; v.rhs = P->rhs; v.x = P->x; v.pnew = P->pnew;
    v.r = P->r; v.rhat = P->rhat; v.p = P->p; v.phat = P->phat; v.v = P->v; v.s = P->s; v.shat = P->shat; v.t = P->t;
    v.fr = P->fr; v.frhat = P->frhat; v.fp = P->fp; v.fphat = P->fphat; v.fv = P->fv; v.fs = P->fs;
    v.fshat = P->fshat; v.ft = P->ft; v.fx = P->fx; v.frhs = P->frhs;
    v.mgx0 = P->mgx0;
    v.partA = P->partA; v.partB = P->partB;
    v.ltail = P->ltail; v.lrows = P->lrows; v.tail_off = P->tail_off; v.tail_cols = P->tail_cols;
    v.tail_vals = P->tail_vals; v.nlong = P->nlong; v.lsum = P->lsum;
    return v;
}

static int nr_of(const sph_ctx* c) { return c->ppe ? c->ppe->nr : 0; }

static int read_scalars(Eng& e) {
    SPH_HIP(&e, hipMemcpyAsync(e.h_sc, e.d_sc, sizeof(double) * SC_COUNT * e.cs.size(), hipMemcpyDeviceToHost, e.st));
    SPH_HIP(&e, hipStreamSynchronize(e.st));
    return 0;
}
// solve controls of every case: the staging buffer is free once the last
// upload from it has run
static int ctl_wait(Eng& e) {
    SPH_HIP(&e, hipEventSynchronize(e.up_ev));
    return 0;
}
static int ctl_upload(Eng& e) {
    SPH_HIP(&e, hipMemcpyAsync(e.d_ctl, e.h_ctl, sizeof(double) * CTL_COUNT * e.cs.size(), hipMemcpyHostToDevice, e.st));
    SPH_HIP(&e, hipEventRecord(e.up_ev, e.st));
    return 0;
}

// launch over the PPE rows of every case, 256-row blocks
#define RL(name, bytes, kern, ...)                                                                   \
    do {                                                                                            \
        Shape sh_;                                                                                  \
        SPH_TRY(shape(e, 1, BLOCK, &sh_));                                                          \
        if (sh_.grid > 0) {                                                                         \
            SPH_TRY(prof_begin(e, name, (double)(bytes)));                                          \
            hipLaunchKernelGGL(kern, dim3(sh_.grid), dim3(BLOCK), 0, e.st, Seg{e.d_desc, sh_.tab}, ##__VA_ARGS__); \
            SPH_TRY(launch_check(&e, name));                                                        \
            SPH_TRY(prof_end(e));                                                                   \
        }                                                                                           \
    } while (0)
#define KF(op)                                                                                     \
    do {                                                                                           \
        SPH_TRY(prof_begin(e, "kr_fin", 0.0));                                                     \
        hipLaunchKernelGGL(k_kr_fin, dim3((unsigned)e.cs.size()), dim3(FIN_BLOCK), 0, e.st,        \
                           (const CaseDesc*)e.d_desc, (int)(op));                                  \
        SPH_TRY(launch_check(&e, "kr_fin"));                                                       \
        SPH_TRY(prof_end(e));                                                                      \
    } while (0)

// the long-row sums for the product that follows (a case without long
// rows has no waves in the launch)
#define RLONG(V_, X_, WHICH)                                                                      \
    do {                                                                                          \
        Shape sh_;                                                                                \
        SPH_TRY(shape(e, 2, BLOCK, &sh_));                                                        \
        if (sh_.grid > 0) {                                                                       \
            SPH_TRY(prof_begin(e, "kr_long", 0.0));                                               \
            hipLaunchKernelGGL((k_kr_long<V_, X_, WHICH>), dim3(sh_.grid), dim3(BLOCK), 0, e.st,  \
                               Seg{e.d_desc, sh_.tab});                                           \
            SPH_TRY(launch_check(&e, "kr_long"));                                                 \
            SPH_TRY(prof_end(e));                                                                 \
        }                                                                                         \
    } while (0)

// one BiCGSTAB iteration of every unfinished case
// true while some case of the step still iterates (half: or ends with the
// half-step update of x), from the host copy of the scalars
static bool any_live(const Eng& e, bool half) {
    for (size_t k = 0; k < e.cs.size(); k++) {
        if (e.h_ctl[k * CTL_COUNT + CTL_SKIP] != 0.0) continue;
        const double* sc = e.h_sc + k * SC_COUNT;
        if (sc[SC_DONE] == 0.0 || (half && sc[SC_HALF] == 1.0)) return true;
    }
    return false;
}

// One BiCGSTAB iteration of every unfinished case.  mg: the multigrid
// cases of the step (float vectors, tables from mg_prepare) replace the
// diagonal scaling of p-hat and s-hat by one V-cycle.  With the profile on
// (launches not captured) the host reads the scalars after every finalize
// and leaves out the launches whose device gate would return at once, so
// per-launch times count only launches that do work; the results are the
// same.
template <class V, class X>
static int kr_iteration(Eng& e, double spmv_bytes, bool mg) {
    const bool pf = e.prof.on;
    auto live = [&](bool half) { return !pf || any_live(e, half); };
#define KFP(op)                                   \
    do {                                          \
        KF(op);                                   \
        if (pf) SPH_TRY(read_scalars(e));         \
    } while (0)
    constexpr bool wide = sizeof(X) == 8;   // FP64 vectors: the V-cycle works on float copies
    if (live(false)) {
        RL("kr_p", 0, k_kr_p<X>);
        if (mg) SPH_TRY(mg_apply(e, 0, SC_DONE));
        if (mg && wide) RL("kr_widen", 0, k_kr_widen<X>, 0);
        RLONG(V, X, 1);
        RL("spmv", spmv_bytes, (k_kr_spmv<V, X, false>));
    }
    KFP(FIN_ALPHA);
    if (live(false)) RL("kr_s", 0, k_kr_s<X>);
    KFP(FIN_S);
    if (live(false)) {
        if (mg) SPH_TRY(mg_apply(e, 1, SC_DONE));
        if (mg && wide) RL("kr_widen", 0, k_kr_widen<X>, 1);
        RLONG(V, X, 2);
        RL("spmv", spmv_bytes, (k_kr_spmv<V, X, true>));
    }
    KFP(FIN_OMEGA);
    if (live(true)) RL("kr_xr", 0, k_kr_xr<X>);
    KFP(FIN_R);
#undef KFP
    return 0;
}

// The iteration as a HIP graph, captured once per composition of the step
// (its launches read everything else from the device descriptors).
template <class V, class X>
static int kr_graph(Eng& e, int gi, double spmv_bytes, bool mg, hipGraphExec_t* out) {
    if (!e.iter_graph[gi]) {
        Shape sh;
        SPH_TRY(shape(e, 1, BLOCK, &sh));   // segment table exists before the capture
        hipGraph_t g = nullptr;
        SPH_HIP(&e, hipStreamBeginCapture(e.st, hipStreamCaptureModeThreadLocal));
        int rc = kr_iteration<V, X>(e, spmv_bytes, mg);
        hipError_t er = hipStreamEndCapture(e.st, &g);
        if (rc) {
            if (er == hipSuccess && g) (void)hipGraphDestroy(g);
            return rc;
        }
        SPH_HIP(&e, er);
        er = hipGraphInstantiate(&e.iter_graph[gi], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (er != hipSuccess) e.iter_graph[gi] = nullptr;
        SPH_HIP(&e, er);
    }
    *out = e.iter_graph[gi];
    return 0;
}

// BiCGSTAB of every case whose control says not to skip it, on the vector
// set X (double: the FP64 solve; float: the inner solve of the mixed mode)
// with the tolerances and iteration caps already in the controls.
// iters[k] receives case k's iteration count.
template <class V, class X>
static int bicgstab(Eng& e, int gi, int pred_slot, std::vector<int>& iters, bool mg = false, bool x_zero = false) {
    const int nc = (int)e.cs.size();
    double spmv_bytes = 0.0, spmv_bytes0 = 0.0;
    int maxit = 0;
    for (int k = 0; k < nc; k++) {
        if (e.h_ctl[k * CTL_COUNT + CTL_SKIP] != 0.0) continue;
        // compulsory bytes of one SpMV launch: every stored entry (column
        // index, value) once, and per row its count, diagonal, own x, the
        // output and the dot partner (plus the slice group bases of 16-bit
        // columns); the x gathers of the neighbours are cache traffic
        // (stencil slices: values only, their columns implied)
        const PPEState* P = e.cs[k]->ppe;
        const bool i16 = P->ecols != nullptr;
        const double diag_b = sizeof(X) == 4 ? 4.0 : 8.0;
        const double cols_nnz = (double)(P->nnz - P->nnz_stencil);
        const double per_row = (double)P->nr * (4.0 + diag_b + 3.0 * sizeof(X) + (i16 ? 4.0 : 0.0));
        spmv_bytes += (double)P->nnz * sizeof(V) + cols_nnz * (i16 ? 2.0 : 4.0) + per_row;
        // the V-cycle's level-0 passes: 4 bytes per entry (column word and
        // half value), 2 in stencil slices
        spmv_bytes0 += cols_nnz * 4.0 + (double)P->nnz_stencil * 2.0 + per_row;
        maxit = std::max(maxit, (int)e.h_ctl[k * CTL_COUNT + CTL_MAXIT]);
    }
    if (mg) e.mg_plan.spmv_bytes0 = spmv_bytes0;
    if (x_zero) {
        RL("kr_init", 0, (k_kr_init<V, X, true>));
    } else {
        RLONG(V, X, 0);
        RL("kr_init", spmv_bytes, (k_kr_init<V, X, false>));
    }
    KF(FIN_INIT);
    RL("kr_zerox", 0, k_kr_zerox<X>);
    SPH_TRY(read_scalars(e));
    auto all_done = [&]() {
        for (int k = 0; k < nc; k++)
            if (e.h_ctl[k * CTL_COUNT + CTL_SKIP] == 0.0 && e.h_sc[k * SC_COUNT + SC_DONE] == 0.0) return false;
        return true;
    };
    const bool graphs = e.use_graphs && !e.prof.on;
    int& pred = e.pred_iters[gi][std::min(pred_slot, Eng::PRED_SLOTS - 1)];
    int launched = 0;
    while (!all_done() && launched < maxit) {
        // run straight to just before the expected end, then poll often;
        // without a prediction grow the batches geometrically
        int batch;
        if (pred > 0) batch = launched + 4 < (pred * 9) / 10 ? (pred * 9) / 10 - launched : 2;
        else batch = std::max(4, launched / 2);
        batch = std::min(batch, maxit - launched);
        if (graphs) {
            hipGraphExec_t g = nullptr;
            SPH_TRY((kr_graph<V, X>(e, gi, spmv_bytes, mg, &g)));
            for (int b = 0; b < batch; b++) SPH_HIP(&e, hipGraphLaunch(g, e.st));
        } else {
            for (int b = 0; b < batch; b++) SPH_TRY((kr_iteration<V, X>(e, spmv_bytes, mg)));
        }
        launched += batch;
        SPH_TRY(read_scalars(e));
    }
    int most = 0;
    iters.assign(nc, 0);
    for (int k = 0; k < nc; k++) {
        if (e.h_ctl[k * CTL_COUNT + CTL_SKIP] != 0.0) continue;
        const double* sc = e.h_sc + k * SC_COUNT;
        iters[k] = sc[SC_DONE] != 0.0 ? (int)sc[SC_ITERS] : (int)e.h_ctl[k * CTL_COUNT + CTL_MAXIT];
        most = std::max(most, iters[k]);
    }
    if (launched > 0) pred = std::max(4, most);
    return 0;
}

// ||r|| / ||b||: 0 for a zero right-hand side, NaN when either norm is not
// finite (a NaN or Inf in the system)
static double rel_residual(double rn, double bn) {
    if (!std::isfinite(rn) || !std::isfinite(bn)) return NAN;
    return bn > 0 ? rn / bn : 0.0;
}

// PPE of every case of the step: assembly, then one Jacobi sweep or the
// Krylov solve, then the pressures.
int ppe_solve(Eng& e) {
    const int nc = (int)e.cs.size();
    const sph_params& pr0 = e.cs[0]->prm;
    const bool mixed = pr0.precision == SPH_MIXED;
    const bool jacobi = pr0.ppe_solver == SPH_PPE_REF_JACOBI;
    double asm_bytes = 0.0;
    long long rows_total = 0;
    for (sph_ctx* c : e.cs) {
        // a squeezed operator (stencil slices without their zero offsets)
        // keeps its values until the next search builds a new structure
        if (c->ppe && c->ppe->squeezed && (c->prm.ale_positions || c->vals_gen != c->generation))
            return fail(&e, SPH_E_STATE, "PPE values of a squeezed operator would be rewritten");
        c->rep = sph_ppe_report{};
        c->rep.unknowns = nr_of(c);
        c->rep.nnz = c->ppe ? c->ppe->nnz : 0;
        c->rep.ordering = c->ppe ? c->ppe->ordering : SPH_ORDER_CELL;
        c->rep.slots = c->ppe ? c->ppe->nnz_slots : 0;
        c->rep.index16 = c->ppe && c->ppe->ecols ? 1 : 0;
        c->rep.tail_rows = c->ppe ? c->ppe->nlong : 0;
        c->rep.stencil_slices = c->ppe ? c->ppe->nstencil : 0;
        c->rep.stencil_nnz = c->ppe ? c->ppe->nnz_stencil : 0;
        c->rep.heavy_lists = c->n_heavy;
        if (e.tile_lanes) {
            c->rep.tile_chunks = c->tl_nchunk;
            c->rep.tile_raw_chunks = c->tl_raw;
            c->rep.tile_halo_largest = c->tl_hbig;
            c->rep.tile_halo_total = c->tl_halo_total;
        }
        c->rep.precond = SPH_PRECOND_JACOBI;
        if (c->rep.unknowns == 0) c->rep.converged = 1;
        rows_total += c->rep.unknowns;
        asm_bytes += (double)c->rep.unknowns * (15.0 * 8) + (double)c->rep.nnz * (4.0 + (mixed ? 4.0 : 8.0));
    }
    if (rows_total == 0) return 0;
    SPH_TRY(pack_records(e, REC_VEL_PQ));   // velocities after the predictor and its boundary passes
    if (mixed && e.tile_lanes) {   // cell tiles: a workgroup per chunk
        Shape sh;
        SPH_TRY(shape(e, 4, e.tile_lanes, &sh));
        const size_t lds = tile_lds_bytes(e, CREC_ADV_FLOATS);
        SPH_TRY(tile_lds_allow(e, (const void*)k_ppe_assemble_tile, lds));
        SPH_TRY(prof_begin(e, "ppe_assemble", asm_bytes));
        hipLaunchKernelGGL(k_ppe_assemble_tile, dim3(sh.grid), dim3(e.tile_lanes), lds, e.st, Seg{e.d_desc, sh.tab},
                           jacobi ? 1 : 0);
        SPH_TRY(launch_check(&e, "ppe_assemble"));
        SPH_TRY(prof_end(e));
    } else {
        const int blk = tuned_block(e, TK_ASSEMBLE);
        Shape sh;
        SPH_TRY(shape(e, 0, blk, &sh));   // over particles
        SPH_TRY(prof_begin(e, "ppe_assemble", asm_bytes));
        SPH_TRY(tune_begin(e, TK_ASSEMBLE, blk));
        if (mixed)
            hipLaunchKernelGGL((k_ppe_assemble<float, float>), dim3(sh.grid), dim3(blk), 0, e.st, Seg{e.d_desc, sh.tab},
                               jacobi ? 1 : 0);
        else
            hipLaunchKernelGGL((k_ppe_assemble<double, double>), dim3(sh.grid), dim3(blk), 0, e.st,
                               Seg{e.d_desc, sh.tab}, jacobi ? 1 : 0);
        SPH_TRY(launch_check(&e, "ppe_assemble"));
        SPH_TRY(tune_end(e));
        SPH_TRY(prof_end(e));
    }
    // the values of a search's first step are written: its stencil slices
    // drop their zero offsets, and the descriptors take the new layout
    if (mixed && !jacobi) {
        bool any = false;
        for (int k = 0; k < nc; k++) {
            sph_ctx* c = e.cs[k];
            PPEState* P = c->ppe;
            if (!P || !P->spat || P->squeezed || (!c->prm.ale_positions && c->vals_gen == c->generation)) continue;
            if (!any) SPH_HIP(&e, hipStreamSynchronize(e.st));   // the assembly's values
            any = true;
            SPH_TRY(stencil_squeeze(c, P));
            c->rep.nnz = P->nnz;
            c->rep.slots = P->nnz_slots;
            c->rep.stencil_nnz = P->nnz_stencil;
        }
        if (any) {
            SPH_HIP(&e, hipEventSynchronize(e.up_ev));
            for (int k = 0; k < nc; k++) e.h_desc[k].ppe = ppe_view(e.cs[k]);
            SPH_HIP(&e, hipMemcpyAsync(e.d_desc, e.h_desc, sizeof(CaseDesc) * nc, hipMemcpyHostToDevice, e.st));
            SPH_HIP(&e, hipEventRecord(e.up_ev, e.st));
        }
    }
    if (jacobi) {
        RL("ppe_scatter", 0, k_ppe_scatter, 1);
        for (sph_ctx* c : e.cs) {
            if (c->rep.unknowns == 0) continue;
            c->rep.iterations = 1;
            c->rep.converged = 1;
            c->rep.rel_residual = -1.0;
        }
        return 0;
    }
    std::vector<int> iters;
    // the multigrid preconditioner of every case that asks for it; a
    // hierarchy is built once per set of values (with fixed particles, at
    // the first step of a search).  FP64 builds it from a float copy of its
    // operator.
    std::vector<const MGState*> mgs(nc, nullptr);
    bool any_mg = false;
    for (int k = 0; k < nc; k++) {
        sph_ctx* c = e.cs[k];
        PPEState* P = c->ppe;
        if (c->prm.ppe_precond != SPH_PRECOND_MULTIGRID || !P) continue;
        const bool fresh = !c->prm.ale_positions && c->vals_gen == c->generation;
        if (!fresh || P->mg_state == 0) {
            mg_free(c, P->mg);
            P->mg = nullptr;
            P->mg_state = -1;
            P->mg_why = SPH_MG_UNSUITED;
            if (P->nr > MG_COARSE_MAX && (mixed || P->fp)) {
                if (!mixed) {
                    hipLaunchKernelGGL(k_vals_narrow, dim3(grid1(P->nnz_slots + 1)), dim3(BLOCK), 0, e.st,
                                       P->nnz_slots + 1, (const double*)P->vals64, P->vals32);
                    SPH_TRY(launch_check(&e, "vals_narrow"));
                }
                double q = 0.0;
                int ok = 0;
                SPH_TRY(finest_spacing(c, P, &q));
                SPH_TRY(mg_build(c, ppe_view(c), P->nnz_slots, q, &P->mg, &ok, &P->mg_why));
                P->mg_state = ok ? 1 : -1;
            }
        }
        if (P->mg_state == 1) {
            mgs[k] = P->mg;
            any_mg = true;
            c->rep.mg_sorted_rows = mg_sorted_rows(P->mg);
            c->rep.mg_long_rows = mg_long_rows(P->mg);
        }
        c->rep.mg_status = P->mg_state == -1 ? P->mg_why : SPH_MG_OK;
    }
    auto mg_tables = [&]() -> int {
        bool changed = false;
        SPH_TRY(mg_prepare(e, mgs, SC_DONE, &changed));
        for (int gi : {2, 3})
            if (changed && e.iter_graph[gi]) {   // its launches hold the old tables
                (void)hipGraphExecDestroy(e.iter_graph[gi]);
                e.iter_graph[gi] = nullptr;
            }
        return 0;
    };
    if (any_mg) SPH_TRY(mg_tables());
    if (!mixed) {
        // one solve, and for a multigrid case whose solve stalled with
        // finite residuals a second one with the diagonal scaling from where
        // it stopped
        std::vector<char> todo(nc);
        for (int k = 0; k < nc; k++) todo[k] = nr_of(e.cs[k]) > 0;
        for (int pass = 0; pass < 2; pass++) {
            SPH_TRY(ctl_wait(e));
            bool any = false;
            for (int k = 0; k < nc; k++) {
                double* ct = e.h_ctl + k * CTL_COUNT;
                ct[CTL_TOL] = e.cs[k]->prm.ppe_tol;
                ct[CTL_MAXIT] = e.cs[k]->prm.ppe_maxit;
                ct[CTL_SKIP] = todo[k] ? 0.0 : 1.0;
                ct[CTL_MG] = mgs[k] ? 1.0 : 0.0;
                any |= todo[k] != 0;
            }
            SPH_TRY(ctl_upload(e));
            if (!any) break;
            SPH_TRY((bicgstab<double, double>(e, any_mg ? 3 : 0, pass, iters, any_mg)));
            bool stalled = false;
            for (int k = 0; k < nc; k++) {
                if (!todo[k]) continue;
                sph_ctx* c = e.cs[k];
                const double* sc = e.h_sc + k * SC_COUNT;
                c->rep.iterations += iters[k];
                c->rep.converged = sc[SC_CONV] != 0.0;
                c->rep.rel_residual = rel_residual(sc[SC_RNORM], sc[SC_BNORM]);
                todo[k] = 0;
                if (!mgs[k]) continue;
                c->rep.precond = SPH_PRECOND_MULTIGRID;
                c->rep.mg_levels = mg_levels(mgs[k]);
                c->rep.mg_coarsest = mg_coarsest(mgs[k]);
                if (sc[SC_CONV] == 0.0 && std::isfinite(sc[SC_RNORM])) {
                    c->ppe->mg_state = -1;
                    c->ppe->mg_why = SPH_MG_STALLED_SOLVE;
                    c->rep.mg_status = SPH_MG_STALLED_SOLVE;
                    mgs[k] = nullptr;
                    todo[k] = 1;
                    stalled = true;
                }
            }
            if (!stalled) break;
            any_mg = false;
            for (const MGState* m : mgs) any_mg |= m != nullptr;
            if (any_mg) SPH_TRY(mg_tables());
        }
    } else {
        // outer FP64 iterative refinement around float inner solves; each
        // case refines until its own tolerance is met
        std::vector<char> active(nc);
        std::vector<int> refine(nc, 0);
        for (int k = 0; k < nc; k++) active[k] = nr_of(e.cs[k]) > 0;
        double res_bytes = 0.0;
        for (sph_ctx* c : e.cs) {   // float values, the stored columns, per row x, b, diag, count, r
            const PPEState* P = c->ppe;
            if (!P) continue;
            res_bytes += (double)P->nnz * 4.0 + (double)(P->nnz - P->nnz_stencil) * (P->ecols ? 2.0 : 4.0) +
                         (double)P->nr * 28.0;
        }
        for (int outer = 0;; outer++) {
            bool any = false;
            for (int k = 0; k < nc; k++) any |= active[k] != 0;
            if (!any) break;
            SPH_TRY(ctl_wait(e));
            for (int k = 0; k < nc; k++) {
                e.h_ctl[k * CTL_COUNT + CTL_SKIP] = active[k] ? 0.0 : 1.0;
                e.h_ctl[k * CTL_COUNT + CTL_MG] = 0.0;   // float vectors: the V-cycle reads them directly
            }
            SPH_TRY(ctl_upload(e));
            RLONG(float, float, 3);
            RL("outer_res", res_bytes, k_outer_res);
            KF(FIN_NORM);
            SPH_TRY(read_scalars(e));
            any = false;
            SPH_TRY(ctl_wait(e));
            for (int k = 0; k < nc; k++) {
                if (!active[k]) continue;
                sph_ctx* c = e.cs[k];
                const sph_params& pr = c->prm;
                const double* sc = e.h_sc + k * SC_COUNT;
                const double rn = sc[SC_RNORM], bn = sc[SC_BNORM];
                c->rep.rel_residual = rel_residual(rn, bn);
                c->rep.refinements = refine[k];
                double* ct = e.h_ctl + k * CTL_COUNT;
                if (rn <= pr.ppe_tol * bn) { c->rep.converged = 1; active[k] = 0; }
                else if (refine[k] == pr.ppe_max_refine || !std::isfinite(rn) || !std::isfinite(bn)) active[k] = 0;
                else {
                    // aim the correction at the final absolute tolerance, but
                    // not below what a float Krylov solve reliably attains
                    ct[CTL_TOL] = std::max(pr.ppe_inner_tol, 0.5 * pr.ppe_tol * bn / rn);
                    ct[CTL_MAXIT] = pr.ppe_maxit;
                    any = true;
                }
                ct[CTL_SKIP] = active[k] ? 0.0 : 1.0;
            }
            SPH_TRY(ctl_upload(e));
            if (!any) break;
            RL("zero_fx", 0, k_zero_fx);
            SPH_TRY((bicgstab<float, float>(e, any_mg ? 2 : 1, outer, iters, any_mg, true)));
            bool stalled = false;
            for (int k = 0; k < nc; k++) {
                if (!active[k]) continue;
                sph_ctx* c = e.cs[k];
                c->rep.iterations += iters[k];
                refine[k]++;
                if (!mgs[k]) continue;
                c->rep.precond = SPH_PRECOND_MULTIGRID;
                c->rep.mg_levels = mg_levels(mgs[k]);
                c->rep.mg_coarsest = mg_coarsest(mgs[k]);
                // an inner solve that stalled (breakdown or the iteration
                // cap) with finite residuals: Jacobi from here on, for the
                // rest of this search
                const double* sc = e.h_sc + k * SC_COUNT;
                if (sc[SC_CONV] == 0.0 && std::isfinite(sc[SC_RNORM])) {
                    c->ppe->mg_state = -1;
                    c->ppe->mg_why = SPH_MG_STALLED_SOLVE;
                    c->rep.mg_status = SPH_MG_STALLED_SOLVE;
                    mgs[k] = nullptr;
                    stalled = true;
                }
            }
            if (stalled) {
                any_mg = false;
                for (const MGState* m : mgs) any_mg |= m != nullptr;
                if (any_mg) SPH_TRY(mg_tables());
            }
            RL("add_corr", 0, k_add_correction);
        }
    }
    RL("ppe_scatter", 0, k_ppe_scatter, 0);
    int rc = 0;
    for (sph_ctx* c : e.cs) {
        if (c->rep.converged) continue;
        if (!std::isfinite(c->rep.rel_residual))
            rc = fail(c, SPH_E_NONFINITE, "non-finite PPE residual: the state holds NaN or Inf");
        else if (rc != SPH_E_NONFINITE)
            rc = fail(c, SPH_E_NOCONV, "PPE did not converge within ppe_maxit");
    }
    return rc;
}

}  // namespace sph
