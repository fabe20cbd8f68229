// Cell tiles of the record-gathering passes.
//
// The MIXED passes over fixed float positions (prep, advection, pressure
// force, PPE assembly) read a 32- or 64-byte record of the neighbour for
// every list entry.  Run per particle, those reads are random gathers, and
// the records one XCD's workgroups reach at a time are several times its
// L2, so each record crosses the fabric about ten times a launch.
//
// Here the particles are grouped by blocks of cells (tile_cells cells a
// block, in the sorted order of the search) and each block is cut into
// chunks of at most tile_particles particles.  A workgroup takes a chunk:
// it stages the records of the chunk's halo (the distinct neighbours of
// its particles, ascending) in LDS with coalesced loads, then each thread
// walks a copy of its particle's neighbour list whose entries hold the halo
// slot instead of the particle index, in the same order and with the same
// selection flags.  The sums see the same operands in the same order, so
// the results equal the per-particle passes bitwise.  A chunk whose halo
// is larger than tile_halo_max keeps particle indices in its entries and
// gathers from memory.
//
// Built once per search (the lists do not change between searches).
#include <algorithm>
#include <climits>
#include <vector>

#include "flags.h"
#include "sph_internal.h"

namespace sph {

namespace {

constexpr int TB_HASH = 4096;   // slots of the LDS hash of one chunk's halo (halo_max <= TB_HASH / 2)

__global__ void __launch_bounds__(BLOCK) k_tb_set(int* a, long long n, int v) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}

__device__ __forceinline__ unsigned tb_hash(int j) { return ((unsigned)j * 2654435761u) >> (32 - 12); }

struct BlockGeo {
    int NI, NJ, NK;          // cell grid
    int cx, cy, cz;          // cells per block
    int nbx, nby, nbz;       // blocks per axis
};

// Particle range [cstart(c0), cstart(c1 + 1)) of the cells c0..c1 of one
// grid row, visited by rows of the block (y, then z), then the particles
// whose key lies past the grid (the last "block").
template <class F>
__device__ void block_particles(const BlockGeo& G, const int* cstart, int n, int b, F f) {
    const int nb = G.nbx * G.nby * G.nbz;
    if (b == nb) {
        const int p0 = cstart[G.NI * G.NJ * G.NK];
        for (int p = p0; p < n; p++) f(p);
        return;
    }
    const int bx = b % G.nbx, by = (b / G.nbx) % G.nby, bz = b / (G.nbx * G.nby);
    const int i0 = bx * G.cx, i1 = min(i0 + G.cx, G.NI) - 1;
    for (int ck = bz * G.cz; ck < min(bz * G.cz + G.cz, G.NK); ck++)
        for (int cj = by * G.cy; cj < min(by * G.cy + G.cy, G.NJ); cj++) {
            const int c0 = i0 + G.NI * (cj + G.NJ * ck), c1 = i1 + G.NI * (cj + G.NJ * ck);
            const int p1 = cstart[c1 + 1];
            for (int p = cstart[c0]; p < p1; p++) f(p);
        }
}

// per block: its particles and chunks
__global__ void __launch_bounds__(BLOCK) k_tb_count(BlockGeo G, const int* cstart, int n, int cap, int* cnt,
                                                    int* nch) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = G.nbx * G.nby * G.nbz + 1;
    if (b >= nb) return;
    int m = 0;
    if (b == nb - 1) {
        m = n - cstart[G.NI * G.NJ * G.NK];
    } else {
        const int bx = b % G.nbx, by = (b / G.nbx) % G.nby, bz = b / (G.nbx * G.nby);
        const int i0 = bx * G.cx, i1 = min(i0 + G.cx, G.NI) - 1;
        for (int ck = bz * G.cz; ck < min(bz * G.cz + G.cz, G.NK); ck++)
            for (int cj = by * G.cy; cj < min(by * G.cy + G.cy, G.NJ); cj++)
                m += cstart[i1 + G.NI * (cj + G.NJ * ck) + 1] - cstart[i0 + G.NI * (cj + G.NJ * ck)];
    }
    cnt[b] = m;
    nch[b] = (m + cap - 1) / cap;
}

// lanes of the chunks: a block's particles split evenly over its chunks
__global__ void __launch_bounds__(BLOCK) k_tb_fill(BlockGeo G, const int* cstart, int n, int cap, const int* cnt,
                                                   const int* nch, const int* chunk0, int* part) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = G.nbx * G.nby * G.nbz + 1;
    if (b >= nb || nch[b] == 0) return;
    const int per = (cnt[b] + nch[b] - 1) / nch[b];
    const long long base = (long long)chunk0[b] * cap;
    int q = 0;
    block_particles(G, cstart, n, b, [&](int p) {
        part[base + (long long)(q / per) * cap + q % per] = p;
        q++;
    });
}

__global__ void __launch_bounds__(BLOCK) k_tl_cnt(long long nlanes, const int* part, const int* nl_cnt, int* cnt) {
    const long long L = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= nlanes) return;
    const int i = part[L];
    cnt[L] = i >= 0 ? nl_cnt[i] : 0;
}

__global__ void __launch_bounds__(BLOCK) k_tl_width(long long nlanes, long long nslices, const int* cnt, long long* w) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int m = 0;
    for (int l = 0; l < SLICE; l++) {
        const long long L = s * SLICE + l;
        if (L < nlanes) m = max(m, cnt[L]);
    }
    w[s] = (long long)m * SLICE;
}

// Inserts j into the chunk's LDS hash; counts new keys.  Every probe
// sequence ends: a key takes a free slot or finds itself, and the table
// (TB_HASH slots) is never full below the halo cap (TB_HASH / 2).
__device__ __forceinline__ void tb_insert(int* hkey, int* ndist, int j) {
    unsigned h = tb_hash(j);
    for (int p = 0; p < TB_HASH; p++) {
        const int old = atomicCAS(&hkey[h], -1, j);
        if (old == -1) { atomicAdd(ndist, 1); return; }
        if (old == j) return;
        h = (h + 1) & (TB_HASH - 1);
    }
}
__device__ __forceinline__ int tb_find(const int* hkey, int j) {
    unsigned h = tb_hash(j);
    for (int p = 0; p < TB_HASH; p++) {
        const int k = hkey[h];
        if (k == j) return (int)h;
        if (k == -1) return -1;
        h = (h + 1) & (TB_HASH - 1);
    }
    return -1;
}

// One workgroup (cap threads, a thread per lane) per chunk.
// FILL == false: the size of each chunk's halo, or -1 above hmax.
// FILL == true: the sorted halo, each lane's own slot, and the tile list
// (slot | flags; particle indices for the chunks above the cap).
template <bool FILL>
__global__ void __launch_bounds__(256) k_tb_halo(int cap, int hmax, const int* part, NList nl, int* hcnt,
                                                 const int2* chunk, int* halo, int* self_slot, long long* tl_off,
                                                 unsigned* tl_ent) {
    __shared__ int hkey[TB_HASH];
    __shared__ int hval[TB_HASH];
    __shared__ int srt[TB_HASH / 2];
    __shared__ int ndist, ncomp, over;
    const int c = blockIdx.x, t = threadIdx.x;
    const long long L = (long long)c * cap + t;
    const int i = part[L];
    if (FILL && chunk[c].y < 0) {   // above the cap: the entries keep particle indices
        if (i >= 0) {
            const NLRow row = nl_row(nl, i);
            unsigned* out = tl_ent + tl_off[L >> 6] + (L & 63);
            for (int k = 0; k < row.cnt; k++) out[(size_t)k * SLICE] = row.p[(size_t)k * SLICE];
        }
        self_slot[L] = -1;
        return;
    }
    for (int s = t; s < TB_HASH; s += blockDim.x) hkey[s] = -1;
    if (t == 0) { ndist = 0; ncomp = 0; over = 0; }
    __syncthreads();
    if (i >= 0) {
        const NLRow row = nl_row(nl, i);
        for (int k = 0; k < row.cnt; k++) {
            if (*(volatile int*)&over) break;
            tb_insert(hkey, &ndist, (int)(row.p[(size_t)k * SLICE] & NB_IDX));
            if (*(volatile int*)&ndist > hmax) over = 1;
        }
    }
    __syncthreads();
    if (!FILL) {
        if (t == 0) hcnt[c] = (over || ndist > hmax) ? -1 : ndist;
        return;
    }
    // FILL: the chunk's halo fits (ndist == chunk[c].y <= hmax <= TB_HASH / 2)
    const int nd = ndist;
    int P = 1;
    while (P < nd) P <<= 1;
    for (int s = t; s < TB_HASH; s += blockDim.x)
        if (hkey[s] != -1) srt[atomicAdd(&ncomp, 1)] = hkey[s];
    __syncthreads();
    for (int s = nd + t; s < P; s += blockDim.x) srt[s] = INT_MAX;
    __syncthreads();
    // bitonic sort of srt[0, P), ascending
    for (int k = 2; k <= P; k <<= 1)
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int a = t; a < P; a += blockDim.x) {
                const int b = a ^ jj;
                if (b > a) {
                    const int x = srt[a], y = srt[b];
                    if ((x > y) == ((a & k) == 0)) { srt[a] = y; srt[b] = x; }
                }
            }
            __syncthreads();
        }
    const int h0 = chunk[c].x;
    for (int s = t; s < nd; s += blockDim.x) {
        halo[h0 + s] = srt[s];
        hval[tb_find(hkey, srt[s])] = s;
    }
    __syncthreads();
    if (i < 0) { self_slot[L] = -1; return; }
    const int ps = tb_find(hkey, i);
    self_slot[L] = ps >= 0 ? hval[ps] : -1;
    const NLRow row = nl_row(nl, i);
    unsigned* out = tl_ent + tl_off[L >> 6] + (L & 63);
    for (int k = 0; k < row.cnt; k++) {
        const unsigned e = row.p[(size_t)k * SLICE];
        out[(size_t)k * SLICE] = (e & ~NB_IDX) | (unsigned)hval[tb_find(hkey, (int)(e & NB_IDX))];
    }
}

}  // namespace

static void tile_free(sph_ctx* c) {
    dfree(c, c->tl_part); dfree(c, c->tl_chunk); dfree(c, c->tl_halo); dfree(c, c->tl_self);
    dfree(c, c->tl_slice_off); dfree(c, c->tl_cnt); dfree(c, c->tl_ent);
    c->tl_part = nullptr; c->tl_chunk = nullptr; c->tl_halo = nullptr; c->tl_self = nullptr;
    c->tl_slice_off = nullptr; c->tl_cnt = nullptr; c->tl_ent = nullptr;
    c->tl_nchunk = 0; c->tl_gen = -1;
}

int tile_build(sph_ctx* c) {
    const sph_params& pr = c->prm;
    if (c->tl_gen == c->generation && c->tl_code == pr.tile_cells && c->tl_cap == pr.tile_particles &&
        c->tl_hcap == pr.tile_halo_max)
        return 0;
    if (pr.tile_cells == 0) { tile_free(c); return 0; }
    tile_free(c);
    const int n = c->n, cap = pr.tile_particles ? pr.tile_particles : 128;
    const int hmax = pr.tile_halo_max ? pr.tile_halo_max : 1536;
    const sph_domain& d = c->dom;
    BlockGeo G{d.NI, d.NJ, d.NK, pr.tile_cells & 0xff, (pr.tile_cells >> 8) & 0xff, (pr.tile_cells >> 16) & 0xff, 0, 0, 0};
    G.nbx = (G.NI + G.cx - 1) / G.cx; G.nby = (G.NJ + G.cy - 1) / G.cy; G.nbz = (G.NK + G.cz - 1) / G.cz;
    const long long nbl = (long long)G.nbx * G.nby * G.nbz + 1;
    if (nbl > INT_MAX / 2) return fail(c, SPH_E_ARG, "tile_cells: too many blocks");
    const int nb = (int)nbl;
    int *cnt = nullptr, *nch = nullptr, *chunk0 = nullptr;
    SPH_TRY(dalloc(c, &cnt, nb));
    SPH_TRY(dalloc(c, &nch, nb));
    SPH_TRY(dalloc(c, &chunk0, nb));
    hipLaunchKernelGGL(k_tb_count, dim3(grid1(nb)), dim3(BLOCK), 0, c->stream, G, (const int*)c->cstart, n, cap, cnt, nch);
    SPH_TRY(launch_check(c, "tile_count"));
    int nchunk = 0;
    SPH_TRY(exclusive_scan_i32(c, nch, chunk0, nb, &nchunk));
    const long long nlanes = (long long)nchunk * cap;
    if (nlanes >= INT_MAX) {   // lanes are int-indexed: such a case gathers per particle
        dfree(c, cnt); dfree(c, nch); dfree(c, chunk0);
        return 0;
    }
    SPH_TRY(dalloc(c, &c->tl_part, (size_t)nlanes + 1));
    hipLaunchKernelGGL(k_tb_set, dim3(grid1(nlanes + 1)), dim3(BLOCK), 0, c->stream, c->tl_part, nlanes + 1, -1);
    hipLaunchKernelGGL(k_tb_fill, dim3(grid1(nb)), dim3(BLOCK), 0, c->stream, G, (const int*)c->cstart, n, cap,
                       (const int*)cnt, (const int*)nch, (const int*)chunk0, c->tl_part);
    SPH_TRY(launch_check(c, "tile_fill"));
    dfree(c, cnt); dfree(c, nch); dfree(c, chunk0);
    // the lane lists: SELL-64 over the lanes, widths from the particles' counts
    const long long nslices = nlanes / SLICE;   // cap is a multiple of 64
    SPH_TRY(dalloc(c, &c->tl_cnt, (size_t)nlanes + 1));
    SPH_TRY(dalloc(c, &c->tl_slice_off, (size_t)nslices + 1));
    long long* w = nullptr;
    SPH_TRY(dalloc(c, &w, (size_t)nslices + 1));
    hipLaunchKernelGGL(k_tl_cnt, dim3(grid1(nlanes)), dim3(BLOCK), 0, c->stream, nlanes, (const int*)c->tl_part,
                       (const int*)c->nl_cnt, c->tl_cnt);
    hipLaunchKernelGGL(k_tl_width, dim3(grid1(nslices)), dim3(BLOCK), 0, c->stream, nlanes, nslices,
                       (const int*)c->tl_cnt, w);
    SPH_TRY(launch_check(c, "tile_width"));
    long long total = 0;
    SPH_TRY(exclusive_scan_i64(c, w, c->tl_slice_off, nslices, &total));
    dfree(c, w);
    SPH_TRY(dalloc(c, &c->tl_ent, (size_t)total + 1));
    // halo sizes, then their offsets on the host
    int* hcnt = nullptr;
    SPH_TRY(dalloc(c, &hcnt, (size_t)nchunk + 1));
    if (nchunk > 0)
        hipLaunchKernelGGL(k_tb_halo<false>, dim3(nchunk), dim3(cap), 0, c->stream, cap, hmax,
                           (const int*)c->tl_part, c->nlist(), hcnt, (const int2*)nullptr, (int*)nullptr,
                           (int*)nullptr, (long long*)nullptr, (unsigned*)nullptr);
    SPH_TRY(launch_check(c, "tile_halo_count"));
    std::vector<int> h(nchunk);
    if (nchunk > 0) {
        SPH_HIP(c, hipMemcpyAsync(h.data(), hcnt, sizeof(int) * nchunk, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(c, hipStreamSynchronize(c->stream));
    }
    dfree(c, hcnt);
    std::vector<int2> tab(nchunk);
    long long hsum = 0;
    int hbig = 0, raw = 0;
    for (int k = 0; k < nchunk; k++) {
        tab[k] = make_int2((int)hsum, h[k]);
        if (h[k] < 0) { raw++; continue; }
        hsum += h[k];
        hbig = std::max(hbig, h[k]);
    }
    if (hsum > INT_MAX) return fail(c, SPH_E_ARG, "tile halos exceed 2^31 entries");
    SPH_TRY(dalloc(c, &c->tl_chunk, (size_t)nchunk + 1));
    SPH_TRY(dalloc(c, &c->tl_halo, (size_t)hsum + 1));
    SPH_TRY(dalloc(c, &c->tl_self, (size_t)nlanes + 1));
    if (nchunk > 0) {
        SPH_HIP(c, hipMemcpyAsync(c->tl_chunk, tab.data(), sizeof(int2) * nchunk, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_tb_halo<true>, dim3(nchunk), dim3(cap), 0, c->stream, cap, hmax,
                           (const int*)c->tl_part, c->nlist(), (int*)nullptr, (const int2*)c->tl_chunk, c->tl_halo,
                           c->tl_self, c->tl_slice_off, c->tl_ent);
        SPH_TRY(launch_check(c, "tile_halo_fill"));
        SPH_HIP(c, hipStreamSynchronize(c->stream));   // the host table above is released on return
    }
    c->tl_nchunk = nchunk;
    c->tl_hbig = hbig;
    c->tl_raw = raw;
    c->tl_halo_total = hsum;
    c->tl_code = pr.tile_cells;
    c->tl_cap = pr.tile_particles;
    c->tl_hcap = pr.tile_halo_max;
    c->tl_lanes = cap;
    c->tl_gen = c->generation;
    return 0;
}

void tile_release(sph_ctx* c) { tile_free(c); }

}  // namespace sph
