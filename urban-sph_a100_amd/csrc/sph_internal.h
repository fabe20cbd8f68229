// Internal declarations of the MI355X ISPH library: device views, the
// context, launch helpers.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/urbansph.h"

namespace sph {

constexpr int BLOCK = 256;        // threads per workgroup for 1-D passes
constexpr int SLICE = 64;         // SELL slice height = one wave64
constexpr int RED_CHUNK = 256;    // canonical reduction chunk

// Persistent particle state (the reference's P1 / SP1), structure of arrays.
// Immutable fields of the step-working copy alias the persistent arrays.
struct PView {
    const int *i_type, *buffer_type, *MOST_buffer, *p_type;
    int *ncell, *source;
    double *x, *y, *z, *ux, *uy, *uz, *m, *h, *temp, *pres, *pres0, *rho, *flt_s,
        *enthalpy, *concn, *concentration, *k_turb, *e_turb, *eli, *vol, *m_ref,
        *h_ref, *vis_t, *fbx, *fby, *fbz, *fmx, *fmy, *hmz, *float1, *float2, *float3;
};
// part2 (P2 / SP2) fields that the step touches.
struct P2View {
    double *rho_ref, *SR, *x0, *y0, *z0, *ux0, *uy0, *uz0, *temp0, *concn0;
};
// part3 (per-step scratch) plus the persistent IB normals.
struct QView {
    double *ftotalx, *ftotaly, *ftotalz, *ftotal, *fpx, *fpy, *fpz, *dtemp, *dconcn;
    double *inv_xx, *inv_yy, *inv_zz, *inv_xy, *inv_yz, *inv_zx;
    double *nx, *ny, *nz;
};
// Cell grid of the count==0 search.
struct Grid {
    sph_domain d;
    const int* gstr;
    const int* gend;
    const int* cstart;   // first particle (sorted order) with cell key >= c, c = 0 .. num_cells + 1
    const int* keys;     // sorted cell keys of the search
};
// Neighbour list, SELL-64 over particles, ascending j per particle.
struct NList {
    const long long* slice_off;   // first entry of each slice
    const int* cnt;               // entries per particle
    const unsigned* ent;          // j | selection flags (flags.h)
};
// one particle's (or tile lane's) entries: entry k at p[k * SLICE]
struct NLRow { const unsigned* p; int cnt; };
__device__ __forceinline__ NLRow nl_row(const NList& nl, long long i) {
    return {nl.ent + nl.slice_off[i >> 6] + (i & 63), nl.cnt[i]};
}
#define NB_AT(row, k) ((row).p[(size_t)(k) * SLICE])

// Cell tiles of the record-gathering passes (tiles.hip): chunks of at most
// tile_particles particles of one block of cells, each with its halo.
constexpr int TILE_MAX = 256;   // most lanes of a tile chunk (the tile passes' workgroup size)
struct TileView {
    const int* part;     // lane -> particle (chunk c: lanes c * cap ..), -1 on padding
    const int2* chunk;   // per chunk: first halo entry, halo size (-1: entries hold particle indices)
    const int* halo;     // the chunks' halo particles, ascending within a chunk
    const int* self;     // per lane: the halo slot of its own particle, or -1
    NList nl;            // the lanes' neighbour lists: halo slot | selection flags
};

// Mutable P1 fields: the per-step working copy holds its own arrays for
// these; the others alias the persistent state.
enum MutField {
    MF_UX, MF_UY, MF_UZ, MF_PRES, MF_TEMP, MF_VIS_T, MF_K_TURB, MF_E_TURB, MF_FLT_S,
    MF_FMX, MF_FMY, MF_HMZ, MF_FLOAT1, MF_FLOAT2, MF_FLOAT3, MF_FBX, MF_FBY, MF_FBZ,
    MF_CONCENTRATION, MF_CONCN, MF_ENTHALPY, MF_COUNT
};

struct Profile {
    bool on = false;
    struct Rec { std::string name; long long launches = 0; double ms = 0, bytes = 0; };
    std::vector<Rec> recs;
    struct Pending { int rec; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
};

struct KConst {
    int type;          // SPH_GAUSSIAN..SPH_WENDLAND6
    double kappa;
    double c_pi15;     // pow(PI, 1.5), evaluated once on the host
};

// PPE storage of one case (ppe.hip).  Rows are the unknowns (fluid
// particles); SELL-64 over rows.
struct PPEView {
    int nr;                       // unknowns
    int nb;                       // 256-row chunks = reduction partials
    const int* rows;              // row -> particle
    const int* rowof;             // particle -> row or -1
    const int* cnt;               // entries per row (off-diagonal)
    const long long* sl_off;      // first entry of each 64-row slice
    const int* cols;              // int32 columns (null when the case uses ecols)
    const unsigned short* ecols;  // MIXED: 16-bit columns (window group, offset), or null
    const int* gbase;             // MIXED: first column of each window group, 64 per slice
    const unsigned char* slotmap; // MIXED lattice order: slot of each entry, by its rank in the neighbour list
    // MIXED lattice order, fixed particles: per 64-row slice the first of its
    // column offsets in poff, or -1.  A stencil slice stores every row over
    // the same ascending offsets (col = row + poff[spat + k]; an offset the
    // row lacks holds the value 0), so its products read no columns.
    const int* spat;
    const int* poff;
    double* vals64;
    float* vals32;
    double *diag, *rhs, *x, *pnew;
    float* fdiag;                 // MIXED: float diagonal of the inner (float) operator
    double *r, *rhat, *p, *phat, *v, *s, *shat, *t;
    float *fr, *frhat, *fp, *fphat, *fv, *fs, *fshat, *ft, *fx, *frhs;
    float* mgx0;                  // the multigrid V-cycle's finest-level x (allocated with the PPE when the
                                  // multigrid may run), or null: k_kr_p / k_kr_s write its pre-smoothing
    double *partA, *partB;
    // MIXED: a row of more than LONG_ROW entries (a coarse row beside
    // refined particles) keeps its first LONG_ROW entries in the SELL slices
    // and the rest contiguously in a tail; one wave per such row sums its
    // tail just before each product, which adds that sum
    const int* ltail;             // row -> long-row index, or -1 (null: no long rows)
    const int* lrows;             // long-row index -> row
    const long long* tail_off;    // long-row index -> first tail entry (nlong + 1)
    const int* tail_cols;
    float* tail_vals;
    int nlong;
    double* lsum;                 // long-row index -> tail sum for the next product
};
constexpr int LONG_ROW = 256;

// Everything a pass needs to know about one case.  The engine keeps one per
// case of a step in device memory; kernels find theirs through the segment
// table of the launch.
struct CaseDesc {
    int n, count;
    int gc_fresh;        // the correction matrices inv_cm hold this search's values (fixed positions)
    int vals_fresh;      // the PPE values and diagonal hold this search's values (fixed positions)
    const int* porder;   // particle passes: thread -> particle (long neighbour lists last), or null
    int n_heavy;         // MIXED: the last n_heavy particles of porder (lists of more than 512 entries),
                         // which prep, advection and the pressure force give a wave each; 0 otherwise
    const int* nm_list;  // particles the Neumann pass updates (p_type != 1), in sorted order
    int nm_count;
    const long long* nmp_off;   // per nm_list particle: its first Neumann pair (nm_count + 1)
    const int* nmp_ent;         // Neumann pairs: the fluid neighbours inside the 2h search range, scan order
    // MIXED with float-exact positions: the distinct particles of those pairs
    // (nm_set, ascending), the pairs as indices into it, and a record of each
    // (NMREC_FLOATS: x y z m | rho ux uy uz | pres as a double, 0) packed by
    // the snapshot before each Neumann pass; nm_rec null otherwise
    const int* nmp_cent;
    const int* nm_set;
    int nm_nset;
    float* nm_rec;
    double dt, time;
    KConst kc;
    sph_params pr;
    PView P, S;          // persistent state, step working copy
    P2View P2, S2;
    QView Q;
    Grid g;
    NList nl;
    const int* keys;     // sorted cell keys of the count==0 search
    double* snap[5];     // ux, uy, uz, pres, temp snapshots (snap[4] doubles as MOST deposits)
    PPEView ppe;
    double* sc;          // Krylov scalars, computed on the device
    const double* ctl;   // solve controls written by the host: tol, maxit, skip
    double* xnew[3];     // ALE positions of the corrector (null unless ale_positions)
    double* rec;         // neighbour records (REC_DOUBLES per particle), refreshed before each gathering pass
    float* crec;         // MIXED with positions still float values: compact records (CREC_FLOATS per
                         // particle) of the pressure force, or null
    float* crec_adv;     // the same case's compact records of advection and assembly (CREC_ADV_FLOATS), or null
    const unsigned char* ibm_near;   // per particle: 1 if an IB marker is in its spreading list (this search)
    TileView tile;       // the cell tiles of this search (tile.part null: the passes gather per particle)
};
// CTL_MG: an FP64 solve whose p-hat and s-hat come from the float V-cycle
// (the double p and s are staged as float for it and widened after)
enum Ctl { CTL_TOL, CTL_MAXIT, CTL_SKIP, CTL_MG, CTL_COUNT = 4 };
constexpr int SCAL_COUNT = 32;   // Krylov scalars per case

// A segmented launch: block b works on case tab[b].x from index tab[b].y on
// (index = particle or PPE row).  tab == nullptr means one case, block b at
// b * blockDim.x.  Segments of different cases never share a block, so a
// case's blocks and their chunk partials are the same alone and in a batch.
struct Seg {
    const CaseDesc* cd;
    const int2* tab;
};
// The passes that give a boundary particle or a heavy list a 64-lane wave
// run four waves (four particles) to a workgroup.  The waves share nothing.
constexpr int WAVE_BLOCK = 256;
// The lane sums of a wave added by a fixed butterfly (every lane gets the
// total), so a particle's sums do not depend on the launch.
__device__ __forceinline__ void wave_fold(double& v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
}
// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b mod 8).  Logical block order is remapped so that each XCD works on one
// contiguous range of blocks: neighbouring rows, whose gathers touch the
// same particles, then share one L2.  A bijection on [0, gridDim.x).
constexpr unsigned NUM_XCD = 8;
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned b = blockIdx.x, nb = gridDim.x;
    const unsigned per = nb / NUM_XCD, rem = nb % NUM_XCD, x = b % NUM_XCD, k = b / NUM_XCD;
    return x * per + min(x, rem) + k;
}
__device__ __forceinline__ int seg_case(const Seg& s, int& first) {
    const unsigned lb = xcd_block();
    if (s.tab == nullptr) { first = (int)(lb * blockDim.x); return 0; }
    const int2 t = s.tab[lb];
    first = t.y;
    return t.x;
}


// launch tuning: the neighbour passes whose workgroup size the autotuner picks
enum TuneKernel {
    TK_GRAD_CORR, TK_PREP, TK_BOUNDARY, TK_NEUMANN, TK_ADVECTION, TK_FREESLIP, TK_PRESSURE, TK_IBM_FORCE,
    TK_IBM_SPREAD, TK_ASSEMBLE, TK_COUNT
};
static_assert(TK_COUNT == SPH_TUNE_PASSES, "tunable passes of the public tuning report");
constexpr int TUNE_CANDIDATES = 4;           // 64, 128, 256, 512 threads
constexpr int TUNE_BLOCKS[TUNE_CANDIDATES] = {64, 128, 256, 512};
constexpr int MAX_BLOCK = 512;

// Neighbour record: the fields a gathering pass reads of its neighbour j,
// one 128-byte line per particle, so a pair costs one line from L2 instead
// of one line per field array (advection, PPE assembly and pressure force;
// the strain pass reads fewer fields and is faster from the field arrays).  Pairs of doubles: 0 (x, y), 1 (z, m),
// 2 (rho, pres - pres0), 3 (ux, uy), 4 (uz, h), 5 (vis_t, p_type | i_type << 32),
// 6 (temp, concn).  The values are copies of the working state taken just
// before the pass (k_pack_records), so every pass computes with exactly the
// values it read from the field arrays before.
constexpr int REC_DOUBLES = 16;
// 16-byte loads of record pieces: explicit vector types in the global
// address space, so each piece is one global_load_dwordx4 (a double2 read
// through a generic pointer becomes two 8-byte flat loads)
typedef double rec_d2 __attribute__((ext_vector_type(2)));
typedef float rec_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ rec_d2 ld_d2(const void* p) {
    return *(const __attribute__((address_space(1))) rec_d2*)p;
}
__device__ __forceinline__ rec_f4 ld_f4(const void* p) {
    return *(const __attribute__((address_space(1))) rec_f4*)p;
}
__device__ __forceinline__ const double* rec_at(const double* rec, int j, int piece) {
    return rec + (size_t)j * REC_DOUBLES + 2 * piece;
}
// x y z m rho pq
__device__ __forceinline__ void rec_pos(const double* rec, int j, double& x, double& y, double& z, double& m,
                                        double& rho, double& pq) {
    const rec_d2 a = ld_d2(rec_at(rec, j, 0)), b = ld_d2(rec_at(rec, j, 1)), c = ld_d2(rec_at(rec, j, 2));
    x = a.x; y = a.y; z = b.x; m = b.y; rho = c.x; pq = c.y;
}
// ux uy uz h
__device__ __forceinline__ void rec_vel(const double* rec, int j, double& ux, double& uy, double& uz, double& h) {
    const rec_d2 a = ld_d2(rec_at(rec, j, 3)), b = ld_d2(rec_at(rec, j, 4));
    ux = a.x; uy = a.y; uz = b.x; h = b.y;
}
// vis_t p_type i_type
__device__ __forceinline__ void rec_aux(const double* rec, int j, double& vis_t, int& p_type, int& i_type) {
    const rec_d2 a = ld_d2(rec_at(rec, j, 5));
    vis_t = a.x;
    const long long t = __double_as_longlong(a.y);
    p_type = (int)(unsigned)(t & 0xffffffffLL);
    i_type = (int)(t >> 32);
}
// temp concn
__device__ __forceinline__ void rec_scalar(const double* rec, int j, double& temp, double& concn) {
    const rec_d2 a = ld_d2(rec_at(rec, j, 6));
    temp = a.x; concn = a.y;
}

// Compact neighbour records of the MIXED gathering passes, used while the
// positions are float values: the float pair math then reads
// exactly the values it reads from the double state ((float)(x_i - x_j) of
// two float-exact doubles is the rounded exact difference either way; m, rho
// and p - p0 enter the pressure force rounded to float; h is float-exact; the
// neighbour's viscosity is stored as the float the kernels round it to;
// velocities and the assembly's p - p0 stay double), so the results are bit
// for bit those of the 128-byte double records.
// Pressure force, 32 bytes: float4 (x, y, z, m), float4 (rho, pres - pres0, h,
// p_type as bits).
constexpr int CREC_FLOATS = 8;
// Advection and PPE assembly, 64 bytes: float4 (x, y, z, m), float4 (rho, h,
// (float)(viscosity2 + vis_t rho), p_type | i_type << 16 as bits), double2
// (ux, uy), double2 (uz, pres - pres0).
constexpr int CREC_ADV_FLOATS = 16;
// type word: p_type in the low 16 bits, i_type in bits 16..29, bit 30 set
// when the particle is an unknown of the case's PPE (it has a row)
constexpr int CREC_ROW_BIT = 1 << 30;
// Neumann-pass record (MIXED, float-exact positions): x y z m, rho ux uy uz as
// floats, then the pressure snapshot as a double and a pad double
constexpr int NMREC_FLOATS = 12;

// Where a gathering pass reads its neighbour records: in memory, indexed by
// particle (base null: the FP64 field arrays), or staged in LDS by a tile,
// indexed by halo slot.
template <int STRIDE>   // floats per record
struct GSrc {
    static constexpr bool lds = false;
    float* const* ref;   // the descriptor's record pointer (read where the loop starts), or null
    __device__ bool ok() const { return ref && *ref; }
    __device__ const float* at(int j) const { return *ref + (size_t)j * STRIDE; }
    __device__ const float* piece(const float* rec, int q) const { return rec + 4 * q; }   // 16-byte piece q
    __device__ static rec_f4 f4(const float* p) { return ld_f4(p); }
    __device__ static rec_d2 d2(const float* p) { return ld_d2(p); }
    __device__ int gidx(int j) const { return j; }           // particle of an entry
    __device__ bool self(int j, int i) const { return j == i; }
};
// The staged halo is piece-major: the first 16-byte piece of every record,
// then the second, ...: lanes reading one piece of random slots then spread
// over all LDS bank quads (whole records would use a quarter or half of them).
template <int STRIDE>
struct LSrc {
    static constexpr bool lds = true;
    const float* base;   // LDS
    int nrec;            // records staged (the stride between pieces, in 16-byte units)
    const int* halo;     // the chunk's halo: slot -> particle
    int self_slot;       // the slot of this thread's particle, or -1
    __device__ bool ok() const { return true; }
    __device__ const float* at(int j) const { return base + 4 * j; }
    __device__ const float* piece(const float* rec, int q) const { return rec + 4 * q * nrec; }
    __device__ static rec_f4 f4(const float* p) { return *(const __attribute__((address_space(3))) rec_f4*)p; }
    __device__ static rec_d2 d2(const float* p) { return *(const __attribute__((address_space(3))) rec_d2*)p; }
    __device__ int gidx(int j) const { return halo[j]; }
    __device__ bool self(int j, int) const { return j == self_slot; }
};

// Cell-tile passes (tiles.hip): a workgroup per chunk, blockDim = the chunk
// width.  Stages the chunk's halo records (RECF floats each, from RECS) in
// LDS, then maps the thread to its lane (a row of cd.tile.nl) and the
// lane's particle; padding lanes return after the barrier.
// A thread copies whole records (Q float4 pieces), TILE_U of them a round,
// all loads of a round issued before its stores; piece q of record r goes
// to float4 q * cnt + r.
constexpr int TILE_U = 4;
template <int Q>
__device__ __forceinline__ void tile_stage_q(float4* lds, const float* recs, const int* halo, int cnt) {
    const int B = blockDim.x;
    for (int r0 = threadIdx.x; r0 < cnt; r0 += TILE_U * B) {
        int idx[TILE_U];
#pragma unroll
        for (int u = 0; u < TILE_U; u++) idx[u] = r0 + u * B < cnt ? halo[r0 + u * B] : -1;
        rec_f4 v[TILE_U][Q];
#pragma unroll
        for (int u = 0; u < TILE_U; u++)
            if (idx[u] >= 0)
#pragma unroll
                for (int q = 0; q < Q; q++) v[u][q] = ld_f4(recs + ((size_t)idx[u] * Q + q) * 4);
#pragma unroll
        for (int u = 0; u < TILE_U; u++)
            if (idx[u] >= 0)
#pragma unroll
                for (int q = 0; q < Q; q++)
                    lds[q * cnt + r0 + u * B] = make_float4(v[u][q].x, v[u][q].y, v[u][q].z, v[u][q].w);
    }
}
__device__ __forceinline__ void tile_stage(float4* lds, const float* recs, const int* halo, int cnt, int recf) {
    if (recf == 8) tile_stage_q<2>(lds, recs, halo, cnt);
    else tile_stage_q<4>(lds, recs, halo, cnt);
}
#define TILE_PARTICLE(sg, RECS, RECF)                                                            \
    extern __shared__ float4 tile_lds_[];                                                       \
    int first_;                                                                                 \
    const int cid_ = seg_case((sg), first_);                                                    \
    const CaseDesc& cd = (sg).cd[cid_];                                                         \
    const int2 chunk_ = cd.tile.chunk[first_ / (int)blockDim.x];                                \
    if (chunk_.y > 0) tile_stage(tile_lds_, (RECS), cd.tile.halo + chunk_.x, chunk_.y, (RECF)); \
    __syncthreads();                                                                            \
    const int lane_ = first_ + (int)threadIdx.x;                                                \
    const int i = cd.tile.part[lane_];                                                          \
    if (i < 0) return;
__device__ __forceinline__ int crec_ptype(float w) { return (int)(short)(__float_as_int(w) & 0xffff); }
__device__ __forceinline__ int crec_itype(float w) { return (__float_as_int(w) >> 16) & 0x3fff; }
__device__ __forceinline__ bool crec_has_row(float w) { return (__float_as_int(w) & CREC_ROW_BIT) != 0; }

struct PPEState;   // ppe.hip

// The step engine: one stream, one profile, the cases of the current step
// and their device descriptors.  A single sph_step runs a batch of one.
struct Eng {
    hipStream_t st = nullptr;
    bool own_stream = false;
    Profile prof;
    std::string err;
    std::vector<sph_ctx*> cs;
    std::vector<long long> comp;            // composition signature of cs
    // descriptors (pinned staging, device copy) and per-case scalars
    CaseDesc* h_desc = nullptr;
    CaseDesc* d_desc = nullptr;
    double* h_sc = nullptr;  double* d_sc = nullptr;
    double* h_ctl = nullptr; double* d_ctl = nullptr;
    int cap = 0;
    hipEvent_t up_ev = nullptr;             // last upload from the staging buffers
    // segment tables: kind 0 = particles, 1 = PPE rows; one per block size
    struct Tab { int kind = 0, block = 0; unsigned nblk = 0; int2* d = nullptr; size_t cap = 0; bool valid = false; };
    std::vector<Tab> tabs;
    // autotuner: per size class (log2 of the batch's particle count)
    struct Tune {
        int phase = 0;
        int block[TK_COUNT];
        double ms[TK_COUNT][TUNE_CANDIDATES];
        int calls[TK_COUNT][TUNE_CANDIDATES];
    };
    std::vector<std::pair<int, Tune>> tune;
    bool autotune = true;
    struct TuneEv { int cls, kid, cand; hipEvent_t a, b; };
    std::vector<TuneEv> tune_ev;
    bool tune_open = false;
    bool tune_skip = false;   // this step starts a search for some case: not representative, not timed
    int tune_class = 0;
    // graph of one BiCGSTAB iteration per vector precision, valid for comp
    bool use_graphs = true;
    // FP64, float, float with multigrid, FP64 with multigrid
    static constexpr int NGRAPH = 4;
    hipGraphExec_t iter_graph[NGRAPH] = {nullptr, nullptr, nullptr, nullptr};
    // Krylov iterations of the last solves (per vector precision and
    // refinement pass): the host polls for convergence only near them
    static constexpr int PRED_SLOTS = 4;
    int pred_iters[NGRAPH][PRED_SLOTS] = {};
    // the physics before (0) and after (1) the pressure solve as graphs, for
    // the composition of the step and the variant of its launch sequence
    // (key): the first step of a variant runs directly (which builds every
    // table it uses), the next one is captured, later ones replay it
    struct StepGraph { long long key = -1, seen = -1; hipGraphExec_t exec = nullptr; };
    StepGraph step_graph[2];
    int tile_lanes = 0;   // this step's record passes run on cell tiles of this width (0: per particle)
    std::unordered_map<const void*, size_t> lds_allowed;   // kernels allowed more than 64 KB of dynamic LDS
    // multigrid V-cycles of the step's cases (mg.hip): per case and level a
    // descriptor, and block tables of the level rows (mg_plan)
    void* mg_lv = nullptr;            // device MGLv[case * MG_MAX_LEVELS + level]
    size_t mg_lv_cap = 0;
    int2* mg_tab = nullptr;           // device block tables (case, first row)
    size_t mg_tab_cap = 0;
    std::vector<long long> mg_sig;    // hierarchies the tables describe
    struct MGPlan {
        int lmax = 0;                 // most levels of any case
        int down_off[16], down_cnt[16];   // rows of level l, cases with a level below l
        int lvl_off[16], lvl_cnt[16];     // rows of level l, cases that have level l
        int dense_off = 0, dense_cnt = 0; // coarsest rows, four per block
        int long_off = 0, long_cnt = 0;   // level-0 long rows, four per block
        int clong_off[16], clong_cnt[16]; // coarse level l's long rows, four per block
        double spmv_bytes0 = 0.0;         // level-0 bytes of one pass (profile)
    } mg_plan;
};

}  // namespace sph

struct sph_ctx {
    sph_params prm{};
    int cap = 0, n = 0;
    int device = 0;                     // HIP device current at sph_create
    bool uploaded = false, searched = false;
    sph_domain dom{};
    double c_pi15 = 0.0;
    hipStream_t stream = nullptr;       // stream of the current call (the engine's during a step)
    hipStream_t own_stream = nullptr;   // stream of single-case calls
    sph::Profile* profp = nullptr;      // profile of the engine driving the step
    sph::Eng* solo = nullptr;           // engine of sph_step (a batch of one)
    std::string err;
    long long bytes = 0;                // device memory held: live blocks and cached free blocks
    long long generation = 0;           // bumped whenever device buffers move
    long long gc_gen = -1;              // generation whose gradient correction inv_cm holds
    long long vals_gen = -1;            // generation whose PPE values the matrix holds
    // device blocks: live ones with their sizes, and freed ones kept for
    // reuse by later allocations of about the same size (every use of a
    // block is ordered on the context's stream, so reuse needs no wait)
    std::unordered_map<void*, size_t> live;
    std::multimap<size_t, void*> cache;

    // persistent state
    double* pf[SPH_F_COUNT_F64 + 3] = {};  // P1 fields in SPH_F_* numbering, then eli, m_ref, h_ref
    int* pi[SPH_I_COUNT] = {};
    double* sm[sph::MF_COUNT] = {};      // working copies of mutable fields
    double* p2[10] = {};                 // rho_ref, SR, x0, y0, z0, ux0, uy0, uz0, temp0, concn0
    double* s2[10] = {};
    double* q[18] = {};                  // part3 scratch + normals (last 3 persistent)
    double* snap[5] = {};                // snapshot buffers (ux, uy, uz, pres, temp)

    // neighbour search
    int* keys = nullptr;  int* perm = nullptr;
    int* keys_tmp = nullptr; int* perm_tmp = nullptr;
    int* gstr = nullptr;  int* gend = nullptr;
    int* cstart = nullptr;
    int* nm_list = nullptr;   // non-fluid particles of the search (Neumann pass)
    int nm_count = 0;
    long long* nmp_off = nullptr;   // Neumann pairs of nm_list (built by neumann_pairs)
    int* nmp_ent = nullptr;
    int* nmp_cent = nullptr;        // MIXED, float-exact positions: pairs as indices into nm_set
    int* nm_set = nullptr;          // the distinct particles of the pairs, ascending
    int nm_nset = 0;
    float* nm_rec = nullptr;        // their records (NMREC_FLOATS each)
    long long nmp_gen = -1;         // generation, list and precision they were built for
    const int* nmp_for = nullptr;
    int nmp_prec = -1;
    int* sort_hist = nullptr; int sort_hist_n = 0;
    long long* nl_slice_off = nullptr; int* nl_cnt = nullptr; unsigned* nl_ent = nullptr;
    long long nl_total = 0;
    long long* nl_wprev = nullptr;   // slice widths of the last search (room for the next)
    int* porder = nullptr;           // order of the particle passes (CaseDesc::porder), or null
    int n_heavy = 0;                 // particles of that search with lists of more than 512 entries
    unsigned char* ibm_near = nullptr;   // per particle of the search: an IB marker in its spreading list
    long long ibm_near_gen = -1;         // generation the flags were built for
    // cell tiles of the record passes (tiles.hip), built for tl_gen with the
    // tile parameters tl_code / tl_cap / tl_hcap
    int* tl_part = nullptr; int2* tl_chunk = nullptr; int* tl_halo = nullptr; int* tl_self = nullptr;
    long long* tl_slice_off = nullptr; int* tl_cnt = nullptr; unsigned* tl_ent = nullptr;
    int tl_nchunk = 0, tl_hbig = 0, tl_raw = 0;   // chunks, largest staged halo, chunks above the cap
    long long tl_halo_total = 0;
    int tl_code = 0, tl_cap = 0, tl_hcap = 0;     // the parameters they were built with
    int tl_lanes = 0;                             // lanes (threads) of a chunk
    long long tl_gen = -1;
    int nl_wprev_n = 0;              // its slices
    int* scan_tmp = nullptr; size_t scan_tmp_n = 0;

    float* stat_buf = nullptr;           // stats reduction scratch
    double* apr_target = nullptr;        // APR target level per particle (the reference's M_num)
    double* xnew[3] = {};                // ALE: corrector positions, committed at the end of the step
    double* rec = nullptr;               // neighbour records of the gathering passes (cap * REC_DOUBLES)
    float* crec = nullptr;               // compact pressure-force records (cap * CREC_FLOATS), MIXED only
    float* crec_adv = nullptr;           // compact advection / assembly records (cap * CREC_ADV_FLOATS), MIXED only
    bool pos_f32 = true;                 // positions are float values (uploaded or refined; no motion)

    sph::PPEState* ppe = nullptr;
    sph_ppe_report rep{};

    sph::PView pview() const;
    sph::PView sview() const;
    sph::P2View p2view() const;
    sph::P2View s2view() const;
    sph::QView qview() const;
    sph::Grid grid() const;
    sph::NList nlist() const;
};

namespace sph {

// error helpers ------------------------------------------------------------
template <class T>
int fail(T* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
#define SPH_HIP(ctx, call)                                                          \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::sph::fail((ctx), SPH_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define SPH_TRY(expr)            \
    do {                         \
        int r_ = (expr);         \
        if (r_ != 0) return r_;  \
    } while (0)

int dmalloc(sph_ctx* c, void** p, size_t bytes);
template <class T>
int dalloc(sph_ctx* c, T** p, size_t count) { return dmalloc(c, (void**)p, count * sizeof(T)); }
void dfree(sph_ctx* c, void* p);
void dcache_release(sph_ctx* c);

// profiling around launches (the profile of the driving engine)
int prof_begin(sph_ctx* c, const char* name, double alg_bytes);
int prof_end(sph_ctx* c);
int prof_begin(Eng& e, const char* name, double alg_bytes);
int prof_end(Eng& e);
int prof_collect(Eng& e);
template <class T>
int launch_check(T* c, const char* what) {
    hipError_t er = hipGetLastError();
    (void)c;
    if (er != hipSuccess) return fail(c, SPH_E_HIP, std::string("launch of ") + what + ": " + hipGetErrorString(er));
    return 0;
}

inline unsigned grid1(long long n, int block = BLOCK) { return (unsigned)((n + block - 1) / block); }

// nnps.hip
int nnps_run(sph_ctx* c, bool build_list = true);
int nnps_list(sph_ctx* c);
int radix_sort_pairs(sph_ctx* c, int** keys, int** vals, int** ktmp, int** vtmp, int n, int max_key);
int set_ncell_run(sph_ctx* c);
int exclusive_scan_i32(sph_ctx* c, const int* in, int* out, long long n, int* total_out);
int exclusive_scan_i64(sph_ctx* c, const long long* in, long long* out, long long n, long long* total_out);

// engine.hip: the step over a set of cases, launch shapes, tuning
int eng_init(Eng& e, hipStream_t st);   // st == nullptr: the engine creates its own stream
void eng_free(Eng& e);
int engine_step(Eng& e, sph_ctx* const* cases, int nc, const int* count, const double* dt, const double* time);
struct Shape { unsigned grid; int block; const int2* tab; };
int shape(Eng& e, int kind, int block, Shape* out);
int tuned_block(Eng& e, int kid);
int tune_begin(Eng& e, int kid, int block);
int tune_end(Eng& e);
size_t tile_lds_bytes(const Eng& e, int recf);
int tile_lds_allow(Eng& e, const void* kern, size_t bytes);

// apr.hip
int apr_refine(sph_ctx* c, const double* target, int max_passes, int* n_out);

// physics.hip
int step_physics_front(Eng& e);
int step_physics_back(Eng& e);
// the choices of those launch sequences that vary from step to step, as keys
long long step_physics_front_key(const Eng& e);
long long step_physics_back_key(const Eng& e);
bool tune_pending(Eng& e);   // the autotuner still times candidates for this step's size class
int stats_run(sph_ctx* c, float* umax, float* drho, float* fmax);
int stats_run_noalloc(sph_ctx* c, float* umax, float* drho, float* fmax);

// ppe.hip
int ppe_build(sph_ctx* c);
PPEView ppe_view(const sph_ctx* c);
int ppe_nlong(const sph_ctx* c);   // rows with tails of the case's PPE
// neighbour records from the working state (physics.hip): REC_ALL, or only
// the pairs that hold velocities and pressures (REC_VEL_PQ) or pressures (REC_PQ)
// REC_PRE: before the strain pass, the compact records only (their viscosity
// slot is refreshed by REC_ALL after that pass has computed vis_t)
enum { REC_ALL = 0, REC_VEL_PQ = 1, REC_PQ = 2, REC_PRE = 3 };
int pack_records(Eng& e, int which);
// Neumann pairs of a case for its current search and precision (physics.hip)
int neumann_pairs(sph_ctx* c);
// the IBM spreading targets of a case's current search (physics.hip)
int ibm_near_build(sph_ctx* c);
// tiles.hip: the cell tiles of a case's current search
int tile_build(sph_ctx* c);
// MIXED: the particles whose neighbour lists the record passes give a wave
// each (the heavy ones at the end of the pass order); 0 in FP64, whose sums
// keep the serial order
int heavy_waves(const sph_ctx* c);
void tile_release(sph_ctx* c);
// aggregation multigrid preconditioner of the MIXED inner solve (mg.hip)
struct MGState;
constexpr int MG_COARSE_MAX = 1024;   // unknowns of the densely inverted coarsest level
constexpr int MG_MAX_LEVELS = 12;
// Builds the hierarchy of the PPE in pp (float values, `slots` stored SELL
// slots) with lattice spacing q; *ok = 0 (and *out = null) when the case
// does not suit it.
int mg_build(sph_ctx* c, const PPEView& pp, long long slots, double q, MGState** out, int* ok, int* why);
// Tables and descriptors of the V-cycles of the step's cases (mgs[k]: case
// k's hierarchy or null for a case that uses Jacobi); *changed when they
// differ from the previous call's.
int mg_prepare(Eng& e, const std::vector<const MGState*>& mgs, int sc_done, bool* changed);
// One V-cycle for every multigrid case of the step: which = 0 p -> p-hat,
// 1 s -> s-hat (float vectors); a case's launches return at once when its
// done flag (scalar sc_done) is set.
int mg_apply(Eng& e, int which, int sc_done);
void mg_free(sph_ctx* c, MGState* mg);
int mg_levels(const MGState* mg);
int mg_coarsest(const MGState* mg);
long long mg_sorted_rows(const MGState* mg);   // coarse rows formed by the sorting Galerkin kernels
int mg_long_rows(const MGState* mg);           // coarse rows of more than 128 entries (a wave each)
int ppe_solve(Eng& e);   // every case of the step; 0, or SPH_E_NOCONV when a case did not converge
void ppe_free(sph_ctx* c);

}  // namespace sph
