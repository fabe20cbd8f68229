// Step engine of the MI355X ISPH library: the projection step of one case
// or of many independent cases as one launch sequence.
//
// Every pass is launched once over the particles (or PPE rows) of all cases
// of the step.  A segment table maps each workgroup to (case, first index);
// segments start at multiples of the workgroup size, so no workgroup mixes
// two cases and a case's arithmetic, its reductions included, is the same
// alone and in any batch.  Case descriptors (buffer pointers, parameters,
// count / dt / time) go to the device once per step from pinned staging
// memory; the Krylov iterations read their scalars and controls from device
// memory, which lets one captured HIP graph replay any iteration.
//
// Autotuner: the neighbour passes run at 64, 128, 256 and 512 threads per
// workgroup on four consecutive steps of each batch size class (log2 of the
// batch's particle count), timed with HIP events; afterwards each pass uses
// the size that was fastest per call.  One thread per particle and no
// cross-thread arithmetic in those passes: the choice never changes results.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "sph_internal.h"

namespace sph {

// ---------------------------------------------------------------- profile
static int prof_rec(Profile& P, const char* name) {
    for (size_t k = 0; k < P.recs.size(); k++)
        if (P.recs[k].name == name) return (int)k;
    P.recs.push_back(Profile::Rec{name, 0, 0.0, 0.0});
    return (int)P.recs.size() - 1;
}
template <class T>
static int prof_begin_on(T* owner, Profile& P, hipStream_t st, const char* name, double bytes) {
    Profile::Pending p{prof_rec(P, name), nullptr, nullptr, bytes};
    SPH_HIP(owner, hipEventCreate(&p.a));
    SPH_HIP(owner, hipEventCreate(&p.b));
    SPH_HIP(owner, hipEventRecord(p.a, st));
    P.pending.push_back(p);
    return 0;
}
int prof_begin(Eng& e, const char* name, double bytes) {
    if (!e.prof.on) return 0;
    return prof_begin_on(&e, e.prof, e.st, name, bytes);
}
int prof_end(Eng& e) {
    if (!e.prof.on) return 0;
    SPH_HIP(&e, hipEventRecord(e.prof.pending.back().b, e.st));
    return 0;
}
int prof_begin(sph_ctx* c, const char* name, double bytes) {
    if (!c->profp || !c->profp->on) return 0;
    return prof_begin_on(c, *c->profp, c->stream, name, bytes);
}
int prof_end(sph_ctx* c) {
    if (!c->profp || !c->profp->on) return 0;
    SPH_HIP(c, hipEventRecord(c->profp->pending.back().b, c->stream));
    return 0;
}
// Every pending pair is timed and destroyed, also after a failure (the
// first failure is returned), so no event outlives the list.
int prof_collect(Eng& e) {
    hipError_t first = hipSuccess;
    for (auto& p : e.prof.pending) {
        float ms = 0.f;
        hipError_t er = hipEventSynchronize(p.b);
        if (er == hipSuccess) er = hipEventElapsedTime(&ms, p.a, p.b);
        if (er == hipSuccess) {
            auto& r = e.prof.recs[p.rec];
            r.launches++;
            r.ms += ms;
            r.bytes += p.bytes;
        } else if (first == hipSuccess) {
            first = er;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    e.prof.pending.clear();
    SPH_HIP(&e, first);
    return 0;
}

// ---------------------------------------------------------------- tuning
static Eng::Tune& tune_for(Eng& e, int cls) {
    for (auto& t : e.tune)
        if (t.first == cls) return t.second;
    Eng::Tune t{};
    for (int k = 0; k < TK_COUNT; k++) t.block[k] = BLOCK;
    t.block[TK_NEUMANN] = 64;   // one wave per boundary particle, not tuned
    e.tune.emplace_back(cls, t);
    return e.tune.back().second;
}
int tuned_block(Eng& e, int kid) {
    if (!e.autotune) return BLOCK;
    Eng::Tune& t = tune_for(e, e.tune_class);
    if (t.phase < TUNE_CANDIDATES) return TUNE_BLOCKS[t.phase];
    return t.block[kid];
}
bool tune_pending(Eng& e) {
    return e.autotune && tune_for(e, e.tune_class).phase < TUNE_CANDIDATES;
}
int tune_begin(Eng& e, int kid, int block) {
    e.tune_open = false;
    if (!e.autotune) return 0;
    Eng::Tune& t = tune_for(e, e.tune_class);
    if (e.tune_skip || t.phase >= TUNE_CANDIDATES || TUNE_BLOCKS[t.phase] != block) return 0;
    Eng::TuneEv ev{e.tune_class, kid, t.phase, nullptr, nullptr};
    SPH_HIP(&e, hipEventCreate(&ev.a));
    SPH_HIP(&e, hipEventCreate(&ev.b));
    SPH_HIP(&e, hipEventRecord(ev.a, e.st));
    e.tune_ev.push_back(ev);
    e.tune_open = true;
    return 0;
}
int tune_end(Eng& e) {
    if (!e.tune_open) return 0;
    e.tune_open = false;
    SPH_HIP(&e, hipEventRecord(e.tune_ev.back().b, e.st));
    return 0;
}
// Fold the timings of the previous step into its size class; after the
// last candidate, pick each pass's fastest size per call.
static int tune_fold(Eng& e) {
    if (e.tune_ev.empty()) return 0;
    std::vector<int> classes;
    for (auto& ev : e.tune_ev) {
        float ms = 0.f;
        SPH_HIP(&e, hipEventSynchronize(ev.b));
        SPH_HIP(&e, hipEventElapsedTime(&ms, ev.a, ev.b));
        Eng::Tune& t = tune_for(e, ev.cls);
        t.ms[ev.kid][ev.cand] += ms;
        t.calls[ev.kid][ev.cand]++;
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
        if (std::find(classes.begin(), classes.end(), ev.cls) == classes.end()) classes.push_back(ev.cls);
    }
    e.tune_ev.clear();
    for (int cls : classes) {
        Eng::Tune& t = tune_for(e, cls);
        if (t.phase >= TUNE_CANDIDATES) continue;
        if (++t.phase < TUNE_CANDIDATES) continue;
        for (int k = 0; k < TK_COUNT; k++) {
            if (k == TK_NEUMANN) continue;
            int best = -1;
            double best_ms = 0.0;
            for (int q = 0; q < TUNE_CANDIDATES; q++) {
                if (t.calls[k][q] == 0) continue;
                double per = t.ms[k][q] / t.calls[k][q];
                if (best < 0 || per < best_ms) { best = q; best_ms = per; }
            }
            t.block[k] = best < 0 ? BLOCK : TUNE_BLOCKS[best];
        }
    }
    return 0;
}

// ---------------------------------------------------------------- shapes
// the cell key of the first particle of every block of `block` particles
__global__ void k_block_keys(const int* keys, long long nb, int block, int* out) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) out[b] = keys[b * block];
}

// Morton code (10 bits an axis) of the cell of each block's first particle,
// and the block's first particle
__device__ __forceinline__ unsigned spread10(unsigned v) {
    v &= 0x3ff;
    v = (v | v << 16) & 0x030000ffu;
    v = (v | v << 8) & 0x0300f00fu;
    v = (v | v << 4) & 0x030c30c3u;
    v = (v | v << 2) & 0x09249249u;
    return v;
}
__global__ void __launch_bounds__(256) k_block_morton(const int* keys, int nb, int block, int NI, int NJ, int* code,
                                                     int* first) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const long long cell = keys[(size_t)b * block];
    const unsigned ci = (unsigned)(cell % NI), cj = (unsigned)((cell / NI) % NJ),
                   ck = (unsigned)(cell / ((long long)NI * NJ));
    code[b] = (int)(spread10(ci) | spread10(cj) << 1 | spread10(ck) << 2);
    first[b] = b * block;
}
__global__ void __launch_bounds__(256) k_block_table(const int* first, int nb, int k, int2* out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) out[b] = make_int2(k, first[b]);
}

// The particle-block table of shape() built on the device: per case, a stable
// radix sort of the blocks by the Morton code of their first cell (the order
// the host's stable sort gives), written at the case's offset.
template <class U>
static int shape_device(Eng& e, int block, long long total, const U& units, Eng::Tab* t, Shape* out) {
    const int nc = (int)e.cs.size();
    if (!t) {
        e.tabs.push_back(Eng::Tab{});
        t = &e.tabs.back();
        t->kind = 0;
        t->block = block;
    }
    if ((size_t)total > t->cap) {
        if (t->d) (void)hipFree(t->d);
        t->d = nullptr;
        t->cap = std::max<size_t>((size_t)total, 2 * t->cap);
        if (hipMalloc(&t->d, sizeof(int2) * t->cap) != hipSuccess) {
            t->d = nullptr;
            t->cap = 0;
            (void)hipGetLastError();
            return fail(&e, SPH_E_NOMEM, "segment table allocation failed");
        }
    }
    long long off = 0;
    for (int k = 0; k < nc; k++) {
        sph_ctx* c = e.cs[k];
        const int nb = (int)((units(k) + block - 1) / block);
        if (nb == 0) continue;
        int *code = nullptr, *first = nullptr, *ct = nullptr, *ft = nullptr;
        SPH_TRY(dalloc(c, &code, nb));
        SPH_TRY(dalloc(c, &first, nb));
        SPH_TRY(dalloc(c, &ct, nb));
        SPH_TRY(dalloc(c, &ft, nb));
        hipLaunchKernelGGL(k_block_morton, dim3(grid1(nb)), dim3(256), 0, e.st, (const int*)c->keys, nb, block,
                           c->dom.NI, c->dom.NJ, code, first);
        SPH_TRY(launch_check(&e, "block_morton"));
        SPH_TRY(radix_sort_pairs(c, &code, &first, &ct, &ft, nb, (1 << 30) - 1));
        hipLaunchKernelGGL(k_block_table, dim3(grid1(nb)), dim3(256), 0, e.st, (const int*)first, nb, k, t->d + off);
        SPH_TRY(launch_check(&e, "block_table"));
        dfree(c, code);
        dfree(c, first);
        dfree(c, ct);
        dfree(c, ft);
        off += nb;
    }
    t->nblk = (unsigned)total;
    t->valid = true;
    out->grid = t->nblk;
    out->tab = t->d;
    return 0;
}

int heavy_waves(const sph_ctx* c) { return c->prm.precision == SPH_MIXED && c->porder ? c->n_heavy : 0; }

int shape(Eng& e, int kind, int block, Shape* out) {
    const int nc = (int)e.cs.size();
    auto units = [&](int k) -> long long {
        const sph_ctx* c = e.cs[k];
        // kind 0: particles; 1: PPE rows; 2: one 64-lane wave per long PPE row;
        // 3: one 64-lane wave per boundary particle of the Neumann pass;
        // 4: the lanes of the cell-tile chunks (a block per chunk); 5: one
        // 64-lane wave per particle of a heavy neighbour list
        return kind == 0 ? c->n : kind == 1 ? ppe_view(c).nr : kind == 2 ? 64LL * ppe_view(c).nlong
             : kind == 3 ? 64LL * c->nm_count : kind == 5 ? 64LL * heavy_waves(c)
             : (long long)c->tl_nchunk * c->tl_lanes;
    };
    out->block = block;
    if (nc == 1 && kind != 0) {
        out->grid = (unsigned)((units(0) + block - 1) / block);
        out->tab = nullptr;
        return 0;
    }
    Eng::Tab* t = nullptr;
    for (auto& x : e.tabs)
        if (x.kind == kind && x.block == block) { t = &x; break; }
    if (t && t->valid) {
        out->grid = t->nblk;
        out->tab = t->d;
        return 0;
    }
    if (kind == 0) {
        // cases whose cell grid fits 10 bits an axis sort their blocks on the
        // device (no copy to the host and back)
        bool dev = true;
        long long total = 0;
        for (int k = 0; k < nc; k++) {
            const sph_ctx* c = e.cs[k];
            const sph_domain& d = c->dom;
            dev &= c->keys != nullptr && d.NI > 0 && d.NJ > 0 && d.NI <= 1024 && d.NJ <= 1024 && d.NK < 1024;
            total += (units(k) + block - 1) / block;
        }
        if (dev) return shape_device(e, block, total, units, t, out);
    }
    std::vector<int2> h;
    if (kind == 0) {
        // Particle blocks of each case in Morton order of the cell of their
        // first particle: the blocks one XCD runs at a time (a contiguous
        // share of the table, see xcd_block) then cover a compact 3-D region,
        // whose neighbours its L2 holds better than a row-by-row sweep's.
        for (int k = 0; k < nc; k++) {
            const sph_ctx* c = e.cs[k];
            const long long u = units(k);
            // the cell key of each block's first particle
            const size_t nb = (size_t)((u + block - 1) / block);
            std::vector<int> key(c->keys && u > 0 ? nb : 0);
            if (!key.empty()) {
                int* dk = nullptr;
                SPH_TRY(dalloc(e.cs[k], &dk, nb));   // the context's block cache: no device-wide wait
                hipLaunchKernelGGL(k_block_keys, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, e.st,
                                   (const int*)c->keys, (long long)nb, block, dk);
                hipError_t er = hipMemcpyAsync(key.data(), dk, sizeof(int) * nb, hipMemcpyDeviceToHost, e.st);
                if (er == hipSuccess) er = hipStreamSynchronize(e.st);
                dfree(e.cs[k], dk);
                SPH_HIP(&e, er);
            }
            const sph_domain& d = c->dom;
            auto spread = [](unsigned v) {
                unsigned long long x = v & 0x1fffff;
                x = (x | x << 32) & 0x1f00000000ffffULL;
                x = (x | x << 16) & 0x1f0000ff0000ffULL;
                x = (x | x << 8) & 0x100f00f00f00f00fULL;
                x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
                x = (x | x << 2) & 0x1249249249249249ULL;
                return x;
            };
            std::vector<std::pair<unsigned long long, int>> blk;
            for (long long f = 0; f < u; f += block) {
                unsigned long long m = 0;
                if (!key.empty() && d.NI > 0 && d.NJ > 0) {
                    const long long cell = key[(size_t)(f / block)];
                    const unsigned ci = (unsigned)(cell % d.NI), cj = (unsigned)((cell / d.NI) % d.NJ),
                                   ck = (unsigned)(cell / ((long long)d.NI * d.NJ));
                    m = spread(ci) | spread(cj) << 1 | spread(ck) << 2;
                }
                blk.push_back({m, (int)f});
            }
            std::stable_sort(blk.begin(), blk.end(),
                             [](const std::pair<unsigned long long, int>& a, const std::pair<unsigned long long, int>& b) {
                                 return a.first < b.first;
                             });
            for (auto& b : blk) h.push_back(make_int2(k, b.second));
        }
    } else {
        for (int k = 0; k < nc; k++) {
            const long long u = units(k);
            for (long long f = 0; f < u; f += block) h.push_back(make_int2(k, (int)f));
        }
    }
    if (!t) {
        e.tabs.push_back(Eng::Tab{});
        t = &e.tabs.back();
        t->kind = kind;
        t->block = block;
    }
    if (h.size() > t->cap) {
        if (t->d) (void)hipFree(t->d);
        t->d = nullptr;
        t->cap = std::max<size_t>(h.size(), 2 * t->cap);
        hipError_t er = hipMalloc(&t->d, sizeof(int2) * t->cap);
        if (er != hipSuccess) {
            t->d = nullptr;
            t->cap = 0;
            (void)hipGetLastError();
            return fail(&e, SPH_E_NOMEM, "segment table allocation failed");
        }
    }
    // a new table is used only by launches after this point on the engine's
    // stream; tables of an older composition were drained before reuse
    if (!h.empty()) SPH_HIP(&e, hipMemcpy(t->d, h.data(), sizeof(int2) * h.size(), hipMemcpyHostToDevice));
    t->nblk = (unsigned)h.size();
    t->valid = true;
    out->grid = t->nblk;
    out->tab = t->d;
    return 0;
}

// dynamic LDS of a cell-tile launch: the largest staged halo of the step's
// cases, recf floats a record
size_t tile_lds_bytes(const Eng& e, int recf) {
    int h = 0;
    for (const sph_ctx* c : e.cs) h = std::max(h, c->tl_hbig);
    return (size_t)h * recf * sizeof(float);
}
// launches of more than 64 KB of dynamic LDS are allowed per kernel
int tile_lds_allow(Eng& e, const void* kern, size_t bytes) {
    if (bytes <= 65536) return 0;
    auto it = e.lds_allowed.find(kern);
    if (it != e.lds_allowed.end() && it->second >= bytes) return 0;
    SPH_HIP(&e, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    e.lds_allowed[kern] = bytes;
    return 0;
}

// ---------------------------------------------------------------- engine
int eng_init(Eng& e, hipStream_t st) {
    if (st) e.st = st;
    else {
        SPH_HIP(&e, hipStreamCreateWithFlags(&e.st, hipStreamNonBlocking));
        e.own_stream = true;
    }
    SPH_HIP(&e, hipEventCreateWithFlags(&e.up_ev, hipEventDisableTiming));
    SPH_HIP(&e, hipEventRecord(e.up_ev, e.st));
    return 0;
}

static void drop_graphs(Eng& e) {
    for (auto& g : e.iter_graph)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    for (auto& g : e.step_graph) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = Eng::StepGraph{};
    }
}

// One half of the step's physics: replayed from its graph when this
// composition already ran the same variant twice, else run directly (the
// first time) or captured and launched (the second).  Graphs are off while
// the profile is on or the autotuner still times candidates.
static int run_half(Eng& e, int which, long long key, int (*fn)(Eng&)) {
    if (!e.use_graphs || e.prof.on || tune_pending(e)) return fn(e);
    Eng::StepGraph& g = e.step_graph[which];
    if (g.exec && g.key == key) {
        SPH_HIP(&e, hipGraphLaunch(g.exec, e.st));
        return 0;
    }
    if (g.seen != key) {
        g.seen = key;
        return fn(e);
    }
    if (g.exec) {   // another variant's graph: nothing in flight may still run it
        SPH_HIP(&e, hipStreamSynchronize(e.st));
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
    }
    hipGraph_t gr = nullptr;
    SPH_HIP(&e, hipStreamBeginCapture(e.st, hipStreamCaptureModeThreadLocal));
    const int rc = fn(e);
    hipError_t er = hipStreamEndCapture(e.st, &gr);
    if (rc) {
        if (er == hipSuccess && gr) (void)hipGraphDestroy(gr);
        return rc;
    }
    SPH_HIP(&e, er);
    er = hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (er != hipSuccess) g.exec = nullptr;
    SPH_HIP(&e, er);
    g.key = key;
    SPH_HIP(&e, hipGraphLaunch(g.exec, e.st));
    return 0;
}

void eng_free(Eng& e) {
    if (e.st) (void)hipStreamSynchronize(e.st);
    (void)prof_collect(e);
    for (auto& ev : e.tune_ev) { (void)hipEventDestroy(ev.a); (void)hipEventDestroy(ev.b); }
    e.tune_ev.clear();
    drop_graphs(e);
    for (auto& t : e.tabs)
        if (t.d) (void)hipFree(t.d);
    e.tabs.clear();
    if (e.mg_lv) (void)hipFree(e.mg_lv);
    if (e.mg_tab) (void)hipFree(e.mg_tab);
    e.mg_lv = nullptr; e.mg_tab = nullptr; e.mg_lv_cap = e.mg_tab_cap = 0;
    e.mg_sig.clear();
    if (e.d_desc) (void)hipFree(e.d_desc);
    if (e.d_sc) (void)hipFree(e.d_sc);
    if (e.d_ctl) (void)hipFree(e.d_ctl);
    if (e.h_desc) (void)hipHostFree(e.h_desc);
    if (e.h_sc) (void)hipHostFree(e.h_sc);
    if (e.h_ctl) (void)hipHostFree(e.h_ctl);
    e.d_desc = nullptr; e.d_sc = nullptr; e.d_ctl = nullptr;
    e.h_desc = nullptr; e.h_sc = nullptr; e.h_ctl = nullptr;
    e.cap = 0;
    if (e.up_ev) (void)hipEventDestroy(e.up_ev);
    e.up_ev = nullptr;
    if (e.own_stream && e.st) (void)hipStreamDestroy(e.st);
    e.st = nullptr;
}

static int eng_reserve(Eng& e, int nc) {
    if (nc <= e.cap) return 0;
    SPH_HIP(&e, hipStreamSynchronize(e.st));
    const int cap = std::max(nc, 2 * e.cap);
    if (e.d_desc) (void)hipFree(e.d_desc);
    if (e.d_sc) (void)hipFree(e.d_sc);
    if (e.d_ctl) (void)hipFree(e.d_ctl);
    if (e.h_desc) (void)hipHostFree(e.h_desc);
    if (e.h_sc) (void)hipHostFree(e.h_sc);
    if (e.h_ctl) (void)hipHostFree(e.h_ctl);
    e.d_desc = nullptr; e.d_sc = nullptr; e.d_ctl = nullptr;
    e.h_desc = nullptr; e.h_sc = nullptr; e.h_ctl = nullptr;
    e.cap = 0;
    bool ok = hipMalloc(&e.d_desc, sizeof(CaseDesc) * cap) == hipSuccess &&
              hipMalloc(&e.d_sc, sizeof(double) * SCAL_COUNT * cap) == hipSuccess &&
              hipMalloc(&e.d_ctl, sizeof(double) * CTL_COUNT * cap) == hipSuccess &&
              hipHostMalloc((void**)&e.h_desc, sizeof(CaseDesc) * cap, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&e.h_sc, sizeof(double) * SCAL_COUNT * cap, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&e.h_ctl, sizeof(double) * CTL_COUNT * cap, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        return fail(&e, SPH_E_NOMEM, "engine descriptor allocation failed");
    }
    std::memset(e.h_ctl, 0, sizeof(double) * CTL_COUNT * cap);
    std::memset(e.h_sc, 0, sizeof(double) * SCAL_COUNT * cap);
    SPH_HIP(&e, hipMemcpy(e.d_ctl, e.h_ctl, sizeof(double) * CTL_COUNT * cap, hipMemcpyHostToDevice));
    SPH_HIP(&e, hipMemcpy(e.d_sc, e.h_sc, sizeof(double) * SCAL_COUNT * cap, hipMemcpyHostToDevice));
    e.cap = cap;
    return 0;
}

static void fill_desc(Eng& e, int k, sph_ctx* c, int count, double dt, double time) {
    CaseDesc& d = e.h_desc[k];
    std::memset(&d, 0, sizeof(d));
    d.n = c->n;
    d.count = count;
    // inv_cm depends only on positions, h, m, rho, p_type and the neighbour
    // list; with fixed particles it is the same every step of a search
    d.gc_fresh = !c->prm.ale_positions && c->gc_gen == c->generation;
    // the PPE values depend on the same inputs and on the row numbering
    d.vals_fresh = !c->prm.ale_positions && c->vals_gen == c->generation;
    d.nm_list = c->nm_list;
    d.nm_count = c->nm_count;
    d.nmp_off = c->nmp_off;
    d.nmp_ent = c->nmp_ent;
    const bool nmrec = c->prm.precision == SPH_MIXED && c->pos_f32 && !c->prm.ale_positions && c->prm.compact_records &&
                       c->nm_rec;
    d.nmp_cent = nmrec ? c->nmp_cent : nullptr;
    d.nm_set = nmrec ? c->nm_set : nullptr;
    d.nm_nset = nmrec ? c->nm_nset : 0;
    d.nm_rec = nmrec ? c->nm_rec : nullptr;
    d.dt = dt;
    d.time = time;
    d.kc = KConst{c->prm.kernel_type, c->prm.kappa, c->c_pi15};
    d.pr = c->prm;
    d.P = c->pview();
    d.S = c->sview();
    d.P2 = c->p2view();
    d.S2 = c->s2view();
    d.Q = c->qview();
    d.g = c->grid();
    d.nl = c->nlist();
    d.porder = c->porder;
    d.n_heavy = heavy_waves(c);
    d.keys = c->keys;
    for (int f = 0; f < 5; f++) d.snap[f] = c->snap[f];
    d.ppe = ppe_view(c);
    for (int f = 0; f < 3; f++) d.xnew[f] = c->xnew[f];
    d.rec = c->rec;
    const bool compact = c->prm.precision == SPH_MIXED && c->pos_f32 && !c->prm.ale_positions && c->prm.compact_records;
    d.crec = compact ? c->crec : nullptr;
    d.crec_adv = compact ? c->crec_adv : nullptr;
    d.ibm_near = c->ibm_near_gen == c->generation ? c->ibm_near : nullptr;
    d.tile = TileView{};
    if (compact && c->tl_gen == c->generation && c->tl_part && c->tl_nchunk > 0)
        d.tile = TileView{c->tl_part, c->tl_chunk, c->tl_halo, c->tl_self, NList{c->tl_slice_off, c->tl_cnt, c->tl_ent}};
    d.sc = e.d_sc + (size_t)k * SCAL_COUNT;
    d.ctl = e.d_ctl + (size_t)k * CTL_COUNT;
}

static int step_body(Eng& e, sph_ctx* const* cases, int nc, const int* count, const double* dt, const double* time) {
    // count == 0 (every step in ALE mode): the neighbour search and the PPE
    // structure of that case
    for (int k = 0; k < nc; k++) {
        sph_ctx* c = cases[k];
        if (c->prm.ale_positions && !c->xnew[0])
            for (int f = 0; f < 3; f++) SPH_TRY(dalloc(c, &c->xnew[f], c->cap));
        if (!c->rec) SPH_TRY(dalloc(c, &c->rec, (size_t)c->cap * REC_DOUBLES));
        if (!c->crec && c->prm.precision == SPH_MIXED) SPH_TRY(dalloc(c, &c->crec, (size_t)c->cap * CREC_FLOATS));
        if (!c->crec_adv && c->prm.precision == SPH_MIXED)
            SPH_TRY(dalloc(c, &c->crec_adv, (size_t)c->cap * CREC_ADV_FLOATS));
        if (count[k] == 0 || c->prm.ale_positions) {   // moving particles: a search every step
            SPH_TRY(set_ncell_run(c));
            SPH_TRY(nnps_run(c));
            SPH_TRY(ppe_build(c));
        } else if (!c->ppe) {
            SPH_TRY(ppe_build(c));
        }
        // the Neumann pass's pairs depend on positions, types and the search
        const int prec = c->prm.precision;
        if (c->nmp_gen != c->generation || c->nmp_for != c->nm_list || c->nmp_prec != prec || c->prm.ale_positions)
            SPH_TRY(neumann_pairs(c));
        SPH_TRY(ibm_near_build(c));   // the spreading targets of this search
        if (prec == SPH_MIXED && c->pos_f32 && !c->prm.ale_positions && c->prm.compact_records) SPH_TRY(tile_build(c));
    }
    // composition of the step: tables and graphs stay valid while it holds
    std::vector<long long> sig;
    sig.reserve(3 * (size_t)nc);
    long long total = 0;
    for (int k = 0; k < nc; k++) {
        sig.push_back((long long)(intptr_t)cases[k]);
        sig.push_back(cases[k]->generation);
        sig.push_back(cases[k]->n);
        total += cases[k]->n;
    }
    if (sig != e.comp) {
        SPH_HIP(&e, hipStreamSynchronize(e.st));   // nothing in flight reads the old tables
        e.comp = sig;
        for (auto& t : e.tabs) t.valid = false;
        drop_graphs(e);
        SPH_TRY(eng_reserve(e, nc));
    }
    e.cs.assign(cases, cases + nc);
    int cls = 0;
    while ((1LL << (cls + 1)) <= total) cls++;
    e.tune_class = cls;
    // the first step of a search writes the gradient correction and the PPE
    // values and is several times the later steps' work: the candidates are
    // timed on the steps after it
    e.tune_skip = false;
    for (int k = 0; k < nc; k++) e.tune_skip |= count[k] == 0;
    // descriptors
    SPH_HIP(&e, hipEventSynchronize(e.up_ev));
    for (int k = 0; k < nc; k++) fill_desc(e, k, cases[k], count[k], dt[k], time[k]);
    SPH_HIP(&e, hipMemcpyAsync(e.d_desc, e.h_desc, sizeof(CaseDesc) * nc, hipMemcpyHostToDevice, e.st));
    SPH_HIP(&e, hipEventRecord(e.up_ev, e.st));
    // cell tiles: every case of the step has them, with one chunk width
    e.tile_lanes = 0;
    for (int k = 0; k < nc; k++) {
        const int w = e.h_desc[k].tile.part ? cases[k]->tl_lanes : 0;
        if (k == 0) e.tile_lanes = w;
        else if (w != e.tile_lanes) e.tile_lanes = 0;
    }
    Shape sh;
    SPH_TRY(shape(e, 0, BLOCK, &sh));
    SPH_TRY(shape(e, 1, BLOCK, &sh));
    if (e.tile_lanes) SPH_TRY(shape(e, 4, e.tile_lanes, &sh));
    SPH_TRY(run_half(e, 0, step_physics_front_key(e), step_physics_front));
    for (int k = 0; k < nc; k++) cases[k]->gc_gen = cases[k]->generation;
    const int prc = ppe_solve(e);
    for (int k = 0; k < nc; k++) cases[k]->vals_gen = cases[k]->generation;
    // a case whose solve did not converge, or met NaN/Inf, still completes
    // the step with the others; its error is reported at the end
    if (prc != 0 && prc != SPH_E_NOCONV && prc != SPH_E_NONFINITE) return prc;
    SPH_TRY(run_half(e, 1, step_physics_back_key(e), step_physics_back));
    return prc;
}

int engine_step(Eng& e, sph_ctx* const* cases, int nc, const int* count, const double* dt, const double* time) {
    if (nc <= 0 || !cases || !count || !dt || !time) return fail(&e, SPH_E_ARG, "empty batch");
    for (int k = 0; k < nc; k++) {
        sph_ctx* c = cases[k];
        if (!c) return fail(&e, SPH_E_ARG, "null case");
        for (int q = 0; q < k; q++)
            if (cases[q] == c) return fail(&e, SPH_E_ARG, "a case appears twice in one step");
        if (!c->uploaded) return fail(&e, SPH_E_STATE, "sph_step before sph_upload");
        if (count[k] < 0) return fail(&e, SPH_E_ARG, "count < 0");
        if (count[k] > 0 && !c->searched)
            return fail(&e, SPH_E_STATE, "the first step after sph_upload must have count == 0");
        if (!(dt[k] > 0)) return fail(&e, SPH_E_ARG, "dt must be > 0");
        if (c->prm.precision != cases[0]->prm.precision || c->prm.ppe_solver != cases[0]->prm.ppe_solver)
            return fail(&e, SPH_E_ARG, "cases of one step must share precision and ppe_solver");
    }
    e.err.clear();
    SPH_TRY(tune_fold(e));
    // hand the cases over to the engine's stream
    std::vector<hipEvent_t> evs;
    int rc = 0;
    for (int k = 0; k < nc && !rc; k++) {
        sph_ctx* c = cases[k];
        c->err.clear();
        if (c->own_stream == e.st) continue;
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(ev, c->own_stream) != hipSuccess || hipStreamWaitEvent(e.st, ev, 0) != hipSuccess)
            rc = fail(&e, SPH_E_HIP, "stream hand-over failed");
        if (ev) evs.push_back(ev);
    }
    for (int k = 0; k < nc; k++) { cases[k]->stream = e.st; cases[k]->profp = &e.prof; }
    if (!rc) rc = step_body(e, cases, nc, count, dt, time);
    for (int k = 0; k < nc; k++) { cases[k]->stream = cases[k]->own_stream; cases[k]->profp = &cases[k]->solo->prof; }
    // the cases' own streams continue after this step
    hipEvent_t done = nullptr;
    bool handed = false;
    for (int k = 0; k < nc; k++) {
        sph_ctx* c = cases[k];
        if (c->own_stream == e.st) continue;
        if (!handed) {
            handed = true;
            if (hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess || hipEventRecord(done, e.st) != hipSuccess) {
                if (!rc) rc = fail(&e, SPH_E_HIP, "stream hand-back failed");
                (void)hipStreamSynchronize(e.st);
                break;
            }
        }
        if (hipStreamWaitEvent(c->own_stream, done, 0) != hipSuccess && !rc) rc = fail(&e, SPH_E_HIP, "stream hand-back failed");
    }
    if (done) (void)hipEventDestroy(done);
    for (hipEvent_t ev : evs) (void)hipEventDestroy(ev);
    if (rc) {
        // every case carries the reason; a case-specific one wins
        std::string first;
        for (int k = 0; k < nc; k++)
            if (!cases[k]->err.empty()) { first = cases[k]->err; break; }
        if (e.err.empty()) e.err = first;
        for (int k = 0; k < nc; k++)
            if (cases[k]->err.empty() && rc != SPH_E_NOCONV && rc != SPH_E_NONFINITE) cases[k]->err = e.err;
    }
    return rc;
}

}  // namespace sph
