// Serving loop of the MI355X ISPH library: scenario requests arrive at any
// time, are admitted while their device memory fits an HBM budget, stay
// resident, and every tick advances all admitted cases by one co-batched
// projection step (sph_batch_step).  A request that finishes frees its
// memory at once, so the next queued request joins the following tick:
// continuous batching, not a queue in front of a one-case solver.
//
// Built only on the public C ABI.  In dry-run mode no GPU call is made: a
// modelled throughput advances a simulated clock, which makes the admission
// and batching policy testable and plannable on a host without a GPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/urbansph.h"

namespace {

struct Request {
    long long id = 0;
    sph_params prm{};
    std::vector<sph_part1> p1;
    std::vector<float> nrm;
    int n = 0, steps = 0, count = 0;
    double dt = 0.0;
    long long est = 0, reserved = 0;
    double t_arrival = 0.0, t_admit = -1.0, t_finish = -1.0;
    int state = SPH_REQ_QUEUED;
    int rc = 0;
    int noconv_steps = 0;
    long long ppe_iterations = 0;
    sph_ctx* ctx = nullptr;
    // results, input order
    std::vector<double> f64[4];   // UX, UY, UZ, PRES
    std::vector<sph_part1> out;
};

}  // namespace

struct sph_server {
    long long budget = 0;
    int max_cases = 0;
    bool dry = false;
    double dry_rate = 0.0;           // modelled particle-steps per second
    double sim_clock = 0.0;
    std::chrono::steady_clock::time_point t0;
    int precision = -1, solver = -1;
    long long next_id = 1;
    long long reserved = 0;
    long long steps_done = 0, particle_steps = 0;
    double t_admit = 0.0, t_step = 0.0, t_retire = 0.0;   // host seconds per phase
    sph_batch* batch = nullptr;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::map<long long, Request*> all;
    std::string err;

    double now() const {
        if (dry) return sim_clock;
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

int sfail(sph_server* s, int code, const std::string& m) {
    s->err = m;
    return code;
}

void drop_inputs(Request* r) {
    std::vector<sph_part1>().swap(r->p1);
    std::vector<float>().swap(r->nrm);
}

void retire(sph_server* s, Request* r, int state) {
    if (r->ctx) {
        sph_destroy(r->ctx);
        r->ctx = nullptr;
    }
    s->reserved -= r->reserved;
    r->reserved = 0;
    r->state = state;
    r->t_finish = s->now();
    drop_inputs(r);
}

// Final state of a finished case into the request, in input order.
int collect(sph_server* s, Request* r) {
    const int n = r->n;
    std::vector<int> perm(n);
    int rc = sph_nnps_debug(r->ctx, nullptr, perm.data(), nullptr, nullptr);
    if (rc) return sfail(s, rc, sph_last_error(r->ctx));
    static const int fields[4] = {SPH_F_UX, SPH_F_UY, SPH_F_UZ, SPH_F_PRES};
    std::vector<double> tmp(n);
    for (int f = 0; f < 4; f++) {
        rc = sph_get_field_f64(r->ctx, fields[f], tmp.data(), n);
        if (rc) return sfail(s, rc, sph_last_error(r->ctx));
        r->f64[f].assign(n, 0.0);
        for (int i = 0; i < n; i++) r->f64[f][perm[i]] = tmp[i];
    }
    std::vector<sph_part1> dev(n);
    rc = sph_download(r->ctx, dev.data(), nullptr, nullptr, n);
    if (rc) return sfail(s, rc, sph_last_error(r->ctx));
    r->out.assign(n, sph_part1{});
    for (int i = 0; i < n; i++) r->out[perm[i]] = dev[i];
    return 0;
}

double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// FIFO admission while the head fits the budget and a slot is free.
int admit(sph_server* s) {
    const double t0 = wall();
    while (!s->queue.empty() && (int)s->active.size() < s->max_cases) {
        Request* r = s->queue.front();
        if (s->reserved + r->est > s->budget) break;
        s->queue.pop_front();
        r->t_admit = s->now();
        r->reserved = r->est;
        s->reserved += r->est;
        if (!s->dry) {
            int rc = sph_create(&r->prm, r->n, &r->ctx);
            if (!rc) rc = sph_upload(r->ctx, r->p1.data(), nullptr, r->nrm.empty() ? nullptr : r->nrm.data(), r->n);
            if (rc) {
                r->rc = rc;
                retire(s, r, SPH_REQ_FAILED);
                continue;
            }
            drop_inputs(r);
        }
        r->state = SPH_REQ_RUNNING;
        s->active.push_back(r);
    }
    s->t_admit += wall() - t0;
    return 0;
}

}  // namespace

extern "C" {

long long sph_estimate_bytes(const sph_params* prm, int n) {
    if (!prm || n <= 0) return 0;
    // Resident state: 96 float64 per particle (persistent, working copy,
    // old values, scratch, snapshots), 6 int32, search keys/permutation;
    // neighbour list: <= 160 SELL slots of 4 bytes; PPE: <= 128 slots of
    // column index + value, plus the Krylov vectors.  Upper bounds for the
    // 1.5h / kappa=2 lattices the workloads use; the real figure replaces
    // the estimate after the first step.
    const bool mixed = prm->precision == SPH_MIXED;
    const double val = mixed ? 4.0 : 8.0;
    const double krylov = mixed ? 10 * 4.0 : 8 * 8.0;
    double per = 96 * 8.0 + 6 * 4.0 + 16.0 + 160 * 4.0 + 128 * (4.0 + val) + 6 * 8.0 + krylov;
    // neighbour records of the gathering passes: 128 bytes, plus the compact
    // 32- and 64-byte records in MIXED
    per += 128.0 + (mixed ? 96.0 : 0.0);
    // Neumann pairs of the boundary particles (a few hundred each, a few per
    // cent of the particles): about 32 four-byte pairs per particle, kept
    // twice in MIXED (particle and set indices) with the set's 48-byte records
    per += 32 * 4.0 + (mixed ? 32 * 4.0 + 8.0 : 0.0);
    // cell tiles, when on: a copy of the list, lanes, halos
    if (mixed && prm->tile_cells != 0) per += 160 * 4.0 + 3 * 4.0 + 16 * 4.0;
    long long fixed = 16LL << 20;
    if (prm->ppe_precond == SPH_PRECOND_MULTIGRID && prm->ppe_solver == SPH_PPE_BICGSTAB) {
        // the level-0 compact copy (4 bytes a slot), the coarse levels (about
        // a seventh of the fine entries at 8 bytes), the Galerkin terms of the
        // set-up (8 bytes a fine entry, kept by the block cache), aggregation
        // keys and links; FP64 adds the float copy of its operator and four
        // float vectors; the dense coarsest inverse and its double build
        per += 128 * 4.0 + 160.0 + 128 * 8.0 + 64.0 + (mixed ? 0.0 : 128 * 4.0 + 4 * 4.0 + 4.0);
        fixed += 12LL << 20;
    }
    return (long long)(per * n) + fixed;
}

int sph_server_create(long long hbm_budget_bytes, int max_cases, double dry_run_rate, sph_server** out) {
    if (!out) return SPH_E_ARG;
    *out = nullptr;
    if (hbm_budget_bytes <= 0 || max_cases <= 0 || dry_run_rate < 0) return SPH_E_ARG;
    sph_server* s = new sph_server();
    s->budget = hbm_budget_bytes;
    s->max_cases = max_cases;
    s->dry = dry_run_rate > 0;
    s->dry_rate = dry_run_rate;
    s->t0 = std::chrono::steady_clock::now();
    if (!s->dry) {
        int rc = sph_batch_create(&s->batch);
        if (rc) { delete s; return rc; }
    }
    *out = s;
    return 0;
}

int sph_server_submit(sph_server* s, const sph_params* prm, const sph_part1* P1, const float* ib_normals, int n,
                      int steps, double dt, double t_arrival, long long* id_out) {
    if (!s || !prm || !P1 || !id_out || n <= 0 || steps <= 0 || !(dt > 0)) return s ? sfail(s, SPH_E_ARG, "bad request") : SPH_E_ARG;
    if (s->precision < 0) { s->precision = prm->precision; s->solver = prm->ppe_solver; }
    if (prm->precision != s->precision || prm->ppe_solver != s->solver)
        return sfail(s, SPH_E_ARG, "requests of one server share precision and ppe_solver");
    Request* r = new Request();
    r->id = s->next_id++;
    r->prm = *prm;
    r->n = n;
    r->steps = steps;
    r->dt = dt;
    r->est = sph_estimate_bytes(prm, n);
    r->t_arrival = t_arrival >= 0 ? t_arrival : s->now();
    if (!s->dry) {
        r->p1.assign(P1, P1 + n);
        if (ib_normals) r->nrm.assign(ib_normals, ib_normals + 3 * (size_t)n);
    }
    s->all[r->id] = r;
    *id_out = r->id;
    if (r->est > s->budget) {
        r->rc = SPH_E_NOMEM;
        r->state = SPH_REQ_FAILED;
        r->t_finish = s->now();
        drop_inputs(r);
        return 0;
    }
    s->queue.push_back(r);
    return 0;
}

int sph_server_tick(sph_server* s, long long* finished, int max_finished, int* n_finished) {
    if (!s) return SPH_E_ARG;
    if (n_finished) *n_finished = 0;
    admit(s);
    if (s->active.empty()) return 0;
    const int k = (int)s->active.size();
    long long ps = 0;
    for (Request* r : s->active) ps += r->n;
    const double t_step0 = wall();
    if (s->dry) {
        s->sim_clock += (double)ps / s->dry_rate;
    } else {
        std::vector<sph_ctx*> cs(k);
        std::vector<int> cnt(k);
        std::vector<double> dts(k), tms(k);
        for (int q = 0; q < k; q++) {
            Request* r = s->active[q];
            cs[q] = r->ctx;
            cnt[q] = r->count;
            dts[q] = r->dt;
            tms[q] = r->count * r->dt;
        }
        int rc = sph_batch_step(s->batch, cs.data(), k, cnt.data(), dts.data(), tms.data());
        if (rc && rc != SPH_E_NOCONV && rc != SPH_E_NONFINITE) {
            std::string why = sph_batch_last_error(s->batch);
            for (Request* r : s->active) { r->rc = rc; retire(s, r, SPH_REQ_FAILED); }
            s->active.clear();
            return sfail(s, rc, why);
        }
        for (Request* r : s->active) {
            sph_ppe_report rep{};
            sph_ppe_info(r->ctx, &rep);
            r->ppe_iterations += rep.iterations;
            if (!rep.converged) r->noconv_steps++;
            // a case whose state went NaN/Inf ends now, failed; the others go on
            if (!rep.converged && !std::isfinite(rep.rel_residual)) r->rc = SPH_E_NONFINITE;
            // the real footprint replaces the estimate once the first step has
            // built the neighbour list and the PPE
            if (r->count == 0) {
                long long b = sph_device_bytes(r->ctx);
                s->reserved += b - r->reserved;
                r->reserved = b;
            }
        }
    }
    s->steps_done++;
    s->particle_steps += ps;
    const double t_ret0 = wall();
    s->t_step += t_ret0 - t_step0;
    int nf = 0;
    std::vector<Request*> keep;
    for (Request* r : s->active) {
        r->count++;
        if (r->count < r->steps && r->rc == 0) { keep.push_back(r); continue; }
        int state = r->rc ? SPH_REQ_FAILED : SPH_REQ_DONE;
        if (!s->dry && state == SPH_REQ_DONE) {
            int rc = sph_synchronize(r->ctx);
            if (!rc) rc = collect(s, r);
            if (rc) { r->rc = rc; state = SPH_REQ_FAILED; }
        }
        retire(s, r, state);
        if (finished && nf < max_finished) finished[nf] = r->id;
        nf++;
    }
    s->active.swap(keep);
    s->t_retire += wall() - t_ret0;
    if (n_finished) *n_finished = nf;
    return 0;
}

int sph_server_idle_until(sph_server* s, double t) {
    if (!s) return SPH_E_ARG;
    if (s->dry && t > s->sim_clock) s->sim_clock = t;
    return 0;
}

double sph_server_clock(const sph_server* s) { return s ? s->now() : 0.0; }

int sph_server_status(const sph_server* s, sph_server_info* info) {
    if (!s || !info) return SPH_E_ARG;
    std::memset(info, 0, sizeof(*info));
    info->queued = (int)s->queue.size();
    info->running = (int)s->active.size();
    for (auto& kv : s->all)
        if (kv.second->state == SPH_REQ_DONE || kv.second->state == SPH_REQ_FAILED) info->finished++;
    info->reserved_bytes = s->reserved;
    info->budget_bytes = s->budget;
    info->ticks = s->steps_done;
    info->particle_steps = s->particle_steps;
    info->clock = s->now();
    info->admit_s = s->t_admit;
    info->step_s = s->t_step;
    info->retire_s = s->t_retire;
    return 0;
}

int sph_server_request(const sph_server* s, long long id, sph_request_info* info) {
    if (!s || !info) return SPH_E_ARG;
    auto it = s->all.find(id);
    if (it == s->all.end()) return SPH_E_ARG;
    const Request* r = it->second;
    std::memset(info, 0, sizeof(*info));
    info->state = r->state;
    info->rc = r->rc;
    info->n = r->n;
    info->steps = r->steps;
    info->steps_done = r->count;
    info->noconv_steps = r->noconv_steps;
    info->ppe_iterations = r->ppe_iterations;
    info->est_bytes = r->est;
    info->t_arrival = r->t_arrival;
    info->t_admit = r->t_admit;
    info->t_finish = r->t_finish;
    return 0;
}

int sph_server_result(const sph_server* s, long long id, int field, double* out, int n) {
    if (!s || !out) return SPH_E_ARG;
    auto it = s->all.find(id);
    if (it == s->all.end()) return SPH_E_ARG;
    const Request* r = it->second;
    if (r->state != SPH_REQ_DONE || n != r->n || r->f64[0].empty()) return SPH_E_STATE;
    int slot = field == SPH_F_UX ? 0 : field == SPH_F_UY ? 1 : field == SPH_F_UZ ? 2 : field == SPH_F_PRES ? 3 : -1;
    if (slot < 0) return SPH_E_ARG;
    std::memcpy(out, r->f64[slot].data(), sizeof(double) * (size_t)n);
    return 0;
}

int sph_server_result_part1(const sph_server* s, long long id, sph_part1* out, int n) {
    if (!s || !out) return SPH_E_ARG;
    auto it = s->all.find(id);
    if (it == s->all.end()) return SPH_E_ARG;
    const Request* r = it->second;
    if (r->state != SPH_REQ_DONE || n != r->n || r->out.empty()) return SPH_E_STATE;
    std::memcpy(out, r->out.data(), sizeof(sph_part1) * (size_t)n);
    return 0;
}

int sph_server_release(sph_server* s, long long id) {
    if (!s) return SPH_E_ARG;
    auto it = s->all.find(id);
    if (it == s->all.end()) return SPH_E_ARG;
    Request* r = it->second;
    if (r->state != SPH_REQ_DONE && r->state != SPH_REQ_FAILED) return SPH_E_STATE;
    delete r;
    s->all.erase(it);
    return 0;
}

const char* sph_server_last_error(const sph_server* s) { return s ? s->err.c_str() : "null server"; }

void sph_server_destroy(sph_server* s) {
    if (!s) return;
    for (Request* r : s->active)
        if (r->ctx) sph_destroy(r->ctx);
    for (auto& kv : s->all) delete kv.second;
    if (s->batch) sph_batch_destroy(s->batch);
    delete s;
}

}  // extern "C"
