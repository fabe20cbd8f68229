// Internal declarations of the MI355X ISPH library: device views, the
// context, launch helpers.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
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
};
// Neighbour list, SELL-64 over particles, ascending j per particle.
struct NList {
    const long long* slice_off;   // first entry of each slice
    const int* cnt;               // entries per particle
    const unsigned* ent;          // j | selection flags (flags.h)
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

struct PPEState;   // ppe.hip

}  // namespace sph

struct sph_ctx {
    sph_params prm{};
    int cap = 0, n = 0;
    bool uploaded = false, searched = false;
    sph_domain dom{};
    double c_pi15 = 0.0;
    hipStream_t stream = nullptr;
    std::string err;
    long long bytes = 0;

    // persistent state
    double* pf[SPH_F_COUNT_F64 + 3] = {};  // P1 fields in SPH_F_* numbering, then eli, m_ref, h_ref
    int* pi[SPH_I_COUNT] = {};
    double* sm[sph::MF_COUNT] = {};      // working copies of mutable fields
    double* p2[10] = {};                 // rho_ref, SR, x0, y0, z0, ux0, uy0, uz0, temp0, concn0
    double* s2[10] = {};
    double* q[18] = {};                  // part3 scratch + normals (last 3 persistent)
    double* snap[5] = {};                // snapshot buffers (ux, uy, uz, pres, temp)

    // upload staging (original order)
    void* aos1 = nullptr;
    void* aos2 = nullptr;
    float* nrm_in = nullptr;

    // neighbour search
    int* keys = nullptr;  int* perm = nullptr;
    int* keys_tmp = nullptr; int* perm_tmp = nullptr;
    int* gstr = nullptr;  int* gend = nullptr;
    int* sort_hist = nullptr; int sort_hist_n = 0;
    long long* nl_slice_off = nullptr; int* nl_cnt = nullptr; unsigned* nl_ent = nullptr;
    long long nl_total = 0;
    int* scan_tmp = nullptr; size_t scan_tmp_n = 0;

    // reductions / stats
    double* red_part = nullptr; size_t red_part_n = 0;
    double* dscal = nullptr;       // device scalars
    double* hscal = nullptr;       // pinned host mirror
    float* stat_buf = nullptr;

    sph::PPEState* ppe = nullptr;
    sph_ppe_report rep{};
    sph::Profile prof;

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
int fail(sph_ctx* c, int code, const std::string& msg);
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

// profiling around launches
int prof_begin(sph_ctx* c, const char* name, double alg_bytes);
int prof_end(sph_ctx* c);
int launch_check(sph_ctx* c, const char* what);

inline unsigned grid1(long long n, int block = BLOCK) { return (unsigned)((n + block - 1) / block); }

// nnps.hip
int nnps_run(sph_ctx* c);
int exclusive_scan_i32(sph_ctx* c, const int* in, int* out, long long n, int* total_out);
int exclusive_scan_i64(sph_ctx* c, const long long* in, long long* out, long long n, long long* total_out);

// physics.hip
int step_physics_front(sph_ctx* c, int count, double dt, double time);
int step_physics_back(sph_ctx* c, int count, double dt, double time);
int stats_run(sph_ctx* c, float* umax, float* drho, float* fmax);

// ppe.hip
int ppe_solve(sph_ctx* c, double dt);
void ppe_free(sph_ctx* c);
// canonical reduction of partials already in c->red_part (count nb) into *dst
int reduce_finalize(sph_ctx* c, const double* part, long long nb, double* dst);

}  // namespace sph
