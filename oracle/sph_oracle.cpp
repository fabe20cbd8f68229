/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in double precision, of the reference's ISPH projection
 * step SOPHIA_single_ISPH (ISPH_Calc.cuh:4-233) and of every kernel it
 * launches.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only as the checker.  The HIP library never
 * links, loads or calls it.
 *
 * Each function names the reference lines it restates.  Expressions keep the
 * reference's operand order so that the HIP FP64 path, which restates the
 * same lines, produces the same IEEE results (both are built with
 * -ffp-contract=off).  The arithmetic deviations from the reference, shared
 * with the HIP path, are:
 *   - all physics in double (the reference is float with silent double
 *     promotions, Variable_Type.cuh:2);
 *   - geometry bookkeeping stays float exactly as the reference computes it:
 *     domain bounds, dcell, NI/NJ/NK, both cell-index formulas, set_ncell;
 *   - integer powers pow(x,k) are repeated multiplications;
 *   - races are removed: PPE Jacobi reads the previous iterate
 *     (functions_PPE.cuh:93 vs :166); Neumann, periodic and outlet copies
 *     read a snapshot taken before the kernel; MOST neighbour deposits are
 *     serialised in ascending particle order (functions_MOST.cuh:298-299);
 *   - IB normals are inputs that persist (the reference computes them at
 *     step 0 for one hard-coded building and wipes them every step,
 *     functions_PREP.cuh:958-963, ISPH_Calc.cuh:61).
 *
 * What pins this restatement (tests/test_oracle_*.py): closed-form kernel
 * normalisation and derivatives; brute-force O(n^2) neighbour sets;
 * std::stable_sort against a numpy stable argsort; gradient correction
 * = identity on a lattice and exact for linear fields; PPE operator row
 * sums, b = 0 for uniform u*, Laplacian of x^2; BiCGSTAB against scipy's
 * sparse LU; hand-built predictor and corrector cases; committed golden
 * fixtures; APR split geometry and mass conservation; ALE trapezoid
 * positions.  Each remaining pass is run on its own (oracle_run_pass) and
 * pinned in tests/test_oracle_pins.py: pressure force and the Cleary and
 * Euler terms by lattice moments of the kernel for linear and quadratic
 * fields; advection (Cleary, Monaghan, artificial viscosity, Euler),
 * pressure force, IBM interpolation and spreading, the periodic and outlet
 * copies and the Neumann pass with its inlet ghost by all-pairs numpy sums
 * under the window rule; the MOST Obukhov length by brentq on the
 * Businger-Dyer relation; the LES Shepard sum, kernel length, Smagorinsky
 * viscosity and the Deardorff k-equation by lattice closed forms.  The
 * planted mistakes below (g_mutant) show that each of those pins fails on a
 * one-line slip.
 *
 * Only the IBM target-velocity table values (functions_IBM.cuh:289-306) are
 * data with nothing to check them against beyond their transcription; the
 * interpolation over them is pinned.
 */
#include "../include/urbansph.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double PI_REF = 3.141592653;          // Parameters.cuh:41 (truncated)
const double K_VONKARMAN = 0.4;             // Parameters.cuh:113
const double H_COEFF = 1.5;                 // Parameters.cuh:90
const double MOST_TOL = 1e-5;               // Parameters.cuh:112
const int MOST_MAXIT = 200;                 // Parameters.cuh:111
const double BOUNDARY_VELOCITY = 1.0;       // Parameters.cuh:98
const int MOVING = 9;                       // Parameters.cuh:17

// Planted mistakes for the mutation tests (tests/test_oracle_pins.py): 0 =
// none.  Each other value changes one coefficient, operand or selection rule
// of one pass, the kind of slip a transcription makes, so that the tests can
// show that the pin of that pass detects it.  Set only by
// oracle_set_mutant(); every product run has 0.
int g_mutant = 0;
inline bool MUT(int k) { return g_mutant == k; }

struct P1d {
    int i_type, buffer_type, MOST_buffer, p_type;
    double x, y, z, ux, uy, uz, m, h, temp, pres, pres0, rho, flt_s, enthalpy,
        concn, concentration, k_turb, e_turb, eli, vol, m_ref, h_ref, vis_t,
        fbx, fby, fbz, fmx, fmy, hmz, float1, float2, float3;
    int ncell, source;
};
struct P2d {
    double rho_ref, SR, x0, y0, z0, ux0, uy0, uz0, rho0, drho0, temp0, concn0,
        enthalpy0;
};
struct P3d {
    double drho, dconcn, denthalpy, ftotalx, ftotaly, ftotalz, ftotal, fpx,
        fpy, fpz, fbody, dtemp, dk_turb, de_trub, cc;
    double inv_cm_xx, inv_cm_yy, inv_cm_zz, inv_cm_xy, inv_cm_yz, inv_cm_zx;
};

struct Oracle {
    sph_params prm;
    int cap = 0, n = 0;
    sph_domain d{};
    bool uploaded = false, searched = false;
    double c_pi15 = 0.0;
    std::vector<P1d> P1, SP1;
    std::vector<P2d> P2, SP2;
    std::vector<P3d> P3;
    std::vector<double> nx, ny, nz;            // persistent IB normals
    std::vector<int> keys, perm, gstr, gend;
    std::vector<double> apr_target;            // APR target level (the reference's M_num)
    std::vector<double> xn, yn, zn;            // ALE: positions of the corrector, committed at step end
    // last PPE system (BICGSTAB), rows in ascending particle order
    std::vector<int> sys_rows, sys_rowptr, sys_cols;
    std::vector<double> sys_diag, sys_rhs, sys_vals, sys_x0, sys_x;
    sph_ppe_report rep{};
};

// ------------------------------------------------------------ helpers ----
double canon_sum(const double* v, long n) {
    // Canonical reduction shared with the HIP path: chunks of 256 values,
    // pairwise tree with offsets 128..1 inside a chunk, chunk partials
    // reduced again the same way.
    if (n <= 0) return 0.0;
    long nb = (n + 255) / 256;
    std::vector<double> part(nb);
    for (long c = 0; c < nb; c++) {
        double t[256];
        for (int th = 0; th < 256; th++) {
            long idx = c * 256 + th;
            t[th] = idx < n ? v[idx] : 0.0;
        }
        for (int off = 128; off > 0; off >>= 1)
            for (int th = 0; th < off; th++) t[th] += t[th + off];
        part[c] = t[0];
    }
    if (nb == 1) return part[0];
    return canon_sum(part.data(), nb);
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    std::vector<double> p(a.size());
    for (size_t i = 0; i < a.size(); i++) p[i] = a[i] * b[i];
    return canon_sum(p.data(), (long)p.size());
}

inline double p4(double a) { return a * a * a * a; }
inline double p5(double a) { return a * a * a * a * a; }

// calc_tmpA, functions_KNL.cuh:2-26 (3-D branch)
double calc_tmpA(int type, double kappa, double c_pi15, double tH) {
    switch (type) {
    case SPH_GAUSSIAN: return 1.0 / (c_pi15 * (tH * tH * tH));
    case SPH_QUINTIC: return 3.0 / (359.0 * PI_REF * (tH * tH * tH));
    case SPH_QUARTIC: return 315.0 / (208.0 * PI_REF * (tH * tH * tH));
    case SPH_WENDLAND2: return 3.342253804929802 / (kappa * tH) / (kappa * tH) / (kappa * tH);
    case SPH_WENDLAND4: return 4.923856051905513 / (kappa * tH) / (kappa * tH) / (kappa * tH);
    case SPH_WENDLAND6: return 6.788953041263660 / (kappa * tH) / (kappa * tH) / (kappa * tH);
    }
    return 0.0;
}

// calc_kernel_wij, functions_KNL.cuh:41-71 (3-D branch)
double kernel_w(int type, double kappa, double tA, double tH, double rr) {
    double tR = rr / tH / kappa;
    switch (type) {
    case SPH_GAUSSIAN: { double a = 2.0 * tR; return tA * exp(-(a * a)); }
    case SPH_QUINTIC: {
        double a = 3.0 - 2.0 * tR, b = 2.0 - 2.0 * tR, c = 1.0 - 2.0 * tR;
        if (tR < 0.5) return tA * (p5(a) - 6.0 * p5(b) + 15.0 * p5(c));
        else if (0.5 <= tR && tR < 1) return tA * (p5(a) - 6.0 * p5(b));
        else if (1 <= tR && tR < 1.5) return tA * p5(a);
        return 0.0;
    }
    case SPH_QUARTIC:
        return (tR < 1) * tA * (2.0 / 3.0 - 9.0 / 2.0 * (tR * tR) + 19.0 / 3.0 * (tR * tR * tR) - 5.0 / 2.0 * (tR * tR * tR * tR));
    case SPH_WENDLAND2:
        return (tR < 1) * tA * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 + 4 * tR);
    case SPH_WENDLAND4:
        return (tR < 1) * tA * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 + 6 * tR + 11.666666666666666 * tR * tR);
    case SPH_WENDLAND6:
        return (tR < 1) * tA * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 + 8 * tR + 25 * tR * tR + 32 * tR * tR * tR);
    }
    return 0.0;
}

// calc_kernel_dwij, functions_KNL.cuh:73-103 (3-D branch)
double kernel_dw(int type, double kappa, double tA, double tH, double rr) {
    double tR = rr / tH / kappa;
    switch (type) {
    case SPH_GAUSSIAN: { double a = 2.0 * tR; return (1.0 / tH) * tA * (-2.0) * 2.0 * tR * exp(-(a * a)); }
    case SPH_QUINTIC: {
        double a = 3.0 - 2.0 * tR, b = 2.0 - 2.0 * tR, c = 1.0 - 2.0 * tR;
        if (tR < 0.5) return (1.0 / tH) * tA * (-5.0 * p4(a) + 30.0 * p4(b) - 75.0 * p4(c));
        else if (0.5 <= tR && tR < 1.0) return (1.0 / tH) * tA * (-5.0 * p4(a) + 30.0 * p4(b));
        else if (1.0 <= tR && tR < 1.5) return (1.0 / tH) * tA * (-5.0 * p4(a));
        return 0.0;
    }
    case SPH_QUARTIC:
        return (tR < 1) * (1.0 / tH) * tA * (-9.0 / 2.0 * 2 * tR + 19.0 / 3.0 * 3.0 * (tR * tR) - 5.0 / 2.0 * 4.0 * (tR * tR * tR));
    case SPH_WENDLAND2:
        return (tR < 1) * (1 / (kappa * tH)) * tA * (-20 * tR * (1 - tR) * (1 - tR) * (1 - tR));
    case SPH_WENDLAND4:
        return (tR < 1) * (1 / (kappa * tH)) * tA * (-18.666666666666668 * tR * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (5 * tR + 1));
    case SPH_WENDLAND6:
        return (tR < 1) * (1 / (kappa * tH)) * tA * (-22 * tR * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (16 * tR * tR + 7 * tR + 1));
    }
    return 0.0;
}

// calc_kernel_wij_pressure / _dwij_pressure, functions_KNL.cuh:340-369:
// Wendland-6 shape with the configured kernel's normalisation tA.
double kernel_wp(double kappa, double tA, double tH, double rr) {
    double tR = rr / tH / kappa;
    return (tR < 1) * tA * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 + 8 * tR + 25 * tR * tR + 32 * tR * tR * tR);
}
double kernel_dwp(double kappa, double tA, double tH, double rr) {
    double tR = rr / tH / kappa;
    return (tR < 1) * (1 / (kappa * tH)) * tA * (-22 * tR * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (1 - tR) * (16 * tR * tR + 7 * tR + 1));
}

// viscosity2, functions_PROP.cuh:121-131 (constants made runtime)
inline double viscosity2(const sph_params& p, int p_type) {
    if (p_type == 1) return p.mu_fluid;
    if (p_type == 0) return p.mu_wall;
    if (p_type == 1000) return p.mu_wall;
    return p.mu_fluid;
}

// Binning formula, functions_NNPS.cuh:891-900 (float arithmetic).
inline int bin_axis(float x, float mn, float mx, int N) {
    if (mx == mn) return 0;
    float t = floorf((x - mn) / (mx - mn) * (float)N);
    float r = (t < (float)(N - 1)) ? t : (float)(N - 1);
    int c = (int)r;
    return c < 0 ? 0 : c;
}
// Query formula, e.g. functions_PPE.cuh:52-57 (float arithmetic).
inline int query_axis(float x, float mn, float mx, float dcell, int N) {
    if (mx == mn) return 0;
    float t = floorf((x - mn) / dcell);
    float r = (t < (float)(N - 1)) ? t : (float)(N - 1);
    int c = (int)r;
    return c < 0 ? 0 : c;
}

struct Cell { int i, j, k; };
inline Cell bin_cell(const sph_domain& d, double x, double y, double z) {
    return {bin_axis((float)x, d.x_min, d.x_max, d.NI), bin_axis((float)y, d.y_min, d.y_max, d.NJ),
            bin_axis((float)z, d.z_min, d.z_max, d.NK)};
}
inline Cell query_cell(const sph_domain& d, double x, double y, double z) {
    return {query_axis((float)x, d.x_min, d.x_max, d.dcell, d.NI), query_axis((float)y, d.y_min, d.y_max, d.dcell, d.NJ),
            query_axis((float)z, d.z_min, d.z_max, d.dcell, d.NK)};
}

// Window scan in the reference's loop order (z, y, x; then j ascending).
template <class F>
void window(const Oracle& o, Cell c, int nc, F&& f) {
    const sph_domain& d = o.d;
    for (int z = -nc; z <= nc; z++)
        for (int y = -nc; y <= nc; y++)
            for (int x = -nc; x <= nc; x++) {
                if ((c.i + x) < 0 || (c.i + x) > (d.NI - 1) || (c.j + y) < 0 || (c.j + y) > (d.NJ - 1) ||
                    (c.k + z) < 0 || (c.k + z) > (d.NK - 1))
                    continue;
                int k = (c.i + x) + d.NI * (c.j + y) + d.NI * d.NJ * (c.k + z);
                if (o.gstr[k] != -1) {
                    int fend = o.gend[k];
                    for (int j = o.gstr[k]; j < fend; j++) f(j);
                }
            }
}

inline double dist3(double xi, double yi, double zi, double xj, double yj, double zj) {
    return sqrt((xi - xj) * (xi - xj) + (yi - yj) * (yi - yj) + (zi - zj) * (zi - zj));
}

// Device_Interpolate_ux_caseH, functions_IBM.cuh:288-338
const double CASEH_Z[30] = {0.0050000, 0.0100000, 0.0200000, 0.0400000, 0.0600000, 0.0800000,
                            0.1000000, 0.1200000, 0.1400000, 0.1600000, 0.1650000, 0.1700000,
                            0.1800000, 0.1900000, 0.2100000, 0.2300000, 0.2500000, 0.3000000,
                            0.3500000, 0.4000000, 0.4500000, 0.5000000, 0.5500000, 0.6000000,
                            0.6528736, 0.7724136, 0.8390808, 0.8689656, 0.8804600, 0.8919544};
const double CASEH_U[30] = {2.745, 2.935, 3.175, 3.435, 3.627, 3.824, 4.021, 4.210, 4.362, 4.491,
                            4.502, 4.586, 4.606, 4.712, 4.854, 4.993, 5.132, 5.449, 5.782, 6.077,
                            6.338, 6.588, 6.693, 6.751, 6.751, 6.751, 6.751, 6.751, 6.751, 6.751};
double interp_ux_caseH(double z_inp, double scale) {
    const int N = 30;
    const double z = scale * z_inp;
    if (z <= CASEH_Z[0]) {
        const double dz = CASEH_Z[1] - CASEH_Z[0];
        const double du = CASEH_U[1] - CASEH_U[0];
        return CASEH_U[0] + (z - CASEH_Z[0]) * (du / (dz + 1e-300));
    }
    if (z >= CASEH_Z[N - 1]) {
        const double dz = CASEH_Z[N - 1] - CASEH_Z[N - 2];
        const double du = CASEH_U[N - 1] - CASEH_U[N - 2];
        return CASEH_U[N - 2] + (z - CASEH_Z[N - 2]) * (du / (dz + 1e-300));
    }
    int lo = 0, hi = N - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (z < CASEH_Z[mid]) hi = mid; else lo = mid;
    }
    const double z0 = CASEH_Z[lo], z1 = CASEH_Z[hi];
    const double u0 = CASEH_U[lo], u1 = CASEH_U[hi];
    const double t = (z - z0) / (z1 - z0 + 1e-300);
    return u0 + t * (u1 - u0);
}

// psi / Phi, functions_MOST.cuh:151-173, evaluated in double.
double psi_m(double zeta) {
    if (zeta < 0.0) {
        double x = sqrt(sqrt(1.0 - 16.0 * zeta));
        return M_PI * 0.5 - 2.0 * atan(x) + log((1.0 + x) * (1.0 + x) * (1.0 + x * x) * 0.125);
    }
    return -5.0 * zeta;
}
double psi_h(double zeta) {
    if (zeta < 0.0) return 2.0 * log(0.5 * (1.0 + sqrt(1.0 - (MUT(25) ? 9.0 : 16.0) * zeta)));
    return -5.0 * zeta;
}
double Phi_m(double L, double z, double z0) { return log(z / z0) - psi_m(z / L) + psi_m(z0 / L); }
double Phi_h(double L, double z, double z0h) { return log(z / z0h) - psi_h(z / L) + psi_h(z0h / L); }

inline double wall_temperature(const sph_params& p, double dt, int count) {
    // functions_TIME_ISPH.cuh:287-291 and functions_MOST.cuh:199-202
    double t = dt * count;
    if (t >= p.T_start_t && t < p.T_con_t) return p.T_ini + dt * count * (p.T_con - p.T_ini) / (p.T_con_t - p.T_start_t);
    else if (t < p.T_start_t) return p.T_ini;
    return p.T_con;
}

// ------------------------------------------------------------ kernels ----

// find_minmax (function_init.cuh:678-708, skips the last particle) and cell
// sizing (ISPH.cuh:75-101, functions_NNPS.cuh:534-546), float arithmetic.
void size_domain(Oracle& o) {
    const auto& P = o.P1;
    int n = o.n;
    float min_x = (float)P[0].x, max_x = (float)P[0].x;
    float min_y = (float)P[0].y, max_y = (float)P[0].y;
    float min_z = (float)P[0].z, max_z = (float)P[0].z;
    for (int i = 0; i < n - 1; i++) {
        float tx = (float)P[i].x, ty = (float)P[i].y, tz = (float)P[i].z;
        if (tx < min_x) min_x = tx;
        if (tx > max_x) max_x = tx;
        if (ty < min_y) min_y = ty;
        if (ty > max_y) max_y = ty;
        if (tz < min_z) min_z = tz;
        if (tz > max_z) max_z = tz;
    }
    sph_domain& d = o.d;
    d.x_min = min_x; d.x_max = max_x; d.y_min = min_y; d.y_max = max_y; d.z_min = min_z; d.z_max = max_z;
    float h0 = (float)P[0].h_ref;
    float crf = 1.1f;
    d.dcell = crf * (float)o.prm.kappa * h0 / 2.0f;
    d.NI = (int)((d.x_max - d.x_min) / d.dcell) + 1;
    d.NJ = (int)((d.y_max - d.y_min) / d.dcell) + 1;
    d.NK = (int)((d.z_max - d.z_min) / d.dcell) + 1;
    d.num_cells = d.NI * d.NJ * d.NK + 1;
}

// KERNEL_set_ncell, functions_NNPS.cuh:943-955
void set_ncell(Oracle& o) {
    for (int i = 0; i < o.n; i++) {
        P1d& p = o.P1[i];
        if (p.i_type == 3) continue;
        p.ncell = (int)(1.0 * ceilf(2.0f / (float)(p.h_ref / p.h)));
    }
}

// Neighbour search of step 0: KERNEL_index_particle_to_cell
// (functions_NNPS.cuh:879-907), SortPairs (ISPH_Calc.cuh:33, stable),
// KERNEL_reorder (functions_NNPS.cuh:813-844), P1 <- SP1 (ISPH_Calc.cuh:45).
void nnps(Oracle& o) {
    int n = o.n;
    const sph_domain& d = o.d;
    std::vector<int> key(n);
    for (int i = 0; i < n; i++) {
        const P1d& p = o.P1[i];
        if (p.i_type == 3) key[i] = d.num_cells;
        else {
            Cell c = bin_cell(d, p.x, p.y, p.z);
            key[i] = c.i + d.NI * c.j + d.NI * d.NJ * c.k;
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
    o.keys.resize(n); o.perm = idx;
    for (int i = 0; i < n; i++) o.keys[i] = key[idx[i]];
    o.gstr.assign(d.num_cells + 1, -1);
    o.gend.assign(d.num_cells + 1, 0);
    for (int i = 0; i < n; i++) {
        int h = o.keys[i];
        if (i == 0 || h != o.keys[i - 1]) {
            o.gstr[h] = i;
            if (i > 0) o.gend[o.keys[i - 1]] = i;
        }
        if (i == n - 1) o.gend[h] = i + 1;
    }
    std::vector<P1d> s1(n); std::vector<P2d> s2(n);
    std::vector<double> tnx(n), tny(n), tnz(n);
    for (int i = 0; i < n; i++) {
        s1[i] = o.P1[idx[i]]; s2[i] = o.P2[idx[i]];
        tnx[i] = o.nx[idx[i]]; tny[i] = o.ny[idx[i]]; tnz[i] = o.nz[idx[i]];
    }
    o.SP1 = s1; o.SP2 = s2; o.P1 = s1; o.P2 = s2;
    o.nx = tnx; o.ny = tny; o.nz = tnz;
    if ((int)o.apr_target.size() == n) {
        std::vector<double> t(n);
        for (int i = 0; i < n; i++) t[i] = o.apr_target[idx[i]];
        o.apr_target = t;
    }
    o.searched = true;
}

// KERNEL_clc_gradient_correction_3D, functions_KNL.cuh:216-324
// KERNEL_clc_IB_normal_vector3D, functions_PREP.cuh:923-1015, run at count ==
// 0 (ISPH_Calc.cuh:70-74): a marker (p_type >= 1000) within ib_face_tol of a
// face plane gets -1 / +1 on that axis (the lower plane first, so the upper
// one wins a tie), and the vector is normalised in float.  At count > 0 the
// reference runs with the zero normals its part3 reset leaves
// (ISPH_Calc.cuh:61).  Modes: SPH_IBN_REFERENCE as the reference,
// SPH_IBN_PLANES keeps the count == 0 normals, SPH_IBN_UPLOADED keeps the
// caller's.
void ib_normals(Oracle& o, int count) {
    const sph_params& pr = o.prm;
    if (pr.ib_normals == SPH_IBN_UPLOADED) return;
    if (count > 0 && pr.ib_normals == SPH_IBN_PLANES) return;
    const auto& S = o.SP1;
    for (int i = 0; i < o.n; i++) {
        if (S[i].p_type < 1000) continue;
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        if (count == 0) {
            const double xi = S[i].x, yi = S[i].y, zi = S[i].z, tol = pr.ib_face_tol;
            if (fabs(xi - pr.ib_face_xlo) < tol) nx = -1.0f;
            if (fabs(xi - pr.ib_face_xhi) < tol) nx = MUT(34) ? -1.0f : 1.0f;
            if (fabs(yi - pr.ib_face_ylo) < tol) ny = -1.0f;
            if (fabs(yi - pr.ib_face_yhi) < tol) ny = 1.0f;
            if (fabs(zi - pr.ib_face_zlo) < tol) nz = -1.0f;
            if (fabs(zi - pr.ib_face_zhi) < tol) nz = 1.0f;
        }
        const float mag = MUT(35) ? std::fabs(nx) + std::fabs(ny) + std::fabs(nz) : sqrtf(nx * nx + ny * ny + nz * nz);
        if ((double)mag > 1e-9) {
            o.nx[i] = (double)(nx / mag); o.ny[i] = (double)(ny / mag); o.nz[i] = (double)(nz / mag);
        } else {
            o.nx[i] = 0.0; o.ny[i] = 0.0; o.nz[i] = 0.0;
        }
    }
}

void gradient_correction(Oracle& o) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        int ptype_i = S[i].p_type;
        double hi = S[i].h;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hi);
        double search_range = pr.kappa * hi;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z;
        Cell c = bin_cell(o.d, xi, yi, zi);
        double tmpxx = 0, tmpyy = 0, tmpzz = 0, tmpxy = 0, tmpyz = 0, tmpzx = 0;
        window(o, c, S[i].ncell, [&](int j) {
            double xj = S[j].x, yj = S[j].y, zj = S[j].z;
            int ptype_j = S[j].p_type;
            double tdist = dist3(xi, yi, zi, xj, yj, zj);
            if (tdist > 0 && tdist < search_range) {
                double tdwij = kernel_dw(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                double mj = S[j].m, rhoj = S[j].rho;
                double mtd = mj * tdwij;
                double rtd = 1.0 / (rhoj * tdist);
                double txx = -mtd * (xi - xj) * (xi - xj); txx *= rtd;
                double txy = -mtd * (yi - yj) * (xi - xj); txy *= rtd;
                double tyy = -mtd * (yi - yj) * (yi - yj); tyy *= rtd;
                double tzx = -mtd * (xi - xj) * (zi - zj); tzx *= rtd;
                double tyz = -mtd * (yi - yj) * (zi - zj); tyz *= rtd;
                double tzz = -mtd * (zi - zj) * (zi - zj); tzz *= rtd;
                double f = (double)(ptype_j == 1);
                tmpxx += txx * f; tmpxy += txy * f; tmpyy += tyy * f;
                tmpzx += tzx * f; tmpyz += tyz * f; tmpzz += tzz * f;
            }
        });
        double tmpcmd = tmpxx * (tmpyy * tmpzz - tmpyz * tmpyz);
        tmpcmd -= tmpxy * (tmpxy * tmpzz - tmpyz * tmpzx);
        tmpcmd += tmpzx * (tmpxy * tmpyz - tmpyy * tmpzx);
        P3d& q = o.P3[i];
        if ((fabs(tmpcmd) > DBL_EPSILON) && (ptype_i == 1)) {
            double rtcmd = 1.0 / tmpcmd;
            q.inv_cm_xx = (tmpyy * tmpzz - tmpyz * tmpyz) * rtcmd;
            q.inv_cm_xy = (tmpzx * tmpyz - tmpxy * tmpzz) * rtcmd;
            q.inv_cm_zx = (tmpxy * tmpyz - tmpzx * tmpyy) * rtcmd;
            q.inv_cm_yy = (tmpxx * tmpzz - tmpzx * tmpzx) * rtcmd;
            q.inv_cm_yz = (tmpzx * tmpxy - tmpxx * tmpyz) * rtcmd;
            q.inv_cm_zz = (tmpxx * tmpyy - tmpxy * tmpxy) * rtcmd;
        } else {
            q.inv_cm_xx = 1.0; q.inv_cm_xy = 0.0; q.inv_cm_zx = 0.0;
            q.inv_cm_yy = 1.0; q.inv_cm_yz = 0.0; q.inv_cm_zz = 1.0;
        }
    }
}

// KERNEL_clc_prep3D, functions_PREP.cuh:287-488
void prep(Oracle& o, int count, double dt) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
    auto& S2 = o.SP2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (S[i].i_type > 2) continue;
        if (S[i].p_type >= 1000) continue;
        if (S[i].p_type <= 0) continue;
        double mi = S[i].m, rhoi = S[i].rho;
        double tvis_t = 0.0;
        double ki = fmax(S[i].k_turb, 0.0);
        double nui = S[i].vis_t;
        const P3d& q = o.P3[i];
        double gc11 = q.inv_cm_xx, gc12 = q.inv_cm_xy, gc22 = q.inv_cm_yy, gc23 = q.inv_cm_yz, gc31 = q.inv_cm_zx, gc33 = q.inv_cm_zz;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z;
        double uxi = S[i].ux, uyi = S[i].uy, uzi = S[i].uz;
        double hi = S[i].h;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hi);
        double search_range = pr.kappa * hi;
        Cell c = query_cell(o.d, xi, yi, zi);
        double tmp_flt = 0.0, tmp_SR = 0.0, tmp_sigma = 0.0;
        double st11 = 0, st12 = 0, st22 = 0, st23 = 0, st31 = 0, st33 = 0;
        bool filt = (count % pr.freq_filt) == 0;
        bool les = (pr.fv_solve != 0) && (pr.turbulence_model != SPH_LAMINAR);
        window(o, c, S[i].ncell, [&](int j) {
            double xj = S[j].x, yj = S[j].y, zj = S[j].z;
            double uxj = S[j].ux, uyj = S[j].uy, uzj = S[j].uz;
            double mj = S[j].m, rhoj = S[j].rho;
            int ptypej = S[j].p_type;
            double tdist = dist3(xi, yi, zi, xj, yj, zj) + 1e-20;
            if (tdist < search_range) {
                double tmp_wij = kernel_w(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                double tmp_dwij = kernel_dw(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                double tdwx = tmp_dwij * (xi - xj) / tdist;
                double tdwy = tmp_dwij * (yi - yj) / tdist;
                double tdwz = tmp_dwij * (zi - zj) / tdist;
                if (filt) tmp_flt += MUT(30) ? tmp_wij : mj / rhoj * tmp_wij;
                if (les) {
                    tmp_sigma += tdist * tdist * tmp_wij * mj / rhoj;
                    if (ptypej == 1) {
                        double tdwxc = gc11 * tdwx + gc12 * tdwy + gc31 * tdwz;
                        double tdwyc = gc12 * tdwx + gc22 * tdwy + gc23 * tdwz;
                        double tdwzc = gc31 * tdwx + gc23 * tdwy + gc33 * tdwz;
                        st11 += mj / rhoj * (uxj - uxi) * tdwxc;
                        st12 += 0.5 * mj / rhoj * ((uxj - uxi) * tdwyc + (uyj - uyi) * tdwxc);
                        st22 += mj / rhoj * (uyj - uyi) * tdwyc;
                        st23 += 0.5 * mj / rhoj * ((uyj - uyi) * tdwzc + (uzj - uzi) * tdwyc);
                        st31 += 0.5 * mj / rhoj * ((uxj - uxi) * tdwzc + (uzj - uzi) * tdwxc);
                        st33 += mj / rhoj * (uzj - uzi) * tdwzc;
                    }
                }
            }
        });
        if (les) {
            tmp_SR = 2 * (st11 * st11 + st22 * st22 + st33 * st33 + 2 * st12 * st12 + 2 * st31 * st31 + 2 * st23 * st23);
            tmp_SR = sqrt(tmp_SR);
            S2[i].SR = tmp_SR;
            tmp_sigma = (MUT(28) ? 0.5 : 0.3333) * tmp_sigma;
            tmp_sigma = sqrt(tmp_sigma);
            double tmp_cl = MUT(29) ? fmin(tmp_sigma, hi) : (tmp_sigma > hi ? tmp_sigma : hi);
            if (pr.turbulence_model == SPH_SSM) {
                tvis_t = pr.Cs_sgs * pr.Cs_sgs * tmp_cl * tmp_cl * tmp_SR;
                if (S[i].buffer_type != 1) S[i].vis_t = tvis_t;
            } else if (pr.turbulence_model == SPH_DDF) {
                double p = nui * tmp_SR * tmp_SR;
                double b = 0.0;
                double eps = pr.Ce_sgs * ki * sqrt(ki) / tmp_cl;
                double dk = p + b - eps + 0.0;
                ki += dk * dt;
                const double kfloor = MUT(33) ? 0.0 : 1e-4;
                ki = ki > kfloor ? ki : kfloor;
                if (MUT(31)) eps = pr.Ce_sgs * ki * sqrt(ki) / tmp_cl;
                tvis_t = MUT(32) ? pr.Ck_sgs * tmp_cl * ki : pr.Ck_sgs * tmp_cl * sqrt(ki);
                if (S[i].buffer_type != 1) {
                    S[i].vis_t = tvis_t; S[i].k_turb = ki; S[i].e_turb = eps;
                }
            }
        }
        (void)mi; (void)rhoi;
        if (count == 0) S2[i].rho_ref = S[i].rho;
        if (filt) S[i].flt_s = tmp_flt;
    }
}

// KERNEL_boundary3D, functions_BC.cuh:100-190
void boundary3D(Oracle& o) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (S[i].i_type > 2) continue;
        int p_type_i = S[i].p_type;
        if (p_type_i > 0) continue;
        double hi = S[i].h;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hi);
        double search_range = pr.kappa * hi;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z;
        double uxi = S[i].ux, uyi = S[i].uy, uzi = S[i].uz;
        Cell c = query_cell(o.d, xi, yi, zi);
        double tmpx = 0.0, tmpy = 0.0, tmpz = 0.0, flt = 1.0e-10;
        window(o, c, S[i].ncell, [&](int j) {
            double tdist = dist3(xi, yi, zi, S[j].x, S[j].y, S[j].z) - 1e-20;
            if (tdist < search_range) {
                double twij = kernel_w(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                int p_type_j = S[j].p_type;
                if (pr.noslip_bc == 1) {
                    if ((p_type_j > 0) && (p_type_j != MOVING)) {
                        tmpx += S[j].ux * twij; tmpy += S[j].uy * twij; tmpz += S[j].uz * twij;
                        flt += twij;
                    }
                }
            }
        });
        if (pr.noslip_bc == 1 && (p_type_i == 0 || p_type_i == MOVING)) {
            S[i].ux = 2 * uxi * (p_type_i == MOVING) - tmpx / flt;
            S[i].uy = 2 * uyi * (p_type_i == MOVING) - tmpy / flt;
            S[i].uz = 2 * uzi * (p_type_i == MOVING) - tmpz / flt;
        }
    }
}

// KERNEL_MOST_boundary3D, functions_MOST.cuh:175-322.  Neighbour deposits
// (functions_MOST.cuh:286-303) are serialised in ascending i.
void most(Oracle& o, int count, double dt) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
    for (int i = 0; i < o.n; i++) {
        if (S[i].MOST_buffer != 1) continue;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z;
        double uxi = S[i].ux, uyi = S[i].uy;
        double tempi = S[i].temp, rhoi = S[i].rho, mi = S[i].m;
        double space = S[i].h / H_COEFF;
        double z0 = pr.most_z0, z0h = pr.most_z0h;
        double u_star = 0.0, theta_star = 0.0, OB_length;
        double theta_surface = wall_temperature(pr, dt, count);
        double wind_speed = sqrt(uxi * uxi + uyi * uyi);
        wind_speed = fmax(0.001, fmin(wind_speed, 100.0));
        if (pr.most_mode == SPH_MOST_PRESCRIBED_L) {
            OB_length = pr.most_L;
            u_star = K_VONKARMAN * wind_speed / Phi_m(OB_length, zi, z0);
            theta_star = K_VONKARMAN * (tempi - theta_surface) / Phi_h(OB_length, zi, z0h);
        } else {
            double Ri_b = (pr.gravity / tempi) * (tempi - theta_surface) * zi / (wind_speed * wind_speed + 1e-10);
            if (fabs(Ri_b) > 1e-3) {
                OB_length = zi / Ri_b;
                OB_length = fmax(-500.0, fmin(OB_length, 500.0));
                const double zh = MUT(26) ? z0 : z0h;
                for (int iter = 0; iter < MOST_MAXIT; iter++) {
                    double ol_m = OB_length, ol_l = 0.999 * ol_m, ol_u = 1.001 * ol_m;
                    double pm = Phi_m(ol_m, zi, z0);
                    double fL = Ri_b - (zi / ol_m) * Phi_h(ol_m, zi, zh) / (pm * pm);
                    double dfL = ((-(zi / ol_u) * Phi_h(ol_u, zi, zh) / (Phi_m(ol_u, zi, z0) * Phi_m(ol_u, zi, z0))) -
                                  (-(zi / ol_l) * Phi_h(ol_l, zi, zh) / (Phi_m(ol_l, zi, z0) * Phi_m(ol_l, zi, z0)))) /
                                 (ol_u - ol_l);
                    OB_length -= (MUT(27) ? -1.0 : 1.0) * fL / dfL;
                    if (fabs(OB_length) > 1e-6) {
                        if (fabs((OB_length - ol_m) / OB_length) < MOST_TOL) break;
                        else if (fabs(OB_length - ol_m) < MOST_TOL) break;
                    }
                }
                u_star = K_VONKARMAN * wind_speed / Phi_m(OB_length, zi, z0);
                theta_star = K_VONKARMAN * (tempi - theta_surface) / Phi_h(OB_length, zi, z0h);
            } else {
                OB_length = 1e6;
                u_star = K_VONKARMAN * wind_speed / log(zi / z0);
                theta_star = K_VONKARMAN * (tempi - theta_surface) / log(zi / z0h);
            }
        }
        double momentum_x = -rhoi * u_star * (uxi * u_star / wind_speed);
        double momentum_y = -rhoi * u_star * (uyi * u_star / wind_speed);
        double shf = -rhoi * pr.air_cp * u_star * theta_star * 100.0;
        double dep = 0.5 * shf / (space * rhoi * pr.air_cp);
        double xg = xi, yg = yi, zg = zi + pr.most_ghost_dz;
        Cell c = query_cell(o.d, xg, yg, zg);
        window(o, c, 1, [&](int j) {
            double tdist = dist3(xg, yg, zg, S[j].x, S[j].y, S[j].z) - 1e-20;
            if (tdist < pr.most_ghost_r) {
                o.P3[j].dtemp += dep;
                S[j].hmz = dep;
            }
        });
        S[i].fmx = 1.0 * momentum_x * space * space / mi;
        S[i].fmy = 1.0 * momentum_y * space * space / mi;
        o.P3[i].dtemp += dep;
        S[i].hmz = dep;
        S[i].float1 = u_star; S[i].float2 = theta_star; S[i].float3 = OB_length;
    }
}

// KERNEL_Neumann_boundary3D, functions_BC.cuh:290-446 (pressure read from a
// snapshot; see the header).
void neumann(Oracle& o, std::vector<P1d>& A) {
    const sph_params& pr = o.prm;
    std::vector<double> pold(o.n);
    for (int i = 0; i < o.n; i++) pold[i] = A[i].pres;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (A[i].p_type == 1) continue;
        double search_coeff = MUT(24) ? 1.0 : 2.0;
        double hi = search_coeff * A[i].h;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hi);
        double search_range = pr.kappa * hi;
        double xi = A[i].x, yi = A[i].y, zi = A[i].z;
        double grad_px = 0, grad_py = 0, grad_pz = 0, tpres = 0, tflt_s = 0, pres_g = 0;
        double tux = 0, tuy = 0, tuz = 0;
        double xg = (A[i].buffer_type == 1 || A[i].buffer_type == -1) ? (MUT(21) ? xi - 2.0 * pr.L2 : 2.0 * pr.L2 - xi) : xi;
        double yg = yi, zg = zi;
        Cell c = query_cell(o.d, xg, yg, zg);
        bool got_g = false;
        window(o, c, 1, [&](int j) {
            double dist = dist3(xg, yg, zg, A[j].x, A[j].y, A[j].z) - 1e-20;
            if (dist < pr.neumann_ghost_r && !(MUT(22) && got_g)) { pres_g = pold[j]; got_g = true; }
        });
        window(o, c, A[i].ncell, [&](int j) {
            double xj = A[j].x, yj = A[j].y, zj = A[j].z;
            double tdist = dist3(xg, yg, zg, xj, yj, zj) - 1e-20;
            if (A[j].p_type < 1000) {
                if (tdist < search_range) {
                    double mj = A[j].m, rhoj = A[j].rho, presj = pold[j];
                    double twij = kernel_w(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                    double tdwij = kernel_dw(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                    double tdwx = (xg - xj) / tdist * tdwij;
                    double tdwy = (yg - yj) / tdist * tdwij;
                    double tdwz = (zg - zj) / tdist * tdwij;
                    double neighbor = (double)((A[j].p_type == 1) * (A[j].buffer_type == 0));
                    tpres += mj / rhoj * presj * twij * neighbor;
                    tux += mj / rhoj * A[j].ux * twij * neighbor;
                    tuy += mj / rhoj * A[j].uy * twij * neighbor;
                    tuz += mj / rhoj * A[j].uz * twij * neighbor;
                    tflt_s += mj / rhoj * twij * neighbor;
                    grad_px += mj / rhoj * (presj - pres_g) * tdwx * neighbor;
                    grad_py += mj / rhoj * (presj - pres_g) * tdwy * neighbor;
                    grad_pz += mj / rhoj * (presj - pres_g) * tdwz * neighbor;
                }
            }
        });
        if (tflt_s < 1e-6) {
            A[i].pres = 0.0;
        } else {
            A[i].pres = tpres / tflt_s;
            if (A[i].p_type == -1) {
                A[i].ux = tux / tflt_s; A[i].uy = tuy / tflt_s; A[i].uz = tuz / tflt_s;
            }
            if (A[i].buffer_type == 1 || A[i].buffer_type == -1)
                A[i].pres = pres_g + (MUT(23) ? xg - xi : xi - xg) * grad_px + (yi - yg) * grad_py + (zi - zg) * grad_pz;
        }
    }
}

// KERNEL_advection_force3D, functions_PPE.cuh:191-480
void advection(Oracle& o) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (S[i].i_type > 2) continue;
        if (S[i].p_type == 1000) continue;
        if (S[i].p_type <= 0) continue;
        int ptypei = S[i].p_type;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z;
        double uxi = S[i].ux, uyi = S[i].uy, uzi = S[i].uz;
        double hi = S[i].h, tempi = S[i].temp, mi = S[i].m, rhoi = S[i].rho;
        double eli = S[i].eli;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hi);
        double kci = 0, cpi = 1, concni = 0, diffi = 0;
        if (pr.con_solve) {
            kci = pr.solid_k * (ptypei == 0 || ptypei == 1000) + pr.air_k * (ptypei == 1 || ptypei == -1);
            cpi = pr.solid_cp * (ptypei == 0 || ptypei == 1000) + pr.air_cp * (ptypei == 1 || ptypei == -1);
        }
        if (pr.concn_solve) { concni = S[i].concn; diffi = pr.diff_coef; }
        double visi = viscosity2(pr, ptypei) + S[i].vis_t * rhoi;
        double betai = 1.0 / pr.T_ini;
        Cell c = query_cell(o.d, xi, yi, zi);
        double tmpx = 0, tmpy = 0, tmpz = 0, tmp_Rc = 0, tmp_Rd = 0;
        double eulerx = 0, eulery = 0, eulerz = 0, eulert = 0, tmp_temp = 0, tmp_flt = 0;
        const double Alpha = pr.av_alpha, Beta = pr.av_beta, soundspeed = pr.av_c;
        window(o, c, S[i].ncell, [&](int j) {
            double xj = S[j].x, yj = S[j].y, zj = S[j].z, hj = S[j].h;
            if (S[j].p_type < 1000) {
                double tdist = dist3(xi, yi, zi, xj, yj, zj) + 1e-20;
                double search_range = pr.kappa * fmax(hi, hj);
                if (tdist < search_range) {
                    double twij = kernel_w(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                    double tdwij = kernel_dw(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                    double tdwx = tdwij * (xi - xj) / tdist;
                    double tdwy = tdwij * (yi - yj) / tdist;
                    double tdwz = tdwij * (zi - zj) / tdist;
                    int ptypej = S[j].p_type;
                    double uxj = S[j].ux, uyj = S[j].uy, uzj = S[j].uz;
                    double mj = S[j].m, tempj = S[j].temp, rhoj = S[j].rho;
                    if (pr.fv_solve == SPH_FV_CLEARY) {
                        double visj = viscosity2(pr, ptypej) + (MUT(6) ? S[i].vis_t : S[j].vis_t) * rhoj;
                        double C_v = (xi - xj) * tdwx + (yi - yj) * tdwy + (zi - zj) * tdwz;
                        C_v *= MUT(5) ? 0.5 * (visi + visj) : (visi * visj) / (visi + visj);
                        C_v *= (MUT(4) ? 2 : 4) * (mj / (rhoi * rhoj));
                        C_v /= tdist;
                        C_v /= tdist;
                        tmpx += C_v * (uxi - uxj); tmpy += C_v * (uyi - uyj); tmpz += C_v * (uzi - uzj);
                    } else if (pr.fv_solve == SPH_FV_MONAGHAN) {
                        double uij_xij = (uxi - uxj) * (xi - xj) + (uyi - uyj) * (yi - yj) + (uzi - uzj) * (zi - zj);
                        double visj = viscosity2(pr, ptypej) + S[j].vis_t * rhoj;
                        double phi_ij = uij_xij;
                        phi_ij /= (tdist * tdist + 0.01 * hi * hi);
                        double P_ij = -2.0 * (3 + 2.0) * phi_ij * mj / rhoj;
                        double C_v = 2.0 * (visi * visj) / (visi + visj - 1e-20);
                        double nui = C_v / rhoi;
                        P_ij *= nui;
                        if (ptypei <= 0 || ptypei >= 1000) P_ij = 0.0;
                        if (ptypej <= 0 || ptypei >= 1000) P_ij = 0.0;
                        tmpx += -(P_ij)*tdwx; tmpy += -(P_ij)*tdwy; tmpz += -(P_ij)*tdwz;
                    }
                    if (eli < 1.0) {
                        const double cx = MUT(10) ? uxj : uxi, cy = MUT(10) ? uyj : uyi, cz = MUT(10) ? uzj : uzi;
                        eulerx += (cx * (uxj - uxi) * tdwx + cy * (uxj - uxi) * tdwy + cz * (uxj - uxi) * tdwz) * mj / rhoj * (1 - eli);
                        eulery += (cx * (uyj - uyi) * tdwx + cy * (uyj - uyi) * tdwy + cz * (uyj - uyi) * tdwz) * mj / rhoj * (1 - eli);
                        eulerz += (cx * (uzj - uzi) * tdwx + cy * (uzj - uzi) * tdwy + cz * (uzj - uzi) * tdwz) * mj / rhoj * (1 - eli);
                        if (pr.con_solve)
                            eulert -= 2.0 * mj / (rhoi + rhoj) * (tempj - tempi) * ((uxi - uxj) * tdwx + (uyi - uyj) * tdwy + (uzi - uzj) * tdwz) * (1 - eli);
                        if (pr.concn_solve) {
                            double concnj = S[j].concn;
                            tmp_Rd -= (uxi * (concnj - concni) * tdwx + uyi * (concnj - concni) * tdwy + uzi * (concnj - concni) * tdwz) * mj / rhoj * (1 - eli);
                        }
                    }
                    if (pr.con_solve) {
                        double kcj = pr.solid_k * (ptypej == 0 || ptypej == 1000) + pr.air_k * (ptypej == 1 || ptypej == -1);
                        double rhoj_temp = (ptypej == 1 || ptypej == -1) ? rhoj : pr.solid_rho;
                        double mj_temp = mj * rhoj_temp / rhoj;
                        double sum_con_H = 4.0 * mj_temp * kcj * kci * (tempi - tempj) * tdwij;
                        sum_con_H /= (tdist + 1e-10) * rhoi * rhoj_temp * (kci + kcj);
                        tmp_Rc += sum_con_H;
                    }
                    if (pr.concn_solve) {
                        double concnj = S[j].concn;
                        double diffj = pr.diff_coef;
                        double tmprd = mi * (diffi * rhoi + diffj * rhoj);
                        tmprd *= ((xi - xj) * tdwx + (yi - yj) * tdwy + (zi - zj) * tdwz);
                        tmprd /= (rhoi * rhoj * (tdist * tdist + 0.01 * hi * hi));
                        double tmprd1 = tmprd * (concni - concnj) * (ptypei == ptypej);
                        tmp_Rd += tmprd1;
                    }
                    if (pr.boussinesq_solve) {
                        tmp_temp += (mj / rhoj) * tempj * twij;
                        tmp_flt += (mj / rhoj) * twij;
                    }
                    if (pr.artificial_viscosity) {
                        double uij_xij = (uxi - uxj) * (xi - xj) + (uyi - uyj) * (yi - yj) + (uzi - uzj) * (zi - zj);
                        if (uij_xij < 0) {
                            double h_ij = MUT(7) ? hi : (hi + hj) * 0.5;
                            double phi_ij = h_ij * uij_xij;
                            phi_ij /= (tdist * tdist + (MUT(8) ? 0.1 : 0.01) * h_ij * h_ij);
                            double P_ij = mi * phi_ij * (-Alpha * soundspeed + Beta * phi_ij);
                            P_ij /= (rhoi + rhoj);
                            if (!MUT(9)) P_ij *= 0.5;
                            tmpx += -(P_ij)*tdwx; tmpy += -(P_ij)*tdwy; tmpz += -(P_ij)*tdwz;
                        }
                    }
                }
            }
        });
        if (pr.fg_solve) tmpz += -pr.gravity;
        if ((pr.boussinesq_solve) && (ptypei > 0) && (ptypei < 1000))
            tmpz = betai * pr.gravity * (tmp_temp / tmp_flt - pr.tempb);
        P3d& q = o.P3[i];
        q.ftotalx = MUT(11) ? tmpx + eulerx : tmpx - eulerx;
        q.ftotaly = MUT(11) ? tmpy + eulery : tmpy - eulery;
        q.ftotalz = MUT(11) ? tmpz + eulerz : tmpz - eulerz;
        if (pr.con_solve) q.dtemp += tmp_Rc / cpi - eulert;
        q.ftotal = sqrt(tmpx * tmpx + tmpy * tmpy + tmpz * tmpz);
        if (pr.concn_solve) q.dconcn = tmp_Rd;
    }
}

// KERNEL_clc_projection, functions_TIME_ISPH.cuh:2-112
void projection(Oracle& o, double dt, double time) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
    auto& S2 = o.SP2;
    for (int i = 0; i < o.n; i++) {
        if (S[i].i_type >= 3) continue;
        if (S[i].p_type >= 1000) continue;
        int buffer_type = S[i].buffer_type, MOST = S[i].MOST_buffer, p_typei = S[i].p_type;
        if (p_typei == MOVING) {
            S2[i].x0 = S[i].x; S2[i].y0 = S[i].y; S2[i].z0 = S[i].z;
            S[i].ux = BOUNDARY_VELOCITY; S[i].uy = 0; S[i].uz = 0;
            S2[i].ux0 = BOUNDARY_VELOCITY; S2[i].uy0 = 0; S2[i].uz0 = 0;
        } else {
            double tx0 = S[i].x, ty0 = S[i].y, tz0 = S[i].z;
            double tux0 = S[i].ux, tuy0 = S[i].uy, tuz0 = S[i].uz;
            double tuxp, tuyp, tuzp;
            if (p_typei > 0) {
                const P3d& q = o.P3[i];
                double b = (double)(buffer_type == 0 || MOST == 1);
                double tdux = q.ftotalx * b + S[i].fmx;
                double tduy = q.ftotaly * b + S[i].fmy;
                double tduz = q.ftotalz * b;
                tuxp = tux0 + tdux * (dt);
                tuyp = tuy0 + tduy * (dt);
                tuzp = tuz0 + tduz * (dt);
            } else {
                tuxp = S[i].ux; tuyp = S[i].uy; tuzp = S[i].uz;
            }
            S[i].ux = tuxp; S[i].uy = tuyp; S[i].uz = tuzp;
            S2[i].x0 = tx0; S2[i].y0 = ty0; S2[i].z0 = tz0;
            S2[i].ux0 = tux0; S2[i].uy0 = tuy0; S2[i].uz0 = tuz0;
        }
        if (time == 0) S2[i].rho_ref = S[i].rho;
        if (pr.con_solve == 1) S2[i].temp0 = S[i].temp;
        if (pr.concn_solve == 1) {
            double tconcn = S[i].concn;
            S2[i].concn0 = tconcn;
            S[i].concn = tconcn + o.P3[i].dconcn * (dt);
        }
    }
}

// KERNEL_periodic_boundary3D, functions_BC.cuh:1015-1154 (u, p, T read from
// a snapshot).  The match is the last match inside the first cell, in window
// order, that holds one (flag checked per cell, :1070).
void periodic(Oracle& o, std::vector<P1d>& A) {
    const sph_params& pr = o.prm;
    std::vector<P1d> old = A;
    const sph_domain& d = o.d;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (old[i].buffer_type == 0) continue;
        if (old[i].p_type >= 1000) continue;
        double xi = old[i].x, yi = old[i].y, zi = old[i].z;
        double xg, yg, zg;
        if (old[i].buffer_type == 3) { xg = xi; yg = MUT(17) ? yi - pr.L_right + pr.L_left : yi + pr.L_right - pr.L_left; zg = zi; }
        else if (old[i].buffer_type == 4) { xg = xi; yg = yi - pr.L_right + pr.L_left; zg = zi; }
        else continue;
        int nc = old[i].ncell;
        Cell c = query_cell(d, xg, yg, zg);
        int flag = 0;
        double tux = 0, tuy = 0, tuz = 0, tpres = 0, ttemp = 0;
        for (int z = -nc; z <= nc; z++)
            for (int y = -nc; y <= nc; y++)
                for (int x = -nc; x <= nc; x++) {
                    if (flag == 1) continue;
                    if ((c.i + x) < 0 || (c.i + x) > (d.NI - 1) || (c.j + y) < 0 || (c.j + y) > (d.NJ - 1) ||
                        (c.k + z) < 0 || (c.k + z) > (d.NK - 1))
                        continue;
                    int k = (c.i + x) + d.NI * (c.j + y) + d.NI * d.NJ * (c.k + z);
                    if (o.gstr[k] != -1) {
                        for (int j = o.gstr[k]; j < o.gend[k]; j++) {
                            double tdist = dist3(xg, yg, zg, old[j].x, old[j].y, old[j].z) - 1e-20;
                            if (tdist < pr.copy_tol && !(MUT(18) && flag)) {
                                tux = old[j].ux; tuy = old[j].uy; tuz = old[j].uz;
                                tpres = old[j].pres; ttemp = old[j].temp;
                                flag = 1;
                            }
                        }
                    }
                }
        A[i].ux = tux; A[i].uy = tuy; A[i].uz = tuz; A[i].pres = tpres; A[i].temp = ttemp;
    }
}

// KERNEL_freeslip_boundary3D, functions_BC.cuh:1156-1271
void freeslip(Oracle& o, std::vector<P1d>& A) {
    const sph_params& pr = o.prm;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (A[i].p_type != -1) continue;
        double hi = A[i].h;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hi);
        double search_range = pr.kappa * hi;
        double xg = A[i].x, yg = A[i].y, zg = A[i].z;
        Cell c = query_cell(o.d, xg, yg, zg);
        double tux = 0, tuy = 0, tuz = 0, tpres = 0, ttemp = 0, tflt_s = 0;
        window(o, c, A[i].ncell, [&](int j) {
            double tdist = dist3(xg, yg, zg, A[j].x, A[j].y, A[j].z) - 1e-20;
            if (tdist < search_range) {
                double mj = A[j].m, rhoj = A[j].rho;
                double neighbor = (double)(A[j].p_type == 1);
                double twij = kernel_w(pr.kernel_type, pr.kappa, tmp_A, hi, tdist);
                tux += A[j].ux * (mj / rhoj) * twij * neighbor;
                tuy += A[j].uy * (mj / rhoj) * twij * neighbor;
                tuz += A[j].uz * (mj / rhoj) * twij * neighbor;
                tpres += A[j].pres * (mj / rhoj) * twij * neighbor;
                ttemp += A[j].temp * (mj / rhoj) * twij * neighbor;
                tflt_s += (mj / rhoj) * twij * neighbor;
            }
        });
        if (tflt_s > 1e-5) {
            A[i].ux = tux / tflt_s; A[i].uy = tuy / tflt_s; A[i].uz = tuz / tflt_s;
            A[i].pres = tpres / tflt_s; A[i].temp = ttemp / tflt_s;
        }
    }
}

// KERNEL_donothing_buffer3D, functions_BC.cuh:673-761 (snapshot reads).
void outlet(Oracle& o, std::vector<P1d>& A) {
    const sph_params& pr = o.prm;
    std::vector<P1d> old = A;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (old[i].buffer_type != 2) continue;
        double xi = old[i].x, yi = old[i].y, zi = old[i].z;
        Cell c = query_cell(o.d, xi, yi, zi);
        int nc = 2 * old[i].ncell;
        double tux = 0, tuy = 0, tuz = 0, tpres = 0, ttemp = 0;
        double xg = MUT(19) ? pr.L3 + pr.outlet_copy_dx : pr.L3 - pr.outlet_copy_dx;
        bool got = false;
        window(o, c, nc, [&](int j) {
            double tdist = dist3(xg, yi, zi, old[j].x, old[j].y, old[j].z) - 1e-20;
            if (old[j].p_type < 1000) {
                if (tdist < pr.copy_tol && !(MUT(20) && got)) {
                    got = true;
                    tux = old[j].ux; tuy = old[j].uy; tuz = old[j].uz;
                    tpres = old[j].pres; ttemp = old[j].temp;
                }
            }
        });
        A[i].pres = tpres; A[i].ux = tux; A[i].uy = tuy; A[i].uz = tuz; A[i].temp = ttemp;
    }
}

// ------------------------------------------------------------ PPE --------
struct PPERow {
    double bi, Aij, kn;           // b_i (already /dt), sum c_ij, known part
    std::vector<int> cols;        // unknown neighbours (particle index), j != i
    std::vector<double> vals;     // c_ij
    double AijPij;                // full sum c_ij*(p_j - p0_j)
};

// Assembly of one row of KERNEL_PPE3D, functions_PPE.cuh:2-160.  pres is the
// pressure vector to read (the previous iterate).  is_unknown flags rows of
// the solve.
void ppe_row(const Oracle& o, int i, double dt, const std::vector<double>& pres,
             const std::vector<char>& is_unknown, PPERow& row) {
    const sph_params& pr = o.prm;
    const auto& S = o.SP1;
    double hi = 1.0 * S[i].h;
    double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hi);
    double xi = S[i].x, yi = S[i].y, zi = S[i].z;
    double uxi = S[i].ux, uyi = S[i].uy, uzi = S[i].uz, rhoi = S[i].rho;
    const P3d& q = o.P3[i];
    double gc11 = q.inv_cm_xx, gc12 = q.inv_cm_xy, gc22 = q.inv_cm_yy, gc23 = q.inv_cm_yz, gc31 = q.inv_cm_zx, gc33 = q.inv_cm_zz;
    double divFb = 0.0, bix = 0, biy = 0, biz = 0, Aij = 0, AijPij = 0, kn = 0;
    row.cols.clear(); row.vals.clear();
    Cell c = query_cell(o.d, xi, yi, zi);
    window(o, c, 1 * S[i].ncell, [&](int j) {
        if (S[j].p_type < 1000) {
            if (S[j].i_type != 4) {
                double xj = S[j].x, yj = S[j].y, zj = S[j].z;
                double uxj = S[j].ux, uyj = S[j].uy, uzj = S[j].uz;
                double rhoj = S[j].rho, mj = S[j].m, presj = pres[j];
                double hj = 1.0 * S[j].h;
                double search_range = pr.kappa * fmax(hi, hj);
                double tdist = dist3(xi, yi, zi, xj, yj, zj) + 1e-20;
                if (tdist < search_range) {
                    double tdwij = kernel_dwp(pr.kappa, tmp_A, hi, tdist);
                    double tdwx = tdwij * (xi - xj) / tdist;
                    double tdwy = tdwij * (yi - yj) / tdist;
                    double tdwz = tdwij * (zi - zj) / tdist;
                    double tdwxc = gc11 * tdwx + gc12 * tdwy + gc31 * tdwz;
                    double tdwyc = gc12 * tdwx + gc22 * tdwy + gc23 * tdwz;
                    double tdwzc = gc31 * tdwx + gc23 * tdwy + gc33 * tdwz;
                    bix += mj / rhoj * (uxj - uxi) * tdwxc * (2.0 * rhoi * rhoj / (rhoi + rhoj));
                    biy += mj / rhoj * (uyj - uyi) * tdwyc * (2.0 * rhoi * rhoj / (rhoi + rhoj));
                    biz += mj / rhoj * (uzj - uzi) * tdwzc * (2.0 * rhoi * rhoj / (rhoi + rhoj));
                    double cc = 2.0 * mj / rhoj * ((xi - xj) * tdwxc + (yi - yj) * tdwyc + (zi - zj) * tdwzc) / tdist / tdist;
                    Aij += cc;
                    double qj = presj - S[j].pres0;
                    AijPij += cc * qj;
                    if (j != i) {
                        if (is_unknown[j]) { row.cols.push_back(j); row.vals.push_back(cc); }
                        else kn += cc * qj;
                    }
                }
            }
        }
    });
    double weight = 1.0;
    row.bi = weight * (bix + biy + biz + divFb) / dt;
    row.Aij = Aij;
    row.AijPij = AijPij;
    row.kn = kn;
}

inline bool ppe_active(const P1d& p) { return p.i_type <= 2 && p.p_type < 1000 && p.p_type > 0; }

void spmv(const std::vector<int>& rowptr, const std::vector<int>& cols, const std::vector<double>& vals,
          const std::vector<double>& diag, const std::vector<double>& x, std::vector<double>& y) {
    long nr = (long)diag.size();
#pragma omp parallel for schedule(static)
    for (long r = 0; r < nr; r++) {
        double s = 0.0;
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) s += vals[e] * x[cols[e]];
        y[r] = diag[r] * x[r] - s;
    }
}

// Jacobi-preconditioned BiCGSTAB.  Every dot is canon_sum over the rows.
int bicgstab(Oracle& o, const std::vector<double>& rhs, std::vector<double>& x) {
    const sph_params& pr = o.prm;
    const auto& rp = o.sys_rowptr; const auto& cl = o.sys_cols; const auto& vl = o.sys_vals;
    const auto& dg = o.sys_diag;
    long nr = (long)rhs.size();
    std::vector<double> r(nr), rhat(nr), p(nr, 0.0), v(nr, 0.0), phat(nr), s(nr), shat(nr), t(nr), y(nr);
    double bnorm = sqrt(dot(rhs, rhs));
    o.rep.iterations = 0;
    if (bnorm == 0.0) {
        for (long i = 0; i < nr; i++) x[i] = 0.0;
        o.rep.rel_residual = 0.0; o.rep.converged = 1;
        return 0;
    }
    spmv(rp, cl, vl, dg, x, y);
    for (long i = 0; i < nr; i++) r[i] = rhs[i] - y[i];
    double rnorm = sqrt(dot(r, r));
    if (!(bnorm <= DBL_MAX) || !(rnorm <= DBL_MAX)) {   // NaN/Inf in the system: no iteration
        o.rep.rel_residual = rnorm / bnorm; o.rep.converged = 0;
        return SPH_E_NONFINITE;
    }
    if (rnorm <= pr.ppe_tol * bnorm) { o.rep.rel_residual = rnorm / bnorm; o.rep.converged = 1; return 0; }
    rhat = r;
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    int it;
    int conv = 0;
    for (it = 1; it <= pr.ppe_maxit; it++) {
        double rho = dot(rhat, r);
        if (rho == 0.0) break;
        if (it == 1) {
            for (long i = 0; i < nr; i++) p[i] = r[i];
        } else {
            double beta = (rho / rho_old) * (alpha / omega);
            for (long i = 0; i < nr; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        for (long i = 0; i < nr; i++) phat[i] = p[i] / dg[i];
        spmv(rp, cl, vl, dg, phat, v);
        double rv = dot(rhat, v);
        alpha = rho / rv;
        for (long i = 0; i < nr; i++) s[i] = r[i] - alpha * v[i];
        double snorm = sqrt(dot(s, s));
        if (snorm <= pr.ppe_tol * bnorm) {
            for (long i = 0; i < nr; i++) x[i] = x[i] + alpha * phat[i];
            rnorm = snorm; conv = 1;
            break;
        }
        for (long i = 0; i < nr; i++) shat[i] = s[i] / dg[i];
        spmv(rp, cl, vl, dg, shat, t);
        double tt = dot(t, t);
        double ts = dot(t, s);
        omega = ts / tt;
        for (long i = 0; i < nr; i++) x[i] = x[i] + alpha * phat[i] + omega * shat[i];
        for (long i = 0; i < nr; i++) r[i] = s[i] - omega * t[i];
        rnorm = sqrt(dot(r, r));
        if (rnorm <= pr.ppe_tol * bnorm) { conv = 1; break; }
        if (omega == 0.0 || !(rnorm <= DBL_MAX)) break;
        rho_old = rho;
    }
    o.rep.iterations = it > pr.ppe_maxit ? pr.ppe_maxit : it;
    o.rep.rel_residual = rnorm / bnorm;
    o.rep.converged = conv;
    if (!conv && !(rnorm <= DBL_MAX)) return SPH_E_NONFINITE;
    return conv ? 0 : SPH_E_NOCONV;
}

int ppe(Oracle& o, double dt) {
    auto& S = o.SP1;
    int n = o.n;
    std::vector<double> pres(n);
    for (int i = 0; i < n; i++) pres[i] = S[i].pres;
    std::vector<char> unk(n, 0);
    std::vector<int> rows;
    for (int i = 0; i < n; i++) if (ppe_active(S[i])) { unk[i] = 1; rows.push_back(i); }
    long nr = (long)rows.size();
    std::vector<PPERow> R(nr);
#pragma omp parallel for schedule(static)
    for (long r = 0; r < nr; r++) ppe_row(o, rows[r], dt, pres, unk, R[r]);
    if (o.prm.ppe_solver == SPH_PPE_REF_JACOBI) {
        for (long r = 0; r < nr; r++) {
            int i = rows[r];
            S[i].pres = S[i].pres0 + (R[r].bi + R[r].AijPij) / R[r].Aij;
        }
        o.rep.iterations = 1; o.rep.refinements = 0; o.rep.unknowns = (int)nr;
        long long nnz = 0; for (auto& rr : R) nnz += rr.cols.size();
        o.rep.nnz = nnz; o.rep.rel_residual = -1.0; o.rep.converged = 1;
        return 0;
    }
    // Krylov: rows over unknowns, columns mapped to row numbers.
    std::vector<int> rowof(n, -1);
    for (long r = 0; r < nr; r++) rowof[rows[r]] = (int)r;
    o.sys_rows = rows;
    o.sys_rowptr.assign(nr + 1, 0);
    for (long r = 0; r < nr; r++) o.sys_rowptr[r + 1] = o.sys_rowptr[r] + (int)R[r].cols.size();
    o.sys_cols.resize(o.sys_rowptr[nr]); o.sys_vals.resize(o.sys_rowptr[nr]);
    o.sys_diag.resize(nr); o.sys_rhs.resize(nr); o.sys_x0.resize(nr);
    for (long r = 0; r < nr; r++) {
        int e0 = o.sys_rowptr[r];
        for (size_t k = 0; k < R[r].cols.size(); k++) {
            o.sys_cols[e0 + k] = rowof[R[r].cols[k]];
            o.sys_vals[e0 + k] = R[r].vals[k];
        }
        o.sys_diag[r] = R[r].Aij;
        o.sys_rhs[r] = R[r].bi + R[r].kn;
        int i = rows[r];
        o.sys_x0[r] = S[i].pres - S[i].pres0;
    }
    std::vector<double> x = o.sys_x0;
    int rc = bicgstab(o, o.sys_rhs, x);
    o.sys_x = x;
    for (long r = 0; r < nr; r++) { int i = rows[r]; S[i].pres = S[i].pres0 + x[r]; }
    o.rep.refinements = 0; o.rep.unknowns = (int)nr; o.rep.nnz = o.sys_rowptr[nr];
    return rc;
}

// KERNEL_pressureforce3D, functions_PPE.cuh:482-647
void pressure_force(Oracle& o) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (S[i].i_type > 2) continue;
        if (S[i].p_type >= 1000) continue;
        if (S[i].p_type <= 0) continue;
        if (S[i].buffer_type != 0) continue;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z, hi = S[i].h;
        double pi = S[i].pres - S[i].pres0, rhoi = S[i].rho;
        Cell c = query_cell(o.d, xi, yi, zi);
        double tmpx = 0, tmpy = 0, tmpz = 0;
        window(o, c, S[i].ncell, [&](int j) {
            double xj = S[j].x, yj = S[j].y, zj = S[j].z, hj = S[j].h;
            double tdist = dist3(xi, yi, zi, xj, yj, zj) - 1e-20;
            double search_range = pr.kappa * fmax(hi, hj);
            if (S[j].p_type < 1000) {
                if (tdist < search_range) {
                    double hij = MUT(2) ? fmax(hi, hj) : 0.5 * (hi + hj);
                    double tmp_Aij = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, hij);
                    double tdwij = kernel_dw(pr.kernel_type, pr.kappa, tmp_Aij, hij, tdist);
                    double tdwx = tdwij * (xi - xj) / tdist;
                    double tdwy = tdwij * (yi - yj) / tdist;
                    double tdwz = tdwij * (zi - zj) / tdist;
                    double mj = S[j].m, rhoj = S[j].rho, pj = S[j].pres - S[j].pres0;
                    if (pr.fp_solve) {
                        const double Pb = 0.0;
                        double C_p = (MUT(1) ? -1.0 : -2.0) * mj / (rhoi * rhoj) *
                                     (MUT(3) ? (pi + Pb) * rhoi + (pj + Pb) * rhoj : (pi + Pb) * rhoj + (pj + Pb) * rhoi) / (rhoi + rhoj);
                        tmpx += C_p * tdwx; tmpy += C_p * tdwy; tmpz += C_p * tdwz;
                    }
                }
            }
        });
        o.P3[i].fpx = tmpx; o.P3[i].fpy = tmpy; o.P3[i].fpz = tmpz;
    }
}

// IBM_force_interpolation3D, functions_IBM.cuh:342-490 (persistent normals)
void ibm_force(Oracle& o, double dt) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (S[i].p_type < 1000) continue;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z;
        double tmp_h = (MUT(12) ? 1.0 : 0.5) * S[i].h;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, tmp_h);
        double search_range = pr.kappa * tmp_h;
        Cell c = bin_cell(o.d, xi, yi, zi);
        double tmpux = 0, tmpuy = 0, tmpuz = 0, filt = 0;
        window(o, c, S[i].ncell, [&](int j) {
            if (S[j].i_type != 4) {
                if (S[j].p_type == 1) {
                    double tdist = dist3(xi, yi, zi, S[j].x, S[j].y, S[j].z) + 1e-20;
                    if (tdist < search_range) {
                        double twij = kernel_w(pr.kernel_type, pr.kappa, tmp_A, tmp_h, tdist);
                        double tmprho = S[j].m / S[j].rho;
                        tmpux += (S[j].ux) * tmprho * twij;
                        tmpuy += (S[j].uy) * tmprho * twij;
                        tmpuz += (S[j].uz) * tmprho * twij;
                        filt += tmprho * twij;
                    }
                }
            }
        });
        S[i].concentration = filt;
        double ufx = tmpux / (filt + 1e-10), ufy = tmpuy / (filt + 1e-10), ufz = tmpuz / (filt + 1e-10);
        double nx = o.nx[i], ny = o.ny[i], nz = o.nz[i];
        double nmag = sqrt(nx * nx + ny * ny + nz * nz) + 1e-20;
        nx = nx / nmag; ny = ny / nmag; nz = nz / nmag;
        double un = ufx * nx + ufy * ny + ufz * nz;
        double utx = ufx - un * nx, uty = ufy - un * ny, utz = ufz - un * nz;
        double Ubtx = MUT(13) ? 0.0 : pr.ibm_ut_factor * interp_ux_caseH(zi, 1.0);
        double Ubty = MUT(13) ? pr.ibm_ut_factor * interp_ux_caseH(zi, 1.0) : 0.0, Ubtz = 0.0;
        double beta_t = 1.0;
        const double mun = MUT(14) ? un : -un;
        double corrx = (mun) * nx + beta_t * (Ubtx - utx);
        double corry = (mun) * ny + beta_t * (Ubty - uty);
        double corrz = (mun) * nz + beta_t * (Ubtz - utz);
        S[i].fbx = pr.rho_air * corrx / dt;
        S[i].fby = pr.rho_air * corry / dt;
        S[i].fbz = pr.rho_air * corrz / dt;
    }
}

// IBM_spreading_interpolation3D, functions_IBM.cuh:492-613
void ibm_spread(Oracle& o) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        if (S[i].p_type != 1) continue;
        if (S[i].buffer_type == 2) continue;
        double xi = S[i].x, yi = S[i].y, zi = S[i].z;
        double tmp_h = (MUT(16) ? 1.0 : 0.5) * S[i].h;
        double tmp_A = calc_tmpA(pr.kernel_type, pr.kappa, o.c_pi15, tmp_h);
        double search_range = pr.kappa * tmp_h;
        Cell c = bin_cell(o.d, xi, yi, zi);
        double filt = 0, fb_x = 0, fb_y = 0, fb_z = 0;
        window(o, c, S[i].ncell, [&](int j) {
            if (S[j].i_type != 4) {
                if (S[j].p_type == 1000) {
                    double tdist = dist3(xi, yi, zi, S[j].x, S[j].y, S[j].z) + 1e-20;
                    if (tdist < search_range) {
                        double twij = kernel_w(pr.kernel_type, pr.kappa, tmp_A, tmp_h, tdist);
                        double tmprho = S[j].m / S[j].rho;
                        fb_x += S[j].fbx * twij * tmprho;
                        fb_y += S[j].fby * twij * tmprho;
                        fb_z += S[j].fbz * twij * tmprho;
                        filt += tmprho * twij;
                    }
                }
            }
        });
        if (filt < 1e-12) {
            S[i].fbx = 0.0; S[i].fby = 0.0; S[i].fbz = 0.0; S[i].concentration = 0.0;
            continue;
        }
        S[i].fbx = MUT(15) ? fb_x : fb_x / (filt + 1e-10);
        S[i].fby = MUT(15) ? fb_y : fb_y / (filt + 1e-10);
        S[i].fbz = MUT(15) ? fb_z : fb_z / (filt + 1e-10);
        S[i].concentration = filt;
    }
}

// KERNEL_time_update_projection, functions_TIME_ISPH.cuh:115-296
void time_update(Oracle& o, double dt, int count) {
    const sph_params& pr = o.prm;
    auto& S = o.SP1; auto& P = o.P1; auto& S2 = o.SP2; auto& P2 = o.P2;
    if (pr.ale_positions) {
        o.xn.resize(o.n); o.yn.resize(o.n); o.zn.resize(o.n);
        for (int i = 0; i < o.n; i++) { o.xn[i] = S[i].x; o.yn[i] = S[i].y; o.zn[i] = S[i].z; }
    }
    for (int i = 0; i < o.n; i++) {
        if (S[i].i_type >= 3) continue;
        if (S[i].p_type >= 1000) continue;
        int buffer_type = S[i].buffer_type, p_type_i = S[i].p_type;
        const P3d& q = o.P3[i];
        if (p_type_i == MOVING) {
            P[i].ux = S[i].ux; P[i].uy = S[i].uy; P[i].uz = S[i].uz;
        } else {
            double tux0, tuy0, tuz0, dux_dt, duy_dt, duz_dt;
            if (p_type_i > 0) {
                tux0 = S[i].ux; tuy0 = S[i].uy; tuz0 = S[i].uz;
                if (buffer_type == 1 || buffer_type == -1) { tux0 = S2[i].ux0; tuy0 = S2[i].uy0; tuz0 = S2[i].uz0; }
                double b = (double)(buffer_type == 0);
                dux_dt = q.fpx * b + S[i].fbx / S[i].rho + S[i].fmx;
                duy_dt = q.fpy * b + S[i].fby / S[i].rho + S[i].fmy;
                duz_dt = q.fpz * b + S[i].fbz / S[i].rho;
            } else {
                tux0 = S2[i].ux0; tuy0 = S2[i].uy0; tuz0 = S2[i].uz0;
                dux_dt = duy_dt = duz_dt = 0.0;
            }
            double uxc = tux0 + dux_dt * (dt), uyc = tuy0 + duy_dt * (dt), uzc = tuz0 + duz_dt * (dt);
            if ((uxc * uxc + uyc * uyc + uzc * uzc) >= pr.u_limit * pr.u_limit) { uxc = tux0; uyc = tuy0; uzc = tuz0; }
            P[i].ux = uxc; P[i].uy = uyc; P[i].uz = uzc;
            if (pr.ale_positions) {
                // functions_TIME_ISPH.cuh:193-195 (computed there, written here)
                const double mv = (double)(p_type_i > 0) * S[i].eli;
                o.xn[i] = S[i].x + (S2[i].ux0 + uxc) / 2.0 * (dt) * mv;
                o.yn[i] = S[i].y + (S2[i].uy0 + uyc) / 2.0 * (dt) * mv;
                o.zn[i] = S[i].z + (S2[i].uz0 + uzc) / 2.0 * (dt) * mv;
            }
        }
        if (pr.con_solve == 1) P[i].temp = S2[i].temp0 + q.dtemp * dt * (buffer_type == 0);
        else { P[i].enthalpy = S[i].enthalpy; P[i].temp = S[i].temp; }
        if (pr.concn_solve == 1) {
            P[i].concn = S2[i].concn0 + q.dconcn * dt;
            if (P[i].concn < 0.0) P[i].concn = 0.0;
        } else P[i].concn = S[i].concn;
        P[i].pres = S[i].pres; P[i].flt_s = S[i].flt_s; P[i].m = S[i].m; P[i].ncell = S[i].ncell;
        P[i].h = S[i].h; P[i].k_turb = S[i].k_turb; P[i].e_turb = S[i].e_turb; P[i].eli = S[i].eli;
        P[i].vol = S[i].vol; P2[i].rho_ref = S2[i].rho_ref; P[i].vis_t = S[i].vis_t;
        P[i].fbx = S[i].fbx; P[i].fby = S[i].fby; P[i].fbz = S[i].fbz;
        P[i].fmx = S[i].fmx; P[i].fmy = S[i].fmy; P[i].hmz = S[i].hmz;
        P[i].float1 = S[i].float1; P[i].float2 = S[i].float2; P[i].float3 = S[i].float3;
        P[i].concentration = S[i].concentration;
        if (P[i].source == 1) P[i].concn += 0.5 * dt;
        if (P[i].p_type == 0) P[i].temp = wall_temperature(pr, dt, count);
    }
}

// IBM_corrector, functions_IBM.cuh:615-725 (only the copies are live)
void ibm_corrector(Oracle& o) {
    auto& S = o.SP1; auto& P = o.P1;
    for (int i = 0; i < o.n; i++) {
        if (S[i].p_type < 1000) continue;
        P[i].fbx = S[i].fbx; P[i].fby = S[i].fby; P[i].fbz = S[i].fbz;
        P[i].concentration = S[i].concentration; P[i].vol = S[i].vol; P[i].pres = S[i].pres;
    }
}

// SOPHIA_single_ISPH, ISPH_Calc.cuh:4-233
// ------------------------------------------------------------ APR --------
// Cube root by Newton's method on the mantissa, 8 iterations from 1.
double apr_cbrt(double v) {
    int e = 0;
    double m = frexp(v, &e);
    const int r = ((e % 3) + 3) % 3;
    m = ldexp(m, r);
    e -= r;
    double y = 1.0;
    for (int it = 0; it < 8; it++) y = y - (y * y * y - m) / (3.0 * y * y);
    return ldexp(y, e / 3);
}

// Splitting passes of Particle_refinement_cell (functions_APR_etc.cuh:354-375):
// the condition of KERNEL_APR_condition3D_cell (functions_APR.cuh:270-330;
// the first particle of a cell with M_num - log8(m_ref/m) >= 0.75), the
// host prefix sum Cell_index, and KERNEL_cell_APR_3D (functions_APR.cuh:
// 420-500: 8 children at +-cbrt(V)/4, m/8, vol/8, h/2, at n + 8*rank; the
// parent becomes i_type 3).  The search runs again after each pass.
int apr_refine(Oracle& o, const double* target, int max_passes, int* n_out) {
    if (!o.searched) return SPH_E_STATE;
    if ((int)o.apr_target.size() != o.n) o.apr_target.assign(o.n, 0.0);
    if (target) for (int i = 0; i < o.n; i++) o.apr_target[i] = target[i];
    for (int pass = 0; pass < max_passes; pass++) {
        const int ncells = o.d.num_cells - 1;
        std::vector<int> pick;
        for (int cell = 0; cell < ncells; cell++) {
            if (o.gstr[cell] == -1) continue;
            for (int j = o.gstr[cell]; j < o.gend[cell]; j++) {
                const P1d& p = o.P1[j];
                double level = log(p.m_ref / p.m) / log(8.0);
                if (o.apr_target[j] - level >= 0.75) {
                    if (!(p.p_type != 1 && o.apr_target[j] > 2.0)) pick.push_back(j);
                    break;
                }
            }
        }
        const int ns = (int)pick.size();
        if (ns == 0) break;
        if ((long long)o.n + 8LL * ns > o.cap) return SPH_E_NOMEM;
        const int n0 = o.n;
        o.n += 8 * ns;
        o.P1.resize(o.n); o.P2.resize(o.n);
        o.nx.resize(o.n); o.ny.resize(o.n); o.nz.resize(o.n); o.apr_target.resize(o.n);
        for (int k = 0; k < ns; k++) {
            const int j = pick[k];
            const P1d par = o.P1[j];
            const double a = 0.25 * apr_cbrt(par.m / par.rho);
            for (int q = 0; q < 8; q++) {
                const int c = n0 + 8 * k + q;
                P1d ch = par;
                // positions rounded to float: the reference stores them in
                // its float records (functions_APR.cuh:544-590)
                ch.x = (double)(float)(q < 4 ? par.x + a : par.x - a);
                ch.y = (double)(float)((q & 2) ? par.y - a : par.y + a);
                ch.z = (double)(float)((q & 1) ? par.z - a : par.z + a);
                ch.m = 0.125 * par.m;
                ch.vol = 0.125 * par.vol;
                if (o.prm.apr_h_change) ch.h = 0.5 * par.h;
                o.P1[c] = ch;
                o.P2[c] = o.P2[j];
                o.nx[c] = o.nx[j]; o.ny[c] = o.ny[j]; o.nz[c] = o.nz[j];
                o.apr_target[c] = o.apr_target[j];
            }
            o.P1[j].i_type = 3;
        }
        set_ncell(o);
        nnps(o);
    }
    if (n_out) *n_out = o.n;
    return 0;
}

int step(Oracle& o, int count, double dt, double time) {
    const sph_params& pr = o.prm;
    set_ncell(o);
    if (count == 0 || pr.ale_positions) {   // moving particles: a search every step
        nnps(o);
    } else {
        if (!o.searched) return SPH_E_STATE;
        o.SP1 = o.P1; o.SP2 = o.P2;
    }
    o.P3.assign(o.n, P3d{});
    ib_normals(o, count);
    gradient_correction(o);
    prep(o, count, dt);
    if (pr.noslip_bc == 1) boundary3D(o);
    most(o, count, dt);
    if (count == 0) neumann(o, o.SP1);
    advection(o);
    if (pr.time_type == SPH_PRE_COR) projection(o, dt, time);
    periodic(o, o.SP1);
    freeslip(o, o.SP1);
    int rc = ppe(o, dt);
    neumann(o, o.SP1);
    outlet(o, o.SP1);
    pressure_force(o);
    ibm_force(o, dt);
    ibm_spread(o);
    time_update(o, dt, count);
    ibm_corrector(o);
    neumann(o, o.P1);
    outlet(o, o.P1);
    periodic(o, o.P1);
    freeslip(o, o.P1);
    if (pr.ale_positions)
        for (int i = 0; i < o.n; i++) { o.P1[i].x = o.xn[i]; o.P1[i].y = o.yn[i]; o.P1[i].z = o.zn[i]; }
    return rc;
}

double field_value(const Oracle& o, int f, int i) {
    const P1d& p = o.P1[i];
    const P2d& p2 = o.SP2[i];
    const P3d& q = o.P3[i];
    switch (f) {
    case SPH_F_X: return p.x; case SPH_F_Y: return p.y; case SPH_F_Z: return p.z;
    case SPH_F_UX: return p.ux; case SPH_F_UY: return p.uy; case SPH_F_UZ: return p.uz;
    case SPH_F_M: return p.m; case SPH_F_H: return p.h; case SPH_F_TEMP: return p.temp;
    case SPH_F_PRES: return p.pres; case SPH_F_PRES0: return p.pres0; case SPH_F_RHO: return p.rho;
    case SPH_F_FLT_S: return p.flt_s; case SPH_F_CONCN: return p.concn;
    case SPH_F_CONCENTRATION: return p.concentration; case SPH_F_K_TURB: return p.k_turb;
    case SPH_F_E_TURB: return p.e_turb; case SPH_F_VIS_T: return p.vis_t;
    case SPH_F_FBX: return p.fbx; case SPH_F_FBY: return p.fby; case SPH_F_FBZ: return p.fbz;
    case SPH_F_FMX: return p.fmx; case SPH_F_FMY: return p.fmy; case SPH_F_HMZ: return p.hmz;
    case SPH_F_FLOAT1: return p.float1; case SPH_F_FLOAT2: return p.float2; case SPH_F_FLOAT3: return p.float3;
    case SPH_F_VOL: return p.vol; case SPH_F_ENTHALPY: return p.enthalpy;
    case SPH_F_RHO_REF: return p2.rho_ref; case SPH_F_SR: return p2.SR;
    case SPH_F_X0: return p2.x0; case SPH_F_Y0: return p2.y0; case SPH_F_Z0: return p2.z0;
    case SPH_F_UX0: return p2.ux0; case SPH_F_UY0: return p2.uy0; case SPH_F_UZ0: return p2.uz0;
    case SPH_F_TEMP0: return p2.temp0; case SPH_F_CONCN0: return p2.concn0;
    case SPH_F_FTOTALX: return q.ftotalx; case SPH_F_FTOTALY: return q.ftotaly;
    case SPH_F_FTOTALZ: return q.ftotalz; case SPH_F_FTOTAL: return q.ftotal;
    case SPH_F_FPX: return q.fpx; case SPH_F_FPY: return q.fpy; case SPH_F_FPZ: return q.fpz;
    case SPH_F_DTEMP: return q.dtemp; case SPH_F_DCONCN: return q.dconcn;
    case SPH_F_INV_CM_XX: return q.inv_cm_xx; case SPH_F_INV_CM_YY: return q.inv_cm_yy;
    case SPH_F_INV_CM_ZZ: return q.inv_cm_zz; case SPH_F_INV_CM_XY: return q.inv_cm_xy;
    case SPH_F_INV_CM_YZ: return q.inv_cm_yz; case SPH_F_INV_CM_ZX: return q.inv_cm_zx;
    case SPH_F_NX: return o.nx[i]; case SPH_F_NY: return o.ny[i]; case SPH_F_NZ: return o.nz[i];
    }
    return NAN;
}

// Working-state view for the per-pass entry points: SP1 / SP2 / P3.
double* work_ref(Oracle& o, int f, int i) {
    P1d& p = o.SP1[i];
    P2d& p2 = o.SP2[i];
    P3d& q = o.P3[i];
    switch (f) {
    case SPH_F_X: return &p.x; case SPH_F_Y: return &p.y; case SPH_F_Z: return &p.z;
    case SPH_F_UX: return &p.ux; case SPH_F_UY: return &p.uy; case SPH_F_UZ: return &p.uz;
    case SPH_F_M: return &p.m; case SPH_F_H: return &p.h; case SPH_F_TEMP: return &p.temp;
    case SPH_F_PRES: return &p.pres; case SPH_F_PRES0: return &p.pres0; case SPH_F_RHO: return &p.rho;
    case SPH_F_FLT_S: return &p.flt_s; case SPH_F_CONCN: return &p.concn;
    case SPH_F_CONCENTRATION: return &p.concentration; case SPH_F_K_TURB: return &p.k_turb;
    case SPH_F_E_TURB: return &p.e_turb; case SPH_F_VIS_T: return &p.vis_t;
    case SPH_F_FBX: return &p.fbx; case SPH_F_FBY: return &p.fby; case SPH_F_FBZ: return &p.fbz;
    case SPH_F_FMX: return &p.fmx; case SPH_F_FMY: return &p.fmy; case SPH_F_HMZ: return &p.hmz;
    case SPH_F_FLOAT1: return &p.float1; case SPH_F_FLOAT2: return &p.float2; case SPH_F_FLOAT3: return &p.float3;
    case SPH_F_VOL: return &p.vol; case SPH_F_ENTHALPY: return &p.enthalpy;
    case SPH_F_RHO_REF: return &p2.rho_ref; case SPH_F_SR: return &p2.SR;
    case SPH_F_X0: return &p2.x0; case SPH_F_Y0: return &p2.y0; case SPH_F_Z0: return &p2.z0;
    case SPH_F_UX0: return &p2.ux0; case SPH_F_UY0: return &p2.uy0; case SPH_F_UZ0: return &p2.uz0;
    case SPH_F_TEMP0: return &p2.temp0; case SPH_F_CONCN0: return &p2.concn0;
    case SPH_F_FTOTALX: return &q.ftotalx; case SPH_F_FTOTALY: return &q.ftotaly;
    case SPH_F_FTOTALZ: return &q.ftotalz; case SPH_F_FTOTAL: return &q.ftotal;
    case SPH_F_FPX: return &q.fpx; case SPH_F_FPY: return &q.fpy; case SPH_F_FPZ: return &q.fpz;
    case SPH_F_DTEMP: return &q.dtemp; case SPH_F_DCONCN: return &q.dconcn;
    case SPH_F_INV_CM_XX: return &q.inv_cm_xx; case SPH_F_INV_CM_YY: return &q.inv_cm_yy;
    case SPH_F_INV_CM_ZZ: return &q.inv_cm_zz; case SPH_F_INV_CM_XY: return &q.inv_cm_xy;
    case SPH_F_INV_CM_YZ: return &q.inv_cm_yz; case SPH_F_INV_CM_ZX: return &q.inv_cm_zx;
    case SPH_F_NX: return &o.nx[i]; case SPH_F_NY: return &o.ny[i]; case SPH_F_NZ: return &o.nz[i];
    }
    return nullptr;
}

// Passes of the step that oracle_run_pass runs one at a time on the
// working state (the step's SP1 / SP2 / P3), for the per-pass pins.
enum {
    OP_GRAD_CORR = 0, OP_PREP, OP_BOUNDARY, OP_MOST, OP_NEUMANN, OP_ADVECTION, OP_PROJECTION,
    OP_PERIODIC, OP_FREESLIP, OP_PPE, OP_OUTLET, OP_PRESSURE_FORCE, OP_IBM_FORCE, OP_IBM_SPREAD, OP_IB_NORMALS,
    OP_COUNT
};

}  // namespace

// ------------------------------------------------------------ C API ------
extern "C" {

void oracle_params_default(sph_params* p) {
    // Reference defaults: Parameters.cuh, functions_PROP.cuh:121-131,
    // functions_NNPS.cuh:4-14, functions_PPE.cuh:268-271.
    std::memset(p, 0, sizeof(*p));
    p->dim = 3; p->kernel_type = SPH_WENDLAND2; p->time_type = SPH_PRE_COR; p->freq_filt = 1;
    p->fp_solve = 1; p->fv_solve = SPH_FV_CLEARY; p->fg_solve = 0; p->con_solve = 0;
    p->boussinesq_solve = 0; p->kgc_solve = 0; p->turbulence_model = SPH_LAMINAR; p->concn_solve = 0;
    p->noslip_bc = 1; p->artificial_viscosity = 1;
    p->kappa = 2.0; p->u_limit = 1000.0;
    p->mu_fluid = 1.8e-5; p->mu_wall = 1.0e-2; p->gravity = 9.810;
    p->T_ini = 300.0; p->T_con = 300.0; p->T_start_t = 0.0; p->T_con_t = 10.0; p->tempb = 300.0;
    p->Cs_sgs = 0.15; p->Ck_sgs = 0.1; p->Ce_sgs = 0.93; p->rho_air = 1.225;
    p->av_alpha = 0.05; p->av_beta = 0.05; p->av_c = 30.0;
    p->solid_k = 1.0; p->air_k = 0.025; p->solid_cp = 900.0; p->air_cp = 1005.0; p->solid_rho = 2300.0;
    p->diff_coef = 2.0;
    p->L2 = 0.028; p->L3 = 13.304; p->L_left = -0.412; p->L_right = 0.412;
    p->outlet_copy_dx = 0.008; p->copy_tol = 0.001; p->neumann_ghost_r = 0.1;
    p->most_mode = SPH_MOST_REFERENCE; p->most_z0 = 1.0; p->most_z0h = 0.1; p->most_L = 1e6;
    p->most_ghost_dz = 5.0; p->most_ghost_r = 0.1;
    p->ibm_ut_factor = 0.2;
    p->ppe_solver = SPH_PPE_REF_JACOBI; p->precision = SPH_FP64; p->ppe_tol = 1e-8; p->ppe_maxit = 5000;
    p->ppe_inner_tol = 1e-5; p->ppe_max_refine = 20; p->ppe_index16 = 1; p->ppe_reorder = 1; p->apr_h_change = 1;
    p->ale_positions = 0;
    p->ppe_precond = SPH_PRECOND_JACOBI;   // the library's multigrid option; the oracle solves with Jacobi
    // IB normals: the face planes of functions_PREP.cuh:958-963
    p->ib_normals = SPH_IBN_REFERENCE;
    p->ib_face_xlo = 0.764; p->ib_face_xhi = 0.836; p->ib_face_ylo = -0.036; p->ib_face_yhi = 0.036;
    p->ib_face_zlo = 0.0; p->ib_face_zhi = 0.152; p->ib_face_tol = 1e-5;
    // the library's LDS-staged record passes (the oracle has no such choice)
    p->tile_cells = 0; p->tile_particles = 128; p->tile_halo_max = 1536; p->compact_records = 1; p->ppe_stencil = 1;
}

void* oracle_create(const sph_params* prm, int n) {
    if (!prm || n <= 0) return nullptr;
    Oracle* o = new Oracle();
    o->prm = *prm;
    o->cap = n;
    o->c_pi15 = pow(PI_REF, 1.5);
    return o;
}

int oracle_upload(void* h, const sph_part1* P1, const sph_part2* P2, const float* normals, int n) {
    Oracle& o = *(Oracle*)h;
    if (n <= 0 || n > o.cap) return SPH_E_ARG;
    o.n = n;
    o.P1.assign(n, P1d{}); o.P2.assign(n, P2d{});
    o.nx.assign(n, 0.0); o.ny.assign(n, 0.0); o.nz.assign(n, 0.0);
    for (int i = 0; i < n; i++) {
        const sph_part1& a = P1[i];
        P1d& b = o.P1[i];
        b.i_type = a.i_type; b.buffer_type = a.buffer_type; b.MOST_buffer = a.MOST_buffer; b.p_type = a.p_type;
        b.x = a.x; b.y = a.y; b.z = a.z; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.m = a.m; b.h = a.h;
        b.temp = a.temp; b.pres = a.pres; b.pres0 = a.pres0; b.rho = a.rho; b.flt_s = a.flt_s;
        b.enthalpy = a.enthalpy; b.concn = a.concn; b.concentration = a.concentration;
        b.k_turb = a.k_turb; b.e_turb = a.e_turb; b.eli = a.eli; b.vol = a.vol; b.m_ref = a.m_ref;
        b.h_ref = a.h_ref; b.vis_t = a.vis_t; b.fbx = a.fbx; b.fby = a.fby; b.fbz = a.fbz;
        b.fmx = a.fmx; b.fmy = a.fmy; b.hmz = a.hmz; b.float1 = a.float1; b.float2 = a.float2;
        b.float3 = a.float3; b.ncell = a.ncell; b.source = a.source;
        if (!(a.h_ref > 0.0f) || !(a.h > 0.0f)) return SPH_E_ARG;
        if (P2) {
            const sph_part2& c = P2[i];
            P2d& e = o.P2[i];
            e.rho_ref = c.rho_ref; e.SR = c.SR; e.x0 = c.x0; e.y0 = c.y0; e.z0 = c.z0;
            e.ux0 = c.ux0; e.uy0 = c.uy0; e.uz0 = c.uz0; e.rho0 = c.rho0; e.drho0 = c.drho0;
            e.temp0 = c.temp0; e.concn0 = c.concn0; e.enthalpy0 = c.enthalpy0;
        }
        if (normals) { o.nx[i] = normals[3 * i]; o.ny[i] = normals[3 * i + 1]; o.nz[i] = normals[3 * i + 2]; }
    }
    size_domain(o);
    o.uploaded = true;
    o.searched = false;
    return 0;
}

int oracle_apr_refine(void* h, const double* target, int max_passes, int* n_out) {
    Oracle& o = *(Oracle*)h;
    if (!o.uploaded) return SPH_E_STATE;
    return apr_refine(o, target, max_passes, n_out);
}

int oracle_step(void* h, int count, double dt, double time) {
    Oracle& o = *(Oracle*)h;
    if (!o.uploaded) return SPH_E_STATE;
    return step(o, count, dt, time);
}

// Test hooks: the neighbour search alone (set_ncell, hash, sort, cell
// bounds, reorder; P3 cleared), then single passes on the working state,
// whose fields can be written and read in the sorted order.
int oracle_search(void* h) {
    Oracle& o = *(Oracle*)h;
    if (!o.uploaded) return SPH_E_STATE;
    set_ncell(o);
    nnps(o);
    o.P3.assign(o.n, P3d{});
    return 0;
}

int oracle_run_pass(void* h, int pass, int count, double dt, double time) {
    Oracle& o = *(Oracle*)h;
    if (!o.searched) return SPH_E_STATE;
    switch (pass) {
    case OP_GRAD_CORR: gradient_correction(o); break;
    case OP_PREP: prep(o, count, dt); break;
    case OP_BOUNDARY: boundary3D(o); break;
    case OP_MOST: most(o, count, dt); break;
    case OP_NEUMANN: neumann(o, o.SP1); break;
    case OP_ADVECTION: advection(o); break;
    case OP_PROJECTION: projection(o, dt, time); break;
    case OP_PERIODIC: periodic(o, o.SP1); break;
    case OP_FREESLIP: freeslip(o, o.SP1); break;
    case OP_PPE: return ppe(o, dt);
    case OP_OUTLET: outlet(o, o.SP1); break;
    case OP_PRESSURE_FORCE: pressure_force(o); break;
    case OP_IBM_FORCE: ibm_force(o, dt); break;
    case OP_IBM_SPREAD: ibm_spread(o); break;
    case OP_IB_NORMALS: ib_normals(o, count); break;
    default: return SPH_E_ARG;
    }
    return 0;
}

int oracle_get_work_f64(void* h, int field, double* out, int n) {
    Oracle& o = *(Oracle*)h;
    if (!o.searched || n != o.n || field < 0 || field >= SPH_F_COUNT_F64) return SPH_E_ARG;
    for (int i = 0; i < n; i++) out[i] = *work_ref(o, field, i);
    return 0;
}

int oracle_set_work_f64(void* h, int field, const double* in, int n) {
    Oracle& o = *(Oracle*)h;
    if (!o.searched || n != o.n || field < 0 || field >= SPH_F_COUNT_F64) return SPH_E_ARG;
    for (int i = 0; i < n; i++) *work_ref(o, field, i) = in[i];
    return 0;
}

int oracle_set_work_i32(void* h, int field, const int* in, int n) {
    Oracle& o = *(Oracle*)h;
    if (!o.searched || n != o.n) return SPH_E_ARG;
    for (int i = 0; i < n; i++) {
        P1d& p = o.SP1[i];
        switch (field) {
        case SPH_I_I_TYPE: p.i_type = in[i]; break;
        case SPH_I_BUFFER_TYPE: p.buffer_type = in[i]; break;
        case SPH_I_MOST_BUFFER: p.MOST_buffer = in[i]; break;
        case SPH_I_P_TYPE: p.p_type = in[i]; break;
        default: return SPH_E_ARG;
        }
    }
    return 0;
}

// Planted mistake for the mutation tests (0 = none); returns the previous.
int oracle_set_mutant(int k) {
    int prev = g_mutant;
    g_mutant = k;
    return prev;
}

int oracle_get_field_f64(void* h, int field, double* out, int n) {
    Oracle& o = *(Oracle*)h;
    if (n != o.n || field < 0 || field >= SPH_F_COUNT_F64) return SPH_E_ARG;
    for (int i = 0; i < n; i++) out[i] = field_value(o, field, i);
    return 0;
}

int oracle_get_field_i32(void* h, int field, int* out, int n) {
    Oracle& o = *(Oracle*)h;
    if (n != o.n) return SPH_E_ARG;
    for (int i = 0; i < n; i++) {
        const P1d& p = o.P1[i];
        switch (field) {
        case SPH_I_I_TYPE: out[i] = p.i_type; break;
        case SPH_I_BUFFER_TYPE: out[i] = p.buffer_type; break;
        case SPH_I_MOST_BUFFER: out[i] = p.MOST_buffer; break;
        case SPH_I_P_TYPE: out[i] = p.p_type; break;
        case SPH_I_NCELL: out[i] = p.ncell; break;
        case SPH_I_SOURCE: out[i] = p.source; break;
        default: return SPH_E_ARG;
        }
    }
    return 0;
}

int oracle_nnps(void* h, int* keys, int* perm, int* gstr, int* gend) {
    Oracle& o = *(Oracle*)h;
    if (!o.searched) return SPH_E_STATE;
    if (keys) std::memcpy(keys, o.keys.data(), sizeof(int) * o.n);
    if (perm) std::memcpy(perm, o.perm.data(), sizeof(int) * o.n);
    if (gstr) std::memcpy(gstr, o.gstr.data(), sizeof(int) * o.d.num_cells);
    if (gend) std::memcpy(gend, o.gend.data(), sizeof(int) * o.d.num_cells);
    return 0;
}

int oracle_domain(void* h, sph_domain* d) {
    Oracle& o = *(Oracle*)h;
    if (!o.uploaded) return SPH_E_STATE;
    *d = o.d;
    return 0;
}

// Neighbour lists of the last search by the window rules: kind 0 =
// dcell-query window, self included, tdist = r + 1e-20 < kappa*max(hi,hj);
// kind 1 = binning window, 0 < r < kappa*hi.  Positions are those the search
// sorted (SP1): in ALE mode the step moves particles after its search.
int oracle_neighbors(void* h, int kind, int* off, int* nbr, long long* total) {
    Oracle& o = *(Oracle*)h;
    if (!o.searched) return SPH_E_STATE;
    const auto& S = o.SP1;
    std::vector<std::vector<int>> L(o.n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < o.n; i++) {
        double xi = S[i].x, yi = S[i].y, zi = S[i].z, hi = S[i].h;
        if (kind == 0) {
            Cell c = query_cell(o.d, xi, yi, zi);
            window(o, c, S[i].ncell, [&](int j) {
                double t = dist3(xi, yi, zi, S[j].x, S[j].y, S[j].z) + 1e-20;
                if (t < o.prm.kappa * fmax(hi, S[j].h)) L[i].push_back(j);
            });
        } else {
            Cell c = bin_cell(o.d, xi, yi, zi);
            window(o, c, S[i].ncell, [&](int j) {
                double t = dist3(xi, yi, zi, S[j].x, S[j].y, S[j].z);
                if (t > 0 && t < o.prm.kappa * hi) L[i].push_back(j);
            });
        }
    }
    long long tot = 0;
    for (int i = 0; i < o.n; i++) tot += (long long)L[i].size();
    if (total) *total = tot;
    if (off) {
        long long s = 0;
        for (int i = 0; i < o.n; i++) { off[i] = (int)s; s += (long long)L[i].size(); }
        off[o.n] = (int)s;
    }
    if (nbr) {
        long long s = 0;
        for (int i = 0; i < o.n; i++) for (int j : L[i]) nbr[s++] = j;
    }
    return 0;
}

int oracle_ppe_info(void* h, sph_ppe_report* r) {
    *r = ((Oracle*)h)->rep;
    return 0;
}

// Last Krylov system: rows (particle index per row), diag, rhs, CSR of the
// off-diagonal unknown couplings, initial guess and solution.
int oracle_ppe_system(void* h, int* nrows, long long* nnz, int* rows, double* diag, double* rhs,
                      int* rowptr, int* cols, double* vals, double* x0, double* x) {
    Oracle& o = *(Oracle*)h;
    long nr = (long)o.sys_rows.size();
    if (nrows) *nrows = (int)nr;
    if (nnz) *nnz = nr ? o.sys_rowptr[nr] : 0;
    if (rows) std::memcpy(rows, o.sys_rows.data(), sizeof(int) * nr);
    if (diag) std::memcpy(diag, o.sys_diag.data(), sizeof(double) * nr);
    if (rhs) std::memcpy(rhs, o.sys_rhs.data(), sizeof(double) * nr);
    if (rowptr) std::memcpy(rowptr, o.sys_rowptr.data(), sizeof(int) * (nr + 1));
    if (cols && nr) std::memcpy(cols, o.sys_cols.data(), sizeof(int) * o.sys_rowptr[nr]);
    if (vals && nr) std::memcpy(vals, o.sys_vals.data(), sizeof(double) * o.sys_rowptr[nr]);
    if (x0) std::memcpy(x0, o.sys_x0.data(), sizeof(double) * nr);
    if (x) std::memcpy(x, o.sys_x.data(), sizeof(double) * nr);
    return 0;
}

// kernel_copy_max + Max reductions, functions_OUTPUT.cuh:19-33.
int oracle_stats(void* h, float* umax, float* drho, float* fmax_) {
    Oracle& o = *(Oracle*)h;
    double mu = 0, mr = 0, mf = 0;
    bool first = true;
    for (int i = 0; i < o.n; i++) {
        const P1d& p = o.P1[i];
        double a = 0, b = 0, c = 0;
        if (!(p.p_type == 0 || p.i_type >= 2)) {
            a = sqrt(p.ux * p.ux + p.uy * p.uy + p.uz * p.uz);
            b = (p.rho - o.SP2[i].rho_ref) / o.SP2[i].rho_ref * 100.0;
            c = o.P3[i].ftotal;
        }
        if (first || a > mu) mu = a;
        if (first || b > mr) mr = b;
        if (first || c > mf) mf = c;
        first = false;
    }
    *umax = (float)mu; *drho = (float)mr; *fmax_ = (float)mf;
    return 0;
}

void oracle_destroy(void* h) { delete (Oracle*)h; }

double oracle_canon_sum(const double* v, long n) { return canon_sum(v, n); }

double oracle_tmpA(int type, double kappa, double h) { return calc_tmpA(type, kappa, pow(PI_REF, 1.5), h); }
double oracle_kernel_w(int type, double kappa, double h, double r) {
    return kernel_w(type, kappa, calc_tmpA(type, kappa, pow(PI_REF, 1.5), h), h, r);
}
double oracle_kernel_dw(int type, double kappa, double h, double r) {
    return kernel_dw(type, kappa, calc_tmpA(type, kappa, pow(PI_REF, 1.5), h), h, r);
}
double oracle_kernel_wp(int type, double kappa, double h, double r) {
    return kernel_wp(kappa, calc_tmpA(type, kappa, pow(PI_REF, 1.5), h), h, r);
}
double oracle_kernel_dwp(int type, double kappa, double h, double r) {
    return kernel_dwp(kappa, calc_tmpA(type, kappa, pow(PI_REF, 1.5), h), h, r);
}
double oracle_interp_ux_caseH(double z) { return interp_ux_caseH(z, 1.0); }
double oracle_phi_m(double L, double z, double z0) { return Phi_m(L, z, z0); }

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"
