// Device-side SPH math: kernel functions, cell formulas, Case-H inflow
// table, Monin-Obukhov stability functions.  R is the compute type (double
// in FP64 mode, float for the mixed-precision interaction sums).  Every
// expression keeps the reference's operand order; the library is built with
// -ffp-contract=off so R=double reproduces IEEE double arithmetic of the
// same expression evaluated anywhere else.
#pragma once

#include <hip/hip_runtime.h>

#include "sph_internal.h"

namespace sph {

constexpr double PI_REF = 3.141592653;   // Parameters.cuh:41

template <class R> __device__ __forceinline__ R p4(R a) { return a * a * a * a; }
template <class R> __device__ __forceinline__ R p5(R a) { return a * a * a * a * a; }

// calc_tmpA, functions_KNL.cuh:2-26 (3-D)
template <class R>
__device__ __forceinline__ R calc_tmpA(const KConst& k, R tH) {
    const R kap = (R)k.kappa;
    switch (k.type) {
    case SPH_GAUSSIAN: return (R)1.0 / ((R)k.c_pi15 * (tH * tH * tH));
    case SPH_QUINTIC: return (R)3.0 / ((R)359.0 * (R)PI_REF * (tH * tH * tH));
    case SPH_QUARTIC: return (R)315.0 / ((R)208.0 * (R)PI_REF * (tH * tH * tH));
    case SPH_WENDLAND2: return (R)3.342253804929802 / (kap * tH) / (kap * tH) / (kap * tH);
    case SPH_WENDLAND4: return (R)4.923856051905513 / (kap * tH) / (kap * tH) / (kap * tH);
    case SPH_WENDLAND6: return (R)6.788953041263660 / (kap * tH) / (kap * tH) / (kap * tH);
    }
    return (R)0;
}

// calc_kernel_wij, functions_KNL.cuh:41-71 (3-D)
template <class R>
__device__ __forceinline__ R kernel_w(const KConst& k, R tA, R tH, R rr) {
    const R kap = (R)k.kappa;
    R tR = rr / tH / kap;
    const R one = (R)1;
    switch (k.type) {
    case SPH_GAUSSIAN: { R a = (R)2.0 * tR; return tA * exp(-(a * a)); }
    case SPH_QUINTIC: {
        R a = (R)3.0 - (R)2.0 * tR, b = (R)2.0 - (R)2.0 * tR, c = (R)1.0 - (R)2.0 * tR;
        if (tR < (R)0.5) return tA * (p5(a) - (R)6.0 * p5(b) + (R)15.0 * p5(c));
        else if ((R)0.5 <= tR && tR < one) return tA * (p5(a) - (R)6.0 * p5(b));
        else if (one <= tR && tR < (R)1.5) return tA * p5(a);
        return (R)0;
    }
    case SPH_QUARTIC:
        return (R)(tR < one) * tA * ((R)(2.0 / 3.0) - (R)(9.0 / 2.0) * (tR * tR) + (R)(19.0 / 3.0) * (tR * tR * tR) - (R)(5.0 / 2.0) * (tR * tR * tR * tR));
    case SPH_WENDLAND2:
        return (R)(tR < one) * tA * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one + (R)4 * tR);
    case SPH_WENDLAND4:
        return (R)(tR < one) * tA * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) *
               (one + (R)6 * tR + (R)11.666666666666666 * tR * tR);
    case SPH_WENDLAND6:
        return (R)(tR < one) * tA * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) *
               (one + (R)8 * tR + (R)25 * tR * tR + (R)32 * tR * tR * tR);
    }
    return (R)0;
}

// calc_kernel_dwij, functions_KNL.cuh:73-103 (3-D).  The leading factor
// 1/h (Gaussian, quintic, quartic) or 1/(kappa h) (Wendland) depends on h
// only: dw_factor computes it, kernel_dw_f takes it, so that a caller whose
// pairs share h evaluates it once (the same value in the same position of
// the product).
template <class R>
__device__ __forceinline__ R dw_factor(const KConst& k, R tH) {
    const R one = (R)1;
    return (k.type == SPH_GAUSSIAN || k.type == SPH_QUINTIC || k.type == SPH_QUARTIC) ? one / tH
                                                                                       : one / ((R)k.kappa * tH);
}
template <class R>
__device__ __forceinline__ R kernel_dw_f(const KConst& k, R tA, R tH, R f, R rr) {
    const R kap = (R)k.kappa;
    R tR = rr / tH / kap;
    const R one = (R)1;
    switch (k.type) {
    case SPH_GAUSSIAN: { R a = (R)2.0 * tR; return f * tA * (R)(-2.0) * (R)2.0 * tR * exp(-(a * a)); }
    case SPH_QUINTIC: {
        R a = (R)3.0 - (R)2.0 * tR, b = (R)2.0 - (R)2.0 * tR, c = (R)1.0 - (R)2.0 * tR;
        if (tR < (R)0.5) return f * tA * ((R)(-5.0) * p4(a) + (R)30.0 * p4(b) - (R)75.0 * p4(c));
        else if ((R)0.5 <= tR && tR < one) return f * tA * ((R)(-5.0) * p4(a) + (R)30.0 * p4(b));
        else if (one <= tR && tR < (R)1.5) return f * tA * ((R)(-5.0) * p4(a));
        return (R)0;
    }
    case SPH_QUARTIC:
        return (R)(tR < one) * f * tA * ((R)(-9.0 / 2.0) * (R)2 * tR + (R)(19.0 / 3.0) * (R)3.0 * (tR * tR) - (R)(5.0 / 2.0) * (R)4.0 * (tR * tR * tR));
    case SPH_WENDLAND2:
        return (R)(tR < one) * f * tA * ((R)(-20) * tR * (one - tR) * (one - tR) * (one - tR));
    case SPH_WENDLAND4:
        return (R)(tR < one) * f * tA *
               ((R)(-18.666666666666668) * tR * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * ((R)5 * tR + one));
    case SPH_WENDLAND6:
        return (R)(tR < one) * f * tA *
               ((R)(-22) * tR * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * ((R)16 * tR * tR + (R)7 * tR + one));
    }
    return (R)0;
}
template <class R>
__device__ __forceinline__ R kernel_dw(const KConst& k, R tA, R tH, R rr) {
    return kernel_dw_f(k, tA, tH, dw_factor(k, tH), rr);
}

// calc_kernel_dwij_pressure, functions_KNL.cuh:355-369: Wendland-6 shape,
// caller's normalisation.
template <class R>
__device__ __forceinline__ R kernel_dwp(const KConst& k, R tA, R tH, R rr) {
    const R kap = (R)k.kappa;
    R tR = rr / tH / kap;
    const R one = (R)1;
    return (R)(tR < one) * (one / (kap * tH)) * tA *
           ((R)(-22) * tR * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * (one - tR) * ((R)16 * tR * tR + (R)7 * tR + one));
}

// Cell formulas (float arithmetic, as the reference computes them).
// Binning: functions_NNPS.cuh:891-900.  Query: e.g. functions_PPE.cuh:52-57.
__device__ __forceinline__ int bin_axis(float x, float mn, float mx, int N) {
    if (mx == mn) return 0;
    float t = floorf(__fdiv_rn(x - mn, mx - mn) * (float)N);
    float r = (t < (float)(N - 1)) ? t : (float)(N - 1);
    int c = (int)r;
    return c < 0 ? 0 : c;
}
__device__ __forceinline__ int query_axis(float x, float mn, float mx, float dcell, int N) {
    if (mx == mn) return 0;
    float t = floorf(__fdiv_rn(x - mn, dcell));
    float r = (t < (float)(N - 1)) ? t : (float)(N - 1);
    int c = (int)r;
    return c < 0 ? 0 : c;
}
struct Cell { int i, j, k; };
__device__ __forceinline__ Cell bin_cell(const sph_domain& d, double x, double y, double z) {
    return {bin_axis((float)x, d.x_min, d.x_max, d.NI), bin_axis((float)y, d.y_min, d.y_max, d.NJ),
            bin_axis((float)z, d.z_min, d.z_max, d.NK)};
}
__device__ __forceinline__ Cell query_cell(const sph_domain& d, double x, double y, double z) {
    return {query_axis((float)x, d.x_min, d.x_max, d.dcell, d.NI),
            query_axis((float)y, d.y_min, d.y_max, d.dcell, d.NJ),
            query_axis((float)z, d.z_min, d.z_max, d.dcell, d.NK)};
}
__device__ __forceinline__ bool cell_ok(const sph_domain& d, int i, int j, int k) {
    return !(i < 0 || i > d.NI - 1 || j < 0 || j > d.NJ - 1 || k < 0 || k > d.NK - 1);
}

template <class R>
__device__ __forceinline__ R dist3(R xi, R yi, R zi, R xj, R yj, R zj) {
    return sqrt((xi - xj) * (xi - xj) + (yi - yj) * (yi - yj) + (zi - zj) * (zi - zj));
}

// viscosity2, functions_PROP.cuh:121-131 (constants from parameters)
__device__ __forceinline__ double viscosity2(const sph_params& p, int p_type) {
    if (p_type == 1) return p.mu_fluid;
    if (p_type == 0) return p.mu_wall;
    if (p_type == 1000) return p.mu_wall;
    return p.mu_fluid;
}

// Device_Interpolate_ux_caseH, functions_IBM.cuh:288-338
__device__ __forceinline__ double interp_ux_caseH(double z_inp, double scale) {
    const double CZ[30] = {0.0050000, 0.0100000, 0.0200000, 0.0400000, 0.0600000, 0.0800000,
                           0.1000000, 0.1200000, 0.1400000, 0.1600000, 0.1650000, 0.1700000,
                           0.1800000, 0.1900000, 0.2100000, 0.2300000, 0.2500000, 0.3000000,
                           0.3500000, 0.4000000, 0.4500000, 0.5000000, 0.5500000, 0.6000000,
                           0.6528736, 0.7724136, 0.8390808, 0.8689656, 0.8804600, 0.8919544};
    const double CU[30] = {2.745, 2.935, 3.175, 3.435, 3.627, 3.824, 4.021, 4.210, 4.362, 4.491,
                           4.502, 4.586, 4.606, 4.712, 4.854, 4.993, 5.132, 5.449, 5.782, 6.077,
                           6.338, 6.588, 6.693, 6.751, 6.751, 6.751, 6.751, 6.751, 6.751, 6.751};
    const int N = 30;
    const double z = scale * z_inp;
    if (z <= CZ[0]) {
        const double dz = CZ[1] - CZ[0], du = CU[1] - CU[0];
        return CU[0] + (z - CZ[0]) * (du / (dz + 1e-300));
    }
    if (z >= CZ[N - 1]) {
        const double dz = CZ[N - 1] - CZ[N - 2], du = CU[N - 1] - CU[N - 2];
        return CU[N - 2] + (z - CZ[N - 2]) * (du / (dz + 1e-300));
    }
    int lo = 0, hi = N - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (z < CZ[mid]) hi = mid; else lo = mid;
    }
    const double z0 = CZ[lo], z1 = CZ[hi], u0 = CU[lo], u1 = CU[hi];
    const double t = (z - z0) / (z1 - z0 + 1e-300);
    return u0 + t * (u1 - u0);
}

// Businger-Dyer stability functions, functions_MOST.cuh:151-173 (double)
__device__ __forceinline__ double psi_m(double zeta) {
    if (zeta < 0.0) {
        double x = sqrt(sqrt(1.0 - 16.0 * zeta));
        return M_PI * 0.5 - 2.0 * atan(x) + log((1.0 + x) * (1.0 + x) * (1.0 + x * x) * 0.125);
    }
    return -5.0 * zeta;
}
__device__ __forceinline__ double psi_h(double zeta) {
    if (zeta < 0.0) return 2.0 * log(0.5 * (1.0 + sqrt(1.0 - 16.0 * zeta)));
    return -5.0 * zeta;
}
__device__ __forceinline__ double Phi_m(double L, double z, double z0) { return log(z / z0) - psi_m(z / L) + psi_m(z0 / L); }
__device__ __forceinline__ double Phi_h(double L, double z, double z0h) { return log(z / z0h) - psi_h(z / L) + psi_h(z0h / L); }

// functions_TIME_ISPH.cuh:287-291, functions_MOST.cuh:199-202
__device__ __forceinline__ double wall_temperature(const sph_params& p, double dt, int count) {
    double t = dt * count;
    if (t >= p.T_start_t && t < p.T_con_t) return p.T_ini + dt * count * (p.T_con - p.T_ini) / (p.T_con_t - p.T_start_t);
    else if (t < p.T_start_t) return p.T_ini;
    return p.T_con;
}

}  // namespace sph
