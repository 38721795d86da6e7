// radiation.hip -- the Radiation integrator on the GPU (photoionisation by
// one star; src/Integrators/Radiation.cpp).
//
//   rad_gather_kernel        Q = fixPrimitives(QfromU(U)) of den, pre, hii, adv
//                            (+ Torch::checkValues before a sub-step)
//   rad_init_field_kernel    Radiation::initField          (Radiation.cpp:78-96)
//   rad_gather_heat_kernel   the gather + preTimeStepCalculations (:257-296) in one pass
//   rad_dt_kernel            calculateTimeStep              (:298-351), block min + one atomic
//   rad_sweep_kernel<IMPL>   transferRadiation's causal loop (:569-651) for the implicit
//                            scheme (IMPL) / Radiation::rayTrace (:480-508) for implicit2 and
//                            explicit, one launch per Chebyshev shell around the star
//   rad_hii_kernel           the per-cell update_HIIfrac loop of transferRadiation2 (:510-567)
//   rad_source_kernel        updateSourceTerms (:653-680) + Fluid::advSolution + fixSolution
//
// The causal sweep runs shell by shell: every upstream neighbour of a cell
// lies one Chebyshev shell closer to the star (see raytrace.hip), so each
// launch reads only what the previous one wrote.  Per-cell state R [6][n]
// (RID order) and the working primitives q4 [4][n] (den, pre, hii, adv)
// stay in HBM between calls; the geometry comes from the cell's indices.
// exp / pow / log come from the device math library, so results agree with
// the reference to rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "../../include/torchgpu.h"
#include "hydro_device.hpp"
#include "hydro_math.hpp"
#include "star_device.hpp"
#include "torchgpu_internal.hpp"

namespace tg {

// Scalars of torchgpu_rad_params the kernels use, plus the star's photon
// rate and the first-call cell of calculateTimeStep.
struct RadConsts {
    double K[4];
    double sigma, tau0, T_hi, T_hii;
    int scheme, coupling, collisions_on, use_once;
    long long once_cell;
    double heat_amp, X, mu_c, mh, Rgas, kB, excess;
    double vA, vX, vK, vP, vDE;
    double photon_rate, T_min;
    double rec_slope0, rec_slope1, cool_slope0, cool_slope1;
};

// tables: x [31], rec_y [31], rec_y2 [31], cool_y [31], cool_y2 [31]
constexpr int TAB_X = 0, TAB_REC = 31, TAB_REC2 = 62, TAB_COOL = 93, TAB_COOL2 = 124, TAB_N = 155;
constexpr int R_HII_A = 0, R_TAU = 1, R_TAU_A = 2, R_DTAU = 3, R_DTAU_A = 4, R_HEAT = 5;

// LinearSplineData::interpolate + SplineData::splint (SplineData.cpp:57-81,
// 108-119): cubic spline between the knots (bisection as the reference),
// power laws beyond them.
__device__ __forceinline__ double lin_spline(const double* __restrict__ tab, int ycol, int y2col, double s0, double s1,
                                             double v) {
    const double* x = tab + TAB_X;
    const double* y = tab + ycol;
    const double* y2 = tab + y2col;
    if (v > x[30]) return y[30] * pow(v / x[30], s1);
    if (v < x[0]) return y[0] * pow(v / x[0], s0);
    int klo = 0, khi = 30;
    while (khi - klo > 1) {
        const int k = (khi + klo) >> 1;
        if (x[k] > v) khi = k;
        else klo = k;
    }
    const double h = x[khi] - x[klo];
    const double a = (x[khi] - v) / h;
    const double b = (v - x[klo]) / h;
    return a * y[klo] + b * y[khi] + ((a * a * a - a) * y2[klo] + (b * b * b - b) * y2[khi]) * (h * h) / 6.0;
}

__device__ __forceinline__ double rad_alpha(const RadConsts& r, const double* tab, double T) {
    return lin_spline(tab, TAB_REC, TAB_REC2, r.rec_slope0, r.rec_slope1, T);
}

__device__ __forceinline__ double rad_cooling(const RadConsts& r, const double* tab, double nH, double x, double T) {
    const double rate = lin_spline(tab, TAB_COOL, TAB_COOL2, r.cool_slope0, r.cool_slope1, T);
    return x * x * nH * nH * r.kB * T * rate;
}

// Voronov (1997) collisional ionisation rate (Radiation.cpp:236-241)
__device__ __forceinline__ double rad_collisional(const RadConsts& r, double T) {
    if (!r.collisions_on || T < 5.0e3) return 0.0;
    const double U = r.vDE / T;
    return r.vA * (1. + r.vP * sqrt(U)) * exp(r.vK * log(U) - U) / (r.vX + U);
}

__device__ __forceinline__ double rad_photoion(double nHI, double tau, double dtau, double sv, double S) {
    if (nHI == 0 || sv == 0) return 0.0;
    const double emT = exp(-tau), emdT = exp(-dtau);
    return S * emT * (1.0 - emdT) / (nHI * sv);
}

// The same with exp(-tau) given (loop-invariant in the implicit iteration).
__device__ __forceinline__ double rad_photoion_emT(double nHI, double emT, double dtau, double sv, double S) {
    if (nHI == 0 || sv == 0) return 0.0;
    const double emdT = exp(-dtau);
    return S * emT * (1.0 - emdT) / (nHI * sv);
}

__device__ __forceinline__ double rad_temperature(const RadConsts& r, double hii, double pre, double den) {
    const double mu_inv = r.X * (hii + 1.0) + r.mu_c;
    return (pre / den) * (1.0 / mu_inv) / r.Rgas;
}

// Radiation::doric (Radiation.cpp:140-182): the time-averaged (x_avg) and
// end-of-step (x) ionised fractions of the exact solution of
// dx/dt = (1 - x) (A_pi + n x A_ci) - n x^2 alpha at fixed rates.
__device__ __forceinline__ void rad_doric(double dt, double& x_avg, double& x, double Api, double nHII_aB,
                                          double nHII_Aci) {
    const double eps = 1.0e-8;
    const double old = x;
    const double xeq = nHII_aB == 0.0 ? 1.0 : (Api + nHII_Aci) / (Api + nHII_Aci + nHII_aB);
    const double inv_ti = Api + nHII_Aci + nHII_aB;
    if (dt * inv_ti < 1.0e-8) {
        x = smax(smin(old, 1.0), 0.0);
        if (1.0 - x < eps) x = 1.0 - eps;
        x_avg = x;
    } else {
        const double e = exp(-dt * inv_ti);
        x = xeq + (old - xeq) * e;
        x = smax(smin(x, 1.0), 0.0);
        if (1.0 - x < eps) x = 1.0 - eps;
        x_avg = xeq + (old - xeq) * (1.0 - e) / (dt * inv_ti);
        x_avg = smax(smin(x_avg, 1.0), 0.0);
        if (1.0 - x_avg < eps) x_avg = 1.0 - eps;
    }
}

// Radiation::update_HIIfrac (Radiation.cpp:375-478) for a non-wind cell on
// values in registers: the cell's density, pressure and HII fraction, its
// TAU (explicit scheme: also DTAU) and the time-averaged optical depth of
// the implicit schemes; returns the new HII fraction x and (implicit
// schemes) the time-averaged x_avg, false when the implicit iteration fails.
__device__ __forceinline__ bool rad_update_hii_v(const RadConsts& r, const double* tab, double den, double pre,
                                                 double q_hii, double tau, double dtau, double tau_avg, double ds,
                                                 double sv, double dt, double& x_out, double& x_avg_out) {
    const double n_H = r.X * den / r.mh;
    double x = q_hii, x_avg = x;
    const double T = rad_temperature(r, q_hii, pre, den);
    const double alphaB = rad_alpha(r, tab, T);
    const double A_ci = rad_collisional(r, T);
    if (r.scheme != TORCHGPU_RT_EXPLICIT) {
        const double em_tau_avg = exp(-tau_avg);
        int niter = 0, miter = 0;
        bool converged = false;
        while (!converged) {
            niter++;
            const double old = x_avg;
            x = q_hii;
            const double dtau_avg = (1.0 - x_avg) * n_H * r.sigma * ds;
            const double nHI = (1.0 - x_avg) * n_H;
            const double A_pi = rad_photoion_emT(nHI, em_tau_avg, dtau_avg, sv, r.photon_rate);
            const double nHII_aB = x_avg * n_H * alphaB;
            const double nHII_Aci = x_avg * n_H * A_ci;
            rad_doric(dt, x_avg, x, A_pi, nHII_aB, nHII_Aci);
            if (fabs((x_avg - old) / x_avg) < 1.0e-3 || x_avg < 1.0e-5 || A_pi == 0) converged = true;
            if (niter > 10005) {
                miter++;
                niter = 0;
                x_avg = 0.5 * (x_avg + old);
            }
            if (miter > 5 || x != x) return false;
        }
    } else {
        const double nHI = (1.0 - x) * (r.X * den / r.mh);
        const double A_pi = rad_photoion(nHI, tau, dtau, sv, r.photon_rate);
        const double nHII_aB = x_avg * n_H * alphaB;
        const double nHII_Aci = x_avg * n_H * A_ci;
        double dummy = x;
        rad_doric(dt, x, dummy, A_pi, nHII_aB, nHII_Aci);
    }
    x_out = x;
    x_avg_out = x_avg;
    return true;
}

// The same on the cell's stored state: reads q4 and R, writes HII_A
// (implicit schemes), the HII fraction and the ADV flag.
__device__ __forceinline__ bool rad_update_hii(const RadConsts& r, const double* tab, long long n, long long c,
                                               double ds, double sv, double dt, double* q4, double* R) {
    const bool expl = r.scheme == TORCHGPU_RT_EXPLICIT;
    const double tau_avg = expl ? 0.0 : (r.scheme == TORCHGPU_RT_IMPLICIT2 ? R[R_TAU * n + c] : R[R_TAU_A * n + c]);
    const double tau = expl ? R[R_TAU * n + c] : 0.0, dtau = expl ? R[R_DTAU * n + c] : 0.0;
    double x, x_avg;
    if (!rad_update_hii_v(r, tab, q4[c], q4[n + c], q4[2 * n + c], tau, dtau, tau_avg, ds, sv, dt, x, x_avg))
        return false;
    if (!expl) R[R_HII_A * n + c] = x_avg;
    q4[2 * n + c] = x;
    if (x > 0.001) q4[3 * n + c] = 1;
    return true;
}

__device__ __forceinline__ void cell_ijk(const StarGeom& s, long long c, int ijk[3]) { cell_ijk32(s, c, ijk); }

// The spline tables into LDS for the block (every thread must call it): the
// bisection of lin_spline is a chain of dependent table reads.
__device__ __forceinline__ const double* stage_tables(const double* __restrict__ tab, double* lds) {
    for (int i = threadIdx.x; i < TAB_N; i += blockDim.x) lds[i] = tab[i];
    __syncthreads();
    return lds;
}

// Q = fixPrimitives(QfromU(U)) of cell c into q4 (and q), after
// Torch::checkValues when `check`.
__device__ __forceinline__ void rad_gather_cell(const HydroGeom& g, const double* __restrict__ U,
                                                double* __restrict__ q4, StepState* st, int check, int check_old,
                                                long long n, long long c, double q[7]) {
    double u[7];
#pragma unroll
    for (int v = 0; v < 7; ++v) u[v] = U[v * n + c];
    if (check) {
        // Torch::checkValues: U finite, Fluid's current Q (the previous
        // gather's) with nonzero density and pressure
        bool bad = false;
#pragma unroll
        for (int v = 0; v < 7; ++v) bad = bad || u[v] != u[v] || isinf(u[v]);
        if (check_old) bad = bad || q4[c] == 0 || q4[n + c] == 0;
        if (bad) set_error(st, ERR_RAD_CHECK);
    }
    QfromU(q, u, g.gm1);
    fix_primitives(g, q, u[DEN]);
    q4[c] = q[DEN];
    q4[n + c] = q[PRE];
    q4[2 * n + c] = q[HII];
    q4[3 * n + c] = q[ADV];
}

__global__ void rad_gather_kernel(HydroGeom g, const double* __restrict__ U, double* __restrict__ q4, StepState* st,
                                  int check, int check_old, int gated) {
    if (gated && st->halted) return;
    const long long n = g.ncell;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        double q[7];
        rad_gather_cell(g, U, q4, st, check, check_old, n, c, q);
    }
}

// Torch::checkValues before a sub-step whose Q is the last gather's.
__global__ void rad_check_kernel(HydroGeom g, const double* __restrict__ U, const double* __restrict__ q4,
                                 StepState* st, int gated) {
    if (gated && st->halted) return;
    const long long n = g.ncell;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        bool bad = q4[c] == 0 || q4[n + c] == 0;
#pragma unroll
        for (int v = 0; v < 7; ++v) {
            const double u = U[v * n + c];
            bad = bad || u != u || isinf(u);
        }
        if (bad) set_error(st, ERR_RAD_CHECK);
    }
}

// Radiation::initField: DTAU and DTAU_A from the density and HII fraction
// of Q (q7: [7][n] primitives fixed as Torch::initialise fixes them, with
// U still zero; or NULL: the gathered q4).
__global__ void rad_init_field_kernel(HydroGeom g, StarGeom s, RadConsts r, const double* __restrict__ q7,
                                      const double* __restrict__ q4, double* __restrict__ R) {
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        int ijk[3];
        cell_ijk(s, c, ijk);
        double den = q4[c], hii = q4[2 * n + c];
        if (q7) {
            double q[7];
#pragma unroll
            for (int v = 0; v < 7; ++v) q[v] = q7[v * n + c];
            fix_primitives(g, q, 0.0);
            den = q[DEN];
            hii = q[HII];
        }
        const double ds = star_path_length(s, ijk);
        const double nH = r.X * den / r.mh;
        R[R_DTAU * n + c] = ((1.0 - hii) * nH) * r.sigma * ds;
        R[R_DTAU_A * n + c] = ((1.0 - R[R_HII_A * n + c]) * nH) * r.sigma * ds;
    }
}

// heat (may be null): the grid's kept GridCell::H terms, whose RHII and
// EUVH this pre-step sets (Radiation.cpp:278-279).
// preTimeStepCalculations (Radiation.cpp:257-296) of cell c from its
// density, pressure and HII fraction.
__device__ __forceinline__ void rad_heat_cell(const StarGeom& s, const RadConsts& r, const double* tab, long long n,
                                              long long c, double den, double pre, double hii, double* __restrict__ R,
                                              StepState* st, double* __restrict__ heat) {
    int ijk[3];
    cell_ijk(s, c, ijk);
    if (star_wind_cell(s, ijk)) return;
    {
        const double sv = star_shell_volume(s, ijk, star_path_length(s, ijk));
        const double n_H = (r.X * den / r.mh);
        const double T = rad_temperature(r, hii, pre, den);
        const double nHI = (1.0 - hii) * n_H;
        const double A_pi = rad_photoion(nHI, R[R_TAU * n + c], R[R_DTAU * n + c], sv, r.photon_rate);
        const double photoion = n_H * (1.0 - hii) * A_pi * r.excess;
        const double recomb = rad_cooling(r, tab, n_H, hii, T);
        const double coll = hii * (1.0 - hii) * n_H * n_H * rad_collisional(r, T);
        const double rate = photoion - recomb - coll;
        double soft = rate;
        if (T < r.T_min + 200 && rate < 0.0) soft = smin(0.0, soft * (T - r.T_min) / 200.0);
        soft *= r.heat_amp;
        R[R_HEAT * n + c] = soft;
        if (heat) {
            heat[HEAT_EUVH * n + c] = photoion * (soft / rate);
            heat[HEAT_RHII * n + c] = -recomb * (soft / rate);
        }
        if (rate != rate) set_error(st, ERR_RAD_NAN);
    }
}

// The gather and preTimeStepCalculations in one pass (the heating reads the
// primitives the gather just formed, from registers).
__global__ void rad_gather_heat_kernel(HydroGeom g, StarGeom s, RadConsts r, const double* __restrict__ gtab,
                                       const double* __restrict__ U, double* __restrict__ q4, double* __restrict__ R,
                                       StepState* st, int check, int check_old, int gated, double* __restrict__ heat) {
    if (gated && st->halted) return;
    __shared__ double stab[TAB_N];
    const double* tab = stage_tables(gtab, stab);
    const long long n = g.ncell;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        double q[7];
        rad_gather_cell(g, U, q4, st, check, check_old, n, c, q);
        rad_heat_cell(s, r, tab, n, c, q[DEN], q[PRE], q[HII], R, st, heat);
    }
}

__global__ __launch_bounds__(256) void rad_dt_kernel(StarGeom s, RadConsts r, const double* __restrict__ gtab,
                                                     const double* __restrict__ q4, const double* __restrict__ U,
                                                     const double* __restrict__ R, double dt_max,
                                                     unsigned long long* dtbits, StepState* st, unsigned long long* once,
                                                     int gated) {
    __shared__ double red[4];
    // in the device loop: no time step is chosen on the first step or after
    // the run halted, and dt_max is the run's
    if (gated && (st->halted || st->first)) return;
    __shared__ double stab[TAB_N];
    const double* tab = stage_tables(gtab, stab);
    if (gated) dt_max = st->dt_max;
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    double m = dt_max;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        int ijk[3];
        cell_ijk(s, c, ijk);
        if (star_wind_cell(s, ijk)) continue;
        const double den = q4[c], pre = q4[n + c], hii = q4[2 * n + c];
        double dtc = dt_max, dt1 = dt_max, dt2 = dt_max, dt3 = dt_max, dt4 = dt_max;
        if (r.coupling == TORCHGPU_RT_NEQ && R[R_HEAT * n + c] != 0) dtc = fabs(U[PRE * n + c] / R[R_HEAT * n + c]);
        if (r.K[0] != 0.0) {
            const double T = rad_temperature(r, hii, pre, den);
            if (r.use_once && c == r.once_cell && *once) {
                dt1 = r.K[0] * r.mh / (r.X * den * rad_alpha(r, tab, T));
                *once = 2;  // used: rad_once_commit_kernel clears it
            } else if (hii != 0) dt1 = r.K[0] * r.mh / (r.X * den * hii * rad_alpha(r, tab, T));
        }
        if (r.K[1] != 0.0) dt2 = dt_max;
        if (r.K[2] != 0.0 || r.K[3] != 0.0) {
            const double sv = star_shell_volume(s, ijk, star_path_length(s, ijk));
            const double nH = r.X * den / r.mh;
            const double nHI = (1.0 - hii) * nH;
            const double A_pi = rad_photoion(nHI, R[R_TAU * n + c], R[R_DTAU * n + c], sv, r.photon_rate);
            const double T = rad_temperature(r, hii, pre, den);
            const double A_ci = rad_collisional(r, T);
            const double A_rr = rad_alpha(r, tab, T);
            const double fr = (1.0 - hii) * (A_pi + hii * nH * A_ci) - hii * hii * nH * A_rr;
            if (fr != 0.0) {
                if (r.K[2] != 0.0) dt3 = r.K[2] * smax(0.05, 1.0 - hii) / fabs(fr);
                if (r.K[3] != 0.0) dt4 = r.K[3] * 1.0 / fabs(fr);
            }
        }
        const double mc = smin(dt_max, smin(dtc, smin(dt1, smin(dt2, smin(dt3, dt4)))));
        // the reference keeps the running minimum and throws at 0
        if (!(mc > 0.0)) {
            if (mc == 0.0 || mc < 0.0) set_error(st, ERR_RAD_DT);
            continue;
        }
        m = smin(m, mc);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
        atomicMin(dtbits, (unsigned long long)__double_as_longlong(b));
    }
}

// Face-contiguous copies of TAU, DTAU, TAU_A, DTAU_A for the x-fixed faces
// of the last two shells (see raytrace.hip's coldens_shell_kernel): fields
// [2 parities][2 sides][4][w * w], y fastest, frame centred on the star.
struct RadXFaces {
    double* buf;
    int K, w;
};

__device__ __forceinline__ long long rad_xface_index(const StarGeom& s, const RadXFaces& xf, int parity, int side,
                                                     int field, int y, int z) {
    return (((long long)(parity * 2 + side) * 4 + field) * xf.w + (z - s.s[2] + xf.K)) * xf.w + (y - s.s[1] + xf.K);
}

// One Chebyshev shell of the causal sweep.  IMPL: transferRadiation
// (TAU and TAU_A, the HII update, then DTAU and DTAU_A); otherwise
// Radiation::rayTrace (TAU, then DTAU and DTAU_A from the current HII).
// With tau_0 > 0 a neighbour of zero weight adds exactly 0 (0 / tau_0 x a
// finite tau), so it is not read, and x-face cells read their upstream
// neighbours from the previous shell's x-face copies (`fast`).
template <bool IMPL>
__global__ __launch_bounds__(256) void rad_sweep_kernel(StarGeom s, RadConsts r, const double* __restrict__ gtab,
                                                        ShellFaces f, int k, RadXFaces xf, double dt, double* q4,
                                                        double* R, StepState* st) {
    if (dt < 0.0 && st->halted) return;  // a DT_* code: the device loop's step
    __shared__ double stab[IMPL ? TAB_N : 1];
    const double* tab = IMPL ? stage_tables(gtab, stab) : gtab;
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= f.start[f.nf]) return;
    dt = stage_dtf(dt, st, 0);
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    int ijk[3];
    shell_cell(f, t, ijk);
    const long long c = star_flat(s, ijk[0], ijk[1], ijk[2]);
    const bool fast = r.tau0 > 0.0;
    const bool xface = fast && (ijk[0] - s.s[0] == k || s.s[0] - ijk[0] == k);
    const int side = ijk[0] > s.s[0] ? 1 : 0;
    double out[4];  // TAU, DTAU, TAU_A, DTAU_A of this cell after the sweep
    if (star_wind_cell(s, ijk)) {
        R[R_TAU * n + c] = 0;
        R[R_TAU_A * n + c] = 0;
        R[R_DTAU * n + c] = 0;
        R[R_DTAU_A * n + c] = 0;
        q4[2 * n + c] = 1;
        R[R_HII_A * n + c] = 1;
        out[0] = out[1] = out[2] = out[3] = 0.0;
    } else {
        // the cell's own state, read once and first, so that its (for
        // x-face cells strided) loads overlap the neighbour reads below
        const double den = q4[c];
        const double pre = IMPL ? q4[n + c] : 0.0;
        const double q_hii = q4[2 * n + c];
        const double hii_a_old = R[R_HII_A * n + c];
        double dist2 = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double e = star_offset(s, ijk, a);
            dist2 += e * e;
        }
        int nbc[4][3];
        bool valid[4];
        double w[4];
        star_neighbour_cells(s, ijk, nbc, valid, w);
        // Radiation::updateTauSC (Radiation.cpp:353-373)
#pragma unroll
        for (int pass = 0; pass < (IMPL ? 2 : 1); ++pass) {
            const int T_ = pass ? R_TAU_A : R_TAU, D_ = pass ? R_DTAU_A : R_DTAU;
            double v = 0;
            if (dist2 > 0.95) {
                double tau[4], wr[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    tau[q] = 0.0;
                    if (valid[q] && (!fast || w[q] != 0.0)) {
                        if (xface) {
                            const int p = (k - 1) & 1;
                            tau[q] = xf.buf[rad_xface_index(s, xf, p, side, 2 * pass, nbc[q][1], nbc[q][2])] +
                                     xf.buf[rad_xface_index(s, xf, p, side, 2 * pass + 1, nbc[q][1], nbc[q][2])];
                        } else {
                            const long long i = star_flat(s, nbc[q][0], nbc[q][1], nbc[q][2]);
                            tau[q] = R[T_ * n + i] + R[D_ * n + i];
                        }
                    }
                    wr[q] = w[q] / smax(r.tau0, tau[q]);
                }
                const double sum_w = wr[0] + wr[1] + wr[2] + wr[3];
                v = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) v += (wr[q] / sum_w) * tau[q];
            }
            R[T_ * n + c] = v;
            out[2 * pass] = v;
        }
        const double ds = star_path_length(s, ijk);
        // the new HII values stay in registers
        double hii = q_hii, hii_a = hii_a_old;
        if (IMPL) {
            // the implicit scheme: update_HIIfrac with tau_avg = TAU_A (just
            // computed above); a failed iteration leaves the state as it was
            const double sv = star_shell_volume(s, ijk, ds);
            double x, x_avg;
            if (rad_update_hii_v(r, tab, den, pre, q_hii, 0.0, 0.0, out[2], ds, sv, dt, x, x_avg)) {
                hii = x;
                hii_a = x_avg;
                R[R_HII_A * n + c] = hii_a;
                q4[2 * n + c] = hii;
                if (hii > 0.001) q4[3 * n + c] = 1;
            } else {
                set_error(st, ERR_RAD_ITER);
            }
        }
        const double nH = r.X * den / r.mh;
        out[1] = ((1.0 - hii) * nH) * r.sigma * ds;
        out[3] = ((1.0 - hii_a) * nH) * r.sigma * ds;
        R[R_DTAU * n + c] = out[1];
        R[R_DTAU_A * n + c] = out[3];
    }
    if (xface) {
        // the star's cell (k = 0) feeds both sides of shell 1
        for (int sd = (k == 0 ? 0 : side); sd <= (k == 0 ? 1 : side); ++sd)
#pragma unroll
            for (int fld = 0; fld < (IMPL ? 4 : 2); ++fld)
                xf.buf[rad_xface_index(s, xf, k & 1, sd, fld, ijk[1], ijk[2])] = out[fld];
    }
}

__global__ void rad_hii_kernel(StarGeom s, RadConsts r, const double* __restrict__ gtab, double dt, double* q4,
                               double* R, StepState* st) {
    if (dt < 0.0 && st->halted) return;
    __shared__ double stab[TAB_N];
    const double* tab = stage_tables(gtab, stab);
    dt = stage_dtf(dt, st, 0);
    const long long n = (long long)s.n[0] * s.n[1] * s.n[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        int ijk[3];
        cell_ijk(s, c, ijk);
        if (star_wind_cell(s, ijk)) continue;
        const double ds = star_path_length(s, ijk);
        if (!rad_update_hii(r, tab, n, c, ds, star_shell_volume(s, ijk, ds), dt, q4, R)) set_error(st, ERR_RAD_ITER);
    }
}

__global__ void rad_source_kernel(HydroGeom g, RadConsts r, const double* __restrict__ q4,
                                  const double* __restrict__ R, double* __restrict__ U, double dt, StepState* st) {
    if (dt < 0.0 && st->halted) return;
    dt = stage_dtf(dt, st, 0);
    const long long n = g.ncell;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        double u[7];
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] = U[v * n + c];
        const double den = q4[c];
        double udot[7] = {0, 0, 0, 0, 0, 0, 0};
        if (r.coupling == TORCHGPU_RT_TTI) {
            const double x = q4[2 * n + c];
            const double mu_inv = r.X * (x + 1.0) + r.mu_c;
            const double T_new = r.T_hi + (r.T_hii - r.T_hi) * x;
            const double pre_new = r.Rgas * mu_inv * den * T_new;
            double ke = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) ke += 0.5 * u[VEL + d] * u[VEL + d] / u[DEN];
            const double E_new = pre_new / g.gm1 + ke;
            udot[PRE] += (E_new - u[PRE]) / dt;
        } else if (r.coupling == TORCHGPU_RT_NEQ) {
            udot[PRE] += R[R_HEAT * n + c];
        }
        udot[HII] += (q4[2 * n + c] * den - u[HII]) / dt;
        udot[ADV] += (q4[3 * n + c] * den - u[ADV]) / dt;
#pragma unroll
        for (int v = 0; v < 7; ++v) u[v] += dt * udot[v];
        if (fix_solution(g, u)) set_error(st, ERR_THERMO_FIX);
#pragma unroll
        for (int v = 0; v < 7; ++v) U[v * n + c] = u[v];
    }
}

__global__ void rad_zero_kernel(double* p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void rad_set_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

__global__ void rad_copy_if_running_kernel(const StepState* st, const double* __restrict__ src,
                                           double* __restrict__ dst, long long n) {
    if (st->halted) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// calculateTimeStep's function-static `once` (Radiation.cpp:307-316): set
// at creation, marked 2 by the time-step kernel that used it, cleared here;
// restored when the device loop's step did not run (a checkpoint pause).
__global__ void rad_once_commit_kernel(unsigned long long* once, const StepState* st, int gated) {
    if (*once == 2) *once = (gated && st->halted) ? 1ull : 0ull;
}

}  // namespace tg

struct torchgpu_radiation {
    torchgpu_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    tg::RadConsts rc{};
    tg::StarGeom star{};
    double* tables = nullptr;
    double* R = nullptr;
    double* q4 = nullptr;
    unsigned long long* dtbits = nullptr;  // [0] time-step minimum, [1] `once`
    bool q_valid = false;                  // q4 holds a gather (checkValues' Q)
    double* xface = nullptr;   // the sweep's x-face copies
};

namespace {

unsigned grid_blocks(long long n) {
    long long nb = (n + 255) / 256;
    return (unsigned)(nb > 65536 ? 65536 : nb);
}

// The first cell of the reference's CausalNonWind list: Grid::buildCausal's
// octant snake (Grid.cpp:294-356,568-574) -- octants (+,+,+), (-,+,+),
// (+,-,+), (-,-,+), (+,+,-), (-,+,-), (+,-,-), (-,-,-), each walked z-major,
// then y, then x away from the star (negative octants start one cell off
// it) -- skipping the CausalWind cells (Fluid.cpp:41-54).  -1 if none.
long long first_nonwind_cell(const tg::StarGeom& s) {
    auto wind = [&](int i, int j, int k) {
        const int ijk[3] = {i, j, k};
        double d2 = 0;
        for (int a = 0; a < 3; ++a) d2 += (double)(ijk[a] - s.s[a]) * (((double)ijk[a] + 0.5) - (double)s.s[a]);
        return d2 <= (double)(s.wind_r * s.wind_r) || (i == s.s[0] && j == s.s[1] && k == s.s[2]);
    };
    for (int oct = 0; oct < 8; ++oct) {
        const int dir[3] = {(oct & 1) ? -1 : 1, (oct & 2) ? -1 : 1, (oct & 4) ? -1 : 1};
        int b[3];
        for (int a = 0; a < 3; ++a) b[a] = s.s[a] + (dir[a] < 0 ? -1 : 0);
        for (int k = b[2]; k >= 0 && k < s.n[2]; k += dir[2])
            for (int j = b[1]; j >= 0 && j < s.n[1]; j += dir[1])
                for (int i = b[0]; i >= 0 && i < s.n[0]; i += dir[0])
                    if (!wind(i, j, k)) return i + (long long)s.n[0] * (j + (long long)s.n[1] * k);
    }
    return -1;
}

#define RAD_OK(expr)                                                                   \
    do {                                                                               \
        if ((expr) != hipSuccess) return TORCHGPU_ERR_HIP;                            \
    } while (0)

int rad_gather(torchgpu_radiation* r, int check, int gated) {
    tg::CtxView v = tg::ctx_view(r->ctx);
    hipLaunchKernelGGL(tg::rad_gather_kernel, dim3(grid_blocks(v.g->ncell)), dim3(256), 0, r->stream, *v.g, v.U, r->q4,
                       v.st, check, (int)r->q_valid, gated);
    RAD_OK(hipGetLastError());
    r->q_valid = true;
    return TORCHGPU_OK;
}

// rad_gather then rad_heat, fused.
int rad_gather_heat(torchgpu_radiation* r, int check, int gated) {
    tg::CtxView v = tg::ctx_view(r->ctx);
    hipLaunchKernelGGL(tg::rad_gather_heat_kernel, dim3(grid_blocks(v.g->ncell)), dim3(256), 0, r->stream, *v.g,
                       r->star, r->rc, r->tables, v.U, r->q4, r->R, v.st, check, (int)r->q_valid, gated, v.heat);
    RAD_OK(hipGetLastError());
    r->q_valid = true;
    return TORCHGPU_OK;
}

}  // namespace

namespace tg {

hipStream_t rad_stream(const torchgpu_radiation* r) { return r->stream; }
const unsigned long long* rad_dt_bits(const torchgpu_radiation* r) { return r->dtbits; }

// Radiation::preTimeStepCalculations on the fixed primitives of U.
int rad_pre_step_dev(torchgpu_radiation* r, int gated) {
    return rad_gather_heat(r, 0, gated);
}

// Radiation::calculateTimeStep's minimum into dtbits[0] (DBL_MAX when no
// cell lowers it), no host sync.
int rad_timestep_dev(torchgpu_radiation* r, double dt_max, int gated) {
    tg::CtxView v = tg::ctx_view(r->ctx);
    hipLaunchKernelGGL(rad_set_u64_kernel, dim3(1), dim3(1), 0, r->stream, r->dtbits, 0x7fefffffffffffffull);
    const unsigned nb = grid_blocks(v.g->ncell) < 2048 ? grid_blocks(v.g->ncell) : 2048;
    hipLaunchKernelGGL(rad_dt_kernel, dim3(nb), dim3(256), 0, r->stream, r->star, r->rc, r->tables, r->q4, v.U, r->R,
                       dt_max, r->dtbits, v.st, r->dtbits + 1, gated);
    RAD_OK(hipGetLastError());
    return TORCHGPU_OK;
}

int rad_once_commit(torchgpu_radiation* r, int gated) {
    hipLaunchKernelGGL(rad_once_commit_kernel, dim3(1), dim3(1), 0, r->stream, r->dtbits + 1,
                       tg::ctx_view(r->ctx).st, gated);
    RAD_OK(hipGetLastError());
    return TORCHGPU_OK;
}

// Torch::subStep for the radiation component (Torch.cpp:351-363): dt
// explicit, or a DT_* code of the device loop's step (then every kernel
// returns once the run has halted).
int rad_substep_dev(torchgpu_radiation* r, double dt, int has_heat) {
    tg::CtxView v = tg::ctx_view(r->ctx);
    const long long nc = v.g->ncell;
    const int gated = dt < 0.0;
    int rc;
    if (!has_heat || !r->q_valid) {
        // checkValues on the current U and Q, then Q and
        // preTimeStepCalculations anew
        if ((rc = rad_gather_heat(r, 1, gated))) return rc;
    } else {
        hipLaunchKernelGGL(rad_check_kernel, dim3(grid_blocks(nc)), dim3(256), 0, r->stream, *v.g, v.U, r->q4, v.st,
                           gated);
        RAD_OK(hipGetLastError());
    }
    // Radiation::integrate (Radiation.cpp:58-63)
    const bool impl = r->rc.scheme == TORCHGPU_RT_IMPLICIT;
    const int kmax = shell_max(r->star);
    RadXFaces xf;
    xf.buf = r->xface;
    xf.K = kmax;
    xf.w = 2 * kmax + 1;
    ShellFaces f;
    for (int k = 0; k <= kmax; ++k) {
        shell_faces(r->star, k, f);
        const long long cells = f.start[f.nf];
        if (cells == 0) continue;
        const unsigned nb = (unsigned)((cells + 255) / 256);
        if (impl)
            hipLaunchKernelGGL(rad_sweep_kernel<true>, dim3(nb), dim3(256), 0, r->stream, r->star, r->rc, r->tables, f,
                               k, xf, dt, r->q4, r->R, v.st);
        else
            hipLaunchKernelGGL(rad_sweep_kernel<false>, dim3(nb), dim3(256), 0, r->stream, r->star, r->rc, r->tables,
                               f, k, xf, dt, r->q4, r->R, v.st);
        RAD_OK(hipGetLastError());
    }
    if (!impl) {
        hipLaunchKernelGGL(rad_hii_kernel, dim3(grid_blocks(nc)), dim3(256), 0, r->stream, r->star, r->rc, r->tables,
                           dt, r->q4, r->R, v.st);
        RAD_OK(hipGetLastError());
    }
    hipLaunchKernelGGL(rad_source_kernel, dim3(grid_blocks(nc)), dim3(256), 0, r->stream, *v.g, r->rc, r->q4, r->R,
                       v.U, dt, v.st);
    RAD_OK(hipGetLastError());
    ctx_state_changed(r->ctx);
    return TORCHGPU_OK;
}

// The HII fraction of the radiation's Q (q4) into a printed Q [7][n] unless
// the run has halted: Radiation::integrate leaves it in Fluid::Q.
int rad_print_hii(torchgpu_radiation* r, double* P) {
    tg::CtxView v = tg::ctx_view(r->ctx);
    const long long nc = v.g->ncell;
    hipLaunchKernelGGL(rad_copy_if_running_kernel, dim3(grid_blocks(nc)), dim3(256), 0, r->stream, v.st,
                       r->q4 + 2 * nc, P + (long long)HII * nc, nc);
    RAD_OK(hipGetLastError());
    return TORCHGPU_OK;
}

}  // namespace tg

extern "C" {

int torchgpu_radiation_create(const torchgpu_rad_params* p, torchgpu_ctx* ctx, const torchgpu_star* star,
                              torchgpu_radiation** out) {
    if (!p || !ctx || !star || !out) return TORCHGPU_ERR_INVALID_ARG;
    tg::CtxView v = tg::ctx_view(ctx);
    const tg::HydroGeom& g = *v.g;
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a)
        if (star->cell[a] < 0 || star->cell[a] >= n[a]) return TORCHGPU_ERR_INVALID_ARG;
    if (star->wind_cell_radius < 0 || !(star->photon_rate >= 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    if (g.ncell >= (1LL << 32)) return TORCHGPU_ERR_INVALID_ARG;  // 32-bit cell index math (cell_ijk32)
    torchgpu_radiation* r = new torchgpu_radiation();
    r->ctx = ctx;
    r->stream = v.stream;
    tg::StarGeom& s = r->star;
    s.on = 1;
    for (int a = 0; a < 3; ++a) {
        s.s[a] = star->cell[a];
        s.n[a] = n[a];
    }
    s.wind_r = star->wind_cell_radius;
    s.dx = g.dx;
    s.photon_rate = star->photon_rate;
    s.pi = p->pi;
    s.hydrogen_mass = p->hydrogen_mass;
    tg::RadConsts& c = r->rc;
    for (int i = 0; i < 4; ++i) c.K[i] = p->K[i];
    c.sigma = p->sigma;
    c.tau0 = p->tau0;
    c.T_hi = p->T_hi;
    c.T_hii = p->T_hii;
    c.scheme = p->scheme;
    c.coupling = p->coupling;
    c.collisions_on = p->collisions_on;
    c.heat_amp = p->heating_amplification;
    c.X = p->mass_fraction_h;
    c.mu_c = (1.0 - p->mass_fraction_h) * 0.25;
    c.mh = p->hydrogen_mass;
    c.Rgas = p->spec_gas_const;
    c.kB = p->boltzmann;
    c.excess = p->photon_energy - p->rydberg;
    c.vA = p->voronov_A;
    c.vX = p->voronov_X;
    c.vK = p->voronov_K;
    c.vP = p->voronov_P;
    c.vDE = p->voronov_DE;
    c.photon_rate = star->photon_rate;
    c.T_min = p->T_min;
    c.rec_slope0 = p->rec_slope[0];
    c.rec_slope1 = p->rec_slope[1];
    c.cool_slope0 = p->cool_slope[0];
    c.cool_slope1 = p->cool_slope[1];
    c.once_cell = first_nonwind_cell(s);
    c.use_once = c.K[0] != 0.0 && c.once_cell >= 0;
    double tab[tg::TAB_N];
    std::memcpy(tab + tg::TAB_X, p->rec_x, sizeof p->rec_x);
    std::memcpy(tab + tg::TAB_REC, p->rec_y, sizeof p->rec_y);
    std::memcpy(tab + tg::TAB_REC2, p->rec_y2, sizeof p->rec_y2);
    std::memcpy(tab + tg::TAB_COOL, p->cool_y, sizeof p->cool_y);
    std::memcpy(tab + tg::TAB_COOL2, p->cool_y2, sizeof p->cool_y2);
    const long long nc = g.ncell;
    bool ok = hipMalloc(&r->tables, sizeof tab) == hipSuccess && hipMalloc(&r->R, 6 * nc * sizeof(double)) == hipSuccess &&
              hipMalloc(&r->q4, 4 * nc * sizeof(double)) == hipSuccess &&
              hipMalloc(&r->dtbits, 2 * sizeof(unsigned long long)) == hipSuccess;
    const long long wx = 2LL * tg::shell_max(s) + 1;
    ok = ok && hipMalloc(&r->xface, 16 * wx * wx * sizeof(double)) == hipSuccess;
    ok = ok && hipMemcpyAsync(r->tables, tab, sizeof tab, hipMemcpyHostToDevice, r->stream) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(tg::rad_zero_kernel, dim3(grid_blocks(6 * nc)), dim3(256), 0, r->stream, r->R, 6 * nc);
        hipLaunchKernelGGL(tg::rad_zero_kernel, dim3(grid_blocks(4 * nc)), dim3(256), 0, r->stream, r->q4, 4 * nc);
        hipLaunchKernelGGL(tg::rad_set_u64_kernel, dim3(1), dim3(1), 0, r->stream, r->dtbits + 1, 1ull);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(r->stream) == hipSuccess;
    }
    if (!ok) {
        torchgpu_radiation_destroy(r);
        return TORCHGPU_ERR_HIP;
    }
    *out = r;
    return TORCHGPU_OK;
}

void torchgpu_radiation_destroy(torchgpu_radiation* r) {
    if (!r) return;
    (void)hipStreamSynchronize(r->stream);
    for (void* p : {(void*)r->tables, (void*)r->R, (void*)r->q4, (void*)r->dtbits, (void*)r->xface})
        if (p) (void)hipFree(p);
    delete r;
}

int torchgpu_radiation_init_field(torchgpu_radiation* r, const double* Q, int on_device) {
    if (!r) return TORCHGPU_ERR_INVALID_ARG;
    int rc = rad_gather(r, 0, 0);
    if (rc) return rc;
    tg::CtxView v = tg::ctx_view(r->ctx);
    const long long nc = v.g->ncell;
    double* q7 = nullptr;
    if (Q && !on_device) {
        RAD_OK(hipMalloc(&q7, 7 * nc * sizeof(double)));
        if (hipMemcpyAsync(q7, Q, 7 * nc * sizeof(double), hipMemcpyHostToDevice, r->stream) != hipSuccess) {
            (void)hipFree(q7);
            return TORCHGPU_ERR_HIP;
        }
    }
    hipLaunchKernelGGL(tg::rad_init_field_kernel, dim3(grid_blocks(nc)), dim3(256), 0, r->stream, *v.g, r->star, r->rc,
                       q7 ? q7 : Q, r->q4, r->R);
    hipError_t e = hipGetLastError();
    if (q7) {
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        (void)hipFree(q7);
    }
    RAD_OK(e);
    return TORCHGPU_OK;
}

int torchgpu_radiation_pre_step(torchgpu_radiation* r) {
    if (!r) return TORCHGPU_ERR_INVALID_ARG;
    return tg::rad_pre_step_dev(r, 0);
}

int torchgpu_radiation_timestep(torchgpu_radiation* r, double dt_max, double* dt) {
    if (!r || !dt || !(dt_max > 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    if (!r->q_valid) {
        const int rc = torchgpu_radiation_pre_step(r);
        if (rc) return rc;
    }
    int rc = tg::rad_timestep_dev(r, dt_max, 0);
    if (!rc) rc = tg::rad_once_commit(r, 0);
    if (rc) return rc;
    unsigned long long bits = 0;
    RAD_OK(hipMemcpyAsync(&bits, r->dtbits, sizeof bits, hipMemcpyDeviceToHost, r->stream));
    RAD_OK(hipStreamSynchronize(r->stream));
    double m;
    std::memcpy(&m, &bits, sizeof m);
    *dt = m < dt_max ? m : dt_max;
    return TORCHGPU_OK;
}

int torchgpu_radiation_substep(torchgpu_radiation* r, double dt, int has_heat) {
    if (!r || !(dt > 0.0)) return TORCHGPU_ERR_INVALID_ARG;
    return tg::rad_substep_dev(r, dt, has_heat);
}

int torchgpu_radiation_get_state(torchgpu_radiation* r, double* R, double* q_hii_adv, int on_device) {
    if (!r) return TORCHGPU_ERR_INVALID_ARG;
    const long long nc = tg::ctx_view(r->ctx).g->ncell;
    const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (R) RAD_OK(hipMemcpyAsync(R, r->R, 6 * nc * sizeof(double), k, r->stream));
    if (q_hii_adv) RAD_OK(hipMemcpyAsync(q_hii_adv, r->q4 + 2 * nc, 2 * nc * sizeof(double), k, r->stream));
    RAD_OK(hipStreamSynchronize(r->stream));
    return TORCHGPU_OK;
}

}  // extern "C"
