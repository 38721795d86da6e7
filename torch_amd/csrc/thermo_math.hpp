// thermo_math.hpp -- Henney et al. (2009) heating/cooling functions and the
// subcycled explicit source step of the reference's Thermodynamics operator
// (hgsteggles/TORCH src/Integrators/Thermodynamics.cpp), FP64, for the HIP
// kernels.  Operand order follows the reference except inside the cooling
// sum, whose exponentials share one reciprocal and one log of T (see
// subtract_cooling); exp/log/pow are the device library's (ocml), so
// agreement with the x86 build is to within rounding rather than bitwise.
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

// Constants in code units, as Thermodynamics::initialise computes them
// (Thermodynamics.cpp:24-57) plus the switches of ThermoParameters.
struct ThermoConsts {
    double thermo_hii_switch, heating_amplification, X, gamma, pfloor;
    double hydrogen_mass, spec_gas_const, dust_xsec, pi;
    double z0, T1, T2, T3, T4, imlc, nmlc, ciec_minT, ciec, cxhi_damp, n0, nmc;
    double fuvh_a, fuvh_b, fuvh_c, irh_a, irh_b, crh;
    double spl_min_slope, spl_max_slope;
    int subcycling;
    int pad;
};

// Natural cubic spline in log10 space: LogSplineData::interpolate +
// SplineData::splint (SplineData.cpp:57-81,142-151).  x, y, y2 are 26-entry
// tables (LDS) whose abscissae are log10 T = 3.5, 3.6, ..., 6.0 up to
// rounding: the bracketing interval of the reference's bisection (largest
// klo <= 24 with x[klo] <= x) is found from the spacing and corrected by one
// step against the stored abscissae, which gives the same klo.
__device__ __forceinline__ double log_spline(const ThermoConsts& t, const double* sx, const double* sy,
                                             const double* sy2, double x) {
    if (x > sx[25]) return sy[25] + t.spl_max_slope * (x - sx[25]);
    if (x < sx[0]) return sy[0] + t.spl_min_slope * (x - sx[0]);
    int klo = (int)((x - sx[0]) * 10.0);
    klo = klo < 0 ? 0 : (klo > 24 ? 24 : klo);
    if (sx[klo] > x) --klo;
    else if (klo < 24 && sx[klo + 1] <= x) ++klo;
    const int khi = klo + 1;
    double h = sx[khi] - sx[klo];
    double a = (sx[khi] - x) / h;
    double b = (x - sx[klo]) / h;
    return a * sy[klo] + b * sy[khi] + ((a * a * a - a) * sy2[klo] + (b * b * b - b) * sy2[khi]) * (h * h) / 6.0;
}

__device__ __forceinline__ double soft_landing(double rate, double T, double T_min) {
    double result = rate;
    if (rate < 0.0) {
        if (T <= T_min) result = 0;
        else if (T <= T_min + 200) result = rate * (T - T_min) / 200;
    }
    return result;
}

// Per-cell factors of the cooling functions that do not depend on T; each
// is the left-to-right prefix of the reference's product, so hoisting it
// out of the subcycle loop does not change the rounding.
struct CoolingCell {
    double imlc_k;   // m_imlc*m_z0*ne*ne                (:122)
    double nmlc_k;   // m_nmlc*m_z0*ne*nn                (:127)
    double ciec_k;   // m_ciec*ne*ne*m_z0                (:133)
    double cehi_k;   // HIIFRAC*(1-HIIFRAC)*nH*nH        (:159)
    double nmc_k;    // m_nmc*(1-x)*(1-x)*pow(nH,1.6)    (:144)
    double nmc_T0;   // 70 + 220*pow(nH/m_n0, 0.2)       (:143)
};

__device__ __forceinline__ CoolingCell cooling_cell(const ThermoConsts& t, double nH, double x) {
    CoolingCell c;
    const double ne = nH * (x), nn = nH * (1.0 - x);
    c.imlc_k = t.imlc * t.z0 * ne * ne;
    c.nmlc_k = t.nmlc * t.z0 * ne * nn;
    c.ciec_k = t.ciec * ne * ne * t.z0;
    c.cehi_k = x * (1.0 - x) * nH * nH;
    c.nmc_k = t.nmc * (1.0 - x) * (1.0 - x) * pow(nH, 1.6);
    c.nmc_T0 = 70.0 + 220.0 * pow(nH / t.n0, 0.2);
    return c;
}

// Sum of the five cooling terms, subtracted in the reference's order from
// `rate` (Thermodynamics.cpp:255-259 / 315-319).  The transcendental work is
// shared between the terms: one reciprocal and one natural log of T feed
// every exponent (x/T = x * (1/T); log10 T = ln T / ln 10; sqrt(T) e^-a =
// e^(ln T / 2 - a); pow(1e-5 T, 1.63) = e^(1.63 (ln T + ln 1e-5))), leaving
// six exponentials per evaluation instead of ten libm calls.  The values
// differ from the reference's by rounding of the exponents only (relative
// ~1e-15 on each term).
__device__ __forceinline__ double subtract_cooling(const ThermoConsts& t, const CoolingCell& c, const double* sx,
                                                   const double* sy, const double* sy2, double rate, double T) {
    const double iT = 1.0 / T;
    const double lnT = log(T);
    {
        const double a = t.T2 * iT;
        rate -= c.imlc_k * exp(-t.T1 * iT - a * a);
    }
    {
        const double a = t.T4 * iT;
        rate -= c.nmlc_k * exp(-t.T3 * iT - a * a);
    }
    {
        const double r = log_spline(t, sx, sy, sy2, lnT * 0.43429448190325182765);
        const double a = T * (1.0 / t.cxhi_damp);
        rate -= c.cehi_k * exp((2.302585093 * r) - a * a);
    }
    if (T > t.ciec_minT) {
        const double p163 = exp(1.63 * (lnT - 11.512925464970228420));
        double cie_rate = c.ciec_k * exp(-0.63 * lnT) * (1.0 - exp(-p163));
        double smoothing = (T - 5.0e4) / (2.0e4);
        smoothing = (smoothing < 1.0) ? smoothing : 1.0;
        rate -= cie_rate * smoothing;
    } else {
        rate -= 0;
    }
    rate -= c.nmc_k * exp(0.5 * lnT - c.nmc_T0 * iT);
    return rate;
}

// Heating terms FUVH + IRH + CRH (Thermodynamics.cpp:178-191, 249-251).
__device__ __forceinline__ double heating(const ThermoConsts& t, double nH, double Av, double F) {
    double rate = 0;
    rate += t.fuvh_a * nH * F * exp(-1.9 * Av) / (t.fuvh_b + t.fuvh_c * F * exp(-1.9 * Av) / nH);
    rate += t.irh_a * nH * F * exp(-0.05 * Av) * exp(-2.0 * log(1.0 + t.irh_b / nH));
    rate += t.crh * nH;
    return rate;
}

// The heating and cooling terms one by one, as Thermodynamics::
// fillHeatingArrays writes them into GridCell::H (Thermodynamics.cpp:445-457,
// the term functions :121-191): ionised and neutral metal lines, HI
// collisional excitation, collisional ionisation equilibrium and neutral /
// molecular cooling (negated), FUV, IR and cosmic-ray heating.  Written with
// the reference's own operations (one libm call per exponential), not the
// shared-log form of subtract_cooling.  Order of H: IMLC, NMLC, RHII, CEHI,
// CIEC, NMC, EUVH, FUVH, IRH, CRH, TOT (Common.hpp:32-34); RHII, EUVH and
// TOT are left to the caller.
__device__ __forceinline__ void heating_terms(const ThermoConsts& t, const double* sx, const double* sy,
                                              const double* sy2, double nH, double x, double T, double Av,
                                              double F, double H[11]) {
    const double ne = x * nH, nn = (1.0 - x) * nH;
    H[7] = t.fuvh_a * nH * F * exp(-1.9 * Av) / (t.fuvh_b + t.fuvh_c * F * exp(-1.9 * Av) / nH);
    H[8] = t.irh_a * nH * F * exp(-0.05 * Av) * exp(-2.0 * log(1.0 + t.irh_b / nH));
    H[9] = t.crh * nH;
    H[0] = -(t.imlc * t.z0 * ne * ne * exp(-t.T1 / T - (t.T2 / T) * (t.T2 / T)));
    H[1] = -(t.nmlc * t.z0 * ne * nn * exp(-t.T3 / T - (t.T4 / T) * (t.T4 / T)));
    {
        const double r = log_spline(t, sx, sy, sy2, log10(T));
        H[3] = -(x * (1.0 - x) * nH * nH * exp((2.302585093 * r) - ((T / t.cxhi_damp) * (T / t.cxhi_damp))));
    }
    if (T > t.ciec_minT) {
        const double cie = t.ciec * ne * ne * t.z0 * exp(-0.63 * log(T)) * (1.0 - exp(-pow(1.0e-5 * T, 1.63)));
        double smoothing = (T - 5.0e4) / (2.0e4);
        smoothing = (smoothing < 1.0) ? smoothing : 1.0;
        H[4] = -(cie * smoothing);
    } else {
        H[4] = -0.0;
    }
    {
        const double T0 = 70.0 + 220.0 * pow(nH / t.n0, 0.2);
        H[5] = -(t.nmc * (1.0 - x) * (1.0 - x) * pow(nH, 1.6) * sqrt(T) * exp(-T0 / T));
    }
}

}  // namespace tg
