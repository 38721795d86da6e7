// units.cpp -- code-unit scaling and thermodynamics constants on the host.
//
// Restates the reference's Converter (src/Torch/Converter.cpp:7-63),
// Constants::initialise (src/Torch/Constants.cpp:27-40) and
// Thermodynamics::initialise / initCollisionalExcitationHI
// (src/Integrators/Thermodynamics.cpp:24-80) plus the natural cubic spline of
// SplineData::spline (src/Integrators/SplineData.cpp:15-55), keeping the
// multiplication sequences so that every constant is bitwise the reference's.
#include <cmath>
#include <cstring>

#include "../../include/torchgpu.h"
#include "units.hpp"

namespace tg {

void Converter::set_rho_pressure_time(double rho, double pressure, double time) {
    M = 1.0 * 2.0e33;          // MO_2_G(1.0)
    L = 14.6 * 3.09e18;        // PC_2_CM(14.6)
    T = 10000.0 * 3.15569e7;   // YR_2_S(10000.0)
    double rho_0 = (rho > 0) ? rho : M / (L * L * L);
    double pressure_0 = (pressure > 0) ? pressure : rho_0 * L * L / (T * T);
    double time_0 = (time > 0) ? time : T;
    T = time_0;
    M = std::sqrt((pressure_0 / rho_0) * pressure_0 * pressure_0) * T * T * T;
    L = std::sqrt(pressure_0 / rho_0) * T;
}

double Converter::convert(double val, double mi, double li, double ti, bool to_code) const {
    double MM = mi >= 0 ? M : 1.0 / M;
    MM = to_code ? 1.0 / MM : MM;
    double LL = li >= 0 ? L : 1.0 / L;
    LL = to_code ? 1.0 / LL : LL;
    double TT = ti >= 0 ? T : 1.0 / T;
    TT = to_code ? 1.0 / TT : TT;
    int im = (int)std::fabs(mi), il = (int)std::fabs(li), it = (int)std::fabs(ti);
    double result = val;
    while (im >= 1 || il >= 1 || it >= 1) {
        if (im-- >= 1) result *= MM;
        if (il-- >= 1) result *= LL;
        if (it-- >= 1) result *= TT;
    }
    double left = std::fabs(mi) - (int)std::fabs(mi);
    if (left != 0) result *= std::pow(MM, left);
    left = std::fabs(li) - (int)std::fabs(li);
    if (left != 0) result *= std::pow(LL, left);
    left = std::fabs(ti) - (int)std::fabs(ti);
    if (left != 0) result *= std::pow(TT, left);
    return result;
}

void natural_spline(const double* x, const double* f, int n, double* y2) {
    double u[64];
    y2[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; i++) {
        double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (f[i + 1] - f[i]) / (x[i + 1] - x[i]) - (f[i] - f[i - 1]) / (x[i] - x[i - 1]);
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    const double qn = 0.0, un = 0.0;
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

}  // namespace tg

extern "C" int torchgpu_units(double density_scale, double pressure_scale, double time_scale, double value,
                              double mass_index, double length_index, double time_index, int to_code, double* out) {
    if (!out) return TORCHGPU_ERR_INVALID_ARG;
    tg::Converter c;
    c.set_rho_pressure_time(density_scale, pressure_scale, time_scale);
    *out = c.convert(value, mass_index, length_index, time_index, to_code != 0);
    return TORCHGPU_OK;
}

extern "C" int torchgpu_thermo_params_init(torchgpu_thermo_params* t, double density_scale, double pressure_scale,
                                           double time_scale, double thermo_hii_switch, double heating_amplification,
                                           double mass_fraction_h, double gamma, double density_floor_cgs,
                                           double temperature_floor, int subcycling) {
    if (!t) return TORCHGPU_ERR_INVALID_ARG;
    std::memset(t, 0, sizeof *t);
    tg::Converter c;
    c.set_rho_pressure_time(density_scale, pressure_scale, time_scale);
    auto code = [&](double v, double m, double l, double tt) { return c.convert(v, m, l, tt, true); };
    t->hydrogen_mass = code(1.674e-24, 1, 0, 0);
    t->spec_gas_const = code(8.314462e7, 0, 2, -2);
    t->dust_xsec = code(5.0e-22, 0, 2, 0);
    t->pi = 3.14159265359;
    t->z0 = 5.0e-4;
    t->T1 = 33610; t->T2 = 2180; t->T3 = 28390; t->T4 = 1780;
    t->imlc = code(2.905e-19, 1.0, 5.0, -3.0);
    t->nmlc = code(4.477e-20, 1.0, 5.0, -3.0);
    t->ciec_minT = 5.0e4;
    t->ciec = code(3.485e-15, 1.0, 5.0, -3.0);
    t->cxhi_damp = 5.0e5;
    t->n0 = code(1.0e6, 0.0, -3.0, 0.0);
    t->nmc = code(3.981e-27, 1.0, 3.8, -3.0);
    t->fuvh_a = code(1.9e-26, 1.0, 4.0, -2.0);
    t->fuvh_b = code(1.00000, 0.0, 0.0, -1.0);
    t->fuvh_c = code(6.40000, 0.0, -1.0, 0.0);
    t->irh_a = code(7.7e-32, 1.0, 4.0, -2.0);
    t->irh_b = code(3.0e4, 0.0, -3.0, 0.0);
    t->crh = code(5.0e-27, 1.0, 2.0, -3.0);
    static const double T[26] = {3162.2776602, 3981.0717055, 5011.8723363, 6309.5734448, 7943.2823472,
        10000.0000000, 12589.2541179, 15848.9319246, 19952.6231497, 25118.8643151,
        31622.7766017, 39810.7170553, 50118.7233627, 63095.7344480, 79432.8234724,
        100000.0000000, 125892.5411794, 158489.3192461, 199526.2314969, 251188.6431510,
        316227.7660168, 398107.1705535, 501187.2336273, 630957.3444802, 794328.2347243,
        1000000.0000000};
    static const double R[26] = {1.150800e-34, 2.312065e-31, 9.571941e-29, 1.132400e-26, 4.954502e-25,
        9.794900e-24, 1.035142e-22, 6.652732e-22, 2.870781e-21, 9.036495e-21, 2.218196e-20,
        4.456562e-20, 7.655966e-20, 1.158777e-19, 1.588547e-19, 2.013724e-19, 2.393316e-19,
        2.710192e-19, 2.944422e-19, 3.104560e-19, 3.191538e-19, 3.213661e-19, 3.191538e-19,
        3.126079e-19, 3.033891e-19, 2.917427e-19};
    for (int i = 0; i < 26; ++i) {
        t->spl_x[i] = std::log10(T[i]);
        t->spl_y[i] = std::log10(code(R[i], 1, 5, -3));
    }
    tg::natural_spline(t->spl_x, t->spl_y, 26, t->spl_y2);
    t->spl_min_slope = (t->spl_y[1] - t->spl_y[0]) / (t->spl_x[1] - t->spl_x[0]);
    t->spl_max_slope = (t->spl_y[25] - t->spl_y[24]) / (t->spl_x[25] - t->spl_x[24]);
    t->thermo_hii_switch = thermo_hii_switch;
    t->heating_amplification = heating_amplification;
    t->mass_fraction_h = mass_fraction_h;
    t->gamma = gamma;
    // TorchParameters::initialise (Parameters.cpp:17-18)
    const double dfloor = code(density_floor_cgs, 1, -3, 0);
    t->pfloor = 0.1 * t->spec_gas_const * dfloor * temperature_floor;
    t->subcycling = subcycling;
    return TORCHGPU_OK;
}

extern "C" int torchgpu_rad_params_init(torchgpu_rad_params* r, double density_scale, double pressure_scale,
                                        double time_scale, const double K[4], double photoion_cross_section,
                                        double tau0, double min_hii_fraction, double temperature_hi,
                                        double temperature_hii, int scheme, int coupling, int collisions_on,
                                        double heating_amplification, double mass_fraction_h, double photon_energy) {
    // Radiation::initialise + initRecombinationHummer (Radiation.cpp:24-56,
    // 184-212), Constants::initialise (Constants.cpp:27-40) and the
    // conversions of TorchParameters::initialise (Parameters.cpp:15-19)
    if (!r || !K) return TORCHGPU_ERR_INVALID_ARG;
    if (scheme < TORCHGPU_RT_IMPLICIT || scheme > TORCHGPU_RT_EXPLICIT) return TORCHGPU_ERR_INVALID_ARG;
    if (coupling < TORCHGPU_RT_TTI || coupling > TORCHGPU_RT_OFF) return TORCHGPU_ERR_INVALID_ARG;
    tg::Converter c;
    c.set_rho_pressure_time(density_scale, pressure_scale, time_scale);
    std::memset(r, 0, sizeof *r);
    for (int i = 0; i < 4; ++i) r->K[i] = K[i];
    r->sigma = c.convert(photoion_cross_section, 0, 2, 0, true);
    r->tau0 = tau0;
    r->min_hii = min_hii_fraction;
    r->T_hi = temperature_hi;
    r->T_hii = temperature_hii;
    r->scheme = scheme;
    r->coupling = coupling;
    r->collisions_on = collisions_on != 0;
    r->heating_amplification = heating_amplification;
    r->mass_fraction_h = mass_fraction_h;
    r->hydrogen_mass = c.convert(1.674e-24, 1, 0, 0, true);
    r->spec_gas_const = c.convert(8.314462e7, 0, 2, -2, true);
    r->boltzmann = c.convert(1.3806488e-16, 1, 2, -2, true);
    r->rydberg = c.convert(13.6 * 1.60217646e-12, 1, 2, -2, true);  // EV_2_ERGS(13.6)
    r->pi = 3.14159265359;
    r->voronov_A = c.convert(2.91e-8, 0, 3, -1, true);
    r->voronov_X = 0.232;
    r->voronov_K = 0.39;
    r->voronov_P = 0.0;
    r->voronov_DE = r->rydberg / r->boltzmann;
    r->photon_energy = c.convert(photon_energy, 1, 2, -2, true);
    r->T_min = 100.0;  // Thermodynamics::initialiseMinTempField (Thermodynamics.cpp:52,102-111)
    // Hummer (1994) case-B recombination and recombination-cooling
    // coefficients at log10 T = 1, 1.2, ..., 7, divided by sqrt(T)
    static const double alpha[31] = {9.283e-11, 8.823e-11, 8.361e-11, 7.898e-11, 7.435e-11, 6.973e-11, 6.512e-11,
        6.054e-11, 5.599e-11, 5.147e-11, 4.700e-11, 4.258e-11, 3.823e-11, 3.397e-11, 2.983e-11, 2.584e-11,
        2.204e-11, 1.847e-11, 1.520e-11, 1.226e-11, 9.696e-12, 7.514e-12, 5.710e-12, 4.257e-12, 3.117e-12,
        2.244e-12, 1.590e-12, 1.110e-12, 7.642e-13, 5.199e-13, 3.498e-13};
    static const double cool[31] = {8.287e-11, 7.821e-11, 7.356e-11, 6.892e-11, 6.430e-11, 5.971e-11, 5.515e-11,
        5.062e-11, 4.614e-11, 4.170e-11, 3.734e-11, 3.306e-11, 2.888e-11, 2.484e-11, 2.098e-11, 1.736e-11,
        1.402e-11, 1.103e-11, 8.442e-12, 6.279e-12, 4.539e-12, 3.192e-12, 2.185e-12, 1.458e-12, 9.484e-13,
        6.023e-13, 3.738e-13, 2.268e-13, 1.348e-13, 7.859e-14, 4.499e-14};
    for (int i = 0; i < 31; ++i) {
        const double T = std::exp(std::log(10.0) * (1.0 + 0.2 * static_cast<double>(i)));
        const double sT = std::sqrt(T);
        r->rec_x[i] = T;
        r->rec_y[i] = c.convert(alpha[i] / sT, 0, 3, -1, true);
        r->cool_y[i] = c.convert(cool[i] / sT, 0, 3, -1, true);
    }
    tg::natural_spline(r->rec_x, r->rec_y, 31, r->rec_y2);
    tg::natural_spline(r->rec_x, r->cool_y, 31, r->cool_y2);
    // LinearSplineData's logarithmic end slopes (SplineData.cpp:97-100)
    auto slope = [&](const double* y, int a, int b) {
        return (std::log10(y[b]) - std::log10(y[a])) / (std::log10(r->rec_x[b]) - std::log10(r->rec_x[a]));
    };
    r->rec_slope[0] = slope(r->rec_y, 0, 1);
    r->rec_slope[1] = slope(r->rec_y, 29, 30);
    r->cool_slope[0] = slope(r->cool_y, 0, 1);
    r->cool_slope[1] = slope(r->cool_y, 29, 30);
    return TORCHGPU_OK;
}
