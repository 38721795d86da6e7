// units.hpp -- host-side code-unit converter (reference src/Torch/Converter.cpp).
#pragma once

namespace tg {

struct Converter {
    double M = 0, L = 0, T = 0;
    // Converter::set_rho_pressure_time, Converter.cpp:17-25
    void set_rho_pressure_time(double rho, double pressure, double time);
    // Converter::convertCodeUnits, Converter.cpp:35-63
    double convert(double val, double mass_index, double length_index, double time_index, bool to_code) const;
};

// Natural cubic spline second derivatives (SplineData::spline with both end
// derivatives 1e99, SplineData.cpp:15-55).
void natural_spline(const double* x, const double* f, int n, double* y2);

}  // namespace tg
