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

}  // namespace tg
