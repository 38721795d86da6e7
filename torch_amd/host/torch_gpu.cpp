// torch_gpu.cpp -- host layer over the C ABI (see torch_gpu.hpp).
#include "torch_gpu.hpp"

#include "lua_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <sstream>
#include <vector>

namespace tg {

namespace {

void check_rc(int rc, const char* what, torchgpu_ctx* ctx = nullptr) {
    if (rc == TORCHGPU_OK) return;
    char buf[1024] = {0};
    if (ctx) torchgpu_last_error(ctx, buf, sizeof buf);
    throw std::runtime_error(std::string(what) + ": " + (buf[0] ? buf : ("error " + std::to_string(rc))));
}

// Converter (src/Torch/Converter.cpp:17-63) through the library.
double convert(const TorchParameters& p, double v, double m, double l, double t, bool to_code) {
    double out = 0;
    check_rc(torchgpu_units(p.dscale, p.pscale, p.tscale, v, m, l, t, to_code ? 1 : 0, &out), "units");
    return out;
}

}  // namespace

// ---------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------
int riemannSolverId(const std::string& type) {  // RiemannSolverFactory::create, Riemann.cpp:246-257
    if (type == "HLLC") return TORCHGPU_SOLVER_HLLC;
    if (type == "RotatedHLLC" || type == "default") return TORCHGPU_SOLVER_ROTATED_HLLC;
    if (type == "HLL") return TORCHGPU_SOLVER_HLL;
    throw std::runtime_error("RiemannSolverFactory::create: unknown type: " + type + "\n");
}

int slopeLimiterId(const std::string& type) {  // SlopeLimiterFactory::create, SlopeLimiter.cpp:56-75
    if (type == "monotonised_central") return TORCHGPU_LIM_MC;
    if (type == "superbee") return TORCHGPU_LIM_SUPERBEE;
    if (type == "minmod") return TORCHGPU_LIM_MINMOD;
    if (type == "maxmod") return TORCHGPU_LIM_MAXMOD;
    if (type == "leer") return TORCHGPU_LIM_LEER;
    if (type == "ospre") return TORCHGPU_LIM_OSPRE;
    if (type == "albada" || type == "default") return TORCHGPU_LIM_ALBADA;
    throw std::runtime_error("SlopeLimiterFactory::create: unknown type: " + type + "\n");
}

int conditionId(const std::string& c) {  // Constants::Constants conditionParser, Constants.cpp:9-14
    if (c == "free" || c == "default") return TORCHGPU_BC_FREE;
    if (c == "reflecting") return TORCHGPU_BC_REFLECTING;
    if (c == "outflow") return TORCHGPU_BC_OUTFLOW;
    if (c == "inflow") return TORCHGPU_BC_INFLOW;
    if (c == "periodic") return TORCHGPU_BC_PERIODIC;
    // EnumParser falls back to "default" with a warning (Constants.hpp:28-45)
    return TORCHGPU_BC_FREE;
}

double TorchParameters::toCodeUnits(double v, double m, double l, double t) const { return convert(*this, v, m, l, t, true); }

void TorchParameters::initialise() {
    // Constants::initialise_DPT + TorchParameters::initialise (Parameters.cpp:9-27)
    tmax = toCodeUnits(tmax, 0, 0, 1);
    dt_max = tmax / 10.0;
    sideLength = toCodeUnits(sideLength, 0, 1, 0);
    if (geometry == "spherical" || geometry == "cylindrical") leftBC[0] = "reflecting";
    dfloor = toCodeUnits(dfloor, 1, -3, 0);
    specificGasConstant = toCodeUnits(8.314462e7, 0, 2, -2);
    pfloor = 0.1 * specificGasConstant * dfloor * tfloor;
    photonRate = toCodeUnits(photonRate, 0, 0, -1);
}

torchgpu_params TorchParameters::gridParams() const {
    if (geometry != "cartesian") throw std::runtime_error("only Cartesian geometry runs on the GPU path");
    if (nd != 3) throw std::runtime_error("the GPU path runs 3-D grids (Grid.no_dimensions = 3)");
    if (spatialOrder != 0 && spatialOrder != 1)
        throw std::runtime_error("Hydrodynamics::calcFluxes: invalid order(=" + std::to_string(spatialOrder) +
                                 "). Valid orders = {0, 1}.");
    torchgpu_params g;
    std::memset(&g, 0, sizeof g);
    g.nx = ncells[0]; g.ny = ncells[1]; g.nz = ncells[2];
    g.spatial_order = spatialOrder;
    for (int d = 0; d < 3; ++d) {
        g.bc_left[d] = conditionId(leftBC[d]);
        g.bc_right[d] = conditionId(rightBC[d]);
    }
    // Torch::initialise keeps the default when the factory rejects the
    // string (Torch.cpp:65-77)
    try { g.riemann_solver = riemannSolverId(riemannSolver); } catch (std::exception&) { g.riemann_solver = TORCHGPU_SOLVER_ROTATED_HLLC; }
    try { g.slope_limiter = slopeLimiterId(slopeLimiter); } catch (std::exception&) { g.slope_limiter = TORCHGPU_LIM_ALBADA; }
    g.precision = (precision == "mixed") ? TORCHGPU_MIXED : TORCHGPU_FP64;
    g.ncheckpoints = ncheckpoints;
    g.debug = debug ? 1 : 0;
    g.side_length = sideLength;
    g.gamma = heatCapacityRatio;
    g.dfloor = dfloor;
    g.pfloor = pfloor;
    g.tfloor = tfloor;
    g.mass_fraction_h = massFractionH;
    g.spec_gas_const = specificGasConstant;
    g.tmax = tmax;
    g.dt_max = dt_max;
    return g;
}

namespace {
template <typename T>
void take(const LuaValue& v, T& var, const std::string& name, std::string* warn) {
    if (!v.exists()) {
        if (warn) *warn += "parseLuaVariable: <" + name + "> does not exist, using default\n";
        return;
    }
    if constexpr (std::is_same<T, std::string>::value) {
        var = v.kind == LuaValue::STRING ? v.str : var;
    } else if constexpr (std::is_same<T, bool>::value) {
        var = v.kind == LuaValue::BOOLEAN ? v.boolean : (v.kind == LuaValue::NUMBER && v.num != 0);
    } else {
        var = (T)(v.kind == LuaValue::BOOLEAN ? (v.boolean ? 1 : 0) : v.num);
    }
}
}  // namespace

void parseParameters(const LuaConfig& cfg, TorchParameters& p, std::string* w) {
    const LuaValue& P = cfg["Parameters"];
    const LuaValue& I = P["Integration"];
    const LuaValue& G = P["Grid"];
    const LuaValue& H = P["Hydrodynamics"];
    const LuaValue& R = P["Radiation"];
    const LuaValue& T = P["Thermodynamics"];
    const LuaValue& S = P["Star"];
    take(I["density_scale"], p.dscale, "density_scale", w);
    take(I["pressure_scale"], p.pscale, "pressure_scale", w);
    take(I["time_scale"], p.tscale, "time_scale", w);
    take(I["spatial_order"], p.spatialOrder, "spatial_order", w);
    take(I["temporal_order"], p.temporalOrder, "temporal_order", w);
    take(I["simulation_time"], p.tmax, "simulation_time", w);
    take(I["radiation_on"], p.radiation_on, "radiation_on", w);
    take(I["cooling_on"], p.cooling_on, "cooling_on", w);
    take(I["debug"], p.debug, "debug", w);
    take(I["initial_conditions"], p.initialConditions, "initial_conditions", w);
    take(I["ncheckpoints"], p.ncheckpoints, "ncheckpoints", w);
    take(I["output_directory"], p.outputDirectory, "output_directory", w);
    take(G["no_dimensions"], p.nd, "no_dimensions", w);
    take(G["no_cells_x"], p.ncells[0], "no_cells_x", w);
    take(G["no_cells_y"], p.ncells[1], "no_cells_y", w);
    take(G["no_cells_z"], p.ncells[2], "no_cells_z", w);
    take(G["side_length"], p.sideLength, "side_length", w);
    take(G["geometry"], p.geometry, "geometry", w);
    if (G["Patch"].exists()) {
        const LuaValue& Pt = G["Patch"];
        take(Pt["filename"], p.patchfilename, "filename", w);
        take(Pt["offset_x"], p.patchoffset[0], "offset_x", w);
        take(Pt["offset_y"], p.patchoffset[1], "offset_y", w);
        take(Pt["offset_z"], p.patchoffset[2], "offset_z", w);
    }
    const char* xyz[3] = {"x", "y", "z"};
    for (int d = 0; d < 3; ++d) {
        take(G[std::string("left_boundary_condition_") + xyz[d]], p.leftBC[d], "left_boundary_condition", w);
        take(G[std::string("right_boundary_condition_") + xyz[d]], p.rightBC[d], "right_boundary_condition", w);
    }
    take(H["gamma"], p.heatCapacityRatio, "gamma", w);
    take(H["density_floor"], p.dfloor, "density_floor", w);
    take(H["pressure_floor"], p.pfloor, "pressure_floor", w);
    take(H["temperature_floor"], p.tfloor, "temperature_floor", w);
    take(H["riemann_solver"], p.riemannSolver, "riemann_solver", w);
    take(H["slope_limiter"], p.slopeLimiter, "slope_limiter", w);
    if (H["precision"].exists()) take(H["precision"], p.precision, "precision", w);
    take(R["K1"], p.K1, "K1", w);
    take(R["K2"], p.K2, "K2", w);
    take(R["K3"], p.K3, "K3", w);
    take(R["K4"], p.K4, "K4", w);
    take(R["photoion_cross_section"], p.photoIonCrossSection, "photoion_cross_section", w);
    take(R["case_b_recombination_coeff"], p.alphaB, "case_b_recombination_coeff", w);
    take(R["tau_0"], p.tau0, "tau_0", w);
    take(R["minimum_hii_fraction"], p.minX, "minimum_hii_fraction", w);
    take(R["temperature_hi"], p.THI, "temperature_hi", w);
    take(R["temperature_hii"], p.THII, "temperature_hii", w);
    take(R["mass_fraction_hydrogen"], p.massFractionH, "mass_fraction_hydrogen", w);
    take(R["collisions_on"], p.collisions_on, "collisions_on", w);
    take(R["coupling"], p.rt_coupling, "coupling", w);
    take(R["integration_scheme"], p.rt_scheme, "integration_scheme", w);
    take(T["thermo_hii_switch"], p.thermoHII_Switch, "thermo_hii_switch", w);
    take(T["heating_amplification"], p.heatingAmplification, "heating_amplification", w);
    take(T["thermo_subcycling"], p.thermoSubcycling, "thermo_subcycling", w);
    take(T["min_temp_initial_state"], p.minTempInitialState, "min_temp_initial_state", w);
    take(S["on"], p.star_on, "on", w);
    take(S["cell_position_x"], p.star_position[0], "cell_position_x", w);
    take(S["cell_position_y"], p.star_position[1], "cell_position_y", w);
    take(S["cell_position_z"], p.star_position[2], "cell_position_z", w);
    take(S["snap_to_face_left_x"], p.faceSnap[0], "snap_to_face_left_x", w);
    take(S["snap_to_face_left_y"], p.faceSnap[1], "snap_to_face_left_y", w);
    take(S["snap_to_face_left_z"], p.faceSnap[2], "snap_to_face_left_z", w);
    take(S["photon_energy"], p.photonEnergy, "photon_energy", w);
    take(S["photon_rate"], p.photonRate, "photon_rate", w);
    take(S["wind_radius_in_cells"], p.windCellRadius, "wind_radius_in_cells", w);
    take(S["mass_loss_rate"], p.massLossRate, "mass_loss_rate", w);
    take(S["wind_velocity"], p.windVelocity, "wind_velocity", w);
    take(S["wind_temperature"], p.windTemperature, "wind_temperature", w);
}

// ---------------------------------------------------------------------------
// Fluid
// ---------------------------------------------------------------------------
Fluid::~Fluid() {
    if (ctx_) torchgpu_destroy(ctx_);
}

void Fluid::initialise(const TorchParameters& p, void* stream) {
    torchgpu_params g = p.gridParams();
    heatCapacityRatio = p.heatCapacityRatio;
    massFractionH = p.massFractionH;
    check_rc(torchgpu_create(&g, stream, &ctx_), "torchgpu_create");
}

void Fluid::setPrimitives(const double* Q) { check_rc(torchgpu_set_primitive(ctx_, Q, 0), "set_primitive", ctx_); }
void Fluid::getPrimitives(double* Q, bool fix) const {
    check_rc(torchgpu_get_primitive(ctx_, Q, fix ? 1 : 0, 0), "get_primitive", ctx_);
}
void Fluid::getConserved(double* U) const { check_rc(torchgpu_get_conserved(ctx_, U, 0), "get_conserved", ctx_); }
void Fluid::check() const { check_rc(torchgpu_synchronize(ctx_), "device", ctx_); }

// Q is never materialised: these record which state the next flux
// evaluation derives Q from.
void Fluid::globalQfromU() { qsrc_ = u_is_w_ ? QSource::U : (w_saved_ ? QSource::PREDICTOR : QSource::U); }
void Fluid::fixPrimitives() {}
void Fluid::globalWfromU() {
    w_saved_ = true;
    u_is_w_ = false;
    qsrc_ = QSource::U;
}
void Fluid::globalUfromW() { u_is_w_ = true; }

void Fluid::advSolution(double dt) {
    if (pending == Pending::HYDRO) {
        int stage;
        if (w_saved_ && !u_is_w_) stage = 1;                       // predictor of hydroStep
        else if (w_saved_ && u_is_w_ && qsrc_ == QSource::PREDICTOR) stage = 2;  // corrector
        else stage = 0;                                             // single-stage sub-step
        check_rc(torchgpu_hydro_stage_dt(ctx_, stage, dt), "hydro stage", ctx_);
        if (stage == 2) w_saved_ = u_is_w_ = false;
    } else if (pending == Pending::THERMO) {
        check_rc(torchgpu_thermo_grid_substep(ctx_, cooling, nullptr, nullptr, nullptr, cooling_excluded, dt, 0),
                 "thermo substep",
                 ctx_);
    } else if (pending == Pending::RAD) {
        // Torch::subStep has made preTimeStepCalculations on this state
        check_rc(torchgpu_radiation_substep(radiation, dt, 1), "radiation substep", ctx_);
    }
    pending = Pending::NONE;
}

void Fluid::fixSolution() {}  // fused into the launch of advSolution

// ---------------------------------------------------------------------------
// integrators
// ---------------------------------------------------------------------------
void Hydrodynamics::preTimeStepCalculations(Fluid&) const {}  // sound speeds are recomputed in-kernel

double Hydrodynamics::calculateTimeStep(double dt_max, Fluid& fluid) const {
    double dt = 0;
    check_rc(torchgpu_hydro_timestep(fluid.ctx(), dt_max, &dt), "hydro timestep", fluid.ctx());
    return dt;
}

void Hydrodynamics::integrate(double, Fluid& fluid) const { fluid.pending = Fluid::Pending::HYDRO; }
void Hydrodynamics::updateSourceTerms(double, Fluid&) const {}

Thermodynamics::~Thermodynamics() {
    if (cool_) torchgpu_cooling_destroy(cool_);
}

void Thermodynamics::initialise(const TorchParameters& p, void* stream) {
    torchgpu_thermo_params t;
    // the Lua-level floors (cgs density floor): TorchParameters keeps dfloor
    // in code units after initialise, so convert back for the thermo set-up
    const double dfloor_cgs = convert(p, p.dfloor, 1, -3, 0, false);
    check_rc(torchgpu_thermo_params_init(&t, p.dscale, p.pscale, p.tscale, p.thermoHII_Switch, p.heatingAmplification,
                                         p.massFractionH, p.heatCapacityRatio, dfloor_cgs, p.tfloor,
                                         p.thermoSubcycling ? 1 : 0),
             "thermo params");
    t.pfloor = p.pfloor;
    check_rc(torchgpu_cooling_create(&t, stream, &cool_), "cooling create");
    // Fluid::initialiseGrid (Fluid.cpp:41-54) puts the star's cell, and the
    // cells within the wind radius, on the CausalWind list whether or not the
    // star is on, and Thermodynamics only walks CausalNonWind: with the star
    // off that is the one cell at Star.cell_position (wind radius 0)
    if (!p.star_on) {
        if (p.windCellRadius > 0 && p.cooling_on)
            throw std::runtime_error("Star.wind_radius_in_cells > 0 with Star.on = false: the thermo operator skips "
                                     "one excluded cell only");
        const torchgpu_params g = p.gridParams();
        for (int a = 0; a < 3; ++a) {
            const int na = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
            if (p.star_position[a] < 0 || p.star_position[a] >= na)
                throw std::runtime_error("Star.cell_position outside the grid");
        }
        excluded_ = p.star_position[0] + (int64_t)g.nx * (p.star_position[1] + (int64_t)g.ny * p.star_position[2]);
    }
}

void Thermodynamics::initialiseMinTempField(const TorchParameters& p, const double* Qcgs, size_t n) {
    // Thermodynamics.cpp:102-111, called by Torch::initialise before
    // toCodeUnits (Torch.cpp:113-117): with min_temp_initial_state each
    // cell's T_min is Fluid::calcTemperature of its cgs primitives with the
    // code-unit gas constant, as the reference computes it; else the
    // constant m_T_min = 100 the library uses by default
    if (!p.minTempInitialState) {
        check_rc(torchgpu_cooling_set_min_temp(cool_, nullptr, 0, 0), "min temp field");
        return;
    }
    std::vector<double> T(n);
    for (size_t c = 0; c < n; ++c) {
        const double hii = Qcgs[2 * n + c], pre = Qcgs[1 * n + c], den = Qcgs[0 * n + c];
        const double mu_inv = p.massFractionH * (hii + 1.0) + (1.0 - p.massFractionH) * 0.25;
        T[c] = (pre / den) * (1.0 / mu_inv) / p.specificGasConstant;
    }
    check_rc(torchgpu_cooling_set_min_temp(cool_, T.data(), (int64_t)n, 0), "min temp field");
}

void Thermodynamics::preTimeStepCalculations(Fluid&) const {}  // rates are evaluated with integrate

double Thermodynamics::calculateTimeStep(double dt_max, Fluid& fluid) const {
    double dt = 0;
    check_rc(torchgpu_thermo_grid_timestep(fluid.ctx(), cool_, nullptr, nullptr, nullptr, excluded_, dt_max, &dt),
             "thermo timestep", fluid.ctx());
    return dt;
}

void Thermodynamics::integrate(double, Fluid& fluid) const {
    fluid.cooling = cool_;
    fluid.cooling_excluded = excluded_;
    fluid.pending = Fluid::Pending::THERMO;
}
void Thermodynamics::updateSourceTerms(double, Fluid&) const {}

torchgpu_star starOf(const TorchParameters& p) {
    // Star::initialise (Star.cpp:17-21): snapped stars sit on a cell face,
    // which the shell sweeps of the library do not cover
    if (p.faceSnap[0] || p.faceSnap[1] || p.faceSnap[2])
        throw std::runtime_error("Star.snap_to_face_left_*: stars on a cell face are not supported on the GPU path");
    torchgpu_star s;
    for (int a = 0; a < 3; ++a) s.cell[a] = p.star_position[a];
    s.wind_cell_radius = p.windCellRadius;
    s.photon_rate = p.photonRate;  // code units after TorchParameters::initialise
    return s;
}

std::vector<double> evalLuaSetup(const TorchParameters& p, const std::string& filename) {
    const torchgpu_params g = p.gridParams();
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    std::vector<double> Q(7 * n, 0.0);
    lua::State L;
    try {
        L.load_file(filename);
    } catch (lua::Error& e) {
        throw std::runtime_error("Torch::setUpLua: could not open lua file: " + filename + "\n" + e.what());
    }
    // Grid::initialise: dx = sideLength / ncells[0]; cell centres at index +
    // 0.5, the star at its cell + 0.5 (Star.cpp:17-21), both times dx and
    // converted to cm as setUpLua does
    const double dx = p.sideLength / g.nx;
    double xs[3];
    for (int a = 0; a < 3; ++a) {
        const double mod = (p.faceSnap[a] && p.star_position[a] == 0) ? 0.5 : 0.0;
        xs[a] = convert(p, (p.star_position[a] + 0.5 - mod) * dx, 0, 1, 0, false);
    }
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
                const int idx[3] = {i, j, k};
                std::vector<double> args(6);
                for (int a = 0; a < 3; ++a) {
                    args[a] = convert(p, ((double)idx[a] + 0.5) * dx, 0, 1, 0, false);
                    args[3 + a] = xs[a];
                }
                std::vector<lua::Value> r;
                try {
                    r = L.call("initialise", args);
                } catch (lua::Error& e) {
                    throw std::runtime_error(std::string("Torch::setUpLua: ") + e.what());
                }
                double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (size_t q = 0; q < 9 && q < r.size(); ++q) {
                    if (r[q].kind == lua::Value::NUMBER) v[q] = r[q].num;
                    else if (r[q].kind != lua::Value::NIL)
                        throw std::runtime_error("Torch::setUpLua: initialise returned a non-number");
                }
                if (v[6] != 0 || v[7] != 0 || v[8] != 0)
                    throw std::runtime_error("Torch::setUpLua: gravity (returns 7-9) is not modelled on the GPU path");
                const size_t c = (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
                Q[0 * n + c] = v[0];
                Q[1 * n + c] = v[1];
                Q[2 * n + c] = v[2];
                for (int d = 0; d < 3; ++d) Q[(4 + d) * n + c] = v[3 + d];
            }
    return Q;
}

void Torch::setUpLua(const std::string& filename) {
    const std::vector<double> Q = evalLuaSetup(params, filename);
    setPrimitivesCgs(Q.data());
}

Radiation::~Radiation() {
    if (rad_) torchgpu_radiation_destroy(rad_);
}

void Radiation::initialise(const TorchParameters& p, Fluid& fluid) {
    // Radiation::initialise (Radiation.cpp:24-56) with the Lua-level values
    // (the cross-section and photon energy stay in cgs here; the library
    // converts them as TorchParameters::initialise does)
    torchgpu_rad_params r;
    const double K[4] = {p.K1, p.K2, p.K3, p.K4};
    const int scheme = p.rt_scheme == "implicit2" ? TORCHGPU_RT_IMPLICIT2
                       : p.rt_scheme == "explicit" ? TORCHGPU_RT_EXPLICIT
                                                   : TORCHGPU_RT_IMPLICIT;
    const int coupling = p.rt_coupling == "neq" ? TORCHGPU_RT_NEQ
                         : p.rt_coupling == "tti" ? TORCHGPU_RT_TTI
                                                  : TORCHGPU_RT_OFF;
    check_rc(torchgpu_rad_params_init(&r, p.dscale, p.pscale, p.tscale, K, p.photoIonCrossSection, p.tau0, p.minX,
                                      p.THI, p.THII, scheme, coupling, p.collisions_on ? 1 : 0,
                                      p.heatingAmplification, p.massFractionH, p.photonEnergy),
             "radiation params");
    const torchgpu_star star = starOf(p);
    check_rc(torchgpu_radiation_create(&r, fluid.ctx(), &star, &rad_), "radiation create", fluid.ctx());
}

void Radiation::initField(const double* Q) {
    check_rc(torchgpu_radiation_init_field(rad_, Q, 0), "radiation initField");
}

void Radiation::preTimeStepCalculations(Fluid& fluid) const {
    check_rc(torchgpu_radiation_pre_step(rad_), "radiation preTimeStep", fluid.ctx());
}

double Radiation::calculateTimeStep(double dt_max, Fluid& fluid) const {
    double dt = 0;
    check_rc(torchgpu_radiation_timestep(rad_, dt_max, &dt), "radiation timestep", fluid.ctx());
    return dt;
}

void Radiation::integrate(double, Fluid& fluid) const {
    fluid.radiation = rad_;
    fluid.pending = Fluid::Pending::RAD;
}
void Radiation::updateSourceTerms(double, Fluid&) const {}

// ---------------------------------------------------------------------------
// driver (Torch.cpp)
// ---------------------------------------------------------------------------
void Torch::initialise(const TorchParameters& p_in, void* stream) {
    params = p_in;
    // Integration.initial_conditions (Torch.cpp:38-45): cell counts, side
    // length (cm) and dimensions from the snapshot's header
    double t0 = 0, side = 0;
    int nx = 0, ny = 0, nz = 0, nd = 0;
    const std::string& ic = params.initialConditions;
    if (!ic.empty()) {
        if (torchgpu_snapshot_header(ic.c_str(), &t0, &nx, &ny, &nz, &nd, &side) != TORCHGPU_OK)
            throw std::runtime_error("DataReader::readDataParameters: invalid input file " + ic + ".");
        params.ncells = {{nx, ny, nz}};
        params.sideLength = side;
        params.nd = nd;
    }
    params.initialise();
    fluid.initialise(params, stream);
    // Torch::initialise sets up Thermodynamics for every run (Torch.cpp:80-81):
    // Torch::run fills and prints its heating arrays whether or not cooling is on
    thermodynamics.initialise(params, stream);
    if (params.star_on) {
        // Thermodynamics::preTimeStepCalculations (and fillHeatingArrays) with Star.on
        const torchgpu_star star = starOf(params);
        check_rc(torchgpu_cooling_set_star(thermodynamics.handle(), fluid.ctx(), &star), "cooling star", fluid.ctx());
    }
    if (params.radiation_on) {
        if (!params.star_on)
            throw std::runtime_error("radiation_on without Star.on: the reference's Radiation then changes nothing "
                                     "but the printed HII fraction, which the GPU path does not model");
        radiation.initialise(params, fluid);
    }
    currentTime = 0;
    steps = 0;
    first_time_ = true;
    quitting_ = false;
    stepCounter_ = 0;
    if (!ic.empty()) {
        // DataReader::readGrid, then the tail of Torch::initialise
        // (Torch.cpp:58, 100-103, 117-119)
        std::vector<double> Q((size_t)7 * nx * ny * nz, 0.0);
        if (torchgpu_snapshot_read(ic.c_str(), Q.data(), nx, ny, nz) != TORCHGPU_OK)
            throw std::runtime_error("DataReader::readGrid: invalid input file " + ic + ".");
        setPrimitivesCgs(Q.data());
        currentTime = params.toCodeUnits(t0, 0, 0, 1);
        check_rc(torchgpu_set_time(fluid.ctx(), currentTime), "set time", fluid.ctx());
    }
}

void Torch::setPrimitivesCgs(const double* Qcgs) {
    // Torch::toCodeUnits (Torch.cpp:146-155), then fixPrimitives + globalUfromQ
    const torchgpu_params g = params.gridParams();
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    std::vector<double> Q(Qcgs, Qcgs + 7 * n);
    if (!params.patchfilename.empty()) {
        // Grid.Patch: DataReader::patchGrid after the set-up (Torch.cpp:110-111)
        if (torchgpu_snapshot_patch(params.patchfilename.c_str(), Q.data(), g.nx, g.ny, g.nz, g.side_length / g.nx,
                                    params.patchoffset[0], params.patchoffset[1], params.patchoffset[2]) != TORCHGPU_OK)
            throw std::runtime_error("DataReader::patchGrid: cannot patch the grid with " + params.patchfilename);
    }
    if (params.cooling_on) thermodynamics.initialiseMinTempField(params, Q.data(), n);
    const double kd = params.toCodeUnits(1.0, 1, -3, 0), kp = params.toCodeUnits(1.0, 1, -1, -2),
                 kv = params.toCodeUnits(1.0, 0, 1, -1);
    const bool unit = (kd == 1.0 && kp == 1.0 && kv == 1.0);
    for (size_t c = 0; c < n && !unit; ++c) {
        Q[0 * n + c] = params.toCodeUnits(Q[0 * n + c], 1, -3, 0);
        Q[1 * n + c] = params.toCodeUnits(Q[1 * n + c], 1, -1, -2);
        for (int d = 0; d < 3; ++d) Q[(4 + d) * n + c] = params.toCodeUnits(Q[(4 + d) * n + c], 0, 1, -1);
    }
    fluid.setPrimitives(Q.data());
    // Torch::initialise: radiation.initField after globalUfromQ (Torch.cpp:121)
    if (params.radiation_on) radiation.initField(Q.data());
}

double Torch::calculateTimeStep() {  // Torch.cpp:304-339 (single process: the MPI min is the identity)
    double dt;
    if (first_time_) {
        dt = params.dt_max * 1.0e-20;
        first_time_ = false;
    } else {
        double dt_hydro = hydrodynamics.calculateTimeStep(params.dt_max, fluid);
        double dt_rad = dt_hydro, dt_thermo = dt_hydro;
        if (params.radiation_on) dt_rad = radiation.calculateTimeStep(params.dt_max, fluid);
        if (params.cooling_on) dt_thermo = thermodynamics.calculateTimeStep(params.dt_max, fluid);
        dt = std::min(std::min(dt_hydro, dt_rad), dt_thermo);
        if (params.debug) {
            // the reference's test reads the hydro percentage in all three
            // terms of its condition (Torch.cpp:321-332)
            const double thyd = 100.0 * dt_hydro / params.tmax;
            if (thyd <= 1.0e-6) {
                std::fprintf(stderr, "Integration deltas are too small.\n");
                quitting_ = true;
            }
        }
    }
    return dt;
}

void Torch::subStep(double dt, bool hasCalculatedHeatFlux, Integrator& comp) {  // Torch.cpp:352-364
    if (!hasCalculatedHeatFlux) {
        fluid.globalQfromU();
        fluid.fixPrimitives();
        comp.preTimeStepCalculations(fluid);
    }
    comp.integrate(dt, fluid);
    comp.updateSourceTerms(dt, fluid);
    fluid.advSolution(dt);
    fluid.fixSolution();
}

void Torch::hydroStep(double dt, bool hasCalculatedHeatFlux) {  // Torch.cpp:366-387
    fluid.globalWfromU();
    if (!hasCalculatedHeatFlux) {
        fluid.globalQfromU();
        fluid.fixPrimitives();
        hydrodynamics.preTimeStepCalculations(fluid);
    }
    hydrodynamics.integrate(dt, fluid);
    hydrodynamics.updateSourceTerms(dt, fluid);
    fluid.advSolution(dt / 2.0);
    fluid.fixSolution();
    fluid.globalQfromU();
    fluid.globalUfromW();
    hydrodynamics.integrate(dt, fluid);
    hydrodynamics.updateSourceTerms(dt, fluid);
    fluid.advSolution(dt);
    fluid.fixSolution();
}

double Torch::fullStep(double dt_nextCheckPoint) {  // Torch.cpp:389-418
    fluid.globalQfromU();
    fluid.fixPrimitives();
    if (params.cooling_on) thermodynamics.preTimeStepCalculations(fluid);
    if (params.radiation_on) radiation.preTimeStepCalculations(fluid);
    double dt = std::min(dt_nextCheckPoint, calculateTimeStep());
    // activeComponents of Torch::run (Torch.cpp:256-261)
    std::vector<Integrator*> comps{&hydrodynamics};
    if (params.cooling_on) comps.push_back(&thermodynamics);
    if (params.radiation_on) comps.push_back(&radiation);
    const int ncomps = (int)comps.size();
    if (ncomps == 1) {
        hydroStep(dt, true);
        return dt;
    }
    stepCounter_ = (stepCounter_ + 1) % ncomps;
    for (int i = 0; i < ncomps; ++i) {
        double h = (i == ncomps - 1) ? 1.0 : 0.5;
        subStep(h * dt, i == 0, *comps[(i + stepCounter_) % ncomps]);
    }
    for (int i = ncomps - 2; i >= 0; --i) subStep(dt / 2.0, false, *comps[(i + stepCounter_) % ncomps]);
    return dt;
}

void Torch::print2D(int count) const {
    // DataPrinter::print2D (DataPrinter.cpp:139-174) of the state Fluid::Q
    // holds at this point of Torch::run
    const torchgpu_params g = params.gridParams();
    std::vector<double> Q((size_t)7 * g.nx * g.ny * g.nz);
    check_rc(torchgpu_get_print_state(fluid.ctx(), Q.data(), 0), "print state", fluid.ctx());
    double t = 0, dt = 0;
    int n = 0;
    check_rc(torchgpu_get_clock(fluid.ctx(), &t, &dt, &n), "clock", fluid.ctx());
    std::filesystem::create_directories(params.outputDirectory);
    char name[64];
    std::snprintf(name, sizeof name, "data2D_%06d.txt.gz", count);
    const std::string path = (std::filesystem::path(params.outputDirectory) / name).string();
    if (torchgpu_snapshot_write(path.c_str(), 1, t, Q.data(), g.nx, g.ny, g.nz, params.sideLength, params.dscale,
                                params.pscale, params.tscale) != TORCHGPU_OK)
        throw std::runtime_error("DataPrinter::print2D: unable to open" + path);
}

std::vector<double> Torch::fillHeating() {
    // Thermodynamics::fillHeatingArrays (Thermodynamics.cpp:419-461) on the
    // printed state; the star's ray trace and FUV flux come from the star
    // set on the cooling handle
    const torchgpu_params g = params.gridParams();
    std::vector<double> H((size_t)TORCHGPU_HEAT_N * g.nx * g.ny * g.nz);
    const int rc = torchgpu_fill_heating(fluid.ctx(), thermodynamics.handle(), nullptr, nullptr,
                                         params.star_on ? -1 : thermodynamics.excluded(), H.data(), 0);
    if (rc != TORCHGPU_OK) {
        char msg[256];
        torchgpu_cooling_last_error(thermodynamics.handle(), msg, sizeof msg);
        throw std::runtime_error(std::string("fillHeatingArrays: ") + msg);
    }
    return H;
}

void Torch::printHeating(int count, const std::vector<double>& H) const {
    // DataPrinter::printHeating (DataPrinter.cpp:325-354)
    const torchgpu_params g = params.gridParams();
    double t = 0, dt = 0;
    int n = 0;
    check_rc(torchgpu_get_clock(fluid.ctx(), &t, &dt, &n), "clock", fluid.ctx());
    std::filesystem::create_directories(params.outputDirectory);
    char name[64];
    std::snprintf(name, sizeof name, "heating_%06d.txt.gz", count);
    const std::string path = (std::filesystem::path(params.outputDirectory) / name).string();
    if (torchgpu_heating_write(path.c_str(), 1, t, H.data(), g.nx, g.ny, g.nz, params.dscale, params.pscale,
                               params.tscale) != TORCHGPU_OK)
        throw std::runtime_error("DataPrinter::printHeating: unable to open" + path);
}

void Torch::run(int max_steps, bool print, bool host_loop) {  // Torch.cpp:240-302, capped at max_steps
    torchgpu_ctx* c = fluid.ctx();
    // the whole Torch::run loop on the device: dt, the pre-step
    // calculations and the sub-steps of every active component
    auto device_run = [&](int n) {
        check_rc(torchgpu_run_full(c, params.cooling_on ? thermodynamics.handle() : nullptr,
                                   params.radiation_on ? radiation.handle() : nullptr, nullptr, nullptr, nullptr,
                                   params.cooling_on ? thermodynamics.excluded() : -1, n, 0),
                 "run", c);
    };
    if (print) {
        // the device loop pauses at every checkpoint for the print
        if (!params.star_on && params.windCellRadius > 0)
            throw std::runtime_error("heating output with Star.on = false and Star.wind_radius_in_cells > 0: only "
                                     "the star's own cell is kept out of the heating arrays");
        check_rc(torchgpu_heating_track(c, 1), "heating track", c);
        int count = 0, due = 0;
        check_rc(torchgpu_checkpoint(c, &count, &due), "checkpoint", c);
        print2D(count);
        check_rc(torchgpu_pause_at_checkpoints(c, 1), "pause at checkpoints", c);
        fillHeating();  // Torch.cpp:264
        bool final_print = false;
        int done = 0;
        double t = 0, dt = 0;
        int n0 = 0, n1 = 0;
        while (done < max_steps) {
            check_rc(torchgpu_get_clock(c, &t, &dt, &n0), "clock", c);
            device_run(max_steps - done);
            fluid.check();
            check_rc(torchgpu_get_clock(c, &t, &dt, &n1), "clock", c);
            done += n1 - n0;
            check_rc(torchgpu_checkpoint(c, &count, &due), "checkpoint", c);
            if (!due) break;
            // Torch.cpp:271-276: fillHeatingArrays, printHeating, print2D
            printHeating(count, fillHeating());
            print2D(count);
            final_print = count != params.ncheckpoints;
        }
        check_rc(torchgpu_get_clock(c, &t, &dt, &n1), "clock", c);
        int quit = 0;
        check_rc(torchgpu_is_quitting(c, &quit), "quitting", c);
        quitting_ = quit != 0;
        if (quitting_) std::fprintf(stderr, "Integration deltas are too small.\n");
        // the print after Torch::run's loop (Torch.cpp:295-297)
        if (final_print && (!(t < params.tmax) || quitting_)) print2D(params.ncheckpoints);
        check_rc(torchgpu_pause_at_checkpoints(c, 0), "pause at checkpoints", c);
        currentTime = t;
        steps = n1;
        return;
    }
    if (!host_loop) {
        device_run(max_steps);
        fluid.check();
        int n = 0;
        double t = 0, dt = 0;
        check_rc(torchgpu_get_clock(c, &t, &dt, &n), "clock", c);
        int quit = 0;
        check_rc(torchgpu_is_quitting(c, &quit), "quitting", c);
        quitting_ = quit != 0;
        if (quitting_) std::fprintf(stderr, "Integration deltas are too small.\n");
        currentTime = t;
        steps = n;
        return;
    }
    // host-driven: Torch::fullStep through the library's operator calls
    int count = 0;
    const double delta = 100.0 / params.ncheckpoints;
    auto checkpoint = [&](double t) {  // Checkpointer::update, Checkpointer.cpp:14-26
        double pct = std::min(100.0, 100.0 * t / params.tmax);
        while (pct - count * delta > delta - 1.0e-8) count += 1;
        return ((count + 1) * delta - pct) * params.tmax / 100.0;
    };
    checkpoint(currentTime);
    while (currentTime < params.tmax && !quitting_ && steps < max_steps) {
        double dt_next = checkpoint(currentTime);
        double dt = fullStep(dt_next);
        currentTime += dt;
        ++steps;
    }
    fluid.check();
}

}  // namespace tg
