// ref_harness.cpp -- Lua-free driver around the reference's own hot-path
// translation units (hgsteggles/TORCH src/Integrators, src/Fluid, src/Torch,
// src/MPI, src/IO/Logger, src/IO/ProgressBar), compiled unmodified from
// /root/reference by oracle/Makefile.ref into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: it writes the golden fixtures under tests/golden/
// and times the reference's FP64 CPU step.  It replaces what Torch.cpp does
// with Lua (Torch::initialise, setUpLua, toCodeUnits) and restates the driver
// loop of Torch::run / fullStep / hydroStep / subStep / calculateTimeStep
// (src/Torch/Torch.cpp:240-418) because Torch.cpp itself needs the Lua
// library, which is only shipped prebuilt and is not used here.
//
// Modes:
//   hydro   <case.txt>              run a hydro-only problem, dump U at steps
//   riemann <solver> <gamma> <n> <in.bin> <out.bin>
//   thermo  <case.txt>              per-cell preTimeStep + integrate
//   consts  <dscale> <pscale> <tscale>   print unit-converted constants
//   raytrace <case.txt>             star geometry + thermo column densities
//   radiation <case.txt>            Radiation integrator sub-steps

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

// Thermodynamics keeps its rate constants private; the harness reads them to
// pin the oracle's unit conversion.  Access changes only in this translation
// unit; the reference objects are compiled untouched.
#define private public
#define protected public
#include "Fluid/Fluid.hpp"
#include "Fluid/Grid.hpp"
#include "Fluid/GridCell.hpp"
#include "Integrators/Hydro.hpp"
#include "Integrators/Riemann.hpp"
#include "Integrators/SlopeLimiter.hpp"
#include "Integrators/Thermodynamics.hpp"
#include "Integrators/Radiation.hpp"
#include "MPI/MPI_Wrapper.hpp"
#include "Torch/Constants.hpp"
#include "Torch/Parameters.hpp"
#undef private
#undef protected

typedef std::map<std::string, std::string> KV;

static KV read_case(const std::string& fn) {
    KV kv;
    std::ifstream f(fn);
    if (!f) throw std::runtime_error("cannot open " + fn);
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string k, v;
        ss >> k;
        std::getline(ss, v);
        size_t b = v.find_first_not_of(" \t");
        kv[k] = (b == std::string::npos) ? "" : v.substr(b);
    }
    return kv;
}
static double getd(const KV& kv, const std::string& k, double def) {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::stod(it->second);
}
static int geti(const KV& kv, const std::string& k, int def) {
    auto it = kv.find(k);
    return it == kv.end() ? def : std::stoi(it->second);
}
static std::string gets(const KV& kv, const std::string& k, const std::string& def) {
    auto it = kv.find(k);
    return it == kv.end() ? def : it->second;
}

struct Sim {
    std::shared_ptr<Constants> consts;
    TorchParameters p;
    Fluid fluid;
    Hydrodynamics hydro;
    Thermodynamics thermo;
    Radiation rad;
};

// Torch::initialise (Torch.cpp:29-99) minus IO, Lua, radiation.
static void init_sim(Sim& s, const KV& kv) {
    TorchParameters& p = s.p;
    p.dscale = getd(kv, "density_scale", 1.0);
    p.pscale = getd(kv, "pressure_scale", 1.0);
    p.tscale = getd(kv, "time_scale", 1.0);
    p.spatialOrder = geti(kv, "spatial_order", 1);
    p.tmax = getd(kv, "simulation_time", 1e3);
    p.ncheckpoints = geti(kv, "ncheckpoints", 1);
    p.debug = geti(kv, "debug", 1) != 0;  // TorchParameters default (Parameters.hpp:50)
    p.cooling_on = geti(kv, "cooling_on", 0) != 0;
    p.radiation_on = geti(kv, "radiation_on", 0) != 0;
    p.nd = 3;
    p.ncells = {{geti(kv, "nx", 8), geti(kv, "ny", 8), geti(kv, "nz", 8)}};
    p.sideLength = getd(kv, "side_length", 1.0);
    p.geometry = "cartesian";
    const char* faces[3] = {"x", "y", "z"};
    for (int i = 0; i < 3; ++i) {
        p.leftBC[i] = gets(kv, std::string("left_boundary_condition_") + faces[i], "outflow");
        p.rightBC[i] = gets(kv, std::string("right_boundary_condition_") + faces[i], "outflow");
    }
    p.heatCapacityRatio = getd(kv, "gamma", 1.67);
    p.dfloor = getd(kv, "density_floor", 1e-30);
    p.tfloor = getd(kv, "temperature_floor", 0.0);
    p.riemannSolver = gets(kv, "riemann_solver", "RotatedHLLC");
    p.slopeLimiter = gets(kv, "slope_limiter", "albada");
    p.massFractionH = getd(kv, "mass_fraction_hydrogen", 1.0);
    p.thermoHII_Switch = getd(kv, "thermo_hii_switch", 1e-2);
    p.heatingAmplification = getd(kv, "heating_amplification", 1.0);
    p.thermoSubcycling = geti(kv, "thermo_subcycling", 1) != 0;
    p.minTempInitialState = geti(kv, "min_temp_initial_state", 0) != 0;
    p.star_on = geti(kv, "star_on", 0) != 0;
    p.photonRate = getd(kv, "photon_rate", 0.0);
    p.star_position = {{geti(kv, "star_x", 0), geti(kv, "star_y", 0), geti(kv, "star_z", 0)}};
    p.windCellRadius = geti(kv, "wind_radius", 0);
    p.photonEnergy = getd(kv, "photon_energy", 0.0);
    p.K1 = getd(kv, "K1", 0.0);
    p.K2 = getd(kv, "K2", 0.0);
    p.K3 = getd(kv, "K3", 0.0);
    p.K4 = getd(kv, "K4", 0.0);
    p.photoIonCrossSection = getd(kv, "photoion_cross_section", 0.0);
    p.alphaB = getd(kv, "case_b_recombination_coeff", 0.0);
    p.tau0 = getd(kv, "tau_0", 0.0);
    p.minX = getd(kv, "minimum_hii_fraction", 0.0);
    p.THI = getd(kv, "temperature_hi", 0.0);
    p.THII = getd(kv, "temperature_hii", 0.0);
    p.collisions_on = geti(kv, "collisions_on", 0) != 0;
    p.rt_coupling = gets(kv, "coupling", "off");
    p.rt_scheme = gets(kv, "integration_scheme", "implicit");

    s.consts = std::make_shared<Constants>();
    s.consts->initialise_DPT(p.dscale, p.pscale, p.tscale);
    p.initialise(s.consts);
    s.consts->nd = p.nd;
    s.consts->dfloor = p.dfloor;
    s.consts->pfloor = p.pfloor;
    s.consts->tfloor = p.tfloor;
    s.fluid.initialise(s.consts, p.getFluidParameters());
    s.fluid.initialiseGrid(p.getGridParameters(), p.getStarParameters());
    s.hydro.initialise(s.consts);
    try {
        s.hydro.setRiemannSolver(RiemannSolverFactory::create(p.riemannSolver, p.nd));
    } catch (std::exception& e) {
        std::cerr << "warning: " << e.what();
    }
    try {
        s.hydro.setSlopeLimiter(SlopeLimiterFactory::create(p.slopeLimiter));
    } catch (std::exception& e) {
        std::cerr << "warning: " << e.what();
    }
    s.thermo.initialise(s.consts, p.getThermoParameters());
    s.rad.initialise(s.consts, p.getRadiationParameters());
}

// Torch::toCodeUnits, Torch.cpp:146-155.
static void to_code_units(Sim& s) {
    for (GridCell& cell : s.fluid.getGrid().getIterable("GridCells")) {
        cell.Q[UID::DEN] = s.consts->converter.toCodeUnits(cell.Q[UID::DEN], 1, -3, 0);
        cell.Q[UID::PRE] = s.consts->converter.toCodeUnits(cell.Q[UID::PRE], 1, -1, -2);
        for (int idim = 0; idim < s.consts->nd; ++idim)
            cell.Q[UID::VEL + idim] = s.consts->converter.toCodeUnits(cell.Q[UID::VEL + idim], 0, 1, -1);
        for (int idim = 0; idim < s.consts->nd; ++idim)
            cell.GRAV[idim] = s.consts->converter.toCodeUnits(cell.GRAV[idim], 1, -2, -2);
    }
}

static void check_values(Sim& s, const std::string& name) {  // Torch.cpp:420-443
    for (GridCell& cell : s.fluid.getGrid().getIterable("GridCells"))
        for (int i = 0; i < UID::N; ++i)
            if (cell.U[i] != cell.U[i] || std::isinf(cell.U[i]) || cell.Q[UID::DEN] == 0 || cell.Q[UID::PRE] == 0)
                throw std::runtime_error(name + " produced an error.");
}

// Torch::hydroStep, Torch.cpp:366-387.
static void hydro_step(Sim& s, double dt) {
    Fluid& fluid = s.fluid;
    check_values(s, "hydro before");
    fluid.globalWfromU();
    s.hydro.integrate(dt, fluid);
    s.hydro.updateSourceTerms(dt, fluid);
    fluid.advSolution(dt / 2.0);
    fluid.fixSolution();
    fluid.globalQfromU();
    fluid.globalUfromW();
    s.hydro.integrate(dt, fluid);
    s.hydro.updateSourceTerms(dt, fluid);
    fluid.advSolution(dt);
    fluid.fixSolution();
}

// Torch::subStep, Torch.cpp:352-364.
static void sub_step(Sim& s, double dt, bool has_heat, Integrator& comp) {
    Fluid& fluid = s.fluid;
    check_values(s, comp.getComponentName() + "before");
    if (!has_heat) {
        fluid.globalQfromU();
        fluid.fixPrimitives();
        comp.preTimeStepCalculations(fluid);
    }
    comp.integrate(dt, fluid);
    comp.updateSourceTerms(dt, fluid);
    fluid.advSolution(dt);
    fluid.fixSolution();
    check_values(s, comp.getComponentName() + " after");
}

static void write_state(Sim& s, const std::string& fn, bool prim) {
    Grid& grid = s.fluid.getGrid();
    const int n = grid.ncells[0] * grid.ncells[1] * grid.ncells[2];
    std::vector<double> out(7 * (size_t)n);
    for (int c = 0; c < n; ++c) {
        GridCell& cell = grid.getCell(c);
        for (int v = 0; v < 7; ++v) out[(size_t)v * n + c] = prim ? cell.Q[v] : cell.U[v];
    }
    FILE* f = fopen(fn.c_str(), "wb");
    fwrite(out.data(), sizeof(double), out.size(), f);
    fclose(f);
}

static int run_hydro(const KV& kv) {
    Sim s;
    init_sim(s, kv);
    MPIW& mpi = MPIW::Instance();
    Grid& grid = s.fluid.getGrid();
    const int NX = grid.ncells[0], NY = grid.ncells[1], NZ = grid.ncells[2];
    const long long NG = (long long)NX * NY * NZ;

    // Initial primitives (cgs, SoA [7][NZ][NY][NX]) from a binary file.  Each
    // rank reads the cells of its own x-slab.
    std::string qfile = gets(kv, "input_q", "");
    {
        FILE* f = fopen(qfile.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open input_q " + qfile);
        std::vector<double> row(NX);
        for (GridCell& cell : grid.getIterable("GridCells")) {
            long long gi = (long long)(cell.xc[0] - 0.5), gj = (long long)(cell.xc[1] - 0.5), gk = (long long)(cell.xc[2] - 0.5);
            long long g = gi + NX * (gj + NY * gk);
            for (int v = 0; v < 7; ++v) {
                double val;
                fseek(f, (long)(((long long)v * NG + g) * 8), SEEK_SET);
                if (fread(&val, 8, 1, f) != 1) throw std::runtime_error("short input_q");
                cell.Q[v] = val;
            }
            cell.heatCapacityRatio = s.fluid.heatCapacityRatio;
        }
        fclose(f);
    }
    s.thermo.initialiseMinTempField(s.fluid);  // Torch.cpp:114-119
    to_code_units(s);
    s.fluid.fixPrimitives();
    s.fluid.globalUfromQ();
    if (s.p.radiation_on) s.rad.initField(s.fluid);  // Torch.cpp:121

    const int nsteps = geti(kv, "nsteps", 1);
    std::vector<int> dumps;
    {
        std::istringstream ss(gets(kv, "dump_steps", ""));
        int k;
        while (ss >> k) dumps.push_back(k);
    }
    std::string prefix = gets(kv, "output_prefix", "");
    const bool write = !prefix.empty() && mpi.nProcessors() == 1;
    if (write && std::find(dumps.begin(), dumps.end(), 0) != dumps.end())
        write_state(s, prefix + "_U_0000.bin", false);

    // Torch::run, Torch.cpp:240-302, with a step cap.  Checkpointer logic as
    // src/IO/Checkpointer.cpp:14-26 (printing omitted).
    s.fluid.globalQfromU();
    s.fluid.fixPrimitives();
    std::vector<double> dts;
    int ckpt_count = 0;
    const double delta = 100.0 / s.p.ncheckpoints;
    auto ckpt = [&](double t) {
        double pct = std::min(100.0, 100.0 * t / s.p.tmax);
        while (pct - ckpt_count * delta > delta - 1.0e-8) ckpt_count += 1;
        return ((ckpt_count + 1) * delta - pct) * s.p.tmax / 100.0;
    };
    ckpt(0.0);
    double t = 0;
    bool first_time = true;
    int stepCounter = 0;
    std::vector<Integrator*> comps;
    comps.push_back(&s.hydro);
    if (s.p.cooling_on) comps.push_back(&s.thermo);
    if (s.p.radiation_on) comps.push_back(&s.rad);

    bool quitting = false;  // Torch::m_isQuitting
    mpi.barrier();
    auto t0 = std::chrono::steady_clock::now();
    for (int step = 1; step <= nsteps && t < s.p.tmax && !quitting; ++step) {
        double dt_next = ckpt(t);
        // Torch::fullStep, Torch.cpp:389-418.
        s.fluid.globalQfromU();
        s.fluid.fixPrimitives();
        if (s.p.cooling_on) s.thermo.preTimeStepCalculations(s.fluid);
        if (s.p.radiation_on) s.rad.preTimeStepCalculations(s.fluid);
        double dt;
        if (first_time) {
            dt = s.p.dt_max * 1.0e-20;
            first_time = false;
        } else {
            double dt_hydro = s.hydro.calculateTimeStep(s.p.dt_max, s.fluid);
            double dt_rad = dt_hydro, dt_thermo = dt_hydro;
            if (s.p.radiation_on) dt_rad = s.rad.calculateTimeStep(s.p.dt_max, s.fluid);
            if (s.p.cooling_on) dt_thermo = s.thermo.calculateTimeStep(s.p.dt_max, s.fluid);
            dt = std::min(std::min(dt_hydro, dt_rad), dt_thermo);
            if (s.p.debug) {  // Torch.cpp:321-332, its condition as written
                double thyd = 100.0 * dt_hydro / s.p.tmax;
                thyd = mpi.minimum(thyd);
                if (thyd <= 1.0e-6) quitting = true;
            }
        }
        dt = mpi.minimum(dt);
        dt = std::min(dt_next, dt);
        int ncomps = (int)comps.size();
        if (ncomps == 1) {
            hydro_step(s, dt);
        } else {
            stepCounter = (stepCounter + 1) % ncomps;
            for (int i = 0; i < ncomps; ++i) {
                double h = (i == ncomps - 1) ? 1.0 : 0.5;
                sub_step(s, h * dt, i == 0, *comps[(i + stepCounter) % ncomps]);
            }
            for (int i = ncomps - 2; i >= 0; --i)
                sub_step(s, dt / 2.0, false, *comps[(i + stepCounter) % ncomps]);
        }
        t += dt;
        dts.push_back(dt);
        if (write && std::find(dumps.begin(), dumps.end(), step) != dumps.end()) {
            char buf[32];
            snprintf(buf, sizeof buf, "_U_%04d.bin", step);
            write_state(s, prefix + buf, false);
        }
    }
    mpi.barrier();
    auto t1 = std::chrono::steady_clock::now();
    double secs = std::chrono::duration<double>(t1 - t0).count();
    if (write) {
        FILE* f = fopen((prefix + "_dt.bin").c_str(), "wb");
        fwrite(dts.data(), 8, dts.size(), f);
        fclose(f);
    }
    // GridCell::H and Fluid::Q as the loop left them, then after each of
    // `heating_fills` calls of Thermodynamics::fillHeatingArrays (the calls
    // Torch::run makes at a checkpoint, Torch.cpp:272-273)
    const int fills = geti(kv, "heating_fills", 0);
    if (write && fills > 0) {
        write_state(s, prefix + "_Q.bin", true);
        for (int k = 0; k <= fills; ++k) {
            if (k > 0) s.thermo.fillHeatingArrays(s.fluid);
            std::vector<double> h((size_t)HID::N * NG);
            for (long long c = 0; c < NG; ++c) {
                GridCell& cell = grid.getCell((int)c);
                for (int i = 0; i < HID::N; ++i) h[(size_t)i * NG + c] = cell.H[i];
            }
            char buf[32];
            snprintf(buf, sizeof buf, "_H_%d.bin", k);
            FILE* f = fopen((prefix + buf).c_str(), "wb");
            fwrite(h.data(), 8, h.size(), f);
            fclose(f);
        }
    }
    if (mpi.getRank() == 0) {
        printf("{\"cells\": %lld, \"steps\": %zu, \"seconds\": %.6f, \"mcell_updates_per_s\": %.6f, \"nproc\": %d, \"time\": %.17g, \"quitting\": %d}\n",
               NG, dts.size(), secs, (double)NG * dts.size() / secs / 1e6, mpi.nProcessors(), t, quitting ? 1 : 0);
    }
    return 0;
}

static int run_riemann(int argc, char** argv) {
    std::string solver = argv[2];
    double gamma = std::stod(argv[3]);
    long n = std::stol(argv[4]);
    std::unique_ptr<RiemannSolver> rs = RiemannSolverFactory::create(solver, 3);
    std::vector<double> in(15 * (size_t)n), out(8 * (size_t)n);
    FILE* f = fopen(argv[5], "rb");
    if (fread(in.data(), 8, in.size(), f) != in.size()) throw std::runtime_error("short riemann input");
    fclose(f);
    for (long i = 0; i < n; ++i) {
        FluidArray Ql, Qr, F;
        for (int v = 0; v < 7; ++v) { Ql[v] = in[15 * i + v]; Qr[v] = in[15 * i + 7 + v]; F[v] = 0; }
        int dim = (int)in[15 * i + 14];
        // Hydro.cpp:69-71,101-103 sound speeds.
        double a_l2 = gamma * Ql[UID::PRE] / Ql[UID::DEN];
        double a_r2 = gamma * Qr[UID::PRE] / Qr[UID::DEN];
        double status = 0;
        try {
            rs->solve(F, Ql, Qr, a_l2, a_r2, gamma, dim);
        } catch (std::exception&) {
            status = 1;
        }
        for (int v = 0; v < 7; ++v) out[8 * i + v] = F[v];
        out[8 * i + 7] = status;
    }
    f = fopen(argv[6], "wb");
    fwrite(out.data(), 8, out.size(), f);
    fclose(f);
    return 0;
}

// Per-cell thermodynamics through the reference's own preTimeStepCalculations
// and integrate.  A 2x1x1 grid holds the star cell (0, excluded as
// CausalWind, Fluid.cpp:41-54) and the cell under test (1).  The star is on
// so that F_FUV = fluxFUV(0.5*photonRate, r^2) is non-zero; the column
// density reaches the test cell through rayTrace's neighbour interpolation
// (Thermodynamics.cpp:339-363) from a ghost cell that holds it with weight 1.
static int run_thermo(const KV& kv_in) {
    KV kv = kv_in;
    kv["nx"] = "2"; kv["ny"] = "1"; kv["nz"] = "1";
    kv["star_on"] = "1";
    kv["cooling_on"] = "1";
    Sim s;
    init_sim(s, kv);
    Grid& grid = s.fluid.getGrid();
    const double gam = s.fluid.heatCapacityRatio;
    long n = std::stol(gets(kv, "ncells_in", "0"));
    std::vector<double> in(9 * (size_t)n), out(10 * (size_t)n);
    FILE* f = fopen(gets(kv, "input", "").c_str(), "rb");
    if (fread(in.data(), 8, in.size(), f) != in.size()) throw std::runtime_error("short thermo input");
    fclose(f);
    GridCell& cell = grid.getCell(1);
    const int ghost = 2;  // first ghost cell; never in a causal list
    cell.neighbourIDs = {{ghost, -1, -1, -1}};
    cell.neighbourWeights = {{1.0, 0, 0, 0}};
    const double dx = grid.dx[0];
    double rsqrd = 0;
    for (int id = 0; id < 3; ++id)
        rsqrd += (cell.xc[id] - s.fluid.getStar().xc[id]) * (cell.xc[id] - s.fluid.getStar().xc[id]) * grid.dx[id] * grid.dx[id];
    (void)dx;
    for (long i = 0; i < n; ++i) {
        const double* r = &in[9 * (size_t)i];
        // inputs (code units): den, pre, hii, adv, T_min, col_den, photonRate, dt, E
        cell.Q[UID::DEN] = r[0];
        cell.Q[UID::PRE] = r[1];
        cell.Q[UID::HII] = r[2];
        cell.Q[UID::ADV] = r[3];
        cell.T_min = r[4];
        grid.getCell(ghost).T[TID::COL_DEN] = r[5];
        grid.getCell(ghost).T[TID::DCOL_DEN] = 0;
        s.fluid.getStar().photonRate = r[6];
        cell.U[UID::PRE] = r[8];
        cell.heatCapacityRatio = gam;
        s.thermo.preTimeStepCalculations(s.fluid);
        double heat = cell.T[TID::HEAT], rate_pre = cell.T[TID::RATE];
        double col = cell.T[TID::COL_DEN];
        s.thermo.integrate(r[7], s.fluid);
        double* o = &out[10 * (size_t)i];
        o[0] = heat; o[1] = rate_pre; o[2] = cell.T[TID::RATE]; o[3] = cell.H[HID::TOT];
        o[4] = col;
        o[5] = s.thermo.fluxFUV(0.5 * r[6], rsqrd);
        o[6] = rsqrd;
        o[7] = s.fluid.calcTemperature(r[2], r[1], r[0]);
        o[8] = 0; o[9] = 0;
    }
    f = fopen(gets(kv, "output", "").c_str(), "wb");
    fwrite(out.data(), 8, out.size(), f);
    fclose(f);
    return 0;
}

// The thermo column-density ray trace of a star-on run: Radiation::initField
// (nearest neighbours, weights, path lengths, shell volumes;
// Radiation.cpp:78-96, Grid.cpp:922-951) and Thermodynamics::
// preTimeStepCalculations with its rayTrace (Thermodynamics.cpp:218-262,
// 339-417) on a grid whose primitives (code units) come from input_q.
// Output per cell (15 doubles): neighbour IDs 0-3, weights 0-3, ds,
// shellVol, T[COL_DEN], T[DCOL_DEN], T[RATE], T[HEAT], 1 if in CausalWind.
static int run_raytrace(const KV& kv_in) {
    KV kv = kv_in;
    kv["star_on"] = "1";
    kv["cooling_on"] = "1";
    Sim s;
    init_sim(s, kv);
    Grid& grid = s.fluid.getGrid();
    const int n = grid.ncells[0] * grid.ncells[1] * grid.ncells[2];
    std::vector<double> q(7 * (size_t)n);
    FILE* f = fopen(gets(kv, "input_q", "").c_str(), "rb");
    if (fread(q.data(), 8, q.size(), f) != q.size()) throw std::runtime_error("short raytrace input");
    fclose(f);
    for (int c = 0; c < n; ++c) {
        GridCell& cell = grid.getCell(c);
        for (int v = 0; v < 7; ++v) cell.Q[v] = q[(size_t)v * n + c];
        cell.heatCapacityRatio = s.fluid.heatCapacityRatio;  // Torch.cpp:184,224
    }
    s.fluid.globalUfromQ();
    s.thermo.initialiseMinTempField(s.fluid);
    Radiation rad;
    rad.initialise(s.consts, s.p.getRadiationParameters());
    rad.initField(s.fluid);
    s.thermo.preTimeStepCalculations(s.fluid);
    // time the column-density ray trace alone (a second pass: the same
    // values, every cell's neighbours already done in this causal order)
    const auto t0 = std::chrono::steady_clock::now();
    s.thermo.rayTrace(s.fluid);
    const double ray_trace_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<char> wind(n, 0);
    for (int id : grid.getOrderedIndices("CausalWind")) wind[id] = 1;
    std::vector<double> out(15 * (size_t)n);
    for (int c = 0; c < n; ++c) {
        GridCell& cell = grid.getCell(c);
        double* o = &out[15 * (size_t)c];
        for (int i = 0; i < 4; ++i) o[i] = cell.neighbourIDs[i];
        for (int i = 0; i < 4; ++i) o[4 + i] = cell.neighbourWeights[i];
        o[8] = cell.ds;
        o[9] = cell.shellVol;
        o[10] = cell.T[TID::COL_DEN];
        o[11] = cell.T[TID::DCOL_DEN];
        o[12] = cell.T[TID::RATE];
        o[13] = cell.T[TID::HEAT];
        o[14] = wind[c];
    }
    f = fopen(gets(kv, "output", "").c_str(), "wb");
    fwrite(out.data(), 8, out.size(), f);
    fclose(f);
    printf("{\"photon_rate\": %.17g, \"hydrogen_mass\": %.17g, \"pi\": %.17g, \"dx\": %.17g, "
           "\"ray_trace_s\": %.6g}\n", s.p.photonRate, s.consts->hydrogenMass, s.consts->pi, grid.dx[0], ray_trace_s);
    return 0;
}

// The Radiation integrator on its own (Radiation.cpp): set-up as
// Torch::initialise (primitives of input_q in code units, fixPrimitives,
// globalUfromQ, Radiation::initField, Torch.cpp:116-122), then `nsteps`
// steps of globalQfromU + fixPrimitives + preTimeStepCalculations +
// calculateTimeStep (dt = min(dt_max, dt_rad)), then two Torch::subStep
// calls of the radiation component (the first with the heat flux already
// computed, the second recomputing it; Torch.cpp:352-364).  Writes per step
// U [7][n], R [6][n] and Q[HII], Q[ADV] [2][n] to <prefix>_<step>.bin and
// the dts to <prefix>_dt.bin.
static int run_radiation(const KV& kv_in) {
    KV kv = kv_in;
    kv["star_on"] = gets(kv, "star_on", "1");
    Sim s;
    init_sim(s, kv);
    Grid& grid = s.fluid.getGrid();
    const int n = grid.ncells[0] * grid.ncells[1] * grid.ncells[2];
    std::vector<double> q(7 * (size_t)n);
    FILE* f = fopen(gets(kv, "input_q", "").c_str(), "rb");
    if (fread(q.data(), 8, q.size(), f) != q.size()) throw std::runtime_error("short radiation input");
    fclose(f);
    for (int c = 0; c < n; ++c) {
        GridCell& cell = grid.getCell(c);
        for (int v = 0; v < 7; ++v) cell.Q[v] = q[(size_t)v * n + c];
        cell.heatCapacityRatio = s.fluid.heatCapacityRatio;  // Torch.cpp:184,224
    }
    s.thermo.initialiseMinTempField(s.fluid);  // Torch.cpp:114
    s.fluid.fixPrimitives();
    s.fluid.globalUfromQ();
    Radiation rad;
    rad.initialise(s.consts, s.p.getRadiationParameters());
    rad.initField(s.fluid);
    const int nsteps = geti(kv, "nsteps", 2);
    const std::string prefix = gets(kv, "output_prefix", "rad");
    std::vector<double> dts;
    double substep_s = 0;
    for (int step = 1; step <= nsteps; ++step) {
        s.fluid.globalQfromU();
        s.fluid.fixPrimitives();
        rad.preTimeStepCalculations(s.fluid);
        const double dt = std::min(s.p.dt_max, rad.calculateTimeStep(s.p.dt_max, s.fluid));
        dts.push_back(dt);
        const auto t0 = std::chrono::steady_clock::now();
        sub_step(s, dt, true, rad);
        sub_step(s, dt, false, rad);
        substep_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 2.0;
        std::vector<double> out(15 * (size_t)n);
        for (int c = 0; c < n; ++c) {
            GridCell& cell = grid.getCell(c);
            for (int v = 0; v < 7; ++v) out[(size_t)v * n + c] = cell.U[v];
            for (int r = 0; r < 6; ++r) out[(size_t)(7 + r) * n + c] = cell.R[r];
            out[(size_t)13 * n + c] = cell.Q[UID::HII];
            out[(size_t)14 * n + c] = cell.Q[UID::ADV];
        }
        char name[512];
        snprintf(name, sizeof name, "%s_%04d.bin", prefix.c_str(), step);
        FILE* o = fopen(name, "wb");
        fwrite(out.data(), 8, out.size(), o);
        fclose(o);
    }
    FILE* o = fopen((prefix + "_dt.bin").c_str(), "wb");
    fwrite(dts.data(), 8, dts.size(), o);
    fclose(o);
    printf("{\"photon_rate\": %.17g, \"photon_energy\": %.17g, \"sigma\": %.17g, \"dt_max\": %.17g, "
           "\"substep_s\": %.6g}\n", s.p.photonRate, s.p.photonEnergy, s.p.photoIonCrossSection, s.p.dt_max,
           substep_s / nsteps);
    return 0;
}

static int run_consts(int argc, char** argv) {
    KV kv;
    kv["density_scale"] = argv[2];
    kv["pressure_scale"] = argv[3];
    kv["time_scale"] = argv[4];
    kv["nx"] = "2"; kv["ny"] = "1"; kv["nz"] = "1";
    if (argc > 5) kv["temperature_floor"] = argv[5];
    if (argc > 6) kv["density_floor"] = argv[6];
    Sim s;
    init_sim(s, kv);
    Constants& c = *s.consts;
    Thermodynamics& t = s.thermo;
    const LogSplineData& sp = *t.m_collisionalExcitationHI_CoolingRates;
    printf("{\"M\": %.17g, \"L\": %.17g, \"T\": %.17g,\n", c.converter.M, c.converter.L, c.converter.T);
    printf(" \"hydrogen_mass\": %.17g, \"spec_gas_const\": %.17g, \"boltzmann\": %.17g, \"dust_xsec\": %.17g,\n",
           c.hydrogenMass, c.specificGasConstant, c.boltzmannConst, c.dustExtinctionCrossSection);
    printf(" \"dfloor\": %.17g, \"pfloor\": %.17g, \"tmax\": %.17g, \"dt_max\": %.17g,\n", s.p.dfloor, s.p.pfloor, s.p.tmax, s.p.dt_max);
    printf(" \"imlc\": %.17g, \"nmlc\": %.17g, \"ciec\": %.17g, \"n0\": %.17g, \"nmc\": %.17g,\n", t.m_imlc, t.m_nmlc, t.m_ciec, t.m_n0, t.m_nmc);
    printf(" \"fuvh_a\": %.17g, \"fuvh_b\": %.17g, \"fuvh_c\": %.17g, \"irh_a\": %.17g, \"irh_b\": %.17g, \"crh\": %.17g,\n",
           t.m_fuvh_a, t.m_fuvh_b, t.m_fuvh_c, t.m_irh_a, t.m_irh_b, t.m_crh);
    printf(" \"spl_min_slope\": %.17g, \"spl_max_slope\": %.17g,\n", sp.m_minSlope, sp.m_maxSlope);
    const char* names[3] = {"spl_x", "spl_y", "spl_y2"};
    const std::vector<double>* arrs[3] = {&sp.m_x, &sp.m_y, &sp.m_y2};
    for (int a = 0; a < 3; ++a) {
        printf(" \"%s\": [", names[a]);
        for (size_t i = 0; i < arrs[a]->size(); ++i) printf("%s%.17g", i ? ", " : "", (*arrs[a])[i]);
        printf("]%s\n", a < 2 ? "," : "}");
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: ref_harness hydro|riemann|thermo|consts ...\n");
        return 2;
    }
    MPIW::Instance();
    std::string mode = argv[1];
    try {
        if (mode == "hydro") return run_hydro(read_case(argv[2]));
        if (mode == "riemann") return run_riemann(argc, argv);
        if (mode == "thermo") return run_thermo(read_case(argv[2]));
        if (mode == "consts") return run_consts(argc, argv);
        if (mode == "raytrace") return run_raytrace(read_case(argv[2]));
        if (mode == "radiation") return run_radiation(read_case(argv[2]));
    } catch (std::exception& e) {
        fprintf(stderr, "ref_harness error: %s\n", e.what());
        return 1;
    }
    fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
