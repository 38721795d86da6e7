// torch_gpu.hpp -- C++ host layer of the MI355X hydro + thermodynamics step:
// the reference's parameter struct, operator (Integrator) interface and
// operator-split driver, re-expressed over the C ABI of include/torchgpu.h.
//
// Reference interfaces mirrored here (hgsteggles/TORCH):
//   TorchParameters + initialise       src/Torch/Parameters.{hpp,cpp}
//   parseParameters (Lua keys)         src/main.cpp:126-208
//   Integrator (four virtuals)         src/Integrators/Integrator.hpp:22-35
//   Hydrodynamics / Thermodynamics     src/Integrators/Hydro.hpp, Thermodynamics.hpp
//   Fluid whole-grid operations        src/Fluid/Fluid.hpp:40-55
//   Torch::run/fullStep/hydroStep/subStep/calculateTimeStep  src/Torch/Torch.cpp:240-418
//
// The GPU kernels fuse Hydrodynamics::integrate + updateSourceTerms +
// Fluid::advSolution + fixSolution into one launch.  The classes below keep
// the reference's call sequence and defer the work: integrate() and
// updateSourceTerms() record what is pending, Fluid::advSolution(h) launches
// the fused kernel that realises it, fixSolution() is part of that launch.
#pragma once

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/torchgpu.h"
#include "lua_config.hpp"

namespace tg {

struct TorchParameters {
    // defaults of src/Torch/Parameters.hpp:25-97
    std::string setupFile = "";
    std::string initialConditions = "";
    std::string patchfilename = "";
    std::array<int, 3> patchoffset{{0, 0, 0}};
    int nd = 0;
    double sideLength = 0;
    std::array<int, 3> ncells{{0, 0, 0}};
    std::array<std::string, 3> leftBC{{"free", "free", "free"}};
    std::array<std::string, 3> rightBC{{"free", "free", "free"}};
    std::string geometry = "cartesian";
    std::string riemannSolver = "hll";
    std::string slopeLimiter = "falle";
    int spatialOrder = 0;
    int temporalOrder = 0;
    double tmax = 0;
    double dt_max = 0;
    bool radiation_on = false;
    bool cooling_on = false;
    bool debug = true;
    std::string outputDirectory = "tmp/";
    int ncheckpoints = 100;
    double dfloor = 0, pfloor = 0, tfloor = 0;
    double dscale = 0, pscale = 0, tscale = 0;
    bool star_on = false;
    std::array<int, 3> star_position{{0, 0, 0}};
    std::array<bool, 3> faceSnap{{false, false, false}};
    double photonRate = 0;
    double photonEnergy = 0;
    int windCellRadius = 0;
    double massLossRate = 0, windVelocity = 0, windTemperature = 0;
    double K1 = 0, K2 = 0, K3 = 0, K4 = 0;
    double photoIonCrossSection = 0, alphaB = 0, tau0 = 0, minX = 0, THI = 0, THII = 0;
    bool collisions_on = false;
    std::string rt_coupling = "off";
    std::string rt_scheme = "implicit";
    double thermoHII_Switch = 0;
    double heatingAmplification = 1.0;
    bool thermoSubcycling = true;
    bool minTempInitialState = false;
    double massFractionH = 0;
    double heatCapacityRatio = 0;
    // additions (no reference counterpart), defaults reproduce the reference
    std::string precision = "fp64";   // Hydrodynamics.precision: "fp64" | "mixed"

    // Code-unit constants filled by initialise()
    double specificGasConstant = 0;
    double M = 0, L = 0, T = 0;

    // TorchParameters::initialise + Constants::initialise_DPT (Parameters.cpp:9-27)
    void initialise();
    double toCodeUnits(double v, double m, double l, double t) const;
    // torchgpu_params for the grid context
    torchgpu_params gridParams() const;
};

// parseParameters (main.cpp:126-208): missing keys keep the defaults and are
// reported on `warnings`.
void parseParameters(const LuaConfig& cfg, TorchParameters& p, std::string* warnings = nullptr);

// Solver / limiter / condition strings as the reference's factories accept
// them; unknown strings throw the factories' messages.
int riemannSolverId(const std::string& type);
int slopeLimiterId(const std::string& type);
int conditionId(const std::string& cond);

class Fluid;

class Integrator {
public:
    explicit Integrator(std::string name) : componentName(std::move(name)) {}
    virtual ~Integrator() {}
    virtual void preTimeStepCalculations(Fluid& fluid) const = 0;
    virtual double calculateTimeStep(double dt_max, Fluid& fluid) const = 0;
    virtual void integrate(double dt, Fluid& fluid) const = 0;
    virtual void updateSourceTerms(double dt, Fluid& fluid) const = 0;
    virtual std::string getComponentName() final { return componentName; }
    std::string componentName;
};

// The grid state on the device (a torchgpu_ctx) with the reference's
// whole-grid operations.
class Fluid {
public:
    Fluid() = default;
    ~Fluid();
    Fluid(const Fluid&) = delete;
    Fluid& operator=(const Fluid&) = delete;

    void initialise(const TorchParameters& p, void* stream = nullptr);
    torchgpu_ctx* ctx() const { return ctx_; }

    // Torch::initialise tail: fixPrimitives + globalUfromQ of primitives in
    // code units (host array [7][nz][ny][nx]).
    void setPrimitives(const double* Q);
    void getPrimitives(double* Q, bool fix) const;
    void getConserved(double* U) const;

    // Fluid.hpp:40-55
    void globalQfromU();
    void fixPrimitives();
    void globalWfromU();
    void globalUfromW();
    void advSolution(double dt);
    void fixSolution();

    // pending work recorded by the integrators
    enum class Pending { NONE, HYDRO, THERMO, RAD };
    Pending pending = Pending::NONE;
    torchgpu_cooling* cooling = nullptr;      // set by Thermodynamics
    int64_t cooling_excluded = -1;            // ... with its excluded cell
    torchgpu_radiation* radiation = nullptr;  // set by Radiation

    void check() const;  // reports device errors as std::runtime_error

    double heatCapacityRatio = 0;
    double massFractionH = 1.0;

private:
    torchgpu_ctx* ctx_ = nullptr;
    enum class QSource { U, PREDICTOR } qsrc_ = QSource::U;
    bool w_saved_ = false;
    bool u_is_w_ = false;
};

class Hydrodynamics : public Integrator {
public:
    Hydrodynamics() : Integrator("Hydrodynamics") {}
    void preTimeStepCalculations(Fluid& fluid) const override;
    double calculateTimeStep(double dt_max, Fluid& fluid) const override;
    void integrate(double dt, Fluid& fluid) const override;
    void updateSourceTerms(double dt, Fluid& fluid) const override;
};

class Thermodynamics : public Integrator {
public:
    Thermodynamics() : Integrator("Thermodynamics") {}
    ~Thermodynamics();
    void initialise(const TorchParameters& p, void* stream = nullptr);
    void preTimeStepCalculations(Fluid& fluid) const override;
    double calculateTimeStep(double dt_max, Fluid& fluid) const override;
    void integrate(double dt, Fluid& fluid) const override;
    void updateSourceTerms(double dt, Fluid& fluid) const override;
    // Thermodynamics::initialiseMinTempField on the set-up's primitives
    // (cgs, [7][n], before Torch::toCodeUnits)
    void initialiseMinTempField(const TorchParameters& p, const double* Qcgs, size_t n);
    torchgpu_cooling* handle() const { return cool_; }
    // the CausalWind cell of a star-off run (-1 with the star on: the
    // cooling handle skips the wind cells itself)
    int64_t excluded() const { return excluded_; }

private:
    torchgpu_cooling* cool_ = nullptr;
    int64_t excluded_ = -1;
};

// Radiation (src/Integrators/Radiation.cpp) for a star-on run.
class Radiation : public Integrator {
public:
    Radiation() : Integrator("Radiation") {}
    ~Radiation();
    void initialise(const TorchParameters& p, Fluid& fluid);
    void initField(const double* Q);  // Radiation::initField on the primitives of Torch::initialise
    void preTimeStepCalculations(Fluid& fluid) const override;
    double calculateTimeStep(double dt_max, Fluid& fluid) const override;
    void integrate(double dt, Fluid& fluid) const override;
    void updateSourceTerms(double dt, Fluid& fluid) const override;
    torchgpu_radiation* handle() const { return rad_; }

private:
    torchgpu_radiation* rad_ = nullptr;
};

// The star of Parameters.Star as the library takes it (cell centre; face
// snapping is rejected).
torchgpu_star starOf(const TorchParameters& p);

// Torch::setUpLua (Torch.cpp:192-228) without the GPU: runs the set-up file
// and calls its initialise(x, y, z, xs, ys, zs) for every cell with the cell
// centre and the star position in cm; returns the primitives in cgs
// [7][nz][ny][nx] (DEN, PRE, HII, VEL0..2; ADV zero).  p must have been
// initialised (code-unit side length).  Gravity must come back zero.
std::vector<double> evalLuaSetup(const TorchParameters& p, const std::string& filename);

// The driver of src/Torch/Torch.cpp for hydro-only runs (device-resident
// time loop) and the operator-split sequence with the integrators above.
class Torch {
public:
    void initialise(const TorchParameters& p, void* stream = nullptr);
    void setPrimitivesCgs(const double* Q);   // Torch::toCodeUnits + fixPrimitives + globalUfromQ
    void setUpLua(const std::string& filename);  // Torch::setUpLua + the tail of Torch::initialise
    // Torch::run, capped; with `print`, the data2D snapshots of
    // Integration.output_directory at the start, every checkpoint and the end;
    // `host_loop` (without `print`): fullStep driven from the host
    void run(int max_steps, bool print = false, bool host_loop = false);
    void print2D(int count) const;
    // Thermodynamics::fillHeatingArrays on the printed state (H [11][n]) and
    // DataPrinter::printHeating of it
    std::vector<double> fillHeating();
    void printHeating(int count, const std::vector<double>& H) const;
    double fullStep(double dt_nextCheckpoint);
    void hydroStep(double dt, bool hasCalculatedHeatFlux);
    void subStep(double dt, bool hasCalculatedHeatFlux, Integrator& comp);
    double calculateTimeStep();

    Fluid fluid;
    Hydrodynamics hydrodynamics;
    Thermodynamics thermodynamics;
    Radiation radiation;
    TorchParameters params;
    double currentTime = 0;
    int steps = 0;
    bool isQuitting() const { return quitting_; }  // Torch::isQuitting (Torch.hpp:47)

private:
    bool first_time_ = true;
    bool quitting_ = false;  // Torch::m_isQuitting
    int stepCounter_ = 0;
};

}  // namespace tg
