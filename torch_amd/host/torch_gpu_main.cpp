// torch_gpu_main.cpp -- command-line driver: reads a TORCH Lua parameter
// file, sets the initial state and runs Torch::run on the GPU.
//
//   torch_gpu --paramfile=FILE [--setup=sod|FILE.bin|FILE.txt[.gz]] [--steps=N]
//             [--output=FILE.bin] [--snapshots] [--dump-params]
//
// --setup=sod builds the Sod shock tube along x; FILE.bin holds primitive
// variables in cgs, float64 [7][nz][ny][nx] (UID order), the values a Lua
// set-up function would return (the reference's Torch::setUpLua,
// Torch.cpp:192-228, needs a Lua runtime).  A data2D snapshot (.txt or
// .txt.gz) is read as Integration.initial_conditions is (grid shape, side
// length, state and time from the file); a non-empty initial_conditions in
// the parameter file does the same.  --snapshots writes the data2D files of
// Torch::run to Integration.output_directory: the initial state, every
// checkpoint and the final state.  --setupfile=FILE.lua runs the set-up
// file's initialise(x, y, z, xs, ys, zs) for every cell, as the reference's
// --setupfile (main.cpp:36-50, Torch::setUpLua); --dump-setup=FILE.bin
// writes the primitives it returns (cgs, float64 [7][nz][ny][nx]) and exits
// without touching the GPU.  Torch::run keeps its whole loop on the device
// (torchgpu_run_full); --host-loop drives Torch::fullStep from the host
// through the operator calls instead (no checkpoint printing).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "torch_gpu.hpp"

namespace {

std::string arg(int argc, char** argv, const std::string& key, const std::string& def) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a.rfind(key + "=", 0) == 0) return a.substr(key.size() + 1);
        if (a == key) return "1";
    }
    return def;
}

void dump(const tg::TorchParameters& p) {
    std::printf("{\"nd\": %d, \"ncells\": [%d, %d, %d], \"side_length\": %.17g, \"spatial_order\": %d,\n", p.nd,
                p.ncells[0], p.ncells[1], p.ncells[2], p.sideLength, p.spatialOrder);
    std::printf(" \"riemann_solver\": \"%s\", \"slope_limiter\": \"%s\", \"gamma\": %.17g, \"tmax\": %.17g,\n",
                p.riemannSolver.c_str(), p.slopeLimiter.c_str(), p.heatCapacityRatio, p.tmax);
    std::printf(" \"dfloor\": %.17g, \"tfloor\": %.17g, \"pfloor\": %.17g, \"scales\": [%.17g, %.17g, %.17g],\n",
                p.dfloor, p.tfloor, p.pfloor, p.dscale, p.pscale, p.tscale);
    std::printf(" \"bc\": [\"%s\", \"%s\", \"%s\", \"%s\", \"%s\", \"%s\"], \"cooling_on\": %s, \"ncheckpoints\": %d,\n",
                p.leftBC[0].c_str(), p.rightBC[0].c_str(), p.leftBC[1].c_str(), p.rightBC[1].c_str(),
                p.leftBC[2].c_str(), p.rightBC[2].c_str(), p.cooling_on ? "true" : "false", p.ncheckpoints);
    std::printf(" \"mass_fraction_h\": %.17g, \"thermo_hii_switch\": %.17g, \"precision\": \"%s\", \"debug\": %s,\n",
                p.massFractionH, p.thermoHII_Switch, p.precision.c_str(), p.debug ? "true" : "false");
    std::printf(" \"star_on\": %s, \"star_cell\": [%d, %d, %d], \"wind_radius\": %d, \"photon_rate\": %.17g, "
                "\"photon_energy\": %.17g, \"radiation_on\": %s,\n", p.star_on ? "true" : "false", p.star_position[0],
                p.star_position[1], p.star_position[2], p.windCellRadius, p.photonRate, p.photonEnergy,
                p.radiation_on ? "true" : "false");
    std::printf(" \"K\": [%.17g, %.17g, %.17g, %.17g], \"photoion_cross_section\": %.17g, \"tau_0\": %.17g, "
                "\"temperature_hi\": %.17g, \"temperature_hii\": %.17g, \"coupling\": \"%s\", "
                "\"integration_scheme\": \"%s\", \"collisions_on\": %s, \"heating_amplification\": %.17g,\n",
                p.K1, p.K2, p.K3, p.K4, p.photoIonCrossSection, p.tau0, p.THI, p.THII, p.rt_coupling.c_str(),
                p.rt_scheme.c_str(), p.collisions_on ? "true" : "false", p.heatingAmplification);
    // the same after TorchParameters::initialise (code units)
    tg::TorchParameters c = p;
    c.initialise();
    std::printf(" \"code_units\": {\"side_length\": %.17g, \"tmax\": %.17g, \"dt_max\": %.17g, \"dfloor\": %.17g, "
                "\"pfloor\": %.17g, \"spec_gas_const\": %.17g, \"photon_rate\": %.17g}}\n", c.sideLength, c.tmax,
                c.dt_max, c.dfloor, c.pfloor, c.specificGasConstant, c.photonRate);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string paramfile = arg(argc, argv, "--paramfile", "");
    if (paramfile.empty()) {
        std::fprintf(stderr, "usage: torch_gpu --paramfile=FILE [--setupfile=FILE.lua | --setup=sod|FILE.bin|FILE.txt[.gz]] "
                             "[--steps=N] [--output=FILE.bin] [--snapshots] [--dump-params] [--dump-setup=FILE.bin]\n");
        return 2;
    }
    try {
        tg::TorchParameters p;
        std::string warnings;
        tg::parseParameters(tg::LuaConfig::parse_file(paramfile), p, &warnings);
        if (arg(argc, argv, "--dump-params", "") == "1") {
            dump(p);
            return 0;
        }
        const std::string setupfile = arg(argc, argv, "--setupfile", "");
        const std::string dump_setup = arg(argc, argv, "--dump-setup", "");
        if (!dump_setup.empty()) {
            if (setupfile.empty()) throw std::runtime_error("--dump-setup needs --setupfile");
            tg::TorchParameters c = p;
            c.initialise();
            const std::vector<double> Q = tg::evalLuaSetup(c, setupfile);
            std::ofstream o(dump_setup, std::ios::binary);
            o.write(reinterpret_cast<const char*>(Q.data()), (std::streamsize)(Q.size() * sizeof(double)));
            return o ? 0 : 1;
        }
        const std::string setup = arg(argc, argv, "--setup", p.initialConditions.empty() ? "sod" : "");
        auto ends_with = [](const std::string& a, const std::string& b) {
            return a.size() >= b.size() && a.compare(a.size() - b.size(), b.size(), b) == 0;
        };
        if (ends_with(setup, ".txt") || ends_with(setup, ".txt.gz")) p.initialConditions = setup;
        tg::Torch torch;
        torch.initialise(p);
        const int nx = torch.params.ncells[0], ny = torch.params.ncells[1], nz = torch.params.ncells[2];
        const size_t n = (size_t)nx * ny * nz;
        std::vector<double> Q(7 * n, 0.0);
        if (!torch.params.initialConditions.empty()) {
            // state and time already read by Torch::initialise
        } else if (!setupfile.empty()) {
            torch.setUpLua(setupfile);
        } else if (setup == "sod") {
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        const size_t c = i + (size_t)nx * (j + (size_t)ny * k);
                        const bool left = (i + 0.5) / nx < 0.5;
                        Q[0 * n + c] = left ? 1.0 : 0.125;
                        Q[1 * n + c] = left ? 1.0 : 0.1;
                    }
        } else {
            std::ifstream f(setup, std::ios::binary);
            if (!f.read(reinterpret_cast<char*>(Q.data()), (std::streamsize)(Q.size() * sizeof(double))))
                throw std::runtime_error("cannot read " + setup);
        }
        if (torch.params.initialConditions.empty() && setupfile.empty()) torch.setPrimitivesCgs(Q.data());
        const int steps = std::stoi(arg(argc, argv, "--steps", "100"));
        auto t0 = std::chrono::steady_clock::now();
        torch.run(steps, arg(argc, argv, "--snapshots", "") == "1", arg(argc, argv, "--host-loop", "") == "1");
        auto t1 = std::chrono::steady_clock::now();
        const double secs = std::chrono::duration<double>(t1 - t0).count();
        std::printf("{\"steps\": %d, \"time\": %.17g, \"seconds\": %.6f, \"mcell_updates_per_s\": %.3f, "
                    "\"quitting\": %s}\n", torch.steps, torch.currentTime, secs, (double)n * torch.steps / secs / 1e6,
                    torch.isQuitting() ? "true" : "false");
        const std::string output = arg(argc, argv, "--output", "");
        if (!output.empty()) {
            std::vector<double> U(7 * n);
            torch.fluid.getConserved(U.data());
            std::ofstream o(output, std::ios::binary);
            o.write(reinterpret_cast<const char*>(U.data()), (std::streamsize)(U.size() * sizeof(double)));
        }
    } catch (std::exception& e) {
        std::fprintf(stderr, "torch_gpu: %s\n", e.what());
        return 1;
    }
    return 0;
}
