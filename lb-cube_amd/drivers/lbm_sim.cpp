// lbm_sim: the GPU path of LB-Cube's lbm_sim driver (src/main.cpp) on the
// MI355X kernel library, written against include/lattice_hip.hpp exactly as
// the reference's driver is written against lattice_cuda.cuh: host populations
// in a LatticeMemory-shaped container, device buffers from hipMalloc, one
// launch_step_gpu per step and a pointer swap, a VTK snapshot and a
// diagnostics row at step 0 and every -out_freq steps (src/main.cpp:572-607).
// Without -load the initial state is the reference's, bit for bit: the rest
// state or the 3D Taylor-Green vortex with its density perturbation
// (src/main.cpp:295-411), populations from compute_equilibrium.
//
// Options (src/main.cpp:149-195, same names and checks):
//   -nx N -ny N -nz N      grid (default 64 x 64 x 64)
//   -steps N               time steps (default 1000)
//   -out_freq N            diagnostics interval (default 100)
//   -init rest|taylor_green  initial condition (default taylor_green)
//   -gpu / -cpu            backend; only the GPU backend is built here
// Additions:
//   -tau T                 relaxation time (default 0.8, src/main.cpp:542)
//   -u0 U                  Taylor-Green amplitude (default 0.01, src/main.cpp:370)
//   -collision bgk|trt|mrt|rlbm   (default bgk, the reference's default)
//   -load FILE / -dump FILE  raw float64 populations [19][Z][Y][X] in/out
//   -csv FILE              diagnostics file (default diagnostics.csv)
//   -vtk_dir DIR           folder of the lbm_<step>.vtk snapshots (default .)
//   -no_vtk                no snapshots
//   -slabs                 keep the device populations in the slab order
//                          [Z][Q][Y][X] (lbc_fluid_step_tiled; same results bit
//                          for bit, a faster HBM pattern); they are converted
//                          into the idle ping-pong buffer for diagnostics,
//                          snapshots and -dump
// Diagnostics rows: step,total_mass,total_kinetic_energy,max_velocity_magnitude
// with %.17g (include/lattice_io.hpp:182-234), summed on the device in FP64.
// Snapshots: density, velocity_magnitude, velocity (write_vtk,
// include/lattice_io.hpp:386-451), reconstructed on the device.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "initial_states.hpp"
#include "lattice_hip.hpp"
#include "lbcube_lattices.hpp"
#include "vtk.hpp"

namespace {

using Lattice = lbm::D3Q19;
using Real = double;

struct Config {
    std::size_t nx = 64, ny = 64, nz = 64;
    int steps = 1000;
    int output_frequency = 100;
    std::string init = "taylor_green";
    bool use_gpu = true;
    double tau = 0.8;
    double u0 = 0.01;
    std::string collision = "bgk";
    std::string load, dump, csv = "diagnostics.csv";
    std::string vtk_dir = ".";
    bool vtk = true;
    bool slabs = false;
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

Config parse(int argc, char** argv) {
    Config c;
    auto value = [&](int& i) -> std::string {
        if (i + 1 >= argc) throw std::invalid_argument(std::string("missing value for ") + argv[i]);
        return argv[++i];
    };
    auto size_value = [&](int& i) -> std::size_t {
        const long long v = std::stoll(value(i));
        if (v <= 0) throw std::invalid_argument(std::string(argv[i - 1]) + " must be positive");
        return static_cast<std::size_t>(v);
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-nx") c.nx = size_value(i);
        else if (a == "-ny") c.ny = size_value(i);
        else if (a == "-nz") c.nz = size_value(i);
        else if (a == "-steps") c.steps = std::stoi(value(i));
        else if (a == "-out_freq") c.output_frequency = std::stoi(value(i));
        else if (a == "-init") c.init = value(i);
        else if (a == "-gpu") c.use_gpu = true;
        else if (a == "-cpu") c.use_gpu = false;
        else if (a == "-tau") c.tau = std::stod(value(i));
        else if (a == "-u0") c.u0 = std::stod(value(i));
        else if (a == "-collision") c.collision = value(i);
        else if (a == "-load") c.load = value(i);
        else if (a == "-dump") c.dump = value(i);
        else if (a == "-csv") c.csv = value(i);
        else if (a == "-vtk_dir") c.vtk_dir = value(i);
        else if (a == "-no_vtk") c.vtk = false;
        else if (a == "-slabs") c.slabs = true;
        else throw std::invalid_argument("unknown argument: " + a);
    }
    if (c.steps < 0) throw std::invalid_argument("-steps must be non-negative");
    if (c.output_frequency <= 0) throw std::invalid_argument("-out_freq must be greater than zero");
    if (c.init != "rest" && c.init != "taylor_green") throw std::invalid_argument("-init must be rest or taylor_green");
    if (!c.use_gpu) throw std::invalid_argument("this build has the GPU backend only (the CPU loop is the oracle)");
    return c;
}

void initialize(lbm::HostLattice<Lattice, Real>& mem, const Config& c) {
    for (std::size_t z = 0; z < c.nz; ++z)
        for (std::size_t y = 0; y < c.ny; ++y)
            for (std::size_t x = 0; x < c.nx; ++x) {
                lbc_init::Macro m{1.0, {0.0, 0.0, 0.0}};
                if (c.init == "taylor_green") m = lbc_init::taylor_green_3d(x, y, z, c.nx, c.ny, c.nz, c.u0);
                for (int i = 0; i < Lattice::Q; ++i) mem(i, z, y, x) = lbc_init::equilibrium<Lattice>(i, m.rho, m.u);
            }
}

template <lbm::CollisionType CT>
hipError_t step(const Real* cur, Real* next, const Config& c, Real omega) {
    return lbm::launch_step_gpu<Lattice, Real, CT>(cur, next, c.nx, c.ny, c.nz, omega);
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Config c = parse(argc, argv);
        const Real omega = 1.0 / c.tau;
        hipError_t (*launch)(const Real*, Real*, const Config&, Real) = nullptr;
        int collision_code = LBC_BGK;
        if (c.collision == "bgk") launch = step<lbm::CollisionType::BGK>;
        else if (c.collision == "trt") launch = step<lbm::CollisionType::TRT>, collision_code = LBC_TRT;
        else if (c.collision == "mrt") launch = step<lbm::CollisionType::MRT>, collision_code = LBC_MRT;
        else if (c.collision == "rlbm") launch = step<lbm::CollisionType::RLBM>, collision_code = LBC_RLBM;
        else throw std::invalid_argument("-collision must be bgk, trt, mrt or rlbm");

        lbm::HostLattice<Lattice, Real> memory{c.nx, c.ny, c.nz};
        const std::size_t count = memory.population_count();
        if (c.load.empty()) {
            initialize(memory, c);
        } else {
            std::ifstream in(c.load, std::ios::binary);
            if (!in.read(reinterpret_cast<char*>(memory.current()), count * sizeof(Real)))
                throw std::runtime_error("cannot read " + c.load);
        }
        std::ofstream csv(c.csv);
        if (!csv) throw std::runtime_error("failed to open " + c.csv);
        csv << "step,total_mass,total_kinetic_energy,max_velocity_magnitude\n";

        std::cout << "LB-Cube D3Q19 simulation (MI355X)\nGrid: " << c.nx << " x " << c.ny << " x " << c.nz
                  << "\nSteps: " << c.steps << "\nOutput frequency: " << c.output_frequency
                  << "\nInitial condition: " << (c.load.empty() ? c.init : c.load) << "\nBackend: GPU\nOmega: " << omega
                  << '\n'
                  << std::flush;

        Real *current = nullptr, *next = nullptr;
        check(hipMalloc(reinterpret_cast<void**>(&current), count * sizeof(Real)), "hipMalloc current populations");
        check(hipMalloc(reinterpret_cast<void**>(&next), count * sizeof(Real)), "hipMalloc next populations");
        check(hipMemcpy(current, memory.current(), count * sizeof(Real), hipMemcpyHostToDevice), "hipMemcpy H2D");
        const int code = lbm::lattice_code<Lattice>();
        if (c.slabs) {  // into the slab order, through the idle buffer
            check(static_cast<hipError_t>(lbc_relayout_tiled(code, LBC_F64, current, 0, next, c.ny, c.nx, c.ny, c.nz,
                                                             nullptr)),
                  "lbc_relayout_tiled");
            std::swap(current, next);
        }
        // The populations in the reference's planes: `current` itself, or its
        // slab order converted into the idle buffer (the next step overwrites it).
        auto planes = [&]() -> const Real* {
            if (!c.slabs) return current;
            check(static_cast<hipError_t>(lbc_relayout_tiled(code, LBC_F64, current, c.ny, next, 0, c.nx, c.ny, c.nz,
                                                             nullptr)),
                  "lbc_relayout_tiled");
            return next;
        };

        auto log = [&](int s) {
            double out[3];
            check(static_cast<hipError_t>(lbc_fluid_diagnostics(code, LBC_F64, planes(), c.nx, c.ny, c.nz, out,
                                                                nullptr)),
                  "lbc_fluid_diagnostics");
            char row[160];
            std::snprintf(row, sizeof(row), "%d,%.17g,%.17g,%.17g\n", s, out[0], out[1], out[2]);
            csv << row;
        };
        lbc_vtk::Staging staging(count / Lattice::Q, Lattice::D, false);
        auto snapshot = [&](int s) {
            if (!c.vtk) return;
            lbc_vtk::write_fluid(c.vtk_dir, static_cast<std::size_t>(s),
                                 lbc_vtk::fluid_fields(staging, code, planes(), c.nx, c.ny, c.nz), c.nx, c.ny,
                                 c.nz, Lattice::D);
        };
        snapshot(0);
        log(0);
        const auto start = std::chrono::steady_clock::now();
        for (int s = 1; s <= c.steps; ++s) {
            if (c.slabs)
                check(static_cast<hipError_t>(lbc_fluid_step_tiled(code, collision_code, LBC_F64, 0u, current, next,
                                                                   c.nx, c.ny, c.nz, c.ny, omega, 256, 1, 1, nullptr)),
                      "lbc_fluid_step_tiled");
            else
                check(launch(current, next, c, omega), "launch_step_gpu");
            std::swap(current, next);
            if (s % c.output_frequency == 0) {
                snapshot(s);
                log(s);
                std::cout << "Step " << s << " / " << c.steps << " complete\n" << std::flush;
            }
        }
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const std::chrono::duration<double> elapsed = std::chrono::steady_clock::now() - start;
        const double updates = static_cast<double>(c.nx) * c.ny * c.nz * c.steps;
        std::cout << "Elapsed time: " << elapsed.count() << " s\nPerformance: " << updates / elapsed.count() / 1e6
                  << " MLUPS\n"
                  << std::flush;
        if (!c.dump.empty()) {
            check(hipMemcpy(memory.current(), planes(), count * sizeof(Real), hipMemcpyDeviceToHost), "hipMemcpy D2H");
            std::ofstream out(c.dump, std::ios::binary);
            if (!out.write(reinterpret_cast<const char*>(memory.current()), count * sizeof(Real)))
                throw std::runtime_error("cannot write " + c.dump);
        }
        check(hipFree(current), "hipFree");
        check(hipFree(next), "hipFree");
        return EXIT_SUCCESS;
    } catch (const std::exception& error) {
        std::cerr << "Error: " << error.what() << '\n';
        return EXIT_FAILURE;
    }
}
