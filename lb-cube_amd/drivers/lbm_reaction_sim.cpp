// lbm_reaction_sim: the GPU path of LB-Cube's segregated reactive-mixing
// driver (src/simulate_reaction.cpp) on the MI355X kernel library, through
// include/lattice_hip.hpp: a D2Q9 fluid (BGK) and two D2Q5 species reacting
// A + B -> C, one launch_step_gpu and one launch_reaction_AB_gpu per step
// (src/simulate_reaction.cpp:516-546), a VTK snapshot and a statistics row at
// step 0 and every -out_freq steps.
//
// Options (src/simulate_reaction.cpp:109-168, same names, defaults and checks):
//   -nx N -ny N (256)  -steps N (5000)  -out_freq N (250)  -u0 U (0.04)
//   -tau T (0.6)  -tau_c T (0.8)  -k_react K (0.05)  -gpu / -cpu
// Additions:
//   -fused                 one lbc_reactive_fused_step per step (same results)
//   -load PREFIX / -dump PREFIX   raw float64 populations PREFIX.{f,a,b}
//   -csv FILE              statistics file (default reaction_diagnostics.csv)
//   -vtk_dir DIR           folder of the reaction_<step>.vtk snapshots (default .)
//   -no_vtk                no snapshots
//   -tiles R               the three population sets in tiles of R rows (R
//                          divides ny): the two launches as lbc_fluid_step_tiled
//                          and lbc_reaction_ab_step_tiled, with -fused
//                          lbc_reactive_fused_step_tiled; statistics by
//                          lbc_reactive_stats_tiled.  The same results bit for
//                          bit as the planes, a faster HBM pattern; converted
//                          into the idle ping-pong buffers for snapshots and -dump
// Rows: step,mean_A,mean_B,var_A,var_B,covariance,segregation_intensity,
// true_reaction_rate,mixed_reaction_rate with %.17g (include/lattice_io.hpp:352-368),
// computed on the device with FP64 accumulation.  Snapshots: C_A, C_B,
// reaction_rate, velocity (write_reaction_vtk, src/simulate_reaction.cpp:267-333),
// reconstructed on the device.
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

using FluidLattice = lbm::D2Q9;
using ScalarLattice = lbm::D2Q5;
using Real = double;

struct Config {
    std::size_t nx = 256, ny = 256;
    int steps = 5000;
    int output_frequency = 250;
    double tau = 0.6, tau_c = 0.8, u0 = 0.04, k_react = 0.05;
    bool use_gpu = true, fused = false;
    std::string load, dump, csv = "reaction_diagnostics.csv";
    std::string vtk_dir = ".";
    bool vtk = true;
    std::size_t tiles = 0;
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
        if (v <= 0) throw std::invalid_argument(std::string(argv[i - 1]) + " must be greater than zero");
        return static_cast<std::size_t>(v);
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-nx") c.nx = size_value(i);
        else if (a == "-ny") c.ny = size_value(i);
        else if (a == "-steps") c.steps = std::stoi(value(i));
        else if (a == "-out_freq") c.output_frequency = std::stoi(value(i));
        else if (a == "-u0") c.u0 = std::stod(value(i));
        else if (a == "-tau") c.tau = std::stod(value(i));
        else if (a == "-tau_c") c.tau_c = std::stod(value(i));
        else if (a == "-k_react") c.k_react = std::stod(value(i));
        else if (a == "-gpu") c.use_gpu = true;
        else if (a == "-cpu") c.use_gpu = false;
        else if (a == "-fused") c.fused = true;
        else if (a == "-load") c.load = value(i);
        else if (a == "-dump") c.dump = value(i);
        else if (a == "-csv") c.csv = value(i);
        else if (a == "-vtk_dir") c.vtk_dir = value(i);
        else if (a == "-no_vtk") c.vtk = false;
        else if (a == "-tiles") c.tiles = size_value(i);
        else throw std::invalid_argument("unknown argument: " + a);
    }
    if (c.tiles != 0 && c.ny % c.tiles != 0) throw std::invalid_argument("-tiles must divide -ny");
    if (c.steps < 0) throw std::invalid_argument("-steps must be non-negative");
    if (c.output_frequency <= 0) throw std::invalid_argument("-out_freq must be greater than zero");
    if (c.tau <= 0.5 || c.tau_c <= 0.5) throw std::invalid_argument("relaxation times must be greater than 0.5");
    if (c.k_react < 0.0) throw std::invalid_argument("-k_react must be non-negative");
    if (!c.use_gpu) throw std::invalid_argument("this build has the GPU backend only (the CPU loop is the oracle)");
    return c;
}

// 2D Taylor-Green fluid and species split at x = nx / 2, at equilibrium
// (src/simulate_reaction.cpp:173-232), in the reference's operation order.
void initialize(lbm::HostLattice<FluidLattice, Real>& f, lbm::HostLattice<ScalarLattice, Real>& a,
                lbm::HostLattice<ScalarLattice, Real>& b, const Config& c) {
    for (std::size_t y = 0; y < c.ny; ++y)
        for (std::size_t x = 0; x < c.nx; ++x) {
            const lbc_init::Macro m = lbc_init::taylor_green_2d(x, y, c.nx, c.ny, c.u0);
            for (int i = 0; i < FluidLattice::Q; ++i) f(i, 0, y, x) = lbc_init::equilibrium<FluidLattice>(i, m.rho, m.u);
            const double ca = x < c.nx / 2 ? 1.0 : 0.0, cb = x < c.nx / 2 ? 0.0 : 1.0;
            for (int i = 0; i < ScalarLattice::Q; ++i) {
                a(i, 0, y, x) = lbc_init::scalar_equilibrium<ScalarLattice>(i, ca, m.u);
                b(i, 0, y, x) = lbc_init::scalar_equilibrium<ScalarLattice>(i, cb, m.u);
            }
        }
}

struct Device {
    Real *cur = nullptr, *next = nullptr;
    std::size_t count;
    explicit Device(std::size_t n) : count(n) {
        check(hipMalloc(reinterpret_cast<void**>(&cur), n * sizeof(Real)), "hipMalloc");
        check(hipMalloc(reinterpret_cast<void**>(&next), n * sizeof(Real)), "hipMalloc");
    }
    ~Device() {
        (void)hipFree(cur);
        (void)hipFree(next);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
};

template <class L> void read_raw(lbm::HostLattice<L, Real>& m, const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.read(reinterpret_cast<char*>(m.current()), m.population_count() * sizeof(Real)))
        throw std::runtime_error("cannot read " + path);
}

void write_raw(const Real* device, std::size_t count, const std::string& path) {
    std::vector<Real> host(count);
    check(hipMemcpy(host.data(), device, count * sizeof(Real), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    std::ofstream out(path, std::ios::binary);
    if (!out.write(reinterpret_cast<const char*>(host.data()), host.size() * sizeof(Real)))
        throw std::runtime_error("cannot write " + path);
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Config c = parse(argc, argv);
        const Real omega = 1.0 / c.tau, omega_c = 1.0 / c.tau_c;
        lbm::HostLattice<FluidLattice, Real> fluid{c.nx, c.ny};
        lbm::HostLattice<ScalarLattice, Real> species_a{c.nx, c.ny}, species_b{c.nx, c.ny};
        if (c.load.empty()) {
            initialize(fluid, species_a, species_b, c);
        } else {
            read_raw(fluid, c.load + ".f");
            read_raw(species_a, c.load + ".a");
            read_raw(species_b, c.load + ".b");
        }
        std::ofstream csv(c.csv);
        if (!csv) throw std::runtime_error("failed to open " + c.csv);
        csv << "step,mean_A,mean_B,var_A,var_B,covariance,segregation_intensity,true_reaction_rate,"
               "mixed_reaction_rate\n";
        std::cout << "LB-Cube reactive mixing (MI355X)\nGrid: " << c.nx << " x " << c.ny << "\nSteps: " << c.steps
                  << "\nOutput frequency: " << c.output_frequency << "\nOmega: " << omega << "\nOmega_c: " << omega_c
                  << "\nk_react: " << c.k_react << "\nBackend: GPU"
                  << (c.fused ? " (fused step)" : "") << (c.tiles ? " (tiles)" : "") << '\n'
                  << std::flush;

        Device f(fluid.population_count()), a(species_a.population_count()), b(species_b.population_count());
        check(hipMemcpy(f.cur, fluid.current(), f.count * sizeof(Real), hipMemcpyHostToDevice), "hipMemcpy H2D");
        check(hipMemcpy(a.cur, species_a.current(), a.count * sizeof(Real), hipMemcpyHostToDevice), "hipMemcpy H2D");
        check(hipMemcpy(b.cur, species_b.current(), b.count * sizeof(Real), hipMemcpyHostToDevice), "hipMemcpy H2D");

        const int fluid_code = lbm::lattice_code<FluidLattice>(), scalar_code = lbm::lattice_code<ScalarLattice>();
        // -tiles: each population set into tiles of c.tiles rows, through its idle buffer
        auto relayout = [&](int code, Device& d, std::size_t from, std::size_t to) {
            check(static_cast<hipError_t>(lbc_relayout_tiled(code, LBC_F64, d.cur, from, d.next, to, c.nx, c.ny, 1,
                                                             nullptr)),
                  "lbc_relayout_tiled");
        };
        if (c.tiles != 0) {
            relayout(fluid_code, f, 0, c.tiles);
            relayout(scalar_code, a, 0, c.tiles);
            relayout(scalar_code, b, 0, c.tiles);
            std::swap(f.cur, f.next);
            std::swap(a.cur, a.next);
            std::swap(b.cur, b.next);
        }
        // A set in the reference's planes: its current buffer, or with tiles its
        // planes converted into the idle buffer (the next step overwrites it).
        auto planes = [&](int code, Device& d) -> const Real* {
            if (c.tiles == 0) return d.cur;
            relayout(code, d, c.tiles, 0);
            return d.next;
        };

        auto log = [&](int s) {
            double out[8];
            if (c.tiles != 0)
                check(static_cast<hipError_t>(lbc_reactive_stats_tiled(scalar_code, LBC_F64, a.cur, b.cur, c.nx, c.ny, 1,
                                                                       c.tiles, c.k_react, 0.0, 0.0, out, nullptr)),
                      "lbc_reactive_stats_tiled");
            else
                check(static_cast<hipError_t>(lbc_reactive_stats(scalar_code, LBC_F64, a.cur, b.cur, c.nx, c.ny, 1,
                                                                 c.k_react, 0.0, 0.0, out, nullptr)),
                      "lbc_reactive_stats");
            char row[400];
            std::snprintf(row, sizeof(row), "%d,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", s, out[0], out[1],
                          out[2], out[3], out[4], out[5], out[6], out[7]);
            csv << row;
        };
        lbc_vtk::Staging staging(c.nx * c.ny, 2, true);
        auto snapshot = [&](int s) {
            if (!c.vtk) return;
            char name[40];
            std::snprintf(name, sizeof(name), "/reaction_%06d.vtk", s);
            const auto fields = lbc_vtk::fluid_fields(staging, fluid_code, planes(fluid_code, f), c.nx, c.ny, 1);
            const auto ca = lbc_vtk::concentration(staging, staging.ca, scalar_code, planes(scalar_code, a), c.nx, c.ny, 1);
            const auto cb = lbc_vtk::concentration(staging, staging.cb, scalar_code, planes(scalar_code, b), c.nx, c.ny, 1);
            lbc_vtk::write_reactive(c.vtk_dir + name, "LB-Cube reactive mixing step " + std::to_string(s), false,
                                    fields, ca, cb, c.k_react, c.nx, c.ny);
        };
        snapshot(0);
        log(0);
        const auto start = std::chrono::steady_clock::now();
        for (int s = 1; s <= c.steps; ++s) {
            if (c.tiles != 0 && c.fused) {
                check(static_cast<hipError_t>(lbc_reactive_fused_step_tiled(
                          fluid_code, LBC_BGK, scalar_code, LBC_F64, 0u, f.cur, f.next, a.cur, a.next, b.cur, b.next,
                          c.nx, c.ny, 1, c.tiles, omega, omega_c, c.k_react, 0.0, 0.0, 256, 1, 1, nullptr)),
                      "lbc_reactive_fused_step_tiled");
                std::swap(f.cur, f.next);
            } else if (c.tiles != 0) {
                check(static_cast<hipError_t>(lbc_fluid_step_tiled(fluid_code, LBC_BGK, LBC_F64, 0u, f.cur, f.next,
                                                                   c.nx, c.ny, 1, c.tiles, omega, 256, 1, 1,
                                                                   nullptr)),
                      "lbc_fluid_step_tiled");
                std::swap(f.cur, f.next);
                check(static_cast<hipError_t>(lbc_reaction_ab_step_tiled(
                          fluid_code, scalar_code, LBC_F64, a.cur, a.next, b.cur, b.next, f.cur, c.nx, c.ny, 1,
                          c.tiles, omega_c, c.k_react, 0.0, 0.0, 256, 1, 1, nullptr)),
                      "lbc_reaction_ab_step_tiled");
            } else if (c.fused) {
                check(static_cast<hipError_t>(lbc_reactive_fused_step(
                          lbm::lattice_code<FluidLattice>(), LBC_BGK, lbm::lattice_code<ScalarLattice>(), LBC_F64, f.cur, f.next, a.cur, a.next, b.cur,
                          b.next, c.nx, c.ny, 1, omega, omega_c, c.k_react, 0.0, 0.0, 256, 1, 1, nullptr)),
                      "lbc_reactive_fused_step");
                std::swap(f.cur, f.next);
            } else {
                check(lbm::launch_step_gpu<FluidLattice, Real>(f.cur, f.next, c.nx, c.ny, 1, omega),
                      "launch fluid step");
                std::swap(f.cur, f.next);
                check(lbm::launch_reaction_AB_gpu<FluidLattice, ScalarLattice, Real>(
                          a.cur, a.next, b.cur, b.next, f.cur, c.nx, c.ny, 1, omega_c, c.k_react),
                      "launch reaction step");
            }
            std::swap(a.cur, a.next);
            std::swap(b.cur, b.next);
            if (s % c.output_frequency == 0) {
                snapshot(s);
                log(s);
                std::cout << "Step " << s << " / " << c.steps << " complete\n" << std::flush;
            }
        }
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const std::chrono::duration<double> elapsed = std::chrono::steady_clock::now() - start;
        std::cout << "Elapsed time: " << elapsed.count() << " s\nFluid lattice updates: "
                  << static_cast<double>(c.nx) * c.ny * c.steps / elapsed.count() / 1e6 << " MLUPS\n"
                  << std::flush;
        if (!c.dump.empty()) {
            write_raw(planes(fluid_code, f), f.count, c.dump + ".f");
            write_raw(planes(scalar_code, a), a.count, c.dump + ".a");
            write_raw(planes(scalar_code, b), b.count, c.dump + ".b");
        }
        return EXIT_SUCCESS;
    } catch (const std::exception& error) {
        std::cerr << "Error: " << error.what() << '\n';
        return EXIT_FAILURE;
    }
}
