// lbm_compare_operators: the reference's collision-operator comparison
// (src/compare_operators.cpp, CPU-only there) run on the MI355X kernel library
// through include/lattice_hip.hpp.  The same 2D reactive-mixing case -- D2Q9
// fluid from the Taylor-Green state, two D2Q5 species split at x = nx / 2,
// A + B -> C -- is advanced once per collision operator: one launch_step_gpu
// and one launch_reaction_AB_gpu per step (src/compare_operators.cpp:262-269),
// reactive statistics every -stats_freq steps, ASCII VTK snapshots every
// -vtk_freq steps.
//
// Defaults are the reference's constants (src/compare_operators.cpp:38-47):
// 256 x 256, 5000 steps, statistics every 10 steps, VTK every 500 steps,
// tau 0.6, tau_c 0.8, u0 0.04, k_react 0.05, operators BGK, TRT, MRT, RLBM.
// Options (the reference has none; these expose its constants):
//   -nx N -ny N -steps N -stats_freq N -vtk_freq N (0: no snapshots)
//   -tau T -tau_c T -u0 U -k_react K
//   -operators LIST        comma-separated subset of bgk,trt,mrt,rlbm
//   -fused                 one lbc_reactive_fused_step per step (same results)
//   -load PREFIX           raw float64 populations PREFIX.{f,a,b} instead of
//                          the Taylor-Green state
//   -out DIR               parent of the output_<NAME> folders (default .)
// Per operator NAME: DIR/output_NAME/stats_NAME.csv with the rows of
// log_reactive_diagnostics (include/lattice_io.hpp:352-368, %.17g, FP64 device
// sums) and DIR/output_NAME/comparison_NAME_<step, 6 digits>.vtk with velocity,
// C_A, C_B and k C_A C_B per cell (the fields of src/compare_operators.cpp:150-236),
// reconstructed on the device.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
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
using lbm::CollisionType;

struct Config {
    std::size_t nx = 256, ny = 256;
    int steps = 5000;
    int stats_frequency = 10;
    int vtk_frequency = 500;
    double tau = 0.6, tau_c = 0.8, u0 = 0.04, k_react = 0.05;
    std::vector<CollisionType> operators{CollisionType::BGK, CollisionType::TRT, CollisionType::MRT,
                                         CollisionType::RLBM};
    bool fused = false;
    std::string load, out = ".";
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

const char* operator_name(CollisionType t) {
    switch (t) {
        case CollisionType::BGK: return "BGK";
        case CollisionType::TRT: return "TRT";
        case CollisionType::MRT: return "MRT";
        case CollisionType::RLBM: return "RLBM";
    }
    return "UNKNOWN";
}

std::vector<CollisionType> parse_operators(const std::string& list) {
    std::vector<CollisionType> ops;
    std::stringstream in(list);
    std::string item;
    while (std::getline(in, item, ',')) {
        if (item == "bgk") ops.push_back(CollisionType::BGK);
        else if (item == "trt") ops.push_back(CollisionType::TRT);
        else if (item == "mrt") ops.push_back(CollisionType::MRT);
        else if (item == "rlbm") ops.push_back(CollisionType::RLBM);
        else throw std::invalid_argument("-operators takes bgk, trt, mrt, rlbm: " + item);
    }
    if (ops.empty()) throw std::invalid_argument("-operators must name at least one operator");
    return ops;
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
        else if (a == "-stats_freq") c.stats_frequency = std::stoi(value(i));
        else if (a == "-vtk_freq") c.vtk_frequency = std::stoi(value(i));
        else if (a == "-tau") c.tau = std::stod(value(i));
        else if (a == "-tau_c") c.tau_c = std::stod(value(i));
        else if (a == "-u0") c.u0 = std::stod(value(i));
        else if (a == "-k_react") c.k_react = std::stod(value(i));
        else if (a == "-operators") c.operators = parse_operators(value(i));
        else if (a == "-fused") c.fused = true;
        else if (a == "-load") c.load = value(i);
        else if (a == "-out") c.out = value(i);
        else throw std::invalid_argument("unknown argument: " + a);
    }
    if (c.steps < 0) throw std::invalid_argument("-steps must be non-negative");
    if (c.stats_frequency <= 0) throw std::invalid_argument("-stats_freq must be greater than zero");
    if (c.vtk_frequency < 0) throw std::invalid_argument("-vtk_freq must be non-negative");
    if (c.tau <= 0.5 || c.tau_c <= 0.5) throw std::invalid_argument("relaxation times must be greater than 0.5");
    if (c.k_react < 0.0) throw std::invalid_argument("-k_react must be non-negative");
    return c;
}

// Host populations of one case: fluid [9][ny][nx], species [5][ny][nx].
struct HostState {
    std::vector<Real> f, a, b;
};

// The Taylor-Green fluid and the split species at equilibrium
// (src/compare_operators.cpp:64-118).
HostState initial_state(const Config& c) {
    const std::size_t n = c.nx * c.ny;
    HostState s{std::vector<Real>(FluidLattice::Q * n), std::vector<Real>(ScalarLattice::Q * n),
                std::vector<Real>(ScalarLattice::Q * n)};
    for (std::size_t y = 0; y < c.ny; ++y)
        for (std::size_t x = 0; x < c.nx; ++x) {
            const lbc_init::Macro m = lbc_init::taylor_green_2d(x, y, c.nx, c.ny, c.u0);
            const std::size_t cell = y * c.nx + x;
            for (int i = 0; i < FluidLattice::Q; ++i)
                s.f[i * n + cell] = lbc_init::equilibrium<FluidLattice>(i, m.rho, m.u);
            const double ca = x < c.nx / 2 ? 1.0 : 0.0, cb = x < c.nx / 2 ? 0.0 : 1.0;
            for (int i = 0; i < ScalarLattice::Q; ++i) {
                s.a[i * n + cell] = lbc_init::scalar_equilibrium<ScalarLattice>(i, ca, m.u);
                s.b[i * n + cell] = lbc_init::scalar_equilibrium<ScalarLattice>(i, cb, m.u);
            }
        }
    return s;
}

void read_raw(std::vector<Real>& v, const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(Real))))
        throw std::runtime_error("cannot read " + path);
}

struct Device {
    Real *cur = nullptr, *next = nullptr;
    std::size_t count;
    explicit Device(const std::vector<Real>& host) : count(host.size()) {
        check(hipMalloc(reinterpret_cast<void**>(&cur), count * sizeof(Real)), "hipMalloc");
        check(hipMalloc(reinterpret_cast<void**>(&next), count * sizeof(Real)), "hipMalloc");
        check(hipMemcpy(cur, host.data(), count * sizeof(Real), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    ~Device() {
        (void)hipFree(cur);
        (void)hipFree(next);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
};

template <CollisionType CT>
void run_case(const Config& c, const HostState& init) {
    const char* name = operator_name(CT);
    const std::filesystem::path dir = std::filesystem::path(c.out) / (std::string("output_") + name);
    std::filesystem::create_directories(dir);
    const std::filesystem::path stats_path = dir / (std::string("stats_") + name + ".csv");
    std::ofstream stats(stats_path);
    if (!stats) throw std::runtime_error("failed to open " + stats_path.string());
    stats << "step,mean_A,mean_B,var_A,var_B,covariance,segregation_intensity,true_reaction_rate,"
             "mixed_reaction_rate\n";

    const Real omega = 1.0 / c.tau, omega_c = 1.0 / c.tau_c;
    Device f(init.f), a(init.a), b(init.b);
    std::cout << "Running " << name << " comparison case\n" << std::flush;

    auto log = [&](int s) {
        double out[8];
        check(static_cast<hipError_t>(lbc_reactive_stats(lbm::lattice_code<ScalarLattice>(), LBC_F64, a.cur, b.cur, c.nx, c.ny, 1,
                                                         c.k_react, 0.0, 0.0, out, nullptr)),
              "lbc_reactive_stats");
        char row[400];
        std::snprintf(row, sizeof(row), "%d,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", s, out[0], out[1],
                      out[2], out[3], out[4], out[5], out[6], out[7]);
        stats << row;
    };
    lbc_vtk::Staging staging(c.nx * c.ny, 2, true);
    auto snapshot = [&](int s) {
        char file[96];
        std::snprintf(file, sizeof(file), "comparison_%s_%06d.vtk", name, s);
        const auto fields = lbc_vtk::fluid_fields(staging, lbm::lattice_code<FluidLattice>(), f.cur, c.nx, c.ny, 1);
        const auto ca = lbc_vtk::concentration(staging, staging.ca, lbm::lattice_code<ScalarLattice>(), a.cur, c.nx, c.ny, 1);
        const auto cb = lbc_vtk::concentration(staging, staging.cb, lbm::lattice_code<ScalarLattice>(), b.cur, c.nx, c.ny, 1);
        lbc_vtk::write_reactive((dir / file).string(),
                                std::string("LB-Cube ") + name + " comparison step " + std::to_string(s), true, fields,
                                ca, cb, c.k_react, c.nx, c.ny);
    };
    log(0);
    if (c.vtk_frequency > 0) snapshot(0);

    const auto start = std::chrono::steady_clock::now();
    for (int s = 1; s <= c.steps; ++s) {
        if (c.fused) {
            check(static_cast<hipError_t>(lbc_reactive_fused_step(
                      lbm::lattice_code<FluidLattice>(), static_cast<int>(CT), lbm::lattice_code<ScalarLattice>(), LBC_F64, f.cur, f.next, a.cur,
                      a.next, b.cur, b.next, c.nx, c.ny, 1, omega, omega_c, c.k_react, 0.0, 0.0, 256, 1, 1, nullptr)),
                  "lbc_reactive_fused_step");
            std::swap(f.cur, f.next);
        } else {
            check(lbm::launch_step_gpu<FluidLattice, Real, CT>(f.cur, f.next, c.nx, c.ny, 1, omega),
                  "launch fluid step");
            std::swap(f.cur, f.next);
            check(lbm::launch_reaction_AB_gpu<FluidLattice, ScalarLattice, Real>(a.cur, a.next, b.cur, b.next, f.cur,
                                                                                 c.nx, c.ny, 1, omega_c, c.k_react),
                  "launch reaction step");
        }
        std::swap(a.cur, a.next);
        std::swap(b.cur, b.next);
        if (s % c.stats_frequency == 0) log(s);
        if (c.vtk_frequency > 0 && s % c.vtk_frequency == 0) {
            snapshot(s);
            std::cout << name << ": step " << s << " / " << c.steps << " complete\n" << std::flush;
        }
    }
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const std::chrono::duration<double> elapsed = std::chrono::steady_clock::now() - start;
    std::cout << name << " complete in " << elapsed.count() << " s ("
              << static_cast<double>(c.nx) * c.ny * c.steps / elapsed.count() / 1e6 << " MLUPS fluid step basis)\n"
              << std::flush;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Config c = parse(argc, argv);
        HostState init;
        if (c.load.empty()) {
            init = initial_state(c);
        } else {
            const std::size_t n = c.nx * c.ny;
            init = HostState{std::vector<Real>(9 * n), std::vector<Real>(5 * n), std::vector<Real>(5 * n)};
            read_raw(init.f, c.load + ".f");
            read_raw(init.a, c.load + ".a");
            read_raw(init.b, c.load + ".b");
        }
        std::cout << "LB-Cube collision operator comparison (MI355X)\nGrid: " << c.nx << " x " << c.ny
                  << "\nSteps per operator: " << c.steps << "\nStats frequency: " << c.stats_frequency
                  << "\nVTK frequency: " << c.vtk_frequency << "\nBackend: GPU" << (c.fused ? " (fused step)" : "")
                  << '\n'
                  << std::flush;
        for (const CollisionType t : c.operators) {
            switch (t) {
                case CollisionType::BGK: run_case<CollisionType::BGK>(c, init); break;
                case CollisionType::TRT: run_case<CollisionType::TRT>(c, init); break;
                case CollisionType::MRT: run_case<CollisionType::MRT>(c, init); break;
                case CollisionType::RLBM: run_case<CollisionType::RLBM>(c, init); break;
            }
        }
        return EXIT_SUCCESS;
    } catch (const std::exception& error) {
        std::cerr << "Error: " << error.what() << '\n';
        return EXIT_FAILURE;
    }
}
