// lbm_shear_layer: the reference's high-Reynolds doubly periodic shear-layer
// benchmark (src/turbulent_shear_layer.cpp, CPU-only there) on the MI355X
// kernel library through include/lattice_hip.hpp.  Each D2Q9 operator starts
// from the same Kelvin-Helmholtz state (two tanh shear layers, a sinusoidal
// cross-stream perturbation, species A above y = ny / 2 and B below) and
// advances the fluid plus two passive D2Q5 scalars, one launch_step_gpu and two
// launch_scalar_step_gpu per step (src/turbulent_shear_layer.cpp:268-276).  The
// device stability guard (lbc_max_speed) runs every -check_freq steps; a speed
// above the threshold or not finite writes a last snapshot and stops that
// operator (src/turbulent_shear_layer.cpp:280-297).
//
// Defaults are the reference's constants (src/turbulent_shear_layer.cpp:40-51):
// 512 x 512, 10000 steps, check every 50 steps, VTK every 1000 steps,
// u0 0.04, delta 0.05, kappa 80, tau 0.5005, tau_c 0.8, threshold 1.0,
// operators BGK, TRT, MRT, RLBM.
// Options (the reference has none; these expose its constants):
//   -nx N -ny N -steps N -check_freq N -vtk_freq N (0: no snapshots)
//   -u0 U -delta D -kappa K -tau T -tau_c T -threshold V
//   -operators LIST        comma-separated subset of bgk,trt,mrt,rlbm
//   -load PREFIX           raw float64 populations PREFIX.{f,a,b} instead of
//                          the shear-layer state
//   -out DIR               parent of the output_shear2d_<NAME> folders (default .)
// Per operator NAME (D2Q9_BGK, ...): DIR/output_shear2d_NAME/stability_NAME.csv
// with rows "step,max_velocity" and shear_NAME_<step, 6 digits>.vtk snapshots
// (velocity, velocity_magnitude, C_A, C_B; src/turbulent_shear_layer.cpp:157-230),
// reconstructed on the device.  The step-0 row prints the speed with the
// stream's default 6 significant digits and later rows with %.17g, as the
// reference's rows do (:255 and :282).
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

#include "lattice_hip.hpp"
#include "lbcube_lattices.hpp"
#include "vtk.hpp"

namespace {

using FluidLattice = lbm::D2Q9;
using ScalarLattice = lbm::D2Q5;
using Real = double;
using lbm::CollisionType;

struct Config {
    std::size_t nx = 512, ny = 512;
    int steps = 10000;
    int check_frequency = 50;
    int vtk_frequency = 1000;
    double u0 = 0.04, delta = 0.05, kappa = 80.0, tau = 0.5005, tau_c = 0.8, threshold = 1.0;
    std::vector<CollisionType> operators{CollisionType::BGK, CollisionType::TRT, CollisionType::MRT,
                                         CollisionType::RLBM};
    std::string load, out = ".";
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

const char* operator_name(CollisionType t) {
    switch (t) {
        case CollisionType::BGK: return "D2Q9_BGK";
        case CollisionType::TRT: return "D2Q9_TRT";
        case CollisionType::MRT: return "D2Q9_MRT";
        case CollisionType::RLBM: return "D2Q9_RLBM";
    }
    return "UNKNOWN";
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
        else if (a == "-check_freq") c.check_frequency = std::stoi(value(i));
        else if (a == "-vtk_freq") c.vtk_frequency = std::stoi(value(i));
        else if (a == "-u0") c.u0 = std::stod(value(i));
        else if (a == "-delta") c.delta = std::stod(value(i));
        else if (a == "-kappa") c.kappa = std::stod(value(i));
        else if (a == "-tau") c.tau = std::stod(value(i));
        else if (a == "-tau_c") c.tau_c = std::stod(value(i));
        else if (a == "-threshold") c.threshold = std::stod(value(i));
        else if (a == "-operators") {
            c.operators.clear();
            std::stringstream in(value(i));
            std::string item;
            while (std::getline(in, item, ',')) {
                if (item == "bgk") c.operators.push_back(CollisionType::BGK);
                else if (item == "trt") c.operators.push_back(CollisionType::TRT);
                else if (item == "mrt") c.operators.push_back(CollisionType::MRT);
                else if (item == "rlbm") c.operators.push_back(CollisionType::RLBM);
                else throw std::invalid_argument("-operators takes bgk, trt, mrt, rlbm: " + item);
            }
            if (c.operators.empty()) throw std::invalid_argument("-operators must name at least one operator");
        } else if (a == "-load") c.load = value(i);
        else if (a == "-out") c.out = value(i);
        else throw std::invalid_argument("unknown argument: " + a);
    }
    if (c.steps < 0) throw std::invalid_argument("-steps must be non-negative");
    if (c.check_frequency <= 0) throw std::invalid_argument("-check_freq must be greater than zero");
    if (c.vtk_frequency < 0) throw std::invalid_argument("-vtk_freq must be non-negative");
    if (c.tau <= 0.5 || c.tau_c <= 0.5) throw std::invalid_argument("relaxation times must be greater than 0.5");
    if (!(c.threshold > 0.0)) throw std::invalid_argument("-threshold must be positive");
    return c;
}

constexpr int FX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int FY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr double FW[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
constexpr int SX[5] = {0, 1, 0, -1, 0};
constexpr int SY[5] = {0, 0, 1, 0, -1};
constexpr double SW[5] = {1.0 / 3, 1.0 / 6, 1.0 / 6, 1.0 / 6, 1.0 / 6};

struct HostState {
    std::vector<Real> f, a, b;
};

// Two tanh shear layers at y = ny / 4 and 3 ny / 4, a cross-stream sine of
// amplitude delta u0, rho = 1; A = 1 above the midline, B = 1 on and below it,
// all at equilibrium (src/turbulent_shear_layer.cpp:52-113).
HostState initial_state(const Config& c) {
    const std::size_t n = c.nx * c.ny;
    HostState s{std::vector<Real>(9 * n), std::vector<Real>(5 * n), std::vector<Real>(5 * n)};
    const double pi = std::acos(-1.0);
    for (std::size_t y = 0; y < c.ny; ++y) {
        const double yn = static_cast<double>(y) / c.ny;
        const double ca = yn > 0.5 ? 1.0 : 0.0, cb = yn <= 0.5 ? 1.0 : 0.0;
        for (std::size_t x = 0; x < c.nx; ++x) {
            const double xn = static_cast<double>(x) / c.nx;
            const double ux = yn <= 0.5 ? c.u0 * std::tanh(c.kappa * (yn - 0.25)) : c.u0 * std::tanh(c.kappa * (0.75 - yn));
            const double uy = c.u0 * c.delta * std::sin(2 * pi * (xn + 0.25));
            const double uu = ux * ux + uy * uy;
            const std::size_t cell = y * c.nx + x;
            for (int i = 0; i < 9; ++i) {
                const double cu = FX[i] * ux + FY[i] * uy;
                s.f[i * n + cell] = FW[i] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
            }
            for (int i = 0; i < 5; ++i) {
                const double e = 1.0 + 3.0 * (SX[i] * ux + SY[i] * uy);
                s.a[i * n + cell] = SW[i] * ca * e;
                s.b[i * n + cell] = SW[i] * cb * e;
            }
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
    explicit Device(const std::vector<Real>& host) {
        check(hipMalloc(reinterpret_cast<void**>(&cur), host.size() * sizeof(Real)), "hipMalloc");
        check(hipMalloc(reinterpret_cast<void**>(&next), host.size() * sizeof(Real)), "hipMalloc");
        check(hipMemcpy(cur, host.data(), host.size() * sizeof(Real), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    ~Device() {
        (void)hipFree(cur);
        (void)hipFree(next);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
};

template <CollisionType CT> void run_case(const Config& c, const HostState& init) {
    const std::string name = operator_name(CT);
    const std::filesystem::path dir = std::filesystem::path(c.out) / ("output_shear2d_" + name);
    std::filesystem::create_directories(dir);
    const std::filesystem::path log_path = dir / ("stability_" + name + ".csv");
    std::ofstream stability(log_path);
    if (!stability) throw std::runtime_error("failed to open " + log_path.string());
    stability << "step,max_velocity\n";

    const Real omega = 1.0 / c.tau, omega_c = 1.0 / c.tau_c;
    const std::size_t cells = c.nx * c.ny;
    Device f(init.f), a(init.a), b(init.b);
    std::cout << "Running " << name << " turbulent shear-layer benchmark\nOutput: " << dir.string() << '\n'
              << std::flush;

    lbc_vtk::Staging staging(cells, 2, true);
    auto snapshot = [&](int s) {
        if (c.vtk_frequency == 0) return;
        char file[96];
        std::snprintf(file, sizeof(file), "shear_%s_%06d.vtk", name.c_str(), s);
        const std::string path = (dir / file).string();
        std::ofstream out(path);
        if (!out) throw std::runtime_error("failed to open " + path);
        const auto fl = lbc_vtk::fluid_fields(staging, lbm::lattice_code<FluidLattice>(), f.cur, c.nx, c.ny, 1);
        const auto ca = lbc_vtk::concentration(staging, staging.ca, lbm::lattice_code<ScalarLattice>(), a.cur, c.nx, c.ny, 1);
        const auto cb = lbc_vtk::concentration(staging, staging.cb, lbm::lattice_code<ScalarLattice>(), b.cur, c.nx, c.ny, 1);
        lbc_vtk::header(out, "LB-Cube turbulent shear layer " + name + " step " + std::to_string(s), c.nx, c.ny, 1);
        lbc_vtk::vectors(out, "velocity", fl.u, cells, 2);
        lbc_vtk::scalars(out, "velocity_magnitude", cells,
                         [&](std::size_t i) { return std::sqrt(fl.u[i] * fl.u[i] + fl.u[cells + i] * fl.u[cells + i]); });
        lbc_vtk::scalars(out, "C_A", cells, [&](std::size_t i) { return ca[i]; });
        lbc_vtk::scalars(out, "C_B", cells, [&](std::size_t i) { return cb[i]; });
    };
    auto max_speed = [&]() {
        double v = 0.0;
        lbc_vtk::check(lbc_max_speed(lbm::lattice_code<FluidLattice>(), LBC_F64, f.cur, c.nx, c.ny, 1, &v, nullptr), "lbc_max_speed");
        return v;
    };

    snapshot(0);
    stability << "0," << max_speed() << '\n';
    const auto start = std::chrono::steady_clock::now();
    int completed = 0;
    bool crashed = false;
    for (int s = 1; s <= c.steps; ++s) {
        check(lbm::launch_step_gpu<FluidLattice, Real, CT>(f.cur, f.next, c.nx, c.ny, 1, omega), "launch fluid step");
        std::swap(f.cur, f.next);
        check(lbm::launch_scalar_step_gpu<FluidLattice, ScalarLattice, Real>(a.cur, a.next, f.cur, c.nx, c.ny, 1,
                                                                             omega_c),
              "launch scalar step A");
        std::swap(a.cur, a.next);
        check(lbm::launch_scalar_step_gpu<FluidLattice, ScalarLattice, Real>(b.cur, b.next, f.cur, c.nx, c.ny, 1,
                                                                             omega_c),
              "launch scalar step B");
        std::swap(b.cur, b.next);
        completed = s;
        if (s % c.check_frequency == 0) {
            const double v = max_speed();
            char line[64];
            std::snprintf(line, sizeof(line), "%d,%.17g\n", s, v);
            stability << line;
            if (!std::isfinite(v) || v > c.threshold) {
                std::cout << '[' << name << "] Solver Crashed due to instability! step=" << s << ", Umax=" << v << '\n'
                          << std::flush;
                snapshot(s);
                crashed = true;
                break;
            }
        }
        if (c.vtk_frequency > 0 && s % c.vtk_frequency == 0) {
            snapshot(s);
            std::cout << '[' << name << "] Step " << s << " / " << c.steps << " complete\n" << std::flush;
        }
    }
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const std::chrono::duration<double> elapsed = std::chrono::steady_clock::now() - start;
    std::cout << '[' << name << "] elapsed=" << elapsed.count()
              << " s, MLUPS=" << static_cast<double>(cells) * completed / elapsed.count() / 1e6
              << (crashed ? " (terminated early)" : "") << "\n\n"
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
        std::cout << "LB-Cube turbulent doubly periodic shear-layer benchmark (MI355X)\nGrid: " << c.nx << " x " << c.ny
                  << "\nSteps per operator: " << c.steps << "\nU0: " << c.u0 << "\ndelta: " << c.delta
                  << "\nkappa: " << c.kappa << "\ntau: " << c.tau << "\nnu: " << (c.tau - 0.5) / 3.0 << '\n'
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
