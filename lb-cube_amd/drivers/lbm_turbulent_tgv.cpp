// lbm_turbulent_tgv: the reference's high-Reynolds 3D Taylor-Green stability
// benchmark (src/turbulent_tgv_3d.cpp, CPU-only there) on the MI355X kernel
// library.  Each lattice/operator configuration starts from the 3D TGV
// equilibrium (rho = 1, u0 = 0.05), steps at tau = 0.5005 close to the
// stability limit, and checks the maximum speed on the device every
// -check_freq steps (lbc_max_speed: compute_max_velocity of
// src/turbulent_tgv_3d.cpp:149-173, +inf for a non-finite cell); a speed above
// the threshold or not finite stops that configuration as the reference does
// (src/turbulent_tgv_3d.cpp:302-315).
//
// Defaults are the reference's constants (src/turbulent_tgv_3d.cpp:40-48):
// 128^3, 5000 steps, check every 50 steps, VTK every 1000 steps, threshold
// 1.0, configuration D3Q19_MRT (the one its main() runs).
// Options (the reference has none; these expose its constants):
//   -nx N -ny N -nz N -steps N -check_freq N -vtk_freq N (0: no snapshots)
//   -tau T -u0 U -threshold V
//   -configs LIST          comma-separated LATTICE_OPERATOR names, LATTICE in
//                          D3Q19, D3Q27 and OPERATOR in BGK, TRT, MRT, RLBM
//   -storage f64|f32       population storage (the reference runs double)
//   -fast                  LBC_FLAG_FAST arithmetic (BGK, TRT, MRT)
//   -load FILE             raw float64 populations [Q][Z][Y][X] instead of the TGV
//   -out DIR               parent of the output_tgv3d_<NAME> folders (default .)
// Per configuration NAME: DIR/output_tgv3d_NAME/stability_NAME.csv with rows
// "step,max_velocity" (%.17g) and tgv3d_NAME_<step, 6 digits>.vtk snapshots
// (density, velocity_magnitude, velocity; src/turbulent_tgv_3d.cpp:184-251),
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

#include "lbcube.h"
#include "vtk.hpp"

namespace {

struct Run {
    std::string name;
    int lattice, q, collision;
};

struct Config {
    std::size_t nx = 128, ny = 128, nz = 128;
    int steps = 5000;
    int check_frequency = 50;
    int vtk_frequency = 1000;
    double tau = 0.5005, u0 = 0.05, threshold = 1.0;
    std::vector<Run> runs;
    int storage = LBC_F64;
    bool fast = false;
    std::string load, out = ".";
};

void check(int status, const char* what) {
    if (status != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(static_cast<hipError_t>(status)));
}

Run parse_run(const std::string& name) {
    const auto cut = name.find('_');
    if (cut == std::string::npos) throw std::invalid_argument("-configs entries are LATTICE_OPERATOR: " + name);
    const std::string lat = name.substr(0, cut), op = name.substr(cut + 1);
    Run r{name, 0, 0, 0};
    if (lat == "D3Q19") r.lattice = LBC_D3Q19, r.q = 19;
    else if (lat == "D3Q27") r.lattice = LBC_D3Q27, r.q = 27;
    else throw std::invalid_argument("-configs lattice must be D3Q19 or D3Q27: " + name);
    if (op == "BGK") r.collision = LBC_BGK;
    else if (op == "TRT") r.collision = LBC_TRT;
    else if (op == "MRT") r.collision = LBC_MRT;
    else if (op == "RLBM") r.collision = LBC_RLBM;
    else throw std::invalid_argument("-configs operator must be BGK, TRT, MRT or RLBM: " + name);
    return r;
}

Config parse(int argc, char** argv) {
    Config c;
    c.runs.push_back(parse_run("D3Q19_MRT"));
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
        else if (a == "-nz") c.nz = size_value(i);
        else if (a == "-steps") c.steps = std::stoi(value(i));
        else if (a == "-check_freq") c.check_frequency = std::stoi(value(i));
        else if (a == "-vtk_freq") c.vtk_frequency = std::stoi(value(i));
        else if (a == "-tau") c.tau = std::stod(value(i));
        else if (a == "-u0") c.u0 = std::stod(value(i));
        else if (a == "-threshold") c.threshold = std::stod(value(i));
        else if (a == "-configs") {
            c.runs.clear();
            std::stringstream in(value(i));
            std::string item;
            while (std::getline(in, item, ',')) c.runs.push_back(parse_run(item));
            if (c.runs.empty()) throw std::invalid_argument("-configs must name at least one configuration");
        } else if (a == "-storage") {
            const std::string s = value(i);
            if (s == "f64") c.storage = LBC_F64;
            else if (s == "f32") c.storage = LBC_F32;
            else throw std::invalid_argument("-storage must be f64 or f32");
        } else if (a == "-fast") c.fast = true;
        else if (a == "-load") c.load = value(i);
        else if (a == "-out") c.out = value(i);
        else throw std::invalid_argument("unknown argument: " + a);
    }
    if (c.steps < 0) throw std::invalid_argument("-steps must be non-negative");
    if (c.check_frequency <= 0) throw std::invalid_argument("-check_freq must be greater than zero");
    if (c.vtk_frequency < 0) throw std::invalid_argument("-vtk_freq must be non-negative");
    if (c.tau <= 0.5) throw std::invalid_argument("-tau must be greater than 0.5");
    if (!(c.threshold > 0.0)) throw std::invalid_argument("-threshold must be positive");
    return c;
}

// Velocity sets of the two 3D lattices in the library's direction order.
void velocities(int q, std::vector<int>& cx, std::vector<int>& cy, std::vector<int>& cz, std::vector<double>& w) {
    const int b[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                          {1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1},
                          {-1, 0, -1}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
    cx.clear(), cy.clear(), cz.clear(), w.clear();
    for (const auto& v : b) {
        cx.push_back(v[0]), cy.push_back(v[1]), cz.push_back(v[2]);
        const int c2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        w.push_back(q == 19 ? (c2 == 0 ? 1.0 / 3 : c2 == 1 ? 1.0 / 18 : 1.0 / 36)
                            : (c2 == 0 ? 8.0 / 27 : c2 == 1 ? 2.0 / 27 : 1.0 / 54));
    }
    if (q == 27)
        for (int sz : {1, -1})
            for (int sy : {1, -1})
                for (int sx : {1, -1}) cx.push_back(sx), cy.push_back(sy), cz.push_back(sz), w.push_back(1.0 / 216);
}

// The 3D TGV at equilibrium, rho = 1 (src/turbulent_tgv_3d.cpp:62-110).
std::vector<double> initial_state(const Config& c, int q) {
    std::vector<int> cx, cy, cz;
    std::vector<double> w;
    velocities(q, cx, cy, cz, w);
    const std::size_t n = c.nx * c.ny * c.nz;
    std::vector<double> f(q * n);
    const double pi = std::acos(-1.0);
    for (std::size_t z = 0; z < c.nz; ++z)
        for (std::size_t y = 0; y < c.ny; ++y)
            for (std::size_t x = 0; x < c.nx; ++x) {
                const double px = 2 * pi * x / c.nx, py = 2 * pi * y / c.ny, pz = 2 * pi * z / c.nz;
                const double ux = c.u0 * std::sin(px) * std::cos(py) * std::cos(pz);
                const double uy = -c.u0 * std::cos(px) * std::sin(py) * std::cos(pz);
                const double uu = ux * ux + uy * uy;
                const std::size_t cell = (z * c.ny + y) * c.nx + x;
                for (int i = 0; i < q; ++i) {
                    const double cu = cx[i] * ux + cy[i] * uy;
                    f[i * n + cell] = w[i] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
                }
            }
    return f;
}

// Ping-pong device buffers, freed on every exit path.
struct DeviceBuffers {
    void* cur = nullptr;
    void* next = nullptr;
    explicit DeviceBuffers(std::size_t bytes) {
        check(hipMalloc(&cur, bytes), "hipMalloc populations");
        check(hipMalloc(&next, bytes), "hipMalloc populations");
    }
    ~DeviceBuffers() {
        (void)hipFree(cur);
        (void)hipFree(next);
    }
    DeviceBuffers(const DeviceBuffers&) = delete;
    DeviceBuffers& operator=(const DeviceBuffers&) = delete;
};

void run_config(const Config& c, const Run& r) {
    const std::filesystem::path dir = std::filesystem::path(c.out) / ("output_tgv3d_" + r.name);
    std::filesystem::create_directories(dir);
    const std::filesystem::path log_path = dir / ("stability_" + r.name + ".csv");
    std::ofstream stability(log_path);
    if (!stability) throw std::runtime_error("failed to open " + log_path.string());
    stability << "step,max_velocity\n";

    const std::size_t cells = c.nx * c.ny * c.nz, count = r.q * cells;
    std::vector<double> host(count);
    if (c.load.empty()) {
        host = initial_state(c, r.q);
    } else {
        std::ifstream in(c.load, std::ios::binary);
        if (!in.read(reinterpret_cast<char*>(host.data()), static_cast<std::streamsize>(count * sizeof(double))))
            throw std::runtime_error("cannot read " + c.load);
    }
    const std::size_t elem = c.storage == LBC_F64 ? 8 : 4;
    DeviceBuffers buffers(count * elem);
    void*& cur = buffers.cur;
    void*& next = buffers.next;
    if (c.storage == LBC_F64) {
        check(hipMemcpy(cur, host.data(), count * 8, hipMemcpyHostToDevice), "hipMemcpy H2D");
    } else {
        std::vector<float> narrow(host.begin(), host.end());
        check(hipMemcpy(cur, narrow.data(), count * 4, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }

    const double omega = 1.0 / c.tau;
    std::cout << "Running " << r.name << " 3D turbulent TGV benchmark\nGrid: " << c.nx << " x " << c.ny << " x "
              << c.nz << "\ntau=" << c.tau << ", omega=" << omega << ", nu=" << (c.tau - 0.5) / 3.0
              << "\nOutput: " << dir.string() << '\n'
              << std::flush;

    lbc_vtk::Staging staging(cells, 3, false);
    auto snapshot = [&](int s) {
        if (c.vtk_frequency == 0) return;
        check(lbc_macro_fields(r.lattice, c.storage, cur, c.nx, c.ny, c.nz, staging.rho, staging.u, nullptr),
              "lbc_macro_fields");
        const lbc_vtk::FluidFields f{lbc_vtk::Staging::host(staging.rho, cells),
                                     lbc_vtk::Staging::host(staging.u, 3 * cells)};
        char file[128];
        std::snprintf(file, sizeof(file), "tgv3d_%s_%06d.vtk", r.name.c_str(), s);
        const std::string path = (dir / file).string();
        std::ofstream out(path);
        if (!out) throw std::runtime_error("failed to open " + path);
        lbc_vtk::header(out, "LB-Cube 3D turbulent TGV " + r.name + " step " + std::to_string(s), c.nx, c.ny, c.nz);
        lbc_vtk::scalars(out, "density", cells, [&](std::size_t i) { return f.rho[i]; });
        lbc_vtk::scalars(out, "velocity_magnitude", cells, [&](std::size_t i) {
            return std::sqrt(f.u[i] * f.u[i] + f.u[cells + i] * f.u[cells + i] +
                             f.u[2 * cells + i] * f.u[2 * cells + i]);
        });
        lbc_vtk::vectors(out, "velocity", f.u, cells, 3);
    };
    auto max_speed = [&]() {
        double v = 0.0;
        check(lbc_max_speed(r.lattice, c.storage, cur, c.nx, c.ny, c.nz, &v, nullptr), "lbc_max_speed");
        return v;
    };
    auto row = [&](int s, double v) {
        char line[64];
        std::snprintf(line, sizeof(line), "%d,%.17g\n", s, v);
        stability << line;
    };

    snapshot(0);
    row(0, max_speed());
    const auto start = std::chrono::steady_clock::now();
    int completed = 0;
    bool crashed = false;
    for (int s = 1; s <= c.steps; ++s) {
        check(lbc_fluid_step_ex(r.lattice, r.collision, c.storage, c.fast ? LBC_FLAG_FAST : 0u, cur, next, c.nx, c.ny,
                                c.nz, omega, 256, 1, 1, nullptr),
              "lbc_fluid_step_ex");
        std::swap(cur, next);
        completed = s;
        if (s % c.check_frequency == 0) {
            const double v = max_speed();
            row(s, v);
            if (!std::isfinite(v) || v > c.threshold) {
                std::cout << '[' << r.name << "] Solver Crashed due to instability! step=" << s << ", Umax=" << v
                          << '\n'
                          << std::flush;
                crashed = true;
                break;
            }
        }
        if (c.vtk_frequency > 0 && s % c.vtk_frequency == 0) {
            snapshot(s);
            std::cout << '[' << r.name << "] Step " << s << " / " << c.steps << " complete\n" << std::flush;
        }
    }
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const std::chrono::duration<double> elapsed = std::chrono::steady_clock::now() - start;
    std::cout << '[' << r.name << "] elapsed=" << elapsed.count()
              << " s, MLUPS=" << static_cast<double>(cells) * completed / elapsed.count() / 1e6
              << (crashed ? " (terminated early)" : "") << "\n\n"
              << std::flush;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Config c = parse(argc, argv);
        for (const Run& r : c.runs) run_config(c, r);
        return EXIT_SUCCESS;
    } catch (const std::exception& error) {
        std::cerr << "Fatal benchmark error: " << error.what() << '\n';
        return EXIT_FAILURE;
    }
}
