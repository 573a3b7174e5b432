// Legacy ASCII VTK output shared by the drivers: STRUCTURED_POINTS datasets,
// x fastest, one "%.17g" value per line, as the reference's writers print
// them (write_vtk, include/lattice_io.hpp:386-451; write_reaction_vtk,
// src/simulate_reaction.cpp:267-333; write_comparison_vtk,
// src/compare_operators.cpp:150-236).  The fields come from the device
// (lbc_macro_fields / lbc_scalar_fields): only (1 + D) or 1 doubles per cell
// cross PCIe instead of every population.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbcube.h"

namespace lbc_vtk {

inline void check(int status, const char* what) {
    if (status != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(static_cast<hipError_t>(status)));
}

inline void header(std::ofstream& out, const std::string& title, std::size_t nx, std::size_t ny, std::size_t nz) {
    out << "# vtk DataFile Version 3.0\n" << title << "\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS " << nx << ' '
        << ny << ' ' << nz << "\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA " << nx * ny * nz << '\n';
}

template <class Value> void scalars(std::ofstream& out, const char* name, std::size_t n, Value value) {
    out << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
    char line[40];
    for (std::size_t i = 0; i < n; ++i) {
        std::snprintf(line, sizeof(line), "%.17g\n", static_cast<double>(value(i)));
        out << line;
    }
}

// u holds dims planes of n values; missing components print as 0.
inline void vectors(std::ofstream& out, const char* name, const std::vector<double>& u, std::size_t n, int dims) {
    out << "VECTORS " << name << " double\n";
    char line[96];
    for (std::size_t i = 0; i < n; ++i) {
        const double z = dims == 3 ? u[2 * n + i] : 0.0;
        std::snprintf(line, sizeof(line), "%.17g %.17g %.17g\n", u[i], u[n + i], z);
        out << line;
    }
}

// Device staging for the fields: allocated once per run, copied back per
// snapshot.
struct Staging {
    double* rho = nullptr;
    double* u = nullptr;
    double* ca = nullptr;
    double* cb = nullptr;
    std::size_t cells;
    int dims;
    Staging(std::size_t cells_, int dims_, bool species) : cells(cells_), dims(dims_) {
        check(hipMalloc(reinterpret_cast<void**>(&rho), cells * sizeof(double)), "hipMalloc rho");
        check(hipMalloc(reinterpret_cast<void**>(&u), dims * cells * sizeof(double)), "hipMalloc u");
        if (species) {
            check(hipMalloc(reinterpret_cast<void**>(&ca), cells * sizeof(double)), "hipMalloc C_A");
            check(hipMalloc(reinterpret_cast<void**>(&cb), cells * sizeof(double)), "hipMalloc C_B");
        }
    }
    ~Staging() {
        (void)hipFree(rho);
        (void)hipFree(u);
        (void)hipFree(ca);
        (void)hipFree(cb);
    }
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;

    static std::vector<double> host(const double* dev, std::size_t n) {
        std::vector<double> h(n);
        check(hipMemcpy(h.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy fields D2H");
        return h;
    }
};

struct FluidFields {
    std::vector<double> rho, u;
};

// Density and velocity of a fluid population buffer (FP64 storage).
inline FluidFields fluid_fields(Staging& s, int lattice, const void* pop, std::size_t nx, std::size_t ny,
                                std::size_t nz) {
    check(lbc_macro_fields(lattice, LBC_F64, pop, nx, ny, nz, s.rho, s.u, nullptr), "lbc_macro_fields");
    return FluidFields{Staging::host(s.rho, s.cells), Staging::host(s.u, s.dims * s.cells)};
}

inline std::vector<double> concentration(Staging& s, double* dev, int scalar_lattice, const void* pop, std::size_t nx,
                                         std::size_t ny, std::size_t nz) {
    check(lbc_scalar_fields(scalar_lattice, LBC_F64, pop, nx, ny, nz, 0.0, dev, nullptr), "lbc_scalar_fields");
    return Staging::host(dev, s.cells);
}

// write_vtk (include/lattice_io.hpp:386-451): density, velocity magnitude,
// velocity, to lbm_<step, 6 digits>.vtk in `dir`.
inline void write_fluid(const std::string& dir, std::size_t step, const FluidFields& f, std::size_t nx, std::size_t ny,
                        std::size_t nz, int dims) {
    char name[32];
    std::snprintf(name, sizeof(name), "lbm_%06zu.vtk", step);
    const std::string path = dir + "/" + name;
    std::ofstream out(path);
    if (!out) throw std::runtime_error("failed to open " + path);
    const std::size_t n = nx * ny * nz;
    header(out, "LBM output step " + std::to_string(step), nx, ny, nz);
    scalars(out, "density", n, [&](std::size_t i) { return f.rho[i]; });
    scalars(out, "velocity_magnitude", n, [&](std::size_t i) {
        double s = f.u[i] * f.u[i] + f.u[n + i] * f.u[n + i];
        if (dims == 3) s += f.u[2 * n + i] * f.u[2 * n + i];
        return std::sqrt(s);
    });
    vectors(out, "velocity", f.u, n, dims);
}

// Reactive snapshot: C_A, C_B, reaction_rate = k C_A C_B and the fluid
// velocity, in the order of the writer `title` names.
inline void write_reactive(const std::string& path, const std::string& title, bool velocity_first,
                           const FluidFields& f, const std::vector<double>& ca, const std::vector<double>& cb,
                           double k_react, std::size_t nx, std::size_t ny) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("failed to open " + path);
    const std::size_t n = nx * ny;
    header(out, title, nx, ny, 1);
    if (velocity_first) vectors(out, "velocity", f.u, n, 2);
    scalars(out, "C_A", n, [&](std::size_t i) { return ca[i]; });
    scalars(out, "C_B", n, [&](std::size_t i) { return cb[i]; });
    scalars(out, "reaction_rate", n, [&](std::size_t i) { return k_react * ca[i] * cb[i]; });
    if (!velocity_first) vectors(out, "velocity", f.u, n, 2);
}

}  // namespace lbc_vtk
