"""include/lattice_hip.hpp is a drop-in for the reference's lattice_cuda.cuh.

A reference driver includes LB-Cube's own include/lattice_traits.hpp (its
lattice types and IsLatticeModel concept) before the launcher header
(src/main.cpp:25-36).  These CPU tests compile, with g++ and with ROCm's
clang++, a translation unit that does exactly that and then instantiates the
reference's 36 explicit launcher instantiations (src/lattice_cuda.cu:756-1029:
24 fluid, 6 scalar, 6 reaction) and both LatticeMemory overloads
(include/lattice_cuda.cuh:243-255, :310-319) with a LatticeMemory-shaped class
template constrained like the reference's, the lattice and precision given
explicitly and deduced.  A traits type in another direction order or with
another cs2 must be rejected at compile time.  The tables of
include/lbcube_lattices.hpp (the traits the repository's own drivers use) must
equal the reference's.

The reference headers are read where they lie and never copied; the tests
skip when the reference tree is absent.
"""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/include"
REF_TRAITS = os.path.join(REF_INCLUDE, "lattice_traits.hpp")
COMPILERS = ["g++", "/opt/rocm/llvm/bin/clang++"]

pytestmark = pytest.mark.skipif(not os.path.exists(REF_TRAITS), reason="the reference's headers are not present")

FLUIDS = ["D2Q9", "D3Q19", "D3Q27"]
PAIRS = [("D2Q9", "D2Q5"), ("D3Q19", "D3Q7"), ("D3Q27", "D3Q7")]
REALS = ["float", "double"]
COLLS = ["BGK", "TRT", "MRT", "RLBM"]

STUB = """
namespace stub {
// Shaped like the reference's LatticeMemory (include/lattice_memory.hpp:67-140):
// a class template over a constrained lattice and a floating-point Real that
// exposes its extents; the launchers read nothing else.
template <lbm::IsLatticeModel Lattice, std::floating_point Real> class LatticeMemory {
public:
    LatticeMemory(std::size_t x, std::size_t y, std::size_t z) : x_(x), y_(y), z_(z) {}
    [[nodiscard]] std::size_t x_extent() const noexcept { return x_; }
    [[nodiscard]] std::size_t y_extent() const noexcept { return y_; }
    [[nodiscard]] std::size_t z_extent() const noexcept { return z_; }
private:
    std::size_t x_, y_, z_;
};
}  // namespace stub
"""


def compile_tu(tmp_path, source, compiler, name="tu"):
    src = tmp_path / f"{name}.cpp"
    src.write_text(source)
    return subprocess.run([compiler, "-std=c++23", "-c", "-o", str(tmp_path / f"{name}.o"), f"-I{REF_INCLUDE}",
                           f"-I{os.path.join(ROOT, 'include')}", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                           str(src)], capture_output=True, text=True, timeout=180)


def instantiations():
    lines = []
    for lat in FLUIDS:
        for real in REALS:
            for ct in COLLS:
                lines.append(f"template hipError_t lbm::launch_step_gpu<lbm::{lat}, {real}, lbm::CollisionType::{ct}>("
                             f"const {real}*, {real}*, std::size_t, std::size_t, std::size_t, {real}, dim3, "
                             f"hipStream_t);")
    for fl, sl in PAIRS:
        for real in REALS:
            lines.append(f"template hipError_t lbm::launch_scalar_step_gpu<lbm::{fl}, lbm::{sl}, {real}>("
                         f"const {real}*, {real}*, const {real}*, std::size_t, std::size_t, std::size_t, {real}, "
                         f"{real}, dim3, hipStream_t);")
            lines.append(f"template hipError_t lbm::launch_reaction_AB_gpu<lbm::{fl}, lbm::{sl}, {real}>("
                         f"const {real}*, {real}*, const {real}*, {real}*, const {real}*, std::size_t, std::size_t, "
                         f"std::size_t, {real}, {real}, dim3, hipStream_t);")
    assert len(lines) == 36
    return "\n".join(lines)


@pytest.mark.parametrize("compiler", COMPILERS)
def test_reference_traits_then_launcher_header(tmp_path, compiler):
    uses = []
    for fl, sl in PAIRS:
        for real in REALS:
            one = "1.0f" if real == "float" else "1.0"
            uses.append(f"""
    {{
        stub::LatticeMemory<lbm::{fl}, {real}> fm(8, 6, {1 if fl == 'D2Q9' else 4});
        stub::LatticeMemory<lbm::{sl}, {real}> sm(8, 6, {1 if fl == 'D2Q9' else 4});
        const {real}* c = nullptr;
        {real}* n = nullptr;
        status += lbm::launch_step_gpu<lbm::{fl}, {real}, lbm::CollisionType::TRT>(fm, c, n, {one});
        status += lbm::launch_step_gpu<lbm::{fl}, {real}>(fm, c, n, {one}, dim3{{128, 1, 1}});
        status += lbm::launch_step_gpu(fm, c, n, {one});
        status += lbm::launch_scalar_step_gpu<lbm::{fl}, lbm::{sl}, {real}>(sm, fm, c, n, c, {one});
        status += lbm::launch_scalar_step_gpu(sm, fm, c, n, c, {one}, {real}{{0}}, dim3{{64, 2, 1}});
    }}""")
    tu = f"""
#include "lattice_traits.hpp"   // the reference's lattice types, first, as its drivers include them
// What the reference's lattice_core.hpp:25-43 declares (that header needs Eigen):
// CollisionType behind the guard macro both headers use.
#ifndef LB_CUBE_COLLISION_TYPE_DEFINED
#define LB_CUBE_COLLISION_TYPE_DEFINED
namespace lbm {{
enum class CollisionType {{ BGK, TRT, MRT, RLBM }};
}}
#endif
#include "lattice_hip.hpp"

#include <concepts>
#include <cstddef>

static_assert(lbm::HipLattice<lbm::D2Q9> && lbm::HipLattice<lbm::D3Q19> && lbm::HipLattice<lbm::D3Q27> &&
              lbm::HipLattice<lbm::D2Q5> && lbm::HipLattice<lbm::D3Q7>);
static_assert(lbm::lattice_code<lbm::D2Q9>() == LBC_D2Q9 && lbm::lattice_code<lbm::D3Q19>() == LBC_D3Q19 &&
              lbm::lattice_code<lbm::D3Q27>() == LBC_D3Q27 && lbm::lattice_code<lbm::D2Q5>() == LBC_D2Q5 &&
              lbm::lattice_code<lbm::D3Q7>() == LBC_D3Q7);
{STUB}
{instantiations()}

int use_memory_overloads() {{
    int status = 0;
{''.join(uses)}
    return status;
}}
"""
    r = compile_tu(tmp_path, tu, compiler)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.parametrize("compiler", COMPILERS)
def test_foreign_tables_are_rejected(tmp_path, compiler):
    tu = """
#include "lattice_traits.hpp"
#include "lattice_hip.hpp"

#include <array>

// D3Q19 with two directions swapped: planes 1 and 2 would be misread.
struct SwappedD3Q19 {
    static constexpr int D = 3;
    static constexpr int Q = 19;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr std::array<int, 57> c = [] {
        auto t = lbm::D3Q19::c;
        for (int d = 0; d < 3; ++d) { const int s = t[3 + d]; t[3 + d] = t[6 + d]; t[6 + d] = s; }
        return t;
    }();
};
// D3Q7 tables with the fluid lattices' cs2.
struct D3Q7WrongCs2 {
    static constexpr int D = 3;
    static constexpr int Q = 7;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr auto c = lbm::D3Q7::c;
};
struct D2Q7 { static constexpr int D = 2; static constexpr int Q = 7; static constexpr double cs2 = 1.0 / 3.0; };
static_assert(!lbm::HipLattice<SwappedD3Q19>);
static_assert(!lbm::HipLattice<D3Q7WrongCs2>);
static_assert(!lbm::HipLattice<D2Q7>);
template <class L> concept Launchable = requires(const float* c, float* n) {
    lbm::launch_step_gpu<L, float>(c, n, 4, 4, 4, 1.0f);
};
static_assert(Launchable<lbm::D3Q19> && !Launchable<SwappedD3Q19>);
"""
    r = compile_tu(tmp_path, tu, compiler)
    assert r.returncode == 0, r.stderr[-4000:]


DUMP = """
#include <cstdio>
int main() {
%s
}
"""


def dump_program(header):
    body = []
    for lat in FLUIDS + ["D2Q5", "D3Q7"]:
        body.append(f'    std::printf("{lat} %d %d %.17g\\n", lbm::{lat}::D, lbm::{lat}::Q, lbm::{lat}::cs2);')
        body.append(f'    for (double w : lbm::{lat}::weights) std::printf("%.17g ", w);')
        body.append(f'    for (int v : lbm::{lat}::directions) std::printf("%d ", v);')
        body.append(f'    for (int v : lbm::{lat}::c) std::printf("%d ", v);')
        body.append(f'    for (int v : lbm::{lat}::opposite) std::printf("%d ", v);')
        body.append('    std::printf("\\n");')
    return f'#include "{header}"\n' + DUMP % "\n".join(body)


def test_lbcube_lattices_equal_the_reference_tables(tmp_path):
    outs = []
    for header in ("lattice_traits.hpp", "lbcube_lattices.hpp"):
        src = tmp_path / f"dump_{header.split('.')[0]}.cpp"
        exe = tmp_path / f"dump_{header.split('.')[0]}"
        src.write_text(dump_program(header))
        r = subprocess.run(["g++", "-std=c++23", "-o", str(exe), f"-I{REF_INCLUDE}", f"-I{os.path.join(ROOT, 'include')}",
                            "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", str(src)], capture_output=True, text=True,
                           timeout=180)
        assert r.returncode == 0, r.stderr[-3000:]
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, check=True).stdout)
    assert outs[0] == outs[1]
