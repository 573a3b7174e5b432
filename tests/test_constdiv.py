"""The correctly rounded constant divisions of lb-cube_amd/csrc/lbc_constdiv.hpp
(used by the reference-order D3Q27 MRT inverse in place of IEEE divisions by
the norms 3 and 6, lattice_mrt.hpp:1173-1205) equal the hardware's IEEE
division: every finite float for both divisors, and sampled doubles over every
exponent, the subnormals and the largest finite values
(tests/constdiv_check.cpp runs the header's own code on the host)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_constant_division_is_the_ieee_division(tmp_path):
    exe = tmp_path / "constdiv_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++23", "--offload-arch=gfx950", "-ffp-contract=off",
                    "-x", "hip", os.path.join(HERE, "constdiv_check.cpp"), "-o", str(exe), "-pthread"], check=True,
                   capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    floats, doubles = (int(x) for x in r.stdout.split()[1:3])
    assert floats == 2 * ((1 << 32) - 2 * (1 << 23))  # every finite float, divisors 3 and 6
    assert doubles > 10 ** 8
