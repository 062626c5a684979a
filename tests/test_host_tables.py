"""Host-side tables of the product library (CPU): builds a small C++ driver
against spheresimulator_amd/csrc/shtables.cpp and runs it."""
import os
import subprocess

import pytest

from conftest import ROOT


def test_resample_tables_exact_on_band_limited_fields(tmp_path):
    csrc = os.path.join(ROOT, "spheresimulator_amd", "csrc")
    exe = str(tmp_path / "test_host_tables")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-I" + csrc,
                    os.path.join(ROOT, "tests", "cpp", "test_host_tables.cpp"), os.path.join(csrc, "shtables.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout


@pytest.mark.parametrize("K", [200, 32])
def test_equivalent_source_template_reproduces_far_field(tmp_path, K):
    """equiv.cpp: P is a left inverse of the unit cell's Stokeslet matrix and the
    fitted point forces reproduce the velocity and traction of 16 clustered
    sphere grids with white-noise forces: the FP32 operator's 200-point
    template at 4 cell radii (velocity < 1e-6, traction < 5e-6, total force to
    1e-8), the coarse operator's 32-point one on its check sphere at 2.5 cell
    radii (velocity < 5e-3, traction < 4e-2, total force to 2e-4)."""
    csrc = os.path.join(ROOT, "spheresimulator_amd", "csrc")
    exe = str(tmp_path / "test_equiv_tables")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-I" + csrc,
                    os.path.join(ROOT, "tests", "cpp", "test_equiv_tables.cpp"), os.path.join(csrc, "equiv.cpp"),
                    os.path.join(csrc, "shtables.cpp"), "-o", exe, "-lpthread"], check=True)
    out = subprocess.run([exe, str(K)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
