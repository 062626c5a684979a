"""Host-side tables of the product library (CPU): builds a small C++ driver
against spheresimulator_amd/csrc/shtables.cpp and runs it."""
import os
import subprocess

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
