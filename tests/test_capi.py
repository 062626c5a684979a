"""C ABI checks that run without a GPU: the library loads, exports exactly what
include/sphstk.h declares, and fails loudly (no CPU fallback) without a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import spheresimulator_amd as S
from spheresimulator_amd import _lib


def header_functions():
    txt = open(os.path.join(ROOT, "include", "sphstk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sphstk_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (sphstk_[a-z0-9_]+)", out))
    declared = header_functions()
    assert declared, "no declarations found"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert sorted(declared) == sorted(_lib.EXPORTED)


def test_library_loads_and_reports_version():
    L = S.load()
    assert L.sphstk_version().startswith(b"sphstk")
    assert L.sphstk_device_count() >= 0


@pytest.mark.skipif(S.load().sphstk_device_count() > 0, reason="checks the no-device error path")
def test_no_device_fails_loudly():
    with pytest.raises(S.SphstkError, match="no HIP device"):
        S.SphereSTKSHOperator([[0, 0, 0]], [1.0], [[1, 0, 0, 0]], [4], 0.5, 0, 1, 1)


def test_null_handle_errors():
    L = S.load()
    n = ctypes.c_int64()
    rc = L.sphstk_operator_sizes(None, ctypes.byref(n), None, None, None)
    assert rc == -1
    assert b"invalid operator handle" in L.sphstk_last_error()


def test_invalid_inputs_rejected_before_device_use():
    # validation happens before any device work: wrong sizes are caught in Python
    with pytest.raises(ValueError):
        S.SphereSTKSHOperator(np.zeros((2, 3)), [1.0], None, [4])


def test_header_and_library_from_c(tmp_path):
    """include/sphstk.h is plain C: a C99 program includes it, links
    libsphstk.so and runs the entry points that need no device."""
    exe = tmp_path / "abi_from_c"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "abi_from_c.c"), "-o", str(exe), "-L", libdir, "-lsphstk",
                    "-Wl,-rpath," + libdir, "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
