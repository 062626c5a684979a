"""Host logic of bench.py (CPU): the roofline object is the pair-sum kernel
with the most kernel time per timed solve, the other is returned beside it."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def roofline(st, peak, name):
    return bench.far_roofline(st, peak, name, 7)


def test_roofline_is_the_kernel_with_more_time_per_solve():
    st32 = {"ms": 20 * 2600.0, "launches": 20 * 76, "pairs": 6.1e10}
    st64 = {"ms": 20 * 2880.0, "launches": 20, "pairs": 3.3e12}
    f32, f64 = roofline(st32, bench.FP32_VALU_PEAK, "f32"), roofline(st64, bench.FP64_VALU_PEAK, "f64")
    main, other = bench.dominant(f32, f64, st32, st64, 20)
    assert main is f64 and other is f32
    assert abs(main["ms_per_solve"] - 2880.0) < 1e-9 and abs(other["ms_per_solve"] - 2600.0) < 1e-9
    # achieved = 30 flop per pair over the apply time (ms / launches)
    assert abs(f64["achieved"] - 30 * 3.3e12 / 2.880 / 1e12) < 1e-9
    st64["ms"] = 20 * 1000.0
    f64 = roofline(st64, bench.FP64_VALU_PEAK, "f64")
    main, other = bench.dominant(f32, f64, st32, st64, 20)
    assert main is f32 and other is f64


def test_roofline_without_fp64_launches():
    st32 = {"ms": 100.0, "launches": 7, "pairs": 1e9}
    f32 = roofline(st32, bench.FP32_VALU_PEAK, "f32")
    assert roofline({"ms": 0.0, "launches": 0, "pairs": 0.0}, bench.FP64_VALU_PEAK, "f64") is None
    main, other = bench.dominant(f32, None, st32, None, 1)
    assert main is f32 and other is None
