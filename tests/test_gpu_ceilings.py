"""The two roofline ceiling kernels bench.py reports beside the published
peaks: the hand-written streaming copy (lulesh_hip_stream_bandwidth) and
the FP64 multiply-add issue kernel (lulesh_hip_fp64_fma_rate).  Both must
return a rate between a plausible floor and the MI355X peak (8 TB/s HBM,
78.6 TFLOP/s FP64 vector) and refuse bad arguments."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_stream_copy_rate(pyl):
    g = ctypes.c_double()
    assert pyl.lib().lulesh_hip_stream_bandwidth(ctypes.c_int64(1 << 28), 5,
                                                 ctypes.byref(g)) == 0
    print(f"\n[ceilings] stream copy {g.value:.0f} GB/s", flush=True)
    assert 1000.0 < g.value < 8000.0
    assert pyl.lib().lulesh_hip_stream_bandwidth(ctypes.c_int64(1000), 5, ctypes.byref(g)) != 0
    assert pyl.lib().lulesh_hip_stream_bandwidth(ctypes.c_int64(1 << 28), 0, ctypes.byref(g)) != 0


def test_fp64_fma_rate(pyl):
    t = ctypes.c_double()
    assert pyl.lib().lulesh_hip_fp64_fma_rate(3, ctypes.byref(t)) == 0
    print(f"\n[ceilings] FP64 multiply-add {t.value:.1f} TFLOP/s", flush=True)
    assert 10.0 < t.value <= 78.6 * 1.02
    assert pyl.lib().lulesh_hip_fp64_fma_rate(0, ctypes.byref(t)) != 0
    assert pyl.lib().lulesh_hip_fp64_fma_rate(3, None) != 0
