"""The FP64 rates the benchmark's roofline reports beside the stage kernel
(wasl_hip_fp64_ceilings): both units measured, both below the device's FP64
peak (78.6 TFLOP/s on MI355X, vector and matrix alike)."""

import pytest

pytestmark = pytest.mark.gpu


def test_fp64_ceilings_are_measured_and_below_peak(gpu):
    valu, mfma = gpu.fp64_ceilings()
    print(f"FP64 vector ALU {valu:.1f} TFLOP/s, matrix core {mfma:.1f} TFLOP/s")
    for rate in (valu, mfma):
        assert 1.0 < rate <= 78.6 * 1.05
