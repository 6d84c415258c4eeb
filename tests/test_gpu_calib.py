"""GPU test of the measured ceilings (calib.hip): the MFMA loop, the
streaming copy and the LDS read loop return rates within the hardware's
plausible range (wide bounds: they are measurements, not known answers) and
the shader clock they ran at."""
import pytest

from aces4_amd import sipgpu as sg

pytestmark = pytest.mark.gpu


def test_measured_ceilings_are_plausible():
    mfma, copy = sg.measure_peaks(1 << 28)
    print("mfma TFLOP/s", mfma, "copy GB/s", copy, "clock MHz", sg.measured_clock_mhz(), flush=True)
    assert 10.0 < mfma <= 80.0      # FP64 matrix peak 78.6 TFLOP/s
    assert 500.0 < copy <= 8000.0   # HBM3E 8 TB/s
    assert 500.0 < sg.measured_clock_mhz() <= 2500.0
    lds = sg.measure_lds()
    print("LDS read B/clock/CU", lds, flush=True)
    assert 32.0 < lds <= 512.0
