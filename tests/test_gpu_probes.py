"""The FP32 FMA peak probe (bench's measured peak) computes what it claims:
each thread's 8 chains x <- x * 0.999 + 0.001 over iters x 16 steps, summed
(checked against an FP64 evaluation of the same recurrence)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fma_peak_probe_values():
    from bp_project_amd import _lib
    from bp_project_amd.device import DeviceArray, synchronize
    blocks, iters = 3, 5
    out = DeviceArray.empty((blocks * 256,), np.float32)
    _lib.call("bpmi_probe_fma_peak", _lib.c_void_p(out.ptr), blocks, iters, None)
    synchronize()
    got = out.get()
    t = np.arange(256, dtype=np.float64)
    x = (t[:, None] + np.arange(8)[None, :]) * np.float64(np.float32(1e-3))
    m, c = np.float64(np.float32(0.999)), np.float64(np.float32(1e-3))
    for _ in range(iters * 16):
        x = x * m + c
    want = np.tile(x.sum(axis=1), blocks)
    assert np.allclose(got, want, rtol=2e-6, atol=0), np.abs(got - want).max()
    with pytest.raises(ValueError):
        _lib.call("bpmi_probe_fma_peak", _lib.c_void_p(out.ptr), 0, iters, None)
