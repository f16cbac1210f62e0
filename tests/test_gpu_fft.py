"""HIP Stockham R2C/C2R against numpy FP64 and against rocFFT on the same device."""
import numpy as np
import pytest

import hip_ctypes as H
import tild

pytestmark = pytest.mark.gpu

SHAPES = [(8, 6, 10), (40, 40, 40), (225, 21, 21), (35, 35, 35), (160, 80, 80), (125, 125), (16, 20),
          (27, 25, 24), (7, 9, 11 - 1), (48, 45, 36), (256, 64, 16), (128, 96, 32)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_c2r(shape):
    rng = np.random.default_rng(1)
    batch = 3
    a = rng.standard_normal((batch,) + shape).astype(np.float32)   # numpy order (z, y, x)
    dim = len(shape)
    Nx = list(shape[::-1])
    half = shape[:-1] + (shape[-1] // 2 + 1,)
    d_in = H.DeviceArray.from_numpy(a)
    d_out = H.DeviceArray(batch * np.prod(half) * 8)
    tild.fft_r2c(dim, Nx, batch, d_in.ptr, d_out.ptr)
    got = d_out.to_numpy(np.complex64, (batch,) + half)
    ref = np.fft.rfftn(a.astype(np.float64), axes=tuple(range(1, dim + 1)))
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max() / scale
    d_roc = H.DeviceArray(batch * np.prod(half) * 8)
    H.rocfft_r2c(d_in, d_roc, shape, batch)
    roc = d_roc.to_numpy(np.complex64, (batch,) + half)
    err_roc = np.abs(roc - ref).max() / scale
    diff = np.abs(got - roc).max() / scale
    print(f"{shape}: r2c vs FP64 {err:.2e}; rocFFT vs FP64 {err_roc:.2e}; engine vs rocFFT {diff:.2e}")
    assert err < 5e-6
    assert diff < 1e-5
    d_back = H.DeviceArray(a.nbytes)
    tild.fft_c2r(dim, Nx, batch, d_out.ptr, d_back.ptr)
    back = d_back.to_numpy(np.float32, a.shape) / np.prod(shape)
    err2 = np.abs(back - a).max() / np.abs(a).max()
    print(f"{shape}: c2r(r2c(x))/N vs x {err2:.2e}")
    assert err2 < 5e-6
