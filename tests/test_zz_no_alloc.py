"""tild_fft_c2r_noalloc (the inverse transform on the library's fixed device scratch)
against tild_fft_c2r on the same inputs: bit for bit, on consecutive calls whose second
input is smaller than the first, with no change in free device memory across calls."""
import ctypes

import numpy as np
import pytest

import hip_ctypes as H
import tild

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20


def _free_bytes():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    L = H.hip()
    L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    assert L.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _spectra(shape, batch, seed):
    """Half spectra of random real fields (numpy order (z, y, x)), on the device."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((batch,) + tuple(shape)).astype(np.float32)
    half = tuple(shape[:-1]) + (shape[-1] // 2 + 1,)
    d_in = H.DeviceArray.from_numpy(a)
    d_spec = H.DeviceArray(batch * int(np.prod(half)) * 8)
    tild.fft_r2c(len(shape), list(shape[::-1]), batch, d_in.ptr, d_spec.ptr)
    return a, d_spec


def _both(shape, batch, seed):
    a, d_spec = _spectra(shape, batch, seed)
    dim, Nx = len(shape), list(shape[::-1])
    d_ref = H.DeviceArray(a.nbytes)
    d_got = H.DeviceArray(a.nbytes)
    tild.fft_c2r(dim, Nx, batch, d_spec.ptr, d_ref.ptr)
    tild.fft_c2r_noalloc(dim, Nx, batch, d_spec.ptr, d_got.ptr)
    ref = d_ref.to_numpy(np.float32, a.shape)
    got = d_got.to_numpy(np.float32, a.shape)
    return a, ref, got


@pytest.mark.parametrize("first,second", [
    (((48, 40, 36), 3), ((24, 20, 18), 2)),
    (((160, 80, 80), 1), ((64, 64), 2)),
], ids=["3d-3d", "3d-2d"])
def test_noalloc_c2r_matches_c2r_on_consecutive_calls(first, second):
    for k, (shape, batch) in enumerate((first, second)):
        a, ref, got = _both(shape, batch, seed=10 + k)
        assert np.array_equal(got, ref), (shape, batch)
        back = got / np.prod(shape)
        err = float(np.abs(back - a).max() / np.abs(a).max())
        print(f"noalloc c2r {shape} x{batch}: bit-identical to tild_fft_c2r; c2r(r2c(x))/N vs x {err:.2e}",
              flush=True)
        assert err < 5e-6


def test_noalloc_c2r_allocates_nothing():
    shape, batch = (40, 40, 40), 2
    a, d_spec = _spectra(shape, batch, seed=3)
    d_out = H.DeviceArray(a.nbytes)
    dim, Nx = 3, [40, 40, 40]
    tild.fft_c2r_noalloc(dim, Nx, batch, d_spec.ptr, d_out.ptr)      # first call: scratch address lookup
    before = _free_bytes()
    for _ in range(3):
        tild.fft_c2r_noalloc(dim, Nx, batch, d_spec.ptr, d_out.ptr)
    assert _free_bytes() == before
    back = d_out.to_numpy(np.float32, a.shape) / np.prod(shape)
    assert float(np.abs(back - a).max() / np.abs(a).max()) < 5e-6


def test_noalloc_c2r_rejects_more_than_its_scratch():
    # 256^3, batch 1: 129*256*256 complex values = 67.6 MB of work copy > 32 MiB; the
    # size check runs on the host before any device work
    shape = (256, 256, 256)
    need = (shape[2] // 2 + 1) * shape[1] * shape[0] * 8
    assert need > SCRATCH_BYTES
    d_in = H.DeviceArray(need)
    d_out = H.DeviceArray(int(np.prod(shape)) * 4)
    with pytest.raises(tild.TildError) as ei:
        tild.fft_c2r_noalloc(3, [256, 256, 256], 1, d_in.ptr, d_out.ptr)
    assert ei.value.status == 6 and "scratch" in str(ei.value)
