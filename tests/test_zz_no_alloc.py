"""The denoiser's entry point that allocates nothing
(kdn_denoise_device_static) against kdn_denoise_device, through the C ABI
on device buffers: the same bits on two calls in a row, the second with a
smaller frame (the static working memory keeps the first call's contents),
and frames beyond its 2^20 pixels refused."""
import ctypes as C

import numpy as np
import pytest

import kodanuki_amd as k
from tests import denoise_fields as F
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    """Device memory for the test's inputs and outputs: the HIP runtime the
    library itself runs on (on the CPU build, host memory)."""

    def __init__(self):
        self.keep = []
        self.hip = None if EMULATION else C.CDLL("libamdhip64.so.7")

    def put(self, array):
        array = np.ascontiguousarray(array)
        if self.hip is None:
            copy = array.copy()
            self.keep.append(copy)
            return copy.ctypes.data
        ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(ptr), C.c_size_t(max(array.nbytes, 1))) == 0
        self.keep.append(ptr)
        assert self.hip.hipMemcpy(ptr, C.c_void_p(array.ctypes.data), C.c_size_t(array.nbytes), 1) == 0  # to device
        return ptr.value

    def empty(self, nbytes):
        return self.put(np.zeros(nbytes, np.uint8))

    def get(self, ptr, shape, dtype):
        out = np.zeros(shape, dtype)
        if self.hip is None:
            C.memmove(out.ctypes.data, ptr, out.nbytes)
        else:
            assert self.hip.hipDeviceSynchronize() == 0
            assert self.hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0  # to host
        return out

    def free(self):
        if self.hip is not None:
            for p in self.keep:
                self.hip.hipFree(p)
        self.keep = []


def _both(device, bufs, px, demod=False):
    h, w = px.shape
    pix = bufs.put(np.ascontiguousarray(px, k.DENOISE_PIXEL_DTYPE))
    out = []
    for static in (False, True):
        rgb = bufs.empty(w * h * 3 * 4)
        device.denoise_device(w, h, pix, rgb, F.FORWARD, F.RIGHT, F.UP, F.REACH, demodulated=demod, static=static)
        out.append(bufs.get(rgb, (h, w, 3), np.float32))
    return out


def test_static_entry_point_equals_the_allocating_one(device):
    bufs = DeviceBuffers()
    try:
        # the first call on a larger frame, then a smaller one: the static
        # working memory holds the first call's fields when the second runs
        for px, demod in ((F.noisy_plane(160, 120), False), (F.checker_field(True), True), (F.noisy_plane(37, 23), False)):
            a, b = _both(device, bufs, px, demod)
            print(f"[no-alloc] {px.shape[1]}x{px.shape[0]}: equal bits {np.mean(a.view(np.uint32) == b.view(np.uint32)):.6f}",
                  flush=True)
            assert np.all(np.isfinite(a))
            assert a.tobytes() == b.tobytes()
    finally:
        bufs.free()


def test_static_entry_point_refuses_frames_beyond_its_memory(device):
    f = k.DenoiseFrame(1025, 1024, (C.c_double * 3)(*F.FORWARD), (C.c_double * 3)(*F.RIGHT), (C.c_double * 3)(*F.UP),
                       F.REACH, 1025 / 1024, 0.5, 0.5, 0, 0)
    # refused by the host-side size check before any device work
    rc = k.library().kdn_denoise_device_static(device._h, C.byref(f), C.c_void_p(16), C.c_void_p(16), None)
    assert rc == 1  # KDN_E_ARGUMENT
