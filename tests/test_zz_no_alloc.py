"""The no-allocation entry point (mg_render_device_noalloc) through the C
ABI, against mg_render_device (the entry point it is a variant of), byte for
byte: two calls in a row, the second on a smaller frame, a composition whose
lines are resident and one whose lines are copied into the static scratch,
no change of free device memory across calls after the first, and a frame
beyond the scratch refused with MG_ERR_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

import scenegen

pytestmark = pytest.mark.gpu

MG_ERR_CAPACITY = 9


@pytest.fixture(scope="module")
def gpu():
    import cassia_amd
    assert cassia_amd.device_count() > 0, "no gfx950 device: the HIP path cannot run"
    lib = cassia_amd._lib
    lib.mg_render_device_noalloc.restype = C.c_int
    lib.mg_render_device_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_float),
                                             C.c_void_p, C.c_void_p, C.c_size_t]
    return cassia_amd


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def device_buffer(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    assert hip.hipMemset(p, 0, nbytes) == 0
    return p


def free_memory(hip):
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def noalloc(gpu, comp, w, h, out, nbytes, bg=(1.0, 1.0, 1.0, 1.0)):
    b = (C.c_float * 4)(*bg)
    return gpu._lib.mg_render_device_noalloc(comp._h, w, h, b, None, out, nbytes)


def reference(gpu, comp, w, h, bg=(1.0, 1.0, 1.0, 1.0)):
    return gpu.device_frame(comp.render_device(w, h, list(bg)), w, h)


def test_noalloc_matches_render_device(gpu, hip):
    # A: lines resident (rendered before); B: a smaller frame rendered right
    # after A; C: a composition never rendered (lines copied into the scratch)
    wa, ha, wb, hb = 150, 96, 70, 40
    a = scenegen.build_scene(gpu, 801, wa, ha, 60, transforms=True, curves=True, clips=True)
    b = scenegen.build_scene(gpu, 802, wb, hb, 20)
    ref_a, ref_b = reference(gpu, a, wa, ha), reference(gpu, b, wb, hb)
    ref_c = reference(gpu, scenegen.build_scene(gpu, 803, wa, ha, 40, gradients=True), wa, ha)
    c = scenegen.build_scene(gpu, 803, wa, ha, 40, gradients=True)
    out_a, out_b = device_buffer(hip, wa * ha * 4), device_buffer(hip, wb * hb * 4)
    print("noalloc: first call (sets up the engine)", flush=True)
    assert noalloc(gpu, a, wa, ha, out_a, wa * ha * 4) == 0, gpu._lib.mg_last_error()
    before = free_memory(hip)
    assert noalloc(gpu, b, wb, hb, out_b, wb * hb * 4) == 0, gpu._lib.mg_last_error()
    assert np.array_equal(gpu.device_frame(out_b.value, wb, hb), ref_b)
    assert np.array_equal(gpu.device_frame(out_a.value, wa, ha), ref_a)
    assert noalloc(gpu, c, wa, ha, out_a, wa * ha * 4) == 0, gpu._lib.mg_last_error()
    assert np.array_equal(gpu.device_frame(out_a.value, wa, ha), ref_c)
    assert noalloc(gpu, a, wa, ha, out_a, wa * ha * 4) == 0, gpu._lib.mg_last_error()
    assert np.array_equal(gpu.device_frame(out_a.value, wa, ha), ref_a)
    assert free_memory(hip) == before
    hip.hipFree(out_a)
    hip.hipFree(out_b)


def test_noalloc_refuses_frames_beyond_the_scratch(gpu, hip):
    from cassia_amd import workloads
    big = workloads.polygons(count=20_000).build(gpu)
    w, h = 3840, 2160
    out = device_buffer(hip, w * h * 4)
    st = noalloc(gpu, big, w, h, out, w * h * 4)
    msg = gpu._lib.mg_last_error().decode()
    assert st == MG_ERR_CAPACITY, (st, msg)
    assert "no-allocation" in msg and "MiB" in msg, msg
    # a short output buffer is refused before any work
    assert noalloc(gpu, big, w, h, out, 16) != 0
    # the engine still renders after the refusal
    small = scenegen.build_scene(gpu, 804, 64, 48, 12)
    ref = reference(gpu, small, 64, 48)
    assert noalloc(gpu, small, 64, 48, out, w * h * 4) == 0, gpu._lib.mg_last_error()
    assert np.array_equal(gpu.device_frame(out.value, 64, 48), ref)
    hip.hipFree(out)
