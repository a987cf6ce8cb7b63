"""cdft_system_hessian_dot_noalloc (the allocation-free variant of cdft_system_hessian_dot, work spectra in the
library's fixed per-GPU scratch array) against cdft_system_hessian_dot: two calls in a row, the second on a smaller
system, both through the C ABI on device pointers; and the size limit."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _case(cdft, shape, seed, model, mf):
    dx = 0.1
    u = oracle.splitmix64_uniform(int(np.prod(shape)), seed).reshape(shape)
    rho = 0.5 * (1 + 0.2 * (2 * u - 1))
    v = np.random.default_rng(seed).standard_normal(shape)
    s = cdft.System(shape, dx, 1.0, model=model, ideal=True, mean_field=mf, kT=1.5)
    lib = cdft.cdft.lib()
    r, vv = cdft.DeviceArray.from_host(rho), cdft.DeviceArray.from_host(v)
    o1, o2 = cdft.DeviceArray(shape), cdft.DeviceArray(shape)
    P = ctypes.c_void_p
    rc1 = lib.cdft_system_hessian_dot(s.h, P(r.ptr), P(vv.ptr), P(o1.ptr), cdft.DEVICE_PTRS)
    rc2 = lib.cdft_system_hessian_dot_noalloc(s.h, P(r.ptr), P(vv.ptr), P(o2.ptr), cdft.DEVICE_PTRS)
    assert rc1 == 0 and rc2 == 0, lib.cdft_last_error()
    a, b = o1.download(), o2.download()
    for d in (r, vv, o1, o2):
        d.free()
    s.close()
    return a, b


def test_noalloc_hessian_matches_the_allocating_entry_point(cdft):
    # first a 48^3 White Bear + mean-field system, then a smaller 24 x 16 x 32 Rosenfeld one: the scratch keeps the
    # first call's contents and must not leak into the second
    for shape, seed, model, mf in [((48, 48, 48), 11, cdft.WHITEBEAR_I, True), ((24, 16, 32), 12, cdft.ROSENFELD, False)]:
        a, b = _case(cdft, shape, seed, model, mf)
        print(shape, "max |noalloc - hessian_dot| / max |hessian_dot| =", np.abs(a - b).max() / np.abs(a).max())
        assert np.abs(a - b).max() <= 1e-14 * np.abs(a).max()


def test_noalloc_hessian_rejects_oversize_and_host_pointers(cdft):
    lib = cdft.cdft.lib()
    P = ctypes.c_void_p
    s = cdft.System((128, 128, 128), 0.05, 1.0, model=cdft.WHITEBEAR_I, ideal=True)
    r = cdft.DeviceArray((128, 128, 128))
    assert lib.cdft_system_hessian_dot_noalloc(s.h, P(r.ptr), P(r.ptr), P(r.ptr), cdft.DEVICE_PTRS) == cdft.cdft.CDFT_BAD_ARG
    assert b"bytes of work memory" in lib.cdft_last_error()
    r.free()
    s.close()
    t = cdft.System((16, 16, 16), 0.1, 1.0, model=cdft.WHITEBEAR_I, ideal=True)
    h = np.ones((16, 16, 16))
    assert lib.cdft_system_hessian_dot_noalloc(t.h, h.ctypes.data_as(P), h.ctypes.data_as(P), h.ctypes.data_as(P),
                                               0) == cdft.cdft.CDFT_BAD_ARG
    t.close()
