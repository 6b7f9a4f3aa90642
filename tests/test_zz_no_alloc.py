"""stz_qs_l2_error_noalloc against stz_mg_l2_error through the C ABI.

The allocation-free entry point keeps its partial sums in a fixed device
array; it must give the same mean-square residual as stz_mg_l2_error (same
residual arithmetic and summation order: bit-identical) on the same vel and
src, call after call, including a second call on a smaller grid whose
partials do not cover the first call's, and it must refuse a grid beyond its
fixed scratch without launching anything.  The test (the caller) allocates
its device buffers with the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

from stz_amd import ShearSim, _lib
from stz_amd.problems import simple_shear

pytestmark = pytest.mark.gpu


def hip():
    """the HIP runtime the library is linked against (already loaded)"""
    _lib.load()
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


H2D, D2H = 1, 2


def dev_array(rt, a):
    p = C.c_void_p()
    assert rt.hipMalloc(C.byref(p), a.nbytes) == 0
    assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _case(m, n, seed, viscdt):
    p = simple_shear(m, n)
    rng = np.random.default_rng(seed)
    vel = rng.normal(size=(n + 1, m, 2)) * 1e-3
    src = rng.normal(size=(n + 1, m, 2))
    g = ShearSim(p, "fp64")
    g.mg_setup(viscdt)
    g.set_vel(vel)
    g.mg_set_src(src)
    want = g.mg_l2_error()
    g.close()
    rt = hip()
    dv, ds = dev_array(rt, vel), dev_array(rt, src)
    out = np.full(1, np.nan)
    dout = dev_array(rt, out)
    pp = _lib.make_params(p)
    _lib.check(_lib.load().stz_qs_l2_error_noalloc(C.byref(pp), viscdt, dv, ds, dout, None))
    assert rt.hipDeviceSynchronize() == 0
    assert rt.hipMemcpy(out.ctypes.data, dout, out.nbytes, D2H) == 0
    for q in (dv, ds, dout):
        rt.hipFree(q)
    return want, float(out[0])


def test_noalloc_residual_matches_mg_l2_error_twice():
    print("case noalloc 256x256 then 96x64", flush=True)
    want1, got1 = _case(256, 256, 1, 0.1)
    assert got1 == want1, (got1, want1)
    want2, got2 = _case(96, 64, 2, 0.04)
    assert got2 == want2, (got2, want2)
    assert want1 != want2
    # the first grid again after the smaller one
    want3, got3 = _case(256, 256, 1, 0.1)
    assert got3 == want3 == want1


def test_noalloc_refuses_a_grid_beyond_its_scratch():
    print("case noalloc refusal", flush=True)
    lib = _lib.load()
    pp = _lib.make_params(simple_shear(8192, 2048))  # 8192 * 2049 nodes > 2^24
    dummy = C.c_void_p(1)
    code = lib.stz_qs_l2_error_noalloc(C.byref(pp), 0.1, dummy, dummy, dummy, None)
    assert code == 1, code
    assert b"exceed" in lib.stz_last_error()
