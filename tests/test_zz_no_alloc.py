"""The allocation-free entry point hsmd_mc_config_energies_noalloc against
hsmd_energies (host buffers, FP64): two calls in a row on caller-owned
device buffers, the second with fewer configurations, and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

import classical_mc_and_smd_for_2d_heisenberg_models_amd as H

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-noalloc]", *a, flush=True)
    sys.stdout.flush()


class _Hip:
    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0   # host to device

    def get(self, a, src):
        assert self.h.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0   # device to host

    def free(self, p):
        assert self.h.hipFree(p) == 0


def _noalloc(mc, n, d_spins, d_E):
    mc.config_energies_noalloc(n, d_spins.value, d_E.value)


def _unit(rng, shape):
    v = rng.normal(size=shape)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


@pytest.mark.parametrize("geom,L,kw", [
    ("triang", 24, dict(J1=1.0, J2=0.12, JX=0.05, h=0.3, lambda1=0.9, lambda2=0.9)),
    ("square", 64, dict(J1=1.0, J2=0.6, imp_ratio=0.2, J1ic=0.5, J1ii=0.3, J2ic=0.1, J2ii=0.05)),
])
def test_noalloc_energies_match_hsmd_energies_on_two_calls(geom, L, kw):
    _say("noalloc energies", geom, L)
    c = H.Couplings(**kw)
    lat = H.Lattice.generate(geom, L, c)
    if kw.get("imp_ratio"):
        rng0 = np.random.default_rng(3)
        lat.set_impurities((rng0.random(lat.N) < 0.2).astype(np.int32))
    mc = H.MonteCarlo(lat, np.array([0.5, 1.0]), precision=H.BATCHED_F64, seed=2)
    hip = _Hip()
    rng = np.random.default_rng(L)
    R1, R2 = 37, 5
    d_s = hip.alloc(R1 * lat.N * 3 * 8)
    d_E = hip.alloc(R1 * 8)
    try:
        for R in (R1, R2):
            s = _unit(rng, (R, lat.N, 3))
            hip.put(d_s, s)
            # the output buffer holds the previous call's values beyond R
            _noalloc(mc, R, d_s, d_E)
            assert hip.h.hipDeviceSynchronize() == 0
            got = np.zeros(R)
            hip.get(got, d_E)
            want = lat.energies(s, H.BATCHED_F64)
            _say("R=%d max rel err %.2e" % (R, np.abs(got / want - 1).max()))
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-10)
    finally:
        hip.free(d_s)
        hip.free(d_E)


def test_noalloc_rejects_sizes_beyond_the_scratch_buffer():
    _say("noalloc limit")
    lat = H.Lattice.generate("square", 64, H.Couplings(J1=1.0, J2=0.6))   # 16384 bonds: 4 blocks
    mc = H.MonteCarlo(lat, np.array([1.0]), precision=H.BATCHED_F64, seed=2)
    # 4 partials per configuration: 65537 configurations need 262148 > 2^18
    with pytest.raises(H.HSMDError) as e:
        _noalloc(mc, 65537, ctypes.c_void_p(16), ctypes.c_void_p(16))
    assert "scratch" in str(e.value)
    with pytest.raises(H.HSMDError):
        _noalloc(mc, 0, ctypes.c_void_p(16), ctypes.c_void_p(16))
