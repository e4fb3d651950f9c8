"""oxr_simulate_noalloc (the entry point that allocates no device memory) against
oxr_simulate on the same gensteps, through the C ABI: the same hits bit for bit and the
same photons as the engine's own launch of that event, on two calls in a row, the second
with a smaller event, in both forms; events beyond the scratch limit are refused."""
import ctypes as C

import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def noalloc(L):
    f = L.oxr_simulate_noalloc
    f.restype = C.c_int32
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p,
                  C.c_int64, C.POINTER(C.c_int64), C.POINTER(abi.LaunchStats)]
    return f


class Hip:
    """The caller's device memory: hipMalloc/hipMemcpy/hipFree of the HIP runtime the
    engine library already loaded."""

    def __init__(self):
        try:
            L = C.CDLL('libamdhip64.so')
        except OSError:
            L = C.CDLL('libamdhip64.so.7')
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.L = L

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 256)) == 0
        assert self.L.hipMemset(p, 0xff, max(nbytes, 256)) == 0
        return p

    def get(self, p, shape):
        out = np.zeros(shape, dtype=np.float32)
        if out.nbytes:
            assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def run_noalloc(engine_mod, e, gs, event_index, form, hip):
    n = G.num_photons(gs)
    d_photon, d_hit = hip.alloc(n * 64), hip.alloc(n * 64)
    gs = np.ascontiguousarray(gs, dtype=np.float32)
    nh = C.c_int64(-1)
    st = abi.LaunchStats()
    rc = noalloc(engine_mod.lib())(e.h, abi.ptr(gs), len(gs), event_index, 0, form, d_photon, d_hit, n,
                                   C.byref(nh), C.byref(st))
    assert rc == 0, engine_mod.lib().oxr_last_error()
    ph, hit = hip.get(d_photon, (n, 4, 4)), hip.get(d_hit, (nh.value, 4, 4))
    hip.free(d_photon)
    hip.free(d_hit)
    return ph, hit, st


@pytest.mark.parametrize('form', [abi.OXR_FORM_PERSISTENT, abi.OXR_FORM_PLAIN])
def test_noalloc_equals_simulate(engine_mod, form):
    hip = Hip()
    s = scenes.csi_calorimeter()
    cfg = abi.default_config(**s.config)
    big = s.gensteps(scale_photons=0.2)
    small = s.gensteps(seed=7, scale_photons=0.03)
    assert 0 < G.num_photons(small) < G.num_photons(big) <= abi.OXR_NOALLOC_MAX_PHOTON
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        for k, gs in enumerate((big, small)):
            print('no-alloc call', k, 'form', form, G.num_photons(gs), 'photons', flush=True)
            ph, hit, st = run_noalloc(engine_mod, e, gs, 3 + k, form, hip)
            want_hit, wst = e.simulate(gs, event_index=3 + k, form=form)
            want_ph = e.photons()
            assert st.num_photon == wst['num_photon'] == len(ph)
            assert st.num_hit == wst['num_hit'] == len(hit) > 0
            assert st.bounce_steps == wst['bounce_steps']
            assert np.array_equal(bits(hit), bits(want_hit)), k
            assert np.array_equal(bits(ph), bits(want_ph)), k


def test_noalloc_refuses_beyond_the_scratch_limit(engine_mod):
    hip = Hip()
    s = scenes.csi_calorimeter()
    with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config)) as e:
        gs = G.torch([abi.OXR_NOALLOC_MAX_PHOTON // 2 + 1] * 2, pos=np.zeros((2, 3)), type=G.T_POINT_ISOTROPIC)
        d = hip.alloc(64)
        nh = C.c_int64(-1)
        rc = noalloc(engine_mod.lib())(e.h, abi.ptr(gs), len(gs), 0, 0, abi.OXR_FORM_PERSISTENT, d, d, 1,
                                       C.byref(nh), None)
        hip.free(d)
        assert rc == abi.OXR_ECAPACITY and b'scratch limit' in engine_mod.lib().oxr_last_error()
        # the engine still simulates normally afterwards
        h, st = e.simulate(s.gensteps(scale_photons=0.01))
        assert st['num_photon'] > 0
