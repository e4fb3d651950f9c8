"""CPU checks of the C ABI: the library loads and exports every declared symbol, and the
host-only entry points behave (no GPU calls)."""
import ctypes as C
import os
import re

from oxr import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, 'include', 'oxr.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(oxr_[a-z0-9_]+)\s*\(', text)))


def test_every_declared_symbol_is_exported(engine_mod):
    L = engine_mod.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_sizes_match_header(engine_mod):
    # sizes implied by include/oxr.h on LP64
    assert C.sizeof(abi.Geometry) == 96
    assert C.sizeof(abi.Config) == 120
    assert C.sizeof(abi.LaunchStats) == 88
    assert C.sizeof(abi.Tuning) == 40
    assert C.sizeof(abi.RequestResult) == 64
    assert C.sizeof(abi.ServerConfig) == 32 and abi.ServerConfig.headroom.offset == 28
    assert C.sizeof(abi.EngineInfo) == 192


def test_config_defaults_are_reference_defaults(engine_mod):
    """sysrap/SEventConfig.cc:45-106: MaxBounce 31, PropagateEpsilon(0) 0.05,
    Epsilon0Mask TO,CK,SI,SC,RE, EventSkipahead 100000, HitMask SD, MaxTime 1e27."""
    c = abi.Config()
    engine_mod.lib().oxr_config_defaults(C.byref(c))
    F = abi.FLAGS
    assert c.max_bounce == 31
    assert abs(c.tmin - 0.05) < 1e-7 and abs(c.tmin0 - 0.05) < 1e-7
    assert c.eps0_mask == F['TORCH'] | F['CERENKOV'] | F['SCINTILLATION'] | F['BULK_SCATTER'] | F['BULK_REEMIT']
    assert c.skipahead_event == 100000
    assert c.hitmask == F['SURFACE_DETECT']
    assert c.max_time > 1e26
    # rec domains (SEventConfig.cc:72-73): 1000 mm, 10 ns; rec, prd and records off
    assert c.extent_domain == 1000.0 and c.time_domain == 10.0 and c.max_rec == 0 and c.max_prd == 0
    # PropagateRefine off, PropagateRefineDistance 5000 mm (SEventConfig.cc:105-106); the
    # ray visibility mask sees every instance (Params.cc:156)
    assert c.propagate_refine == 0 and c.refine_distance == 5000.0 and c.vizmask == 0xff
    py = abi.default_config()
    for k, _ in abi.Config._fields_:
        assert getattr(py, k) == getattr(c, k), k


def test_version_and_error_string(engine_mod):
    L = engine_mod.lib()
    assert L.oxr_version().decode().startswith('oxr')
    # a null engine is rejected on the host with a message, before any device call
    rc = L.oxr_engine_set_config(None, None)
    assert rc == -1 and b'null' in L.oxr_last_error()


def test_python_constants_match_header():
    """Every OXR_* enum value the header assigns equals the Python binding's constant of
    the same name (gather kinds, forms, RNG schedules, status codes)."""
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include', 'oxr.h')).read()
    vals = {m.group(1): int(m.group(2), 0) for m in re.finditer(r'\b(OXR_[A-Z0-9_]+)\s*=\s*(-?\w+)', hdr)}
    shared = [k for k in vals if hasattr(abi, k)]
    assert {'OXR_GET_REC', 'OXR_GET_PRD', 'OXR_FORM_PERSISTENT', 'OXR_RNG_ALIGNED', 'OXR_OK'} <= set(shared)
    for k in shared:
        assert getattr(abi, k) == vals[k], k
    # the #define constants (no-alloc limits, autotuner result columns)
    defs = {m.group(1): eval(m.group(2)) for m in re.finditer(r'#define\s+(OXR_[A-Z0-9_]+)\s+([0-9()<x ]+)\n', hdr)}
    assert {'OXR_NOALLOC_MAX_PHOTON', 'OXR_NOALLOC_MAX_GENSTEP', 'OXR_NOALLOC_SCRATCH_BYTES',
            'OXR_AUTOTUNE_COLUMNS'} <= set(defs)
    for k, v in defs.items():
        if k.startswith(('OXR_NOALLOC_', 'OXR_AUTOTUNE_')):
            assert getattr(abi, k) == v, k


def test_tuning_range_checked_on_the_host(engine_mod):
    """oxr_engine_set_tuning validates the tuning before touching the engine: a block of
    512 lanes (beyond the kernels' 256-lane launch bound), a refill batch beyond a wave or
    a stack mode outside 0/1 is rejected as out of range; an in-range tuning reaches the
    null-engine check."""
    L = engine_mod.lib()
    good = dict(block_size=256, blocks_per_cu=4, lds_top_nodes=64, refill_batch=8, stack_mode=1, bvh_mode=0)
    for bad in (dict(block_size=512), dict(block_size=1024), dict(block_size=96), dict(refill_batch=65),
                dict(stack_mode=2), dict(blocks_per_cu=0)):
        t = abi.Tuning(**dict(good, **bad))
        assert L.oxr_engine_set_tuning(None, C.byref(t)) == abi.OXR_EINVAL
        assert b'out of range' in L.oxr_last_error(), bad
    t = abi.Tuning(**good)
    assert L.oxr_engine_set_tuning(None, C.byref(t)) == abi.OXR_EINVAL
    assert b'null' in L.oxr_last_error()
