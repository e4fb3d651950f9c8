"""The refine trace and the instance visibility mask on the GPU, bit for bit against the
oracle (whose semantics tests/test_refine_visibility.py pins on the CPU): simulate in the
production and debug kernels of both forms, simtrace, and the raw trace query."""
import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes

from test_refine_visibility import far_sphere_arrays, saturating_scene

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def run(engine_mod, s, gs, cfg, form):
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_gensteps(gs)
        st = e.launch(form)
        return e.photons(), (e.seq() if cfg.seq_enabled else None), e.hits(), st


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
@pytest.mark.parametrize('seq', [0, 1])
def test_hidden_instances(engine_mod, oracle_lib, form, seq):
    """vizmask without solid 2's bit: GPU == oracle, and == the geometry without those
    instances on the GPU as well."""
    s = scenes.hidden_instances()
    s0 = scenes.hidden_instances(hide=False)
    gs = s.gensteps(scale_photons=0.2)
    cfg = abi.default_config(**s.config, vizmask=0xff & ~(1 << 2), seq_enabled=seq)
    print('hidden instances, form', form, 'seq', seq, flush=True)
    p, q, h, st = run(engine_mod, s, gs, cfg, form)
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs)
    assert np.array_equal(bits(p), bits(r['photon']))
    if seq:
        assert np.array_equal(q, r['seq'])
    p0, _, _, _ = run(engine_mod, s0, gs, abi.default_config(**s0.config, seq_enabled=seq), form)
    assert np.array_equal(bits(p), bits(p0))
    assert st['bounce_steps'] == int(r['bounce'].sum())


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
@pytest.mark.parametrize('seq', [0, 1])
def test_refine_thin_shells(engine_mod, oracle_lib, form, seq):
    """PropagateRefine on the thin shells (shells 0.39 m apart, so a 300 mm refine distance
    is exceeded by most traces) and on instanced CSG (the default 5 m, exceeded by most
    traces across the 19.5 m sphere): GPU == oracle photon by photon."""
    for name, make, frac, dist in (('thin_shells', lambda: scenes.thin_shells(num_inst=300), 0.002, 300.0),
                                   ('instanced_csg', lambda: scenes.instanced_csg(num_inst=2000), 0.002, 5000.0)):
        s = make()
        gs = s.gensteps(scale_photons=frac)
        cfg = abi.default_config(**s.config, propagate_refine=1, refine_distance=dist, seq_enabled=seq)
        print('refine', name, 'form', form, 'seq', seq, flush=True)
        p, q, h, st = run(engine_mod, s, gs, cfg, form)
        o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
        o.set_cull(True)
        r = o.simulate(gs)
        assert np.array_equal(bits(p), bits(r['photon'])), name
        if seq:
            assert np.array_equal(q, r['seq']), name
        # the refined run differs from the unrefined one (the second pass ran)
        p1, _, _, _ = run(engine_mod, s, gs, abi.default_config(**s.config), form)
        assert not np.array_equal(bits(p), bits(p1)), name


def test_refine_simtrace_and_miss(engine_mod, oracle_lib):
    geometry, physics = far_sphere_arrays()
    gs = np.concatenate([G.frame((3000.0, 2000.0, 1000.0, 100.0), (2, 0, 2, 40)),
                         G.frame((40_000.0, 0.0, 0.0, 10.0), (0, 0, 0, 16))])
    for kw in (dict(propagate_refine=1, refine_distance=5000.0), dict(propagate_refine=1, refine_distance=0.5), {}):
        cfg = abi.default_config(**kw)
        want = oracle_lib.Oracle(geometry, physics, cfg).simtrace(gs)
        with engine_mod.Engine(geometry, physics, cfg) as e:
            got = e.simtrace(gs)
        assert np.array_equal(bits(got), bits(want)), kw


def test_visibility_trace_query(engine_mod, oracle_lib):
    f, rays = saturating_scene()
    for mask in (1 | 1 << 7, 1 | 1 << 3, 0xff, 1 << 8):
        cfg = abi.default_config(vizmask=mask)
        want = oracle_lib.Oracle(f.arrays(), f.physics(), cfg).trace(rays)
        with engine_mod.Engine(f.arrays(), f.physics(), cfg) as e:
            got = e.trace(rays)
        assert np.array_equal(bits(got), bits(want)), mask
