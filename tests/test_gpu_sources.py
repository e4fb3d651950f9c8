"""GPU parity of the generation and transport corners: every reference torch type,
re-emission in a wavelength-dependent scintillator (both random-number schedules, Philox
and precooked streams that force it), both engine forms and the production kernels.
Photons, seqhis and hits must equal the oracle bit for bit."""
import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def torch_gensteps():
    """One genstep of each reference torch type inside the 100 mm water cube of the nested
    boxes, oriented so that their photons cross the water, air and lead boundaries."""
    mom = np.array([1.0, -2.0, 2.0]) / 3.0
    return {
        'disc': G.torch(4000, pos=(5, -7, 11), mom=mom, type=G.T_DISC, radius=30.0, zenith=(0.2, 1.0),
                        azimuth=(0.0, 1.0)),
        'sphere_out': G.torch(4000, type=G.T_SPHERE, radius=40.0, distance=0.3, zenith=(0.0, 1.0), azimuth=(0.0, 1.0)),
        'sphere_in': G.torch(4000, type=G.T_SPHERE, radius=-45.0, distance=0.7, zenith=(0.1, 0.9), azimuth=(0.2, 0.6)),
        'sphere_marsaglia': G.torch(4000, type=G.T_SPHERE_MARSAGLIA, radius=35.0, distance=0.25),
        'line': G.torch(4000, pos=(1, 2, 3), mom=mom, type=G.T_LINE, radius=40.0),
        'point': G.torch(4000, pos=(-10, 20, 5), mom=mom, type=G.T_POINT),
        'circle': G.torch(4000, pos=(5, 0, 0), type=G.T_CIRCLE, radius=-30.0, azimuth=(0.0, 1.0)),
        'rectangle': G.torch(4000, pos=(0, 5, 0), type=G.T_RECTANGLE, zenith=(-30.0, 40.0), azimuth=(-20.0, 35.0)),
    }


_nb = {}


def nested():
    if 'nb' not in _nb:
        _nb['nb'] = scenes.nested_boxes()
    return _nb['nb']


@pytest.mark.parametrize('kind', list(torch_gensteps()))
@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_torch_types_bitexact(engine_mod, oracle_lib, kind, form):
    s = nested()
    gs = torch_gensteps()[kind]
    cfg = abi.default_config(**s.config, seq_enabled=1)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        n = e.set_gensteps(gs, event_index=2, slot_offset=1000)
        st = e.launch(form)
        ph, seq = e.photons(), e.seq()
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs, event_index=2, slot_offset=1000)
    print(kind, form, n, 'photons', st['bounce_steps'], 'steps', flush=True)
    assert n == 4000 and st['bounce_steps'] > n
    assert np.array_equal(bits(ph), bits(r['photon']))
    assert np.array_equal(seq, r['seq'])


@pytest.mark.parametrize('kind', ['disc', 'sphere_marsaglia', 'line', 'rectangle'])
def test_torch_types_production_kernels(engine_mod, oracle_lib, kind):
    """Without seqhis the lean kernels run (generation in its own pass): same photons."""
    s = nested()
    gs = torch_gensteps()[kind]
    cfg = abi.default_config(**s.config)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_gensteps(gs, slot_offset=77)
        e.launch(abi.OXR_FORM_PERSISTENT)
        ph = e.photons()
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs, slot_offset=77)
    assert np.array_equal(bits(ph), bits(r['photon']))


_ls = {}


def ls():
    if 'ls' not in _ls:
        _ls['ls'] = scenes.scintillator_sphere()
    return _ls['ls']


@pytest.mark.parametrize('schedule', [abi.OXR_RNG_PRODUCTION, abi.OXR_RNG_ALIGNED])
@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_reemission_scene_bitexact(engine_mod, oracle_lib, schedule, form):
    """Scintillator sphere: UV light absorbed and re-emitted (about 80%) with every table
    lookup between distinct texels; photons, seqhis and hits equal the oracle's."""
    s = ls()
    cfg = abi.default_config(**s.config, seq_enabled=1, rng_schedule=schedule)
    gs = s.gensteps(scale_photons=0.5)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_gensteps(gs)
        st = e.launch(form)
        ph, seq, hit = e.photons(), e.seq(), e.hits()
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs)
    u = bits(r['photon'])[:, 3, 3]
    print('reemission', schedule, form, st['num_photon'], st['bounce_steps'], 'RE',
          int(((u & abi.FLAGS['BULK_REEMIT']) != 0).sum()), 'hits', st['num_hit'], flush=True)
    assert ((u & abi.FLAGS['BULK_REEMIT']) != 0).mean() > 0.5
    assert np.array_equal(bits(ph), bits(r['photon']))
    assert np.array_equal(seq, r['seq'])
    sel = (u & cfg.hitmask) == cfg.hitmask
    assert np.array_equal(bits(hit), bits(r['photon'][sel]))


def test_reemission_production_kernels(engine_mod, oracle_lib):
    s = ls()
    cfg = abi.default_config(**s.config)
    gs = s.gensteps(scale_photons=0.5)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_gensteps(gs)
        e.launch(abi.OXR_FORM_PERSISTENT)
        ph = e.photons()
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs)
    assert np.array_equal(bits(ph), bits(r['photon']))


def test_reemission_forced_by_precooked_streams(engine_mod, oracle_lib):
    """Precooked rows that force absorption then re-emission on the first bounce (the
    layout of the oracle pin of the re-emission draws): GPU == oracle, records included."""
    s = ls()
    cfg = abi.default_config(**s.config, seq_enabled=1, max_record=6)
    n = 2000
    rng = np.random.default_rng(21)
    pre = rng.uniform(1e-6, 1.0, size=(n, 64)).astype(np.float32)
    pre[:, 3] = 1e-6
    pre[:, 4] = rng.uniform(0.3, 0.9, n)
    pre[:, 5] = rng.uniform(0.0, 0.75, n)
    gs = s.gensteps(scale_photons=n / 20000)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_precooked(pre)
        e.set_gensteps(gs)
        e.launch(abi.OXR_FORM_PLAIN)
        ph, seq, rec = e.photons(), e.seq(), e.get(abi.OXR_GET_RECORD)
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs, precooked=pre, max_record=6)
    assert np.all(r['record'].view(np.uint32)[:, 1, 3, 0] & 0xffff == abi.FLAGS['BULK_REEMIT'])
    assert np.array_equal(bits(ph), bits(r['photon']))
    assert np.array_equal(seq, r['seq'])
    assert np.array_equal(bits(rec), bits(r['record']))
