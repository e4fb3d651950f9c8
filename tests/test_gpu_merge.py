"""GPU photonlite output and device hit merging through the C ABI, against the oracle's
photonlite and its numpy restatement of SPM::merge_partial_select."""
import numpy as np
import pytest

from oxr import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def csi():
    s = scenes.csi_calorimeter()
    return s, s.gensteps(scale_photons=0.02)


def run(engine_mod, s, gs, form, **cfgkw):
    cfg = abi.default_config(**s.config, photonlite=1, **cfgkw)
    e = engine_mod.Engine(s.geometry, s.physics, cfg)
    e.set_gensteps(gs)
    e.launch(form)
    return e


@pytest.fixture(scope='module')
def oracle_csi(oracle_lib, csi):
    s, gs = csi
    o = oracle_lib.Oracle(s.geometry, s.physics, abi.default_config(**s.config))
    o.set_cull(True)
    return o.simulate(gs, lite=True, seq=False)


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_photonlite_and_hitlite_bitexact(engine_mod, csi, oracle_csi, form):
    s, gs = csi
    with run(engine_mod, s, gs, form) as e:
        ph, lt, hl = e.photons(), e.get(abi.OXR_GET_PHOTONLITE), e.get(abi.OXR_GET_HITLITE)
    r = oracle_csi
    print('csi photonlite', len(lt), 'photons', len(hl), 'hitlite', flush=True)
    assert np.array_equal(ph.view(np.uint32), r['photon'].view(np.uint32))
    assert np.array_equal(lt, r['lite'])
    sd = abi.FLAGS['SURFACE_DETECT']
    assert len(hl) > 0 and np.array_equal(hl, r['lite'][(r['lite'][:, 3] & sd) == sd])


@pytest.mark.parametrize('tw', [0.0, 1.0, 50.0])
def test_device_merge_matches_oracle(engine_mod, oracle_lib, csi, oracle_csi, tw):
    s, gs = csi
    sd = abi.FLAGS['SURFACE_DETECT']
    with run(engine_mod, s, gs, abi.OXR_FORM_PERSISTENT) as e:
        mh = e.merged(abi.OXR_GET_HIT, tw)
        ml = e.merged(abi.OXR_GET_HITLITE, tw)
    want_h = oracle_lib.merge(oracle_csi['photon'], tw, mask=sd)
    want_l = oracle_lib.merge(oracle_csi['lite'], tw, mask=sd)
    print('merge tw', tw, 'hits', int(((oracle_csi['lite'][:, 3] & sd) != 0).sum()), '->', len(want_h), flush=True)
    assert np.array_equal(mh.view(np.uint32).reshape(len(mh), 16), want_h)
    assert np.array_equal(ml, want_l)
    if tw == 50.0:
        assert len(want_h) < int(((oracle_csi['lite'][:, 3] & sd) != 0).sum())


def test_final_merge_of_slices(engine_mod, oracle_lib, csi, oracle_csi):
    """Per-slice merges concatenated and merged again (QEvt::FinalMerge) equal one merge of
    the whole event's selection."""
    s, gs = csi
    sd = abi.FLAGS['SURFACE_DETECT']
    half = len(gs) // 2
    parts = []
    with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config)) as e:
        n0 = e.set_gensteps(gs[:half], slot_offset=0)
        e.launch(abi.OXR_FORM_PERSISTENT)
        parts.append(e.merged(abi.OXR_GET_HIT, 50.0))
        e.set_gensteps(gs[half:], slot_offset=n0)
        e.launch(abi.OXR_FORM_PERSISTENT)
        parts.append(e.merged(abi.OXR_GET_HIT, 50.0))
        final = e.merge_hits(np.concatenate(parts), abi.OXR_GET_HIT, 50.0)
    want = oracle_lib.merge(oracle_csi['photon'], 50.0, mask=sd)
    assert np.array_equal(final.view(np.uint32).reshape(len(final), 16), want)
