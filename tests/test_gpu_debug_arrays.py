"""The per-bounce debug arrays on the GPU through the C ABI: rec (srec) and prd (quad2) must
equal the oracle's bit for bit, next to the full records, in both launch forms."""
import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes

from test_foundry_io import mixed_scene_arrays

pytestmark = pytest.mark.gpu


class _Scene:
    def __init__(self, geometry, physics, config, gensteps):
        self.geometry, self.physics, self.config, self._gs = geometry, physics, config, gensteps

    def gensteps(self, scale_photons=1.0):
        return self._gs


def _mixed():
    g, p = mixed_scene_arrays()
    return _Scene(g, p, {}, G.torch([2000] * 4, pos=(0.0, 0.0, -200.0), wavelength=440.0))


CASES = {
    'nested_boxes': (lambda: scenes.nested_boxes(), 0.05),
    'csi_calorimeter': (lambda: scenes.csi_calorimeter(), 0.002),
    'instanced_csg': (lambda: scenes.instanced_csg(num_inst=2000), 0.001),
    'list_nodes': (_mixed, 1.0),
}


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
@pytest.mark.parametrize('name', list(CASES))
def test_rec_prd_bitexact(engine_mod, oracle_lib, name, form):
    build, scale = CASES[name]
    s = build()
    cfg = abi.default_config(**s.config, max_record=12, max_rec=12, max_prd=12, time_domain=50.0)
    gs = s.gensteps(scale_photons=scale)
    print('rec/prd', name, 'form', form, flush=True)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_gensteps(gs)
        e.launch(form)
        ph, record = e.photons(), e.get(abi.OXR_GET_RECORD)
        rec, prd = e.get(abi.OXR_GET_REC), e.get(abi.OXR_GET_PRD)
    want = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs, max_record=12, max_rec=12, max_prd=12)
    assert np.array_equal(ph.view(np.uint32), want['photon'].view(np.uint32))
    assert np.array_equal(record.view(np.uint32), want['record'].view(np.uint32))
    assert rec.shape == want['rec'].shape and np.array_equal(rec, want['rec'])
    assert prd.shape == want['prd'].shape and np.array_equal(prd.view(np.uint32), want['prd'].view(np.uint32))
    assert np.any(prd != 0) and np.any(rec != 0)


def test_debug_arrays_need_their_config(engine_mod):
    s = scenes.nested_boxes()
    with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config)) as e:
        e.set_gensteps(s.gensteps(scale_photons=0.01))
        e.launch(abi.OXR_FORM_PLAIN)
        for what in (abi.OXR_GET_REC, abi.OXR_GET_PRD):
            with pytest.raises(engine_mod.OxrError):
                e.get(what)
