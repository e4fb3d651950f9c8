"""An engine created from a CSGFoundry folder read by oxr_foundry_load simulates exactly
as one created from the in-memory arrays the folder was written from."""
import numpy as np
import pytest

from oxr import abi, cfio, gensteps, scenes

from test_foundry_io import mixed_scene_arrays

pytestmark = pytest.mark.gpu


def _photons(engine_mod, geometry, physics, config, gs, form):
    cfg = abi.default_config(**config)
    with engine_mod.Engine(geometry, physics, cfg) as e:
        e.set_gensteps(gs)
        e.launch(form)
        return e.photons(), e.get(abi.OXR_GET_HIT)


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_csi_from_folder_bitexact(engine_mod, tmp_path, form):
    s = scenes.csi_calorimeter()
    gs = s.gensteps(scale_photons=0.01)
    cfio.save_foundry(str(tmp_path), s.geometry, s.physics)
    g, p = cfio.load_foundry(str(tmp_path))
    print('csi from folder, form', form, 'gensteps', len(gs), flush=True)
    a_ph, a_hit = _photons(engine_mod, s.geometry, s.physics, s.config, gs, form)
    b_ph, b_hit = _photons(engine_mod, g, p, s.config, gs, form)
    assert len(a_ph) > 0 and len(a_hit) > 0
    assert np.array_equal(a_ph.view(np.uint32), b_ph.view(np.uint32))
    assert np.array_equal(a_hit.view(np.uint32), b_hit.view(np.uint32))


def test_mixed_scene_from_folder_bitexact(engine_mod, tmp_path):
    """List nodes, convex polyhedron planes and a transformed leaf, through torch photons."""
    geometry, physics = mixed_scene_arrays()
    cfio.save_foundry(str(tmp_path), geometry, physics)
    g, p = cfio.load_foundry(str(tmp_path))
    gs = gensteps.torch([5000] * 4, pos=(0.0, 0.0, -200.0), wavelength=440.0)
    print('mixed scene from folder, photons', 20_000, flush=True)
    a_ph, _ = _photons(engine_mod, geometry, physics, {}, gs, abi.OXR_FORM_PERSISTENT)
    b_ph, _ = _photons(engine_mod, g, p, {}, gs, abi.OXR_FORM_PERSISTENT)
    assert np.array_equal(a_ph.view(np.uint32), b_ph.view(np.uint32))
    assert (a_ph.view(np.uint32)[:, 3, 3] & abi.FLAGS['BOUNDARY_TRANSMIT']).any()
