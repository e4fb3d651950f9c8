"""Geometries imported from the reference's GDML test files, simulated on the GPU through
the C ABI and compared bit for bit with the oracle on the same arrays."""
import os

import numpy as np
import pytest

from oxr import abi, gdml, gensteps as G, sources

from test_gdml import config, geom

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def run_gpu(engine_mod, g, gs, form, input_photons=None, **cfg):
    c = abi.default_config(**cfg)
    with engine_mod.Engine(g.geometry, g.physics, c) as e:
        if input_photons is not None:
            e.set_input_photons(input_photons)
        e.set_gensteps(gs)
        e.launch(form)
        return e.photons(), e.get(abi.OXR_GET_HIT)


def run_oracle(oracle_lib, g, gs, input_photons=None, **cfg):
    o = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config(**cfg))
    return o.simulate(gs, input_photons=input_photons)['photon']


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_csi_sipm_array_bitexact(engine_mod, oracle_lib, form):
    g = gdml.load_gdml(geom('8x8SiPM_w_CSI_optial_grease'))
    gs = sources.torch_config(config('8x8SiPM_crystal')).genstep
    print('8x8 SiPM + CsI GDML, torch config, form', form, flush=True)
    ph, hit = run_gpu(engine_mod, g, gs, form, max_bounce=32)
    want = run_oracle(oracle_lib, g, gs, max_bounce=32)
    assert np.array_equal(ph.view(np.uint32), want.view(np.uint32))
    sd = abi.FLAGS['SURFACE_DETECT']
    assert len(hit) == int(((want.view(np.uint32)[:, 3, 3] & sd) == sd).sum()) > 0


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_pfrich_bitexact(engine_mod, oracle_lib, form):
    """pfRICH: 103 volumes, trap/trd polyhedra, boolean trees; isotropic photons."""
    g = gdml.load_gdml(geom('pfrich_min_FINAL'))
    gs = G.torch([5000] * 4, pos=(0.0, 0.0, 0.0), type=G.T_POINT_ISOTROPIC, wavelength=400.0)
    print('pfRICH GDML, 20000 photons, form', form, flush=True)
    ph, _ = run_gpu(engine_mod, g, gs, form, max_bounce=32)
    want = run_oracle(oracle_lib, g, gs, max_bounce=32)
    assert np.array_equal(ph.view(np.uint32), want.view(np.uint32))


def test_raindrop_photon_file_kat(engine_mod):
    """tests/test_GPUPhotonFileSource.sh on the GPU: 10 photons -> 10 hits."""
    g = gdml.load_gdml(geom('opticks_raindrop'))
    ip = sources.load_photons_txt(os.path.join(GOLDEN, 'raindrop_photons.txt'))
    print('raindrop GDML photon-file KAT', flush=True)
    ph, hit = run_gpu(engine_mod, g, G.input_photon(10), abi.OXR_FORM_PLAIN, input_photons=ip, max_bounce=31)
    assert len(hit) == 10
    wl = np.round(hit[:, 2, 3]).astype(int)
    assert sorted(wl.tolist()) == [420] * 3 + [450] * 2 + [500] * 5


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_raindrop_simg4ox_bitexact(engine_mod, oracle_lib, form):
    """tests/geom/raindrop.gdml (tests/test_simg4ox.sh geometry), isotropic photons from the drop."""
    g = gdml.load_gdml(geom('raindrop'))
    gs = G.torch([5000] * 2, pos=(0.0, 0.0, 0.0), type=G.T_POINT_ISOTROPIC, wavelength=420.0)
    print('raindrop GDML, 10000 photons, form', form, flush=True)
    ph, _ = run_gpu(engine_mod, g, gs, form, max_bounce=32)
    want = run_oracle(oracle_lib, g, gs, max_bounce=32)
    assert np.array_equal(ph.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize('name', ['sphere_leak', 'basic_detector'])
def test_other_geometries_bitexact(engine_mod, oracle_lib, name):
    g = gdml.load_gdml(geom(name))
    gs = sources.torch_config(config('sphere_leak' if name == 'sphere_leak' else 'dev')).genstep.copy()
    gs.view(np.uint32)[0, 0, 3] = min(int(gs.view(np.uint32)[0, 0, 3]), 20000)
    print(name, 'GDML', flush=True)
    ph, _ = run_gpu(engine_mod, g, gs, abi.OXR_FORM_PERSISTENT, max_bounce=32)
    want = run_oracle(oracle_lib, g, gs, max_bounce=32)
    assert np.array_equal(ph.view(np.uint32), want.view(np.uint32))
