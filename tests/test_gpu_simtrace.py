"""simtrace on the GPU through the C ABI (oxr_simtrace) against the oracle, bit for bit,
on geometries imported from the reference's GDML files and on list-node scenes."""
import numpy as np
import pytest

from oxr import abi, gdml, gensteps as G

from test_foundry_io import mixed_scene_arrays
from test_gdml import geom

pytestmark = pytest.mark.gpu


def both(engine_mod, oracle_lib, geometry, physics, gs, **cfg):
    c = abi.default_config(**cfg)
    with engine_mod.Engine(geometry, physics, c) as e:
        got = e.simtrace(gs)
    want = oracle_lib.Oracle(geometry, physics, c).simtrace(gs)
    return got, want


@pytest.mark.parametrize('fp64', [0, 1])
def test_csi_sipm_cross_section(engine_mod, oracle_lib, fp64):
    g = gdml.load_gdml(geom('8x8SiPM_w_CSI_optial_grease'))
    gs = G.frame((0.0, 0.0, 4.0, 10.0), (10, 0, 10, 20), gridscale=0.1)
    print('8x8 SiPM XZ simtrace, rays', 21 * 21 * 20, 'fp64', fp64, flush=True)
    got, want = both(engine_mod, oracle_lib, g.geometry, g.physics, gs, fp64_roots=fp64)
    assert got.shape == want.shape == (21 * 21 * 20, 4, 4)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (got[:, 0, 3] > 0).mean() > 0.99


def test_pfrich_cross_section(engine_mod, oracle_lib):
    g = gdml.load_gdml(geom('pfrich_min_FINAL'))
    gs = G.frame((0.0, 0.0, 0.0, 1000.0), (0, 8, 8, 30), gridscale=0.1)
    print('pfRICH YZ simtrace', flush=True)
    got, want = both(engine_mod, oracle_lib, g.geometry, g.physics, gs)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_list_node_scene_xyz(engine_mod, oracle_lib):
    geometry, physics = mixed_scene_arrays()
    gs = G.frame((0.0, 0.0, 0.0, 100.0), (2, 2, 2, 16), gridscale=0.5)
    print('mixed list-node scene XYZ simtrace', flush=True)
    got, want = both(engine_mod, oracle_lib, geometry, physics, gs)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_rejects_other_gensteps(engine_mod):
    geometry, physics = mixed_scene_arrays()
    with engine_mod.Engine(geometry, physics, abi.default_config()) as e:
        with pytest.raises(engine_mod.OxrError) as ei:
            e.simtrace(G.torch(10))
        assert ei.value.code == abi.OXR_EINVAL
        assert len(e.simtrace(np.zeros((0, 6, 4), dtype=np.float32))) == 0
