"""CSGFoundry folders in the reference layout (CSG/CSGFoundry.cc:2768-2800 save_,
:2977-3050 load; SSim tables under SSim/stree/standard, sysrap/SSim.hh:52-55): a scene
written with oxr.cfio.save_foundry and read back through the library's own NPY reader
(oxr_foundry_load, host only) gives the same arrays, the same oracle photons, and clear
errors for folders the reader cannot take."""
import os

import numpy as np
import pytest

from oxr import abi, cfio, scenes
from oxr.csg import Foundry, box3, sphere, contiguous, convexpolyhedron, overlap, rotate_z_to, translate

GEOM = ('solid', 'prim', 'node', 'plan', 'itra', 'inst')


TET = [np.r_[np.array(p[:3]) / np.linalg.norm(p[:3]), p[3]]
       for p in [(1, 1, 1, 60), (-1, -1, 1, 60), (-1, 1, -1, 60), (1, -1, -1, 60)]]


def mixed_scene_arrays():
    """List nodes, a convex polyhedron (plan rows) and a transformed leaf (itra rows)."""
    f = Foundry()
    f.add_material('Air', 1.0)
    f.add_material('Glass', 1.5)
    b = f.boundary('Air', None, None, 'Glass')
    f.begin_solid('s')
    f.add_prim(sphere(70.0) & overlap(box3(100.0), sphere(60.0)), b)
    f.add_prim(contiguous(box3(40.0), sphere(25.0, 30.0, 0, 0)), b)
    f.add_prim(convexpolyhedron(TET, (-110, -110, -110, 110, 110, 110)), b)
    f.add_prim(box3(80.0).moved(rotate_z_to([1.0, 2.0, 3.0]) @ translate(5.0, -7.0, 9.0)), b)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    return f.arrays(), f.physics()


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _check_round_trip(geometry, physics, g, p):
    for k in GEOM:
        want = geometry.get(k)
        if want is None or len(want) == 0:
            assert len(g[k]) == 0, k
        else:
            assert _same(np.ascontiguousarray(want, dtype=g[k].dtype), g[k]), k
    assert _same(physics['bnd'].astype(np.float32), p['bnd'])
    assert np.array_equal(np.asarray(physics['optical'], dtype=np.uint32), p['optical'])
    assert p['wl_min'] == physics['wl_min'] and p['wl_step'] == physics['wl_step']
    if 'icdf' in physics:
        assert _same(np.asarray(physics['icdf'], dtype=np.float32), p['icdf'])
        assert p['hd_factor'] == physics['hd_factor'] == 20
    else:
        assert 'icdf' not in p and p['hd_factor'] == 0


@pytest.mark.parametrize('name', ['csi_calorimeter', 'raindrop', 'mixed'])
def test_round_trip(engine_mod, tmp_path, name):
    if name == 'mixed':
        geometry, physics = mixed_scene_arrays()
    else:
        s = getattr(scenes, name)()
        geometry, physics = s.geometry, s.physics
    d = cfio.save_foundry(str(tmp_path), geometry, physics)
    assert os.path.exists(os.path.join(d, 'SSim', 'stree', 'standard', 'bnd_meta.txt'))
    g, p = cfio.load_foundry(str(tmp_path))
    _check_round_trip(geometry, physics, g, p)
    if name == 'mixed':
        assert len(g['plan']) == 4 and len(g['itra']) >= 1
        assert np.load(os.path.join(d, 'plan.npy')).shape == (4, 1, 4)     # NP::Write(..., 1, 4)
    # the CSGFoundry folder itself is accepted too
    g2, p2 = cfio.load_foundry(d)
    assert all(_same(g[k], g2[k]) for k in GEOM)


def test_loaded_folder_simulates_like_arrays(engine_mod, oracle_lib, tmp_path):
    """Raindrop known-answer case: the oracle on the loaded arrays returns the photons it
    returns on the in-memory arrays, bit for bit."""
    s = scenes.raindrop()
    cfio.save_foundry(str(tmp_path), s.geometry, s.physics)
    g, p = cfio.load_foundry(str(tmp_path))
    cfg = abi.default_config(**s.config)
    ip = scenes.raindrop_input_photons()
    want = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(s.gensteps(), input_photons=ip)
    got = oracle_lib.Oracle(g, p, cfg).simulate(s.gensteps(), input_photons=ip)
    assert np.array_equal(want['photon'].view(np.uint32), got['photon'].view(np.uint32))


def test_npy_format_versions_and_domain(engine_mod, tmp_path):
    """NPY format 2.0 headers are read; the wavelength domain comes from bnd_meta.txt."""
    s = scenes.raindrop()
    d = cfio.save_foundry(str(tmp_path), s.geometry, s.physics)
    with open(os.path.join(d, 'node.npy'), 'wb') as fh:
        np.lib.format.write_array(fh, np.ascontiguousarray(s.geometry['node']), version=(2, 0))
    meta = os.path.join(d, 'SSim', 'stree', 'standard', 'bnd_meta.txt')
    with open(meta, 'w') as fh:
        fh.write('creator:test\ndomain_low:80.5\ndomain_step:2\n')
    g, p = cfio.load_foundry(str(tmp_path))
    assert _same(s.geometry['node'], g['node'])
    assert p['wl_min'] == 80.5 and p['wl_step'] == 2.0
    os.remove(meta)
    g, p = cfio.load_foundry(str(tmp_path))
    assert p['wl_min'] == 60.0 and p['wl_step'] == 1.0        # sdomain defaults


def _expect_error(engine_mod, path, code, text):
    with pytest.raises(engine_mod.OxrError) as ei:
        cfio.load_foundry(path)
    assert ei.value.code == code and text in str(ei.value), str(ei.value)


def test_errors(engine_mod, tmp_path):
    s = scenes.raindrop()
    _expect_error(engine_mod, str(tmp_path / 'nowhere'), abi.OXR_ENOTFOUND, 'no CSGFoundry folder')
    d = cfio.save_foundry(str(tmp_path), s.geometry, s.physics)

    def rewrite(name, arr, **kw):
        path = os.path.join(d, name)
        with open(path, 'wb') as fh:
            np.lib.format.write_array(fh, arr, **kw)

    node = np.ascontiguousarray(s.geometry['node'])
    rewrite('node.npy', node.reshape(len(node), 16))
    _expect_error(engine_mod, str(tmp_path), abi.OXR_EINVAL, 'node.npy: expected')
    rewrite('node.npy', node.astype('>f4'))
    _expect_error(engine_mod, str(tmp_path), abi.OXR_EINVAL, 'little-endian')
    rewrite('node.npy', node.astype(np.float64))
    _expect_error(engine_mod, str(tmp_path), abi.OXR_EINVAL, 'node.npy')
    rewrite('node.npy', np.asfortranarray(node.reshape(len(node), 16)).reshape(len(node), 4, 4, order='F'))
    _expect_error(engine_mod, str(tmp_path), abi.OXR_EINVAL, 'node.npy')
    rewrite('node.npy', node)
    with open(os.path.join(d, 'prim.npy'), 'r+b') as fh:
        fh.truncate(fh.seek(0, 2) - 8)
    _expect_error(engine_mod, str(tmp_path), abi.OXR_EINVAL, 'truncated')
    rewrite('prim.npy', np.ascontiguousarray(s.geometry['prim']))
    optical = os.path.join(d, 'SSim', 'stree', 'standard', 'optical.npy')
    np.save(optical, np.asarray(s.physics['optical'], dtype=np.uint32)[:-4])
    _expect_error(engine_mod, str(tmp_path), abi.OXR_EINVAL, 'optical.npy')
    os.remove(optical)
    _expect_error(engine_mod, str(tmp_path), abi.OXR_ENOTFOUND, 'optical.npy')
