"""Seeded random scenes simulated end to end: random materials (refractive index,
absorption and scattering lengths), sensor, polished and ground surfaces, random CSG-tree
prims at random places in a world box, and a second solid instanced at random positions
and orientations. Photons (and seqhis) from the GPU, through the C ABI, equal the
oracle's bit for bit in the production and debug kernels and both launch forms."""
import numpy as np
import pytest

from oxr import abi, gensteps as G
from oxr.csg import Foundry, box3, translate, rotate_z_to

from test_gpu_parity import bits
from test_gpu_random_csg import random_tree

pytestmark = pytest.mark.gpu


def random_scene(seed):
    rng = np.random.default_rng(seed)
    f = Foundry()
    f.add_material('World', 1.0)
    mats = []
    for k in range(3):
        name = 'M%d' % k
        f.add_material(name, float(rng.uniform(1.3, 1.8)), absorption_length=float(rng.uniform(200, 5000)),
                       scattering_length=float(rng.uniform(300, 5000)))
        mats.append(name)
    eff = float(rng.uniform(0.2, 0.8))
    f.add_surface('Sens', detect=eff, absorb=1.0 - eff, osn0='S')
    f.add_surface('Mirror', absorb=0.1, specular=0.9)
    f.add_surface('Ground', absorb=0.2, diffuse=0.8)
    surfs = [None, None, 'Sens', 'Mirror', 'Ground']
    f.begin_solid('global')
    f.add_prim(box3(1000.0), f.boundary('World', None, None, 'World'))
    for _ in range(8):
        t = random_tree(rng, int(rng.integers(0, 3))).moved(translate(*[float(x) for x in rng.uniform(-300, 300, 3)]))
        sur = surfs[int(rng.integers(0, len(surfs)))]
        f.add_prim(t, f.boundary('World', sur, None, mats[int(rng.integers(0, 3))]))
    f.end_solid()
    f.begin_solid('inst')
    f.add_prim(random_tree(rng, 2), f.boundary('World', 'Sens', None, mats[0]))
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    for _ in range(40):
        m = rotate_z_to(rng.normal(size=3)) @ translate(*[float(x) for x in rng.uniform(-400, 400, 3)])
        f.add_instance(m, 1)
    wl = float(rng.uniform(400, 600))
    gs = G.torch([3000] * 3, pos=(0.0, 0.0, 0.0), type=G.T_POINT_ISOTROPIC, wavelength=wl)
    return f, gs


@pytest.mark.parametrize('seed', [7, 8, 9])
def test_random_scene_bitexact(engine_mod, oracle_lib, seed):
    f, gs = random_scene(seed)
    g, p = f.arrays(), f.physics()
    cfg = abi.default_config(max_bounce=32, seq_enabled=1)
    print('random scene', seed, flush=True)
    want = oracle_lib.Oracle(g, p, cfg).simulate(gs)
    flags = bits(want['photon'])[:, 3, 3]
    assert len(set(flags.tolist())) > 5          # a mix of histories
    with engine_mod.Engine(g, p, cfg) as e:       # debug kernels: seqhis on
        e.set_gensteps(gs)
        e.launch(abi.OXR_FORM_PLAIN)
        assert np.array_equal(bits(e.photons()), bits(want['photon']))
        assert np.array_equal(e.seq(), want['seq'])
    prod = abi.default_config(max_bounce=32)
    with engine_mod.Engine(g, p, prod) as e:      # production kernels, both forms, every traversal form
        e.set_gensteps(gs)
        for bvh in (0, 1):
            for stack in (0, 1):
                e.set_tuning(bvh_mode=bvh, stack_mode=stack)
                for form in (abi.OXR_FORM_PERSISTENT, abi.OXR_FORM_PLAIN):
                    e.launch(form)
                    assert np.array_equal(bits(e.photons()), bits(want['photon'])), (bvh, stack, form)
