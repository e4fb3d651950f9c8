"""Every traversal form the tuning can select (two-level or flattened BVH, plain or
distance-carrying stack, CSG stacks in registers or private memory, radial culling) and
launch shape (workgroup size, LDS staging, workgroups per CU, refill batch, photons claimed
in index or sorted order, generated in a pass of their own or where they are claimed) gives
the same photons, bit for bit, and the same photons as the oracle."""
import numpy as np
import pytest

from oxr import abi, scenes

from test_foundry_io import mixed_scene_arrays

pytestmark = pytest.mark.gpu

FORMS = [dict(bvh_mode=b, stack_mode=s, csg_stack=c, bvh_radial=r) for b in (0, 1) for s in (0, 1) for c in (0, 1)
         for r in (0, 1)]
# (block size, LDS top nodes, workgroups per CU, refill batch, photon order, generation in place)
SHAPES = [(256, 64, 3, 8, 0, 0), (128, 512, 8, 1, 1, 0), (64, 16, 1, 64, 0, 1)]


def cases():
    s = scenes.csi_calorimeter()
    yield 'csi', s.geometry, s.physics, s.config, s.gensteps(scale_photons=0.02)
    s = scenes.instanced_csg(num_inst=2000, num_photon=40000)
    yield 'instanced', s.geometry, s.physics, s.config, s.gensteps()
    s = scenes.thin_shells(num_shell=10, num_inst=500, num_photon=20000)
    yield 'thin', s.geometry, s.physics, s.config, s.gensteps()
    from oxr import gensteps as G
    g, p = mixed_scene_arrays()
    yield 'mixed', g, p, {}, G.torch([5000] * 4, pos=(0.0, 0.0, -200.0), wavelength=440.0)


@pytest.mark.parametrize('case', list(cases()), ids=lambda c: c[0])
def test_forms_agree(engine_mod, oracle_lib, case):
    name, geometry, physics, config, gs = case
    cfg = abi.default_config(**config)
    want = oracle_lib.Oracle(geometry, physics, cfg)
    want.set_cull(True)
    ref = want.simulate(gs, seq=False)['photon']
    with engine_mod.Engine(geometry, physics, cfg) as e:
        e.set_gensteps(gs)
        for form in FORMS:
            for block, top, bpc, refill, order, inplace in SHAPES:
                e.set_tuning(block_size=block, lds_top_nodes=top, blocks_per_cu=bpc, refill_batch=refill,
                             photon_order=order, generate_inplace=inplace, **form)
                for kind in (abi.OXR_FORM_PERSISTENT, abi.OXR_FORM_PLAIN):
                    print(name, form, block, top, order, inplace, kind, flush=True)
                    e.launch(kind)
                    got = e.photons()
                    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (name, form, block, kind)
