"""The autotuner (oxr_autotune): rows come back for distinct effective tunings, the
engine is left on the fastest row, photons launched with it equal the oracle's, the row
count honours max_results, and a failing call leaves the tuning as it was."""
import ctypes as C

import numpy as np
import pytest

from oxr import abi, scenes

pytestmark = pytest.mark.gpu

KEYS = ['block_size', 'blocks_per_cu', 'lds_top_nodes', 'refill_batch', 'stack_mode', 'bvh_mode', 'csg_stack', 'bvh_radial',
        'photon_order', 'generate_inplace']


def test_autotune_picks_the_fastest_row_and_keeps_parity(engine_mod, oracle_lib):
    s = scenes.instanced_csg(num_inst=2000, num_photon=30000)
    cfg = abi.default_config(**s.config)
    gs = s.gensteps()
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_gensteps(gs)
        print('autotune, instanced CSG, 30k photons', flush=True)
        rows, best = e.autotune(repeat=1)
        assert len(rows) >= 8 and np.all(rows[:, 10] > 0)
        assert len({tuple(r[:10]) for r in rows}) == len(rows)
        k = int(np.argmin(rows[:, 10]))
        assert best == dict(zip(KEYS, (int(v) for v in rows[k, :10])))
        assert e.tuning() == best
        # the CSG stack depth, the radial culling, the photon order and where photons are
        # generated were searched dimensions
        assert set(rows[:, 6].astype(int)) == {0, 1} and set(rows[:, 7].astype(int)) == {0, 1}
        assert set(rows[:, 8].astype(int)) == {0, 1} and set(rows[:, 9].astype(int)) == {0, 1}
        e.launch(abi.OXR_FORM_PERSISTENT)
        got = e.photons()
        few, _ = e.autotune(repeat=1, max_results=3)
        assert len(few) == 3
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    o.set_cull(True)
    want = o.simulate(gs, seq=False)['photon']
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_autotune_without_an_event_keeps_the_tuning(engine_mod):
    s = scenes.nested_boxes()
    with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config)) as e:
        before = e.tuning()
        res = np.zeros((4, abi.OXR_AUTOTUNE_COLUMNS), dtype=np.float64)
        n = C.c_int32(-1)
        rc = engine_mod.lib().oxr_autotune(e.h, 1, abi.ptr(res), 4, C.byref(n))
        assert rc == abi.OXR_ESTATE
        assert e.tuning() == before
