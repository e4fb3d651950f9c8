"""Full-size workloads: photon-by-photon parity on sampled index ranges (photons are
independent, so the oracle can replay any subset of a full event) plus size-independent
properties over the whole event."""
import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _check_event(engine_mod, oracle_lib, s, gs, cfg, ranges):
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        n = e.set_gensteps(gs)
        st = e.launch(abi.OXR_FORM_PERSISTENT)
        p = e.photons()
        q = e.seq()
        h = e.hits()
        st2 = e.launch(abi.OXR_FORM_PLAIN)
        p2 = e.photons()
    print(s.name, n, 'photons', st['bounce_steps'], 'steps', st['num_hit'], 'hits', f"{st['kernel_ms']:.1f} ms")
    # plain == persistent over the whole event
    assert np.array_equal(bits(p), bits(p2))
    assert st['bounce_steps'] == st2['bounce_steps'] and st['num_hit'] == st2['num_hit']
    # hits are exactly the SD photons in order, each carrying its own photon index
    u = bits(p)
    sel = (u[:, 3, 3] & cfg.hitmask) == cfg.hitmask
    assert np.array_equal(bits(h), u[sel])
    assert np.array_equal(u[:, 3, 2], np.arange(n, dtype=np.uint32))
    # every photon ended on a terminal flag or at the bounce limit
    nib = q[:, 0, 0]
    assert np.all(nib & 0xf != 0)
    # sampled ranges replayed on the CPU oracle, bit for bit
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    o.set_cull(True)
    for b, e_ in ranges:
        r = o.simulate(gs, begin=b, end=e_)
        assert np.array_equal(u[b:e_], bits(r['photon'])), (b, e_)
        assert np.array_equal(q[b:e_], r['seq'])
    return n, st


def test_csi_calorimeter_full(engine_mod, oracle_lib):
    s = scenes.csi_calorimeter()
    cfg = abi.default_config(**s.config, seq_enabled=1)
    gs = s.gensteps()
    n = G.num_photons(gs)
    assert 2.0e6 < n < 2.2e6
    _check_event(engine_mod, oracle_lib, s, gs, cfg, [(0, 1500), (n // 2, n // 2 + 1500), (n - 1500, n)])


def test_instanced_csg_full(engine_mod, oracle_lib):
    s = scenes.instanced_csg()
    cfg = abi.default_config(**s.config, seq_enabled=1)
    gs = s.gensteps()
    n = G.num_photons(gs)
    assert n == 10_000_000
    _, st = _check_event(engine_mod, oracle_lib, s, gs, cfg, [(0, 400), (5_000_000, 5_000_400), (n - 400, n)])
    assert st['num_hit'] > 0.05 * n


def test_calorimeter_production_whole_event_bitexact(engine_mod, oracle_lib):
    """The headline configuration itself: production kernels (no debug outputs), the
    committed tuning for the workload (persistent form, flattened BVH, distance stack),
    every one of the event's photons against the oracle, and the hits it selects."""
    import json
    import os
    table = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'eic-opticks_amd', 'oxr',
                         'tuning_gfx950.json')
    with open(table) as f:
        tune = json.load(f)['csi_calorimeter']
    s = scenes.csi_calorimeter()
    cfg = abi.default_config(**s.config)
    gs = s.gensteps()
    print('whole calorimeter event, production kernels, tuning', tune, flush=True)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_tuning(**tune)
        n = e.set_gensteps(gs)
        st = e.launch(abi.OXR_FORM_PERSISTENT)
        ph, hit = e.photons(), e.hits()
    assert n > 2_000_000 and len(ph) == n
    print('oracle on', n, 'photons', flush=True)
    want = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(gs, nthreads=min(16, os.cpu_count() or 1),
                                                                  seq=False)['photon']
    assert np.array_equal(bits(ph), bits(want))
    sel = (bits(want)[:, 3, 3] & cfg.hitmask) == cfg.hitmask
    assert st['num_hit'] == len(hit) == int(sel.sum()) > 0
    assert np.array_equal(bits(hit), bits(want[sel]))
