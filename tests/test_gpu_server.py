"""Serving form: continuous batching keeps each event's result identical to running it
alone, survives a failing request, and admits by the HBM budget."""
import time

import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _events(s, rng, n, lo=500, hi=60000):
    out = []
    for k in range(n):
        m = int(rng.integers(lo, hi))
        pos = rng.uniform(-60, 60, size=(max(1, m // 5000), 3))
        gs = G.torch(np.full(len(pos), max(1, m // len(pos))), pos=pos, wavelength=420.0, type=G.T_POINT_ISOTROPIC)
        out.append(gs)
    return out


@pytest.mark.parametrize('lite', [0, 1])
def test_alone_equals_in_mix(engine_mod, lite):
    """With photonlite on, single launches run the debug kernels (they write the lite
    items) while batched steps run the production kernels: the hits still agree."""
    s = scenes.csi_calorimeter()
    cfg = abi.default_config(**s.config, photonlite=lite)
    rng = np.random.default_rng(12)
    evs = _events(s, rng, 12)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        alone = []
        for k, gs in enumerate(evs):
            h, _ = e.simulate(gs, event_index=100 + k)
            alone.append(h)
        srv = engine_mod.Server(e, max_step_photons=50000, max_active=5)
        for k, gs in enumerate(evs):
            srv.submit(k, gs, 100 + k)
        got = {}
        for _ in evs:
            r = srv.wait(60000)
            assert r is not None and r['status'] == 0
            got[r['request_id']] = srv.take_hits(r['request_id'], r['num_hit'])
            assert r['t_done'] >= r['t_admit'] >= r['t_submit']
        st = srv.stats()
        srv.close()
    for k in range(len(evs)):
        assert np.array_equal(bits(got[k]), bits(alone[k])), k
    assert st['steps'] >= 2 and st['photons'] == sum(G.num_photons(g) for g in evs)


def test_failed_request_does_not_disturb_others(engine_mod):
    s = scenes.csi_calorimeter()
    cfg = abi.default_config(**s.config)
    rng = np.random.default_rng(13)
    evs = _events(s, rng, 3)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        ref = [e.simulate(gs, event_index=k)[0] for k, gs in enumerate(evs)]
        srv = engine_mod.Server(e, max_step_photons=40000)
        bad = G.cerenkov([10], [[0, 0, 0]], [[0, 0, 1.0]], 1.0, 0.9, 99999, 200.0, 700.0, 1.5, 10.0)
        with pytest.raises(engine_mod.OxrError):
            srv.submit(99, bad, 0)                 # rejected on submit by table validation
        with pytest.raises(engine_mod.OxrError):
            srv.submit(98, G.input_photon(10), 0)  # no input-photon array per request
        for k, gs in enumerate(evs):
            srv.submit(k, gs, k)
        srv.submit(50, np.zeros((0, 6, 4), np.float32), 7)   # empty request completes at once
        done = {}
        for _ in range(len(evs) + 1):
            r = srv.wait(60000)
            done[r['request_id']] = r
        for k in range(len(evs)):
            assert np.array_equal(bits(srv.take_hits(k, done[k]['num_hit'])), bits(ref[k]))
        assert done[50]['num_photon'] == 0 and done[50]['status'] == 0
        with pytest.raises(engine_mod.OxrError):
            srv.take_hits(12345, 0)
        srv.close()


def test_admission_budget_serialises_when_tight(engine_mod):
    s = scenes.csi_calorimeter()
    cfg = abi.default_config(**s.config)
    rng = np.random.default_rng(14)
    evs = _events(s, rng, 4, lo=20000, hi=30000)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        per = max(G.num_photons(g) for g in evs) * (4 + 64 + 16) + 2000
        srv = engine_mod.Server(e, max_step_photons=1 << 20, hbm_budget_bytes=per)
        for k, gs in enumerate(evs):
            srv.submit(k, gs, k)
        rs = [srv.wait(60000) for _ in evs]
        srv.close()
    # with room for one event at a time, no event is admitted before the previous one is done
    rs.sort(key=lambda r: r['t_admit'])
    for a, b in zip(rs, rs[1:]):
        assert b['t_admit'] >= a['t_done'] - 1e-6
