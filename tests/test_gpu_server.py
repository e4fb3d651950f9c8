"""Serving form: continuous batching keeps each event's result identical to running it
alone and to the oracle, survives a failing request, and admits by the HBM budget. The
serving workload's own stream (256 log-normal events on the calorimeter geometry) and
events on the instanced-CSG geometry are served and checked event by event."""
import threading
import time

import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def oracle_hits(o, gs, event_index, hitmask, begin=0, end=-1):
    """The oracle's hits of photons [begin, end) of one event, in photon order."""
    r = o.simulate(gs, event_index=event_index, begin=begin, end=end, seq=False)
    u = bits(r['photon'])
    return u[(u[:, 3, 3] & hitmask) == hitmask]


def serve(e, evs, event0, arrivals=None, **kw):
    """Submit every event (from a feeder thread at the given arrival times, in seconds),
    collect every result; returns ({request id: hits}, {request id: result}, stats)."""
    from oxr.engine import Server
    srv = Server(e, **kw)

    def feed():
        t0 = time.perf_counter()
        for k, gs in enumerate(evs):
            if arrivals is not None:
                d = arrivals[k] - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
            srv.submit(k, gs, event0 + k)

    th = threading.Thread(target=feed)
    th.start()
    hits, res = {}, {}
    for _ in evs:
        r = srv.wait(120000)
        assert r is not None and r['status'] == 0, r
        hits[r['request_id']] = srv.take_hits(r['request_id'], r['num_hit'])
        res[r['request_id']] = r
    th.join()
    st = srv.stats()
    srv.close()
    return hits, res, st


def _events(s, rng, n, lo=500, hi=60000):
    out = []
    for k in range(n):
        m = int(rng.integers(lo, hi))
        pos = rng.uniform(-60, 60, size=(max(1, m // 5000), 3))
        gs = G.torch(np.full(len(pos), max(1, m // len(pos))), pos=pos, wavelength=420.0, type=G.T_POINT_ISOTROPIC)
        out.append(gs)
    return out


@pytest.mark.parametrize('lite', [0, 1])
def test_alone_equals_in_mix(engine_mod, oracle_lib, lite):
    """With photonlite on, single launches run the debug kernels (they write the lite
    items) while batched steps run the production kernels: the hits still agree, and
    every served event's hits equal the oracle's hit selection of that event."""
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
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    o.set_cull(True)
    for k, gs in enumerate(evs):
        assert np.array_equal(bits(got[k]), oracle_hits(o, gs, 100 + k, cfg.hitmask)), k


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


def test_hit_rows_beyond_the_host_block(engine_mod):
    """A step whose hits outnumber the rows its pinned host block was sized for (hit
    fraction set far below the geometry's): the rows beyond it are fetched from the
    device buffer, and every event still equals its lone run bit for bit."""
    s = scenes.instanced_csg()
    cfg = abi.default_config(**s.config)
    rng = np.random.default_rng(15)
    evs = []
    for k in range(4):
        ngs = 2
        u = rng.uniform(size=ngs) ** (1.0 / 3.0) * 17_000.0
        v = rng.normal(size=(ngs, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        evs.append(G.torch(np.full(ngs, 60_000), pos=v * u[:, None], wavelength=430.0, type=G.T_POINT_ISOTROPIC))
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        ref = [e.simulate(gs, event_index=100 + k)[0] for k, gs in enumerate(evs)]
        hits, res, st = serve(e, evs, 100, max_step_photons=1 << 20, hit_fraction=1e-4)
    # the first step's block holds about 4,150 rows; the events hit far more than that
    assert sum(len(h) for h in ref) > 20000
    for k in range(len(evs)):
        assert res[k]['num_hit'] == len(ref[k])
        assert np.array_equal(bits(hits[k]), bits(ref[k])), k


def _check_served(s, cfg, evs, hits, res, alone, o, rng, full_below, sample):
    """Per event: served hits == the event run alone; hits in photon order, each carrying
    its own photon index, all flagged with the hit mask; photon totals conserved. Events
    below `full_below` photons are replayed whole on the oracle, larger ones over one
    random range of `sample` photons."""
    for k, gs in enumerate(evs):
        n = G.num_photons(gs)
        h = bits(hits[k])
        assert res[k]['num_photon'] == n and res[k]['num_hit'] == len(h)
        assert np.array_equal(h, bits(alone[k])), k
        idx = h[:, 3, 2].astype(np.int64)
        assert np.all(np.diff(idx) > 0) and (len(idx) == 0 or (idx[0] >= 0 and idx[-1] < n)), k
        assert np.all((h[:, 3, 3] & cfg.hitmask) == cfg.hitmask), k
        if n <= full_below:
            b, e_ = 0, n
        else:
            b = int(rng.integers(0, n - sample))
            e_ = b + sample
        want = oracle_hits(o, gs, 1000 + k, cfg.hitmask, b, e_)
        got = h[(idx >= b) & (idx < e_)]
        assert np.array_equal(got, want), (k, b, e_)


def test_serving_stream_256_events(engine_mod, oracle_lib):
    """The serving workload on the calorimeter geometry: 256 events, photons log-normal
    (median 200k, sigma 1, clipped to 1k-20M), Poisson arrivals compressed 100x (40/s ->
    4000/s), torch gensteps at random interior origins. Every event: served == alone,
    ordered hits with their own indices, totals conserved; the oracle replays every event
    below 20k photons whole and one random 10k-photon range of each larger one."""
    s = scenes.csi_calorimeter()
    cfg = abi.default_config(**s.config)
    evs, counts, arrivals = scenes.serving_stream(np.random.default_rng(42))
    assert len(evs) == 256 and counts.min() >= 1000 and counts.max() <= 20_000_000
    print('serving stream:', int(counts.sum()), 'photons, largest event', int(counts.max()), flush=True)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        hits, res, st = serve(e, evs, 1000, arrivals=arrivals / 100.0, max_step_photons=4 << 20)
        print('served', int(st['steps']), 'steps', flush=True)
        alone = [e.simulate(gs, event_index=1000 + k)[0] for k, gs in enumerate(evs)]
    assert st['photons'] == int(counts.sum())
    assert st['hits'] == sum(len(h) for h in hits.values())
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    o.set_cull(True)
    _check_served(s, cfg, evs, hits, res, alone, o, np.random.default_rng(7), 20_000, 10_000)


def test_serving_instanced_csg_events(engine_mod, oracle_lib):
    """Served events on the instanced-CSG geometry (20,000 instanced 7-node CSG solids):
    eight events of 20k-400k photons from random interior origins, batched into shared
    persistent steps; each equal to its lone run and, over a sampled range, to the oracle."""
    s = scenes.instanced_csg()
    cfg = abi.default_config(**s.config)
    rng = np.random.default_rng(5)
    evs = []
    for k in range(8):
        m = int(rng.integers(20_000, 400_000))
        ngs = max(1, m // 50_000)
        u = rng.uniform(size=ngs) ** (1.0 / 3.0) * 17_000.0
        v = rng.normal(size=(ngs, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        evs.append(G.torch(np.full(ngs, m // ngs), pos=v * u[:, None], wavelength=430.0, type=G.T_POINT_ISOTROPIC))
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        hits, res, st = serve(e, evs, 1000, max_step_photons=1 << 20)
        alone = [e.simulate(gs, event_index=1000 + k)[0] for k, gs in enumerate(evs)]
    assert st['steps'] >= 2 and st['photons'] == sum(G.num_photons(g) for g in evs)
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    o.set_cull(True)
    _check_served(s, cfg, evs, hits, res, alone, o, np.random.default_rng(8), 0, 4000)
