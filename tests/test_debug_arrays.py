"""The per-bounce debug arrays of the oracle: rec (srec, sysrap/srec.h:39-43, filled by
sevent::add_rec, sysrap/sevent.h:614-622) and prd (quad2 of sctx::trace, sysrap/sctx.h:170-173).

rec is pinned against the full-precision record array of the same run through srec's own
decoding (srec.h:148-191): every decoded coordinate lies within half a quantisation step
of the record. prd is pinned against the records around it: a unit normal, the boundary
that propagate then stamps into the next record point, the traced distance to the next
point of a boundary step, and a prd slot exactly where a next record point exists."""
import numpy as np
import pytest

from oxr import abi, scenes

FL = abi.FLAGS
BOUNDARY_FLAGS = (FL['BOUNDARY_TRANSMIT'], FL['BOUNDARY_REFLECT'], FL['SURFACE_DETECT'], FL['SURFACE_ABSORB'],
                  FL['SURFACE_DREFLECT'], FL['SURFACE_SREFLECT'])
BULK_FLAGS = (FL['BULK_ABSORB'], FL['BULK_SCATTER'], FL['BULK_REEMIT'])


def decode_rec(rec, extent, tdom):
    """srec getters: post as int16 x,y,z,t and polw as int8 x,y,z,w."""
    post = np.ascontiguousarray(rec[..., :2]).view(np.int16).astype(np.float64)
    polw = np.ascontiguousarray(rec[..., 2:3]).view(np.int8).astype(np.float64)
    pos = post[..., :3] * extent / 32767.0
    t = post[..., 3] * tdom / 32767.0
    pol = polw[..., :3] / 127.0
    wl = polw[..., 3] * 380.0 / 127.0 + 440.0
    return pos, t, pol, wl


def run(oracle_lib, name, scale, **cfg_kw):
    s = getattr(scenes, name)()
    cfg = abi.default_config(**s.config, **cfg_kw)
    gs = s.gensteps(scale_photons=scale)
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    return o.simulate(gs, max_record=cfg_kw.get('max_record', 0), max_rec=cfg_kw.get('max_rec', 0),
                      max_prd=cfg_kw.get('max_prd', 0)), cfg


@pytest.mark.parametrize('name,scale,tdom', [('nested_boxes', 0.02, 100.0), ('csi_calorimeter', 0.001, 2.0e4)])
def test_rec_decodes_to_records(oracle_lib, name, scale, tdom):
    r, cfg = run(oracle_lib, name, scale, max_record=12, max_rec=12, time_domain=tdom)
    rec, record = r['rec'], r['record'].astype(np.float64)
    live = np.any(r['record'].view(np.uint32).reshape(*record.shape[:2], 16) != 0, axis=2)
    assert live.sum() > 1000
    assert np.all(rec[~live] == 0) and np.all(rec[..., 3] == 0)
    pos, t, pol, wl = decode_rec(rec, cfg.extent_domain, cfg.time_domain)
    assert np.abs(record[..., 0, :3]).max() < cfg.extent_domain and record[..., 0, 3].max() < tdom
    # half a step, plus the float32 rounding of the scaled value before it is rounded to
    # an integer (two roundings of at most 2^-24 relative on values up to 32767: 0.004 step)
    h = 0.5 + 0.004
    assert np.abs(pos - record[..., 0, :3])[live].max() <= h * cfg.extent_domain / 32767.0
    assert np.abs(t - record[..., 0, 3])[live].max() <= h * tdom / 32767.0
    assert np.abs(pol - record[..., 2, :3])[live].max() <= h / 127.0
    assert np.abs(wl - record[..., 2, 3])[live].max() <= h * 380.0 / 127.0


def test_rec_time_wraps_outside_its_domain(oracle_lib):
    """Scintillation times run far past the default 10 ns time domain: the int16 keeps the
    low 16 bits of the rounded value, so the stored and the true step counts agree modulo
    65536 (20 ns). Checked where the float32 scaling is exact to one step (t < 2 us)."""
    r, cfg = run(oracle_lib, 'csi_calorimeter', 0.001, max_record=12, max_rec=12)
    record = r['record'].astype(np.float64)
    live = np.any(r['record'].view(np.uint32).reshape(*record.shape[:2], 16) != 0, axis=2)
    t = record[..., 0, 3]
    k = np.ascontiguousarray(r['rec'][..., 1:2]).view(np.int16)[..., 1].astype(np.int64)
    want = np.rint(t * 32767.0 / cfg.time_domain).astype(np.int64)
    sel = live & (t < 2000.0)
    assert (t[sel] > 4 * cfg.time_domain).sum() > 100
    d = (want - k)[sel]
    assert np.all(np.abs(((d + 32768) % 65536) - 32768) <= 1)


@pytest.mark.parametrize('name,scale', [('nested_boxes', 0.02), ('csi_calorimeter', 0.001)])
def test_prd_matches_the_steps_it_traced(oracle_lib, name, scale):
    r, _ = run(oracle_lib, name, scale, max_record=11, max_prd=10)
    prd, record = r['prd'].astype(np.float64), r['record'].astype(np.float64)
    rbits = r['record'].view(np.uint32)
    live_rec = np.any(rbits.reshape(*record.shape[:2], 16) != 0, axis=2)
    live_prd = np.any(r['prd'].view(np.uint32).reshape(*prd.shape[:2], 8) != 0, axis=2)
    # a bounce is traced with a boundary exactly when propagate writes the next point
    assert np.array_equal(live_prd, live_rec[:, 1:])
    assert live_prd.sum() > 1000
    n = prd[..., 0, :3][live_prd]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=2e-6)
    gpbd = r['prd'][..., 1, 3].view(np.uint32)
    nxt = rbits[:, 1:, 3, 0]                                   # orient_boundary_flag of the next point
    assert np.array_equal((gpbd & 0xffff)[live_prd], ((nxt >> 16) & 0x7fff)[live_prd])
    flag = nxt & 0xffff
    t = prd[..., 0, 3]
    step = np.linalg.norm(record[:, 1:, 0, :3] - record[:, :-1, 0, :3], axis=2)
    at_boundary = live_prd & np.isin(flag, BOUNDARY_FLAGS)
    in_bulk = live_prd & np.isin(flag, BULK_FLAGS)
    assert at_boundary.sum() > 500
    assert np.allclose(step[at_boundary], t[at_boundary], rtol=1e-5, atol=2e-3)
    assert np.all(step[in_bulk] <= t[in_bulk] * (1 + 1e-5) + 2e-3)
    # identity of the hit, as sphoton::set_prd copies it (sysrap/sphoton.h:482-488)
    ii = r['prd'][..., 1, 2].view(np.uint32)
    ident = rbits[:, 1:, 3, 1]
    assert np.array_equal((ii & 0xffff)[at_boundary], ident[at_boundary])
