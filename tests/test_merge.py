"""CPU pins of the hit-merge restatement (oracle.merge) and of the oracle's photonlite
output. The merge cases are small enough to work out by hand from the reference's
key_functor / reduce_op definitions (sysrap/sphoton.h:277-298, sphotonlite.h:68-90)."""
import numpy as np

from oxr import abi, scenes

F = abi.FLAGS
SD, SA, BT = F['SURFACE_DETECT'], F['SURFACE_ABSORB'], F['BOUNDARY_TRANSMIT']


def sphoton(ident, time, fm, index, iindex=3):
    u = np.zeros(16, dtype=np.uint32)
    u[3] = np.float32(time).view(np.uint32)
    u[7] = (1 << 16) | iindex
    u[13] = ident
    u[14] = index
    u[15] = fm
    return u


def lite(ident, time, fm, lpos=0x12345678):
    return np.array([(1 << 16) | ident, np.float32(time).view(np.uint32), lpos, fm], dtype=np.uint32)


def t_of(u, w):
    return u[:, w].view(np.float32)


def test_merge_sphoton_by_hand(oracle_lib):
    items = np.array([
        sphoton(5, 1.2, SD, 0),
        sphoton(5, 0.3, SD | BT, 1),
        sphoton(9, 0.4, BT, 2),          # not selected (no SD bit)
        sphoton(5, 1.9, SD, 3),
        sphoton(7, 0.5, SD, 4),
        sphoton(5, 0.7, SD | SA, 5),
    ])
    out = oracle_lib.merge(items, 1.0, mask=SD)
    # keys in order: (5,b0) (5,b1) (7,b0)
    assert list(out[:, 14]) == [1, 0, 4]                   # earliest item of each run kept
    assert list(out[:, 15]) == [SD | BT | SA, SD, SD]        # flagmasks OR-ed
    assert list(out[:, 7] >> 16) == [2, 2, 1]                # hitcounts summed
    assert list(out[:, 7] & 0xffff) == [3, 3, 3]
    assert np.allclose(t_of(out, 3), [0.3, 1.2, 0.5])
    # window 0: the selection, photon order, unmerged
    sel = oracle_lib.merge(items, 0.0, mask=SD)
    assert list(sel[:, 14]) == [0, 1, 3, 4, 5]


def test_merge_sphoton_ties_and_identity_truncation(oracle_lib):
    # equal times: the earlier item in input order stays (a_first on fminf(a,b)==a);
    # identity bits above 16 fall off the 64-bit key (uint64(id) << 48)
    items = np.array([
        sphoton(5, 2.5, SD, 0),
        sphoton(5 | (1 << 16), 2.5, SD, 1),
        sphoton(5, 2.25, SD, 2),
    ])
    out = oracle_lib.merge(items, 1.0, mask=SD)
    assert len(out) == 1
    assert out[0, 14] == 2 and (out[0, 7] >> 16) == 3
    out = oracle_lib.merge(items[:2], 1.0, mask=SD)
    assert len(out) == 1 and out[0, 14] == 0


def test_merge_lite_by_hand(oracle_lib):
    items = np.array([
        lite(5, 1.2, SD, 0xaaaa0001),
        lite(5, 1.1, SD | BT, 0xbbbb0002),
        lite(4, 3.5, SD, 0xcccc0003),
        lite(5, 0.9, SD, 0xdddd0004),
    ])
    out = oracle_lib.merge(items, 1.0, mask=SD)
    # keys: (4,b3) < (5,b0) < (5,b1)
    assert list(out[:, 0] & 0xffff) == [4, 5, 5]
    assert list(out[:, 0] >> 16) == [1, 1, 2]
    # the run's first item keeps its lpos word; time is the run minimum
    assert list(out[:, 2]) == [0xcccc0003, 0xdddd0004, 0xaaaa0001]
    assert np.allclose(t_of(out, 1), [3.5, 0.9, 1.1])
    assert list(out[:, 3]) == [SD, SD, SD | BT]


def test_merge_idempotent_and_conserves_counts(oracle_lib):
    rng = np.random.default_rng(7)
    n = 3000
    ids = rng.integers(0, 40, size=n)
    times = rng.exponential(20.0, size=n).astype(np.float32)
    fms = np.where(rng.random(n) < 0.7, SD, BT) | np.where(rng.random(n) < 0.5, SA, 0)
    items = np.array([sphoton(int(i), float(t), int(f), k) for k, (i, t, f) in enumerate(zip(ids, times, fms))])
    once = oracle_lib.merge(items, 2.0, mask=SD)
    twice = oracle_lib.merge(once, 2.0)
    assert np.array_equal(once, twice)
    assert int((once[:, 7] >> 16).sum()) == int(((items[:, 15] & SD) != 0).sum())
    keys = (once[:, 13].astype(np.uint64) << np.uint64(48)) | oracle_lib.time_bucket(t_of(once, 3), 2.0)
    assert np.all(keys[1:] > keys[:-1])


def test_photonlite_fields_closed_form(oracle_lib):
    """On the raindrop KAT every photon ends detected on the water cube (identity
    transform): the last intersect's local position is the end position, so the packed
    lposcost/lposfphi equal 0xffff*(z/|p|) and 0xffff*(atan2(y,x)+pi)/2pi rounded."""
    s = scenes.raindrop()
    cfg = abi.default_config(**s.config)
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    r = o.simulate(s.gensteps(), input_photons=scenes.raindrop_input_photons(), lite=True)
    ph, lt = r['photon'], r['lite']
    u = ph.view(np.uint32)
    assert np.all(lt[:, 3] == u[:, 3, 3])
    assert np.all(lt[:, 1] == u[:, 0, 3])
    assert np.all(lt[:, 0] == (0x10000 | (u[:, 3, 1] & 0xffff)))
    p = ph[:, 0, :3].astype(np.float64)
    cost = p[:, 2] / np.linalg.norm(p, axis=1)
    fphi = (np.arctan2(p[:, 1], p[:, 0]) + np.pi) / (2 * np.pi)
    want_c = np.floor(np.clip(cost, 0, None) * 65535 + 0.5)
    want_f = np.floor(fphi * 65535 + 0.5)
    assert np.all(np.abs((lt[:, 2] >> 16).astype(np.int64) - want_c) <= 1)
    assert np.all(np.abs((lt[:, 2] & 0xffff).astype(np.int64) - want_f) <= 1)
