"""The refine trace (PropagateRefine, CSGOptiX/CSGOptiX7.cu:110-216) and the instance
visibility mask (Params vizmask against each instance's mask, SBT.cc:538-545,
Properties.cc:45-55) in the oracle, pinned on the CPU:

* visibility: a launch whose mask drops solid 2's bit gives the photons of the same
  geometry built without solid 2's instances, bit for bit (they are the last instances,
  so every other instance index is unchanged); with the full mask the photons differ;
* refine: a refine distance beyond every hit changes nothing; on a sphere of radius
  19.5 m the refined distance from off-centre starts agrees with the analytic FP64
  distance, and differs from the unrefined one, so the second pass ran and its 0.99
  part was added back; a miss carries the miss program's distance 1, so with a refine
  distance below 0.99 a missing ray is traced again and reports 1 + 0.99."""
import numpy as np

from oxr import abi, gensteps as G, scenes
from oxr.csg import Foundry, sphere


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_hidden_instances_equal_absent_instances(oracle_lib):
    s = scenes.hidden_instances()
    s0 = scenes.hidden_instances(hide=False)
    gs = s.gensteps(scale_photons=0.1)
    full = abi.default_config(**s.config)
    hide = abi.default_config(**s.config, vizmask=0xff & ~(1 << 2))
    a = oracle_lib.Oracle(s.geometry, s.physics, hide).simulate(gs)
    b = oracle_lib.Oracle(s0.geometry, s0.physics, full).simulate(gs)
    c = oracle_lib.Oracle(s.geometry, s.physics, full).simulate(gs)
    assert np.array_equal(bits(a['photon']), bits(b['photon']))
    assert np.array_equal(a['seq'], b['seq'])
    assert not np.array_equal(bits(a['photon']), bits(c['photon']))
    # no photon of the masked launch ever ended on a cube (instances 41..80)
    ii = bits(a['photon'])[:, 1, 3] & 0xffff
    assert not np.any(ii > 40)


def saturating_scene():
    """Nine small spheres, solids 1..9, on the x axis at 100 mm steps inside a world sphere."""
    from oxr.csg import translate
    f = Foundry()
    f.add_material('Rock', 1.0)
    f.add_material('Glass', 1.5)
    b0 = f.boundary('Rock', None, None, 'Rock')
    b1 = f.boundary('Rock', None, None, 'Glass')
    f.begin_solid('r0')
    f.add_prim(sphere(5000.0), b0)
    f.end_solid()
    for k in range(1, 10):
        f.begin_solid(f'r{k}')
        f.add_prim(sphere(10.0 + k), b1)
        f.end_solid()
    f.add_instance(np.eye(4), 0)
    for k in range(1, 10):
        f.add_instance(translate(100.0 * k, 0.0, 0.0), k)
    rays = np.zeros((9, 2, 4), dtype=np.float32)
    rays[:, 0, 0] = 100.0 * np.arange(1, 10)
    rays[:, 0, 1] = -200.0
    rays[:, 1, 1] = 1.0
    rays[:, 1, 3] = 1e6
    return f, rays


def test_visibility_bit_saturates_at_the_eighth_solid(oracle_lib):
    """Solids 7 and above share visibility bit 7 (marker_bit = min(idx, 7)): with the ray
    mask 1 | 1 << 7 a ray through sphere k hits it only for k >= 7, otherwise the world
    (solid 0, bit 0); with 1 | 1 << 3 only sphere 3 is seen; without bit 0 the world is
    hidden too and a ray that sees no sphere misses."""
    f, rays = saturating_scene()
    for mask, seen in ((1 | 1 << 7, {7, 8, 9}), (1 | 1 << 3, {3}), (0xff, set(range(1, 10)))):
        o = oracle_lib.Oracle(f.arrays(), f.physics(), abi.default_config(vizmask=mask))
        prd = o.trace(rays)
        inst = bits(prd)[:, 1, 2] >> 16
        want = [k if k in seen else 0 for k in range(1, 10)]
        assert list(inst) == want, (mask, list(inst))
    o = oracle_lib.Oracle(f.arrays(), f.physics(), abi.default_config(vizmask=1 << 8))
    prd = o.trace(rays)
    assert np.all(bits(prd)[:, 1, 2] == 0xffffffff) and np.all(bits(prd)[:, 1, 3] == 0xffffffff)


def far_sphere_arrays():
    """A water sphere of radius 19.5 m inside a 25 m world sphere."""
    f = Foundry()
    f.add_material('Water', 1.333)
    f.add_material('Rock', 1.0)
    b0 = f.boundary('Rock', None, None, 'Rock')
    b1 = f.boundary('Rock', None, None, 'Water')
    f.begin_solid('r0')
    f.add_prim(sphere(25_000.0), b0)
    f.add_prim(sphere(19_500.0), b1)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    return f.arrays(), f.physics()


def _far_sphere(oracle_lib, **kw):
    geometry, physics = far_sphere_arrays()
    cfg = abi.default_config(**kw)
    return oracle_lib.Oracle(geometry, physics, cfg), cfg


def _analytic(st, R):
    o = st[:, 2, :3].astype(np.float64)
    d = st[:, 3, :3].astype(np.float64)
    d /= np.linalg.norm(d, axis=1)[:, None]
    b = np.sum(o * d, axis=1)
    c = np.sum(o * o, axis=1) - R * R
    return -b + np.sqrt(b * b - c)


def test_refine_off_and_out_of_reach_are_identical(oracle_lib):
    gs = G.frame((3000.0, 2000.0, 1000.0, 100.0), (2, 0, 2, 40))
    a, _ = _far_sphere(oracle_lib)
    b, _ = _far_sphere(oracle_lib, propagate_refine=1, refine_distance=1e9)
    assert np.array_equal(bits(a.simtrace(gs)), bits(b.simtrace(gs)))


def test_refine_distance_matches_the_analytic_sphere(oracle_lib):
    gs = G.frame((3000.0, 2000.0, 1000.0, 100.0), (2, 0, 2, 40))
    plain, _ = _far_sphere(oracle_lib)
    ref, _ = _far_sphere(oracle_lib, propagate_refine=1, refine_distance=5000.0)
    a = plain.simtrace(gs)
    b = ref.simtrace(gs)
    # the same rays and the same surface (boundary word) on both passes
    assert np.array_equal(a[:, 2:], b[:, 2:])
    t = _analytic(b, 19_500.0)
    assert np.all(t > 5000.0 / 0.99)
    err_ref = np.abs(b[:, 0, 3].astype(np.float64) - t)
    err_plain = np.abs(a[:, 0, 3].astype(np.float64) - t)
    # the refined distance is the sum of the second trace's distance and 0.99 of the first:
    # within a few float ulps of 19.5 m of the exact value (a dropped 0.99 part would be off
    # by kilometres, a second trace from the original start would equal the plain one)
    assert err_ref.max() < 0.02
    assert np.mean(b[:, 0, 3] != a[:, 0, 3]) > 0.2
    assert err_plain.max() < 0.05


def test_refined_miss_reports_the_miss_distance(oracle_lib):
    o, cfg = _far_sphere(oracle_lib, propagate_refine=1, refine_distance=0.5)
    gs = G.frame((40_000.0, 0.0, 0.0, 10.0), (0, 0, 0, 16))
    st = o.simtrace(gs)
    # rays from outside the world sphere that point away from it miss twice
    away = st[:, 3, 0] > 0.2
    assert away.sum() > 0
    assert np.all(st[away, 0, 3] == np.float32(1.0) + np.float32(0.99))
    assert np.all(bits(st[away])[:, 2, 3] == 0xffffffff)
