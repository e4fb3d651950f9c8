"""Seeded random CSG trees (union, intersection and difference over spheres, z-cut spheres,
cylinders, cones and boxes, each leaf moved by a random translation and some rotated), up
to depth 3: every prim's GPU intersect, through the C ABI, equals the oracle's bit for
bit for random rays, with single- and double-precision roots."""
import numpy as np
import pytest

from oxr import abi
from oxr.csg import Foundry, box3, sphere, zsphere, cylinder, cone, translate, rotate_z_to

from test_gpu_parity import bits, random_rays

pytestmark = pytest.mark.gpu


def random_leaf(rng):
    k = int(rng.integers(0, 6))
    if k == 0:
        t = sphere(float(rng.uniform(20, 80)))
    elif k == 1:
        r = float(rng.uniform(30, 80))
        t = zsphere(r, -float(rng.uniform(0.1, 0.9)) * r, float(rng.uniform(0.1, 0.9)) * r)
    elif k == 2:
        t = cylinder(float(rng.uniform(10, 50)), -float(rng.uniform(20, 80)), float(rng.uniform(20, 80)))
    elif k == 3:
        t = box3(*[float(x) for x in rng.uniform(20, 120, size=3)])
    elif k == 4:
        t = cone(float(rng.uniform(20, 60)), -float(rng.uniform(20, 60)), float(rng.uniform(5, 40)),
                 float(rng.uniform(20, 60)))
    else:
        t = box3(float(rng.uniform(30, 100))).moved(rotate_z_to(rng.normal(size=3)))
    return t.moved(translate(*[float(x) for x in rng.uniform(-30, 30, size=3)]))


def random_tree(rng, depth):
    if depth == 0 or rng.random() < 0.2:
        return random_leaf(rng)
    a, b = random_tree(rng, depth - 1), random_tree(rng, depth - 1)
    op = int(rng.integers(0, 3))
    return a | b if op == 0 else (a & b if op == 1 else a - b)


def random_foundry(seed, n):
    rng = np.random.default_rng(seed)
    f = Foundry()
    f.add_material('M', 1.0)
    b = f.boundary('M', None, None, 'M')
    f.begin_solid('random')
    for _ in range(n):
        f.add_prim(random_tree(rng, int(rng.integers(1, 4))), b)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    return f


@pytest.mark.parametrize('fp64', [0, 1])
@pytest.mark.parametrize('seed', [101, 202])
def test_random_trees_bitexact(engine_mod, oracle_lib, seed, fp64):
    n = 24
    f = random_foundry(seed, n)
    cfg = abi.default_config(fp64_roots=fp64)
    rng = np.random.default_rng(seed + 1)
    o = oracle_lib.Oracle(f.arrays(), f.physics(), cfg)
    print('random CSG trees, seed', seed, 'fp64', fp64, flush=True)
    hits = 0
    with engine_mod.Engine(f.arrays(), f.physics(), cfg) as e:
        for p in range(n):
            rays = random_rays(rng, 3000, 120.0)
            gi, gv = e.intersect_prim(p, rays)
            oi, ov = o.intersect_prim(p, rays, 0)
            assert np.array_equal(gv, ov), (seed, p)
            assert np.array_equal(bits(gi), bits(oi)), (seed, p)
            hits += int(gv.sum())
    assert hits > n * 3000 * 0.05
