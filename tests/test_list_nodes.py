"""CPU pins of the oracle's list nodes (CSG_CONTIGUOUS, CSG_DISCONTIGUOUS, CSG_OVERLAP of
CSG/csg_intersect_node.h): the first surface crossing found by a brute-force march over
an inside test written from the shape's definition, from outside and from inside, and
agreement with the equivalent boolean tree of the same leaves."""
import numpy as np
import pytest

from oxr import abi
from oxr.csg import Foundry, box3, sphere, cylinder, contiguous, discontiguous, overlap

from test_oracle_pins import _first_crossing, _prim_oracle, _rays


def in_sphere(p, r, c=(0.0, 0.0, 0.0)):
    return np.sum((p - np.asarray(c)) ** 2, axis=1) < r * r


def in_box(p, full):
    return np.all(np.abs(p) < 0.5 * full, axis=1)


def in_cyl(p, r, z1, z2):
    return (p[:, 0] ** 2 + p[:, 1] ** 2 < r * r) & (p[:, 2] > z1) & (p[:, 2] < z2)


SHAPES = {
    # union of overlapping box and sphere
    'contiguous': (lambda: contiguous(box3(100.0), sphere(40.0, 60.0, 0.0, 0.0)),
                   lambda p: in_box(p, 100.0) | in_sphere(p, 40.0, (60.0, 0, 0)), 120.0),
    # union of three leaves that do not touch
    'discontiguous': (lambda: discontiguous(sphere(30.0, -80.0, 0, 0), box3(40.0), sphere(25.0, 90.0, 10.0, 0)),
                      lambda p: in_sphere(p, 30.0, (-80.0, 0, 0)) | in_box(p, 40.0) | in_sphere(p, 25.0, (90.0, 10.0, 0)),
                      130.0),
    # intersection of box, sphere and cylinder
    'overlap': (lambda: overlap(box3(100.0), sphere(60.0), cylinder(35.0, -80.0, 80.0)),
                lambda p: in_box(p, 100.0) & in_sphere(p, 60.0) & in_cyl(p, 35.0, -80.0, 80.0), 80.0),
    # a tree with a list node as one leaf
    'tree_with_list': (lambda: sphere(70.0) & contiguous(box3(100.0), sphere(40.0, 60.0, 0.0, 0.0)),
                       lambda p: in_sphere(p, 70.0) & (in_box(p, 100.0) | in_sphere(p, 40.0, (60.0, 0, 0))), 100.0),
}


def _rays_through(inside, ext, n, seed, from_inside):
    rng = np.random.default_rng(seed)
    if from_inside:
        pts = rng.uniform(-ext, ext, size=(20 * n, 3))
        o = pts[inside(pts)][:n]
    else:
        o = rng.normal(size=(n, 3))
        o = o / np.linalg.norm(o, axis=1)[:, None] * (2.5 * ext)
    tgt = rng.uniform(-0.6 * ext, 0.6 * ext, size=(len(o), 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    return o, d


@pytest.mark.parametrize('name', list(SHAPES))
@pytest.mark.parametrize('from_inside', [False, True])
def test_list_node_matches_brute_force(oracle_lib, name, from_inside):
    make, inside, ext = SHAPES[name]
    O = _prim_oracle(oracle_lib, make())
    o, d = _rays_through(inside, ext, 120, 11, from_inside)
    want = _first_crossing(inside, o, d, 6 * ext)
    for precision in (0, 1):
        isect, valid = O.intersect_prim(0, _rays(o, d), precision)
        got = np.where(valid.astype(bool), isect[:, 3], np.nan)
        agree = (np.isnan(want) & np.isnan(got)) | (np.abs(got - want) < 2e-3)
        assert agree.mean() >= 0.99, (name, precision, agree.mean(), np.c_[want, got][~agree][:5])
        hit = valid.astype(bool)
        if hit.any():
            assert np.allclose(np.linalg.norm(isect[hit, :3], axis=1), 1.0, atol=1e-4)


@pytest.mark.parametrize('pair', [('contiguous', 'union'), ('overlap', 'intersection')])
def test_list_node_equals_boolean_tree(oracle_lib, pair):
    """A contiguous list is the union of its leaves and an overlap list their
    intersection: the same leaf intersects come out as the boolean tree's."""
    kind, op = pair
    leaves = [box3(100.0), sphere(60.0)] if kind == 'overlap' else [box3(100.0), sphere(40.0, 60.0, 0, 0)]
    lst = overlap(*leaves) if kind == 'overlap' else contiguous(*leaves)
    tree = (leaves[0] & leaves[1]) if op == 'intersection' else (leaves[0] | leaves[1])
    A, B = _prim_oracle(oracle_lib, lst), _prim_oracle(oracle_lib, tree)
    rng = np.random.default_rng(5)
    o = rng.normal(size=(3000, 3))
    o = o / np.linalg.norm(o, axis=1)[:, None] * 300.0
    d = rng.uniform(-50, 50, size=(3000, 3)) - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    ia, va = A.intersect_prim(0, _rays(o, d), 0)
    ib, vb = B.intersect_prim(0, _rays(o, d), 0)
    assert (va == vb).mean() > 0.999
    ok = va.astype(bool) & vb.astype(bool)
    same = np.all(ia[ok] == ib[ok], axis=1)
    assert same.mean() > 0.999


def test_list_prim_layout():
    """Leaves of a list node sit after the tree nodes; subOffset counts from the prim's
    first node and numNode includes them (CSG/CSGNode.h:187-211)."""
    f = Foundry()
    f.add_material('M', 1.0)
    b = f.boundary('M', None, None, 'M')
    f.begin_solid('s')
    f.add_prim(sphere(70.0) & overlap(box3(10.0), sphere(7.0)), b)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    a = f.arrays()
    u = a['node'].view(np.uint32).reshape(-1, 16)
    assert list(u[:, 14]) == [2, 101, 13, 110, 101]
    assert list(u[2, :2]) == [2, 3]
    assert list(a['prim'].view(np.int32).reshape(-1, 16)[0, :2]) == [5, 0]


# ---------------------------------------------------------------- hyperboloid leaf
def test_hyperboloid_closed_form(oracle_lib):
    """x^2+y^2 = r0^2((z/zf)^2+1), r0=50, zf=100, between z=-100 and z=150
    (csg_intersect_leaf_hyperboloid.h)."""
    from oxr.csg import hyperboloid
    H = _prim_oracle(oracle_lib, hyperboloid(50.0, 100.0, -100.0, 150.0))
    A = -50.0 ** 2 / 100.0 ** 2
    r100 = np.sqrt(2500.0 * 2.0)
    rays = _rays([[-300, 0, 0], [-300, 0, 100], [0, 0, -400], [0, 0, 400]],
                 [[1, 0, 0], [1, 0, 0], [0, 0, 1], [0, 0, -1]])
    for precision in (0, 1):
        i, v = H.intersect_prim(0, rays, precision)
        assert v.all()
        assert np.allclose(i[0], [-1, 0, 0, 250], atol=1e-4)            # waist radius
        n = np.array([-r100, 0.0, A * 100.0]); n /= np.linalg.norm(n)
        assert np.allclose(i[1, :3], n, atol=1e-5) and abs(i[1, 3] - (300 - r100)) < 1e-3
        assert np.allclose(i[2], [0, 0, -1, 300], atol=1e-4)            # lower cap
        assert np.allclose(i[3], [0, 0, 1, 250], atol=1e-4)             # upper cap


@pytest.mark.parametrize('from_inside', [False, True])
def test_hyperboloid_matches_brute_force(oracle_lib, from_inside):
    from oxr.csg import hyperboloid
    O = _prim_oracle(oracle_lib, hyperboloid(50.0, 100.0, -100.0, 150.0))

    def inside(p):
        return (p[:, 0] ** 2 + p[:, 1] ** 2 < 2500.0 * ((p[:, 2] / 100.0) ** 2 + 1.0)) & (p[:, 2] > -100.0) & (p[:, 2] < 150.0)
    o, d = _rays_through(inside, 160.0, 120, 13, from_inside)
    want = _first_crossing(inside, o, d, 1000.0)
    isect, valid = O.intersect_prim(0, _rays(o, d), 1)
    got = np.where(valid.astype(bool), isect[:, 3], np.nan)
    agree = (np.isnan(want) & np.isnan(got)) | (np.abs(got - want) < 2e-3)
    assert agree.mean() >= 0.99, (agree.mean(), np.c_[want, got][~agree][:5])
