"""The capacities of the register-resident CSG stacks (RegStacks<3> in oxr_device.h) are
enough for every tree of height <= 2.

intersect_tree (CSG/csg_intersect_tree.h:277-672) walks a complete binary tree in
postorder; a leaf pushes a partial intersect, an operator pops two and either pushes a
result or loops: it pushes the other side back and two tranches (the rest of the walk and
the subtree to re-evaluate from an advanced tmin, csg_intersect_tree.h:589-640). Which
happens at each operator depends on the geometry, so this test enumerates every sequence
of decisions (result, loop left, loop right) up to a bound on the number of loops, and
records the deepest stack and tranche stack reached. The node-index arithmetic of the
walk (next node, subtree slices packed as 8-bit begin/end) is the reference's."""
import pytest


def walk_bounds(h, max_loops, zero=frozenset()):
    best = {'stack': 0, 'tranche': 0, 'paths': 0}

    def nxt(i, elevation):
        return i >> 1 if (i & 1) else (i << elevation) + (1 << elevation)

    def pop(tr, sp, loops):
        if not tr:
            best['paths'] += 1
            return
        tr = list(tr)
        begin, end = tr.pop()
        run(tr, begin, end, sp, loops)

    def run(tr, i, end, sp, loops):
        while i != end:
            depth = i.bit_length() - 1
            elevation = h - depth
            if i in zero:
                i = nxt(i, elevation)
                continue
            if depth == h or ((2 * i) in zero and (2 * i + 1) in zero):
                sp += 1
                best['stack'] = max(best['stack'], sp + 1)
                i = nxt(i, elevation)
                continue
            assert sp >= 1
            run(tr, nxt(i, elevation), end, sp - 1, loops)          # LUT gives a result
            if loops < max_loops:
                left, right = 2 * i, 2 * i + 1
                for loop_tree in (((left << (elevation - 1)) & 0xff, (right << (elevation - 1)) & 0xff),
                                  ((right << (elevation - 1)) & 0xff, i)):
                    t2 = list(tr) + [(i, end), loop_tree]
                    best['tranche'] = max(best['tranche'], len(t2))
                    pop(t2, sp - 1, loops + 1)
            return
        pop(tr, sp, loops)

    pop([((1 << h) & 0xff, 0)], -1, 0)
    return best


@pytest.mark.parametrize('h', [0, 1, 2, 3])
def test_stack_depth_is_height_plus_one(h):
    b = walk_bounds(h, 5 if h < 3 else 4)
    assert b['stack'] == h + 1
    assert b['paths'] > 0


@pytest.mark.parametrize('h', [1, 2])
def test_shallow_trees_fit_three_tranches(h):
    for loops in range(0, 7):
        assert walk_bounds(h, loops)['tranche'] <= 3


def test_height_three_reaches_the_reference_tranche_capacity():
    # deeper trees can need all four tranches of the reference (and more, which it
    # reports as an error), so they keep the private-memory stacks
    assert walk_bounds(3, 3)['tranche'] == 4


def test_zero_nodes_only_lower_the_depth():
    # the pmt-like solid of the instanced workload: node 3 a leaf over two CSG_ZERO nodes
    b = walk_bounds(2, 5, zero=frozenset({6, 7}))
    assert b['stack'] <= 3 and b['tranche'] <= 3
