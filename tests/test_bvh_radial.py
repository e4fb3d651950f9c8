"""The BVH builder's radial bounds (eic-opticks_amd/csrc/oxr_bvh.cpp), pinned on the CPU by
a small C++ program built against the builder: for random spheres and boxes, every point
sampled on an item's surface (sphere) or in its box (any other item) lies, for every node
on the item's root path, within the stored [r_in, r_out] of the child holding it, measured
from the node's stored centre; concentric spheres get the exact (widened) radii; and the
split of coincident centroids groups nested spheres by radius."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, 'eic-opticks_amd', 'csrc')

PROGRAM = r'''
#include "oxr_bvh.h"
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
using namespace oxh;

static int check(const std::vector<Box>& boxes, const std::vector<RadItem>& items, unsigned seed) {
    Bvh b = build_bvh(boxes, 0, 0.0, &items);
    if (b.rad.size() != b.nodes.size()) { std::printf("size mismatch\n"); return 1; }
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    // walk every node and every item below each child
    int bad = 0;
    std::vector<std::vector<int>> below(b.nodes.size() * 2);
    // gather items per child by descending from each child
    for (size_t n = 0; n < b.nodes.size(); n++) {
        for (int c = 0; c < 2; c++) {
            std::vector<int> todo{b.nodes[n].ch[c]};
            while (!todo.empty()) {
                int r = todo.back(); todo.pop_back();
                if (r < 0) below[2 * n + c].push_back(~r);
                else { todo.push_back(b.nodes[size_t(r)].ch[0]); todo.push_back(b.nodes[size_t(r)].ch[1]); }
            }
        }
    }
    for (size_t n = 0; n < b.nodes.size(); n++) {
        const RadOut& R = b.rad[n];
        for (int c = 0; c < 2; c++) {
            for (int it : below[2 * n + c]) {
                for (int s = 0; s < 64; s++) {
                    double p[3];
                    if (items[size_t(it)].sphere) {
                        double v[3] = {U(rng), U(rng), U(rng)};
                        double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                        for (int k = 0; k < 3; k++) p[k] = items[size_t(it)].c[k] + items[size_t(it)].r * v[k] / l;
                    } else {
                        for (int k = 0; k < 3; k++) p[k] = boxes[size_t(it)].lo[k] + (0.5 + 0.5 * U(rng)) * boxes[size_t(it)].extent(k);
                    }
                    double d = std::sqrt((p[0] - R.c[0]) * (p[0] - R.c[0]) + (p[1] - R.c[1]) * (p[1] - R.c[1]) +
                                         (p[2] - R.c[2]) * (p[2] - R.c[2]));
                    if (d < R.r[2 * c] || d > R.r[2 * c + 1]) bad++;
                }
            }
        }
    }
    return bad;
}

int main() {
    std::mt19937_64 rng(42);
    std::uniform_real_distribution<double> U(-1000.0, 1000.0), S(1.0, 200.0);
    // random mix of spheres and boxes
    std::vector<Box> boxes;
    std::vector<RadItem> items;
    for (int i = 0; i < 300; i++) {
        RadItem it;
        double c[3] = {U(rng), U(rng), U(rng)}, r = S(rng);
        it.sphere = (i % 3) != 0;
        for (int k = 0; k < 3; k++) it.c[k] = c[k];
        it.r = r;
        Box b;
        for (int k = 0; k < 3; k++) { b.lo[k] = c[k] - r; b.hi[k] = c[k] + r * (it.sphere ? 1.0 : 0.5); }
        boxes.push_back(b);
        items.push_back(it);
    }
    int bad1 = check(boxes, items, 1);
    // 50 concentric spheres: the root's children split by radius, bounds equal the radii
    std::vector<Box> cb;
    std::vector<RadItem> ci;
    for (int i = 0; i < 50; i++) {
        double r = 1000.0 + 390.0 * i;
        RadItem it; it.sphere = true; it.c[0] = it.c[1] = it.c[2] = 0.0; it.r = r;
        Box b;
        for (int k = 0; k < 3; k++) { b.lo[k] = -r; b.hi[k] = r; }
        cb.push_back(b); ci.push_back(it);
    }
    int bad2 = check(cb, ci, 2);
    Bvh bv = build_bvh(cb, 0, 0.0, &ci);
    const RadOut& R = bv.rad[0];
    std::printf("bad %d %d root %.3f %.3f %.3f %.3f %.3f %.3f %.3f\n", bad1, bad2, R.c[0], R.c[1], R.c[2], R.r[0], R.r[1],
                R.r[2], R.r[3]);
    return 0;
}
'''


def test_radial_bounds_hold_every_surface(tmp_path):
    src = tmp_path / 'radial.cpp'
    src.write_text(PROGRAM)
    exe = tmp_path / 'radial'
    subprocess.check_call(['g++', '-O2', '-std=c++17', '-I', CSRC, str(src), os.path.join(CSRC, 'oxr_bvh.cpp'), '-o', str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout.split()
    assert out[0] == 'bad' and out[1] == '0' and out[2] == '0', out
    cx, cy, cz, a0, b0, a1, b1 = map(float, out[4:11])
    assert abs(cx) < 1e-3 and abs(cy) < 1e-3 and abs(cz) < 1e-3
    # the root's children are the inner and the outer radii, each range tight up to its margin
    lo, hi = sorted([(a0, b0), (a1, b1)])
    assert lo[0] <= 1000.0 and lo[0] > 999.0 and hi[1] >= 1000.0 + 390.0 * 49 and hi[1] < 1000.0 + 390.0 * 49 + 2.0
    assert lo[1] < hi[0] + 400.0 and lo[1] <= hi[0] + 1.0   # adjacent radius groups
