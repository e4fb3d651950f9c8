"""simtrace (2-D geometry cross-sections): the FRAME genstep builder
(SFrameGenstep::MakeCenterExtentGenstep) and the oracle's restatement of
qsim::generate_photon_simtrace_frame + trace + sevent::add_simtrace, pinned by closed
forms on a sphere and by the Philox stream."""
import numpy as np
import pytest

from oxr import abi, gensteps as G
from oxr.csg import Foundry, sphere


def sphere_oracle(oracle_lib, r=100.0):
    f = Foundry()
    f.add_material('Air', 1.0)
    f.add_material('Glass', 1.5)
    f.add_material('World', 1.0)
    b_world = f.boundary('World', None, None, 'World')
    b_ball = f.boundary('World', None, None, 'Glass')
    f.begin_solid('r0')
    f.add_prim(sphere(1000.0), b_world)
    f.add_prim(sphere(r), b_ball)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    cfg = abi.default_config()
    return oracle_lib.Oracle(f.arrays(), f.physics(), cfg), cfg


def test_frame_gensteps_layout():
    ce = (10.0, -5.0, 2.0, 50.0)
    gs = G.frame(ce, (3, 0, 2, 7), gridscale=0.2)
    u = gs.view(np.uint32)
    assert gs.shape == (7 * 1 * 5, 6, 4)
    assert np.all(u[:, 0, 0] == G.GS_FRAME) and np.all(u[:, 0, 1] == G.AX_XZ) and np.all(u[:, 0, 3] == 7)
    # grid order ix, iy, iz; the first genstep is (-3, 0, -2)
    ids = u[:, 0, 2].astype(np.uint32).view(np.int8).reshape(-1, 4)
    assert list(ids[0]) == [-3, 0, -2, 0] and list(ids[-1]) == [3, 0, 2, 0]
    # transform: translate(grid * extent * gridscale) then translate(ce)
    step = 50.0 * 0.2
    assert np.allclose(gs[0, 5, :3], [10.0 - 3 * step, -5.0, 2.0 - 2 * step])
    assert np.allclose(gs[:, 2:5, :3], np.eye(3)[None])
    assert np.allclose(gs[:, 1], [0, 0, 0, 1])
    assert G.grid_axes(0, 2, 2) == G.AX_YZ and G.grid_axes(2, 2, 0) == G.AX_XY and G.grid_axes(1, 1, 1) == G.AX_XYZ
    assert G.genstep_id(-1, 0, 2, 0) == 0x000200ff
    # standardised form with high resolution: grid points multiply, spacing divides
    hi = G.frame(ce, (-1, 1, 0, 0, -1, 1, 5, 2))
    assert len(hi) == 5 * 5 and np.allclose(hi[1, 5, :3] - hi[0, 5, :3], [0, 0, 25.0])


def test_simtrace_closed_form_on_sphere(oracle_lib):
    """XZ grid inside a sphere of radius 100: every ray leaves through the sphere, so
    |pos + t dir| = 100 with the normal along the hit position; directions are the unit
    vectors (cos 2 pi u0, 0, sin 2 pi u0) of each ray's own Philox stream."""
    O, cfg = sphere_oracle(oracle_lib)
    gs = G.frame((0.0, 0.0, 0.0, 100.0), (2, 0, 2, 40), gridscale=0.2)
    st = O.simtrace(gs)
    assert st.shape == (25 * 40, 4, 4)
    t = st[:, 0, 3]
    hit = st[:, 1, :3].astype(np.float64)
    start, d = st[:, 2, :3].astype(np.float64), st[:, 3, :3].astype(np.float64)
    assert np.all(t > 0)
    assert np.allclose(np.linalg.norm(hit, axis=1), 100.0, atol=2e-3)
    n = st[:, 0, :3] / np.linalg.norm(st[:, 0, :3], axis=1)[:, None]
    assert np.allclose(n, hit / 100.0, atol=1e-4)
    assert np.allclose(start + t[:, None] * d, hit, atol=1e-3)
    assert np.all(st[:, 1, 3] == np.float32(cfg.tmin))
    assert np.all(d[:, 1] == 0.0) and np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # start points are the grid points, rays of genstep g are slots g*40 .. g*40+39
    assert np.allclose(start[::40, 0].reshape(5, 5)[:, 0], [-40, -20, 0, 20, 40])
    u0 = np.array([oracle_lib.uniforms(2, seed=cfg.rng_seed, photon=i)[0] for i in range(len(st))])
    phi = (2.0 * np.pi) * u0.astype(np.float64)
    assert np.allclose(d[:, 0], np.cos(phi), atol=2e-6) and np.allclose(d[:, 2], np.sin(phi), atol=2e-6)
    # identity words: boundary of the ball in the low 16 bits, global prim 1 above
    assert np.all(st[:, 2, 3].view(np.uint32) == (1 << 16) | 1)


def test_simtrace_xyz_and_miss(oracle_lib):
    """A non-planar grid draws directions over the sphere; rays from outside every solid
    miss: the miss program's q0 (mid-grey background as normal, distance 1,
    CSGOptiX7.cu:655-682 with SBT.cc:180-192), so the "hit" position is one step along the
    direction from the start, and the identity words are all ones."""
    O, cfg = sphere_oracle(oracle_lib)
    gs = G.frame((0.0, 0.0, 0.0, 10.0), (1, 1, 1, 50))
    st = O.simtrace(gs)
    d = st[:, 3, :3].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6) and abs(d[:, 2].mean()) < 0.1
    far = G.frame((5000.0, 0.0, 0.0, 10.0), (0, 1, 1, 20))
    st = O.simtrace(far)
    assert np.all(st[:, 0, :3] == np.float32(0.6)) and np.all(st[:, 0, 3] == 1.0)
    assert np.array_equal(st[:, 1, :3], st[:, 2, :3] + st[:, 3, :3])
    assert np.all(st[:, 2, 3].view(np.uint32) == 0xffffffff) and np.all(st[:, 3, 3].view(np.uint32) == 0xffffffff)
