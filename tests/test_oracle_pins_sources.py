"""Pins of the oracle's generation and transport corners that the bulk pins leave open:
every reference torch type, bulk re-emission, wavelength interpolation of the boundary
table and the halfspace leaf.

Each check compares the oracle with geometry or statistics fixed by the reference's
documented semantics (sysrap/storch.h:189-516, qudarap/qsim.h:177-183,804-825,
qudarap/qbnd.h:103-125, CSG/csg_intersect_leaf_halfspace.h:155-187), computed here with
numpy from the Philox stream (itself pinned by the Random123 vectors) or with textbook
routines (np.interp, ray marching), so that a dropped term, a swapped draw or a transposed
operand in the oracle fails.
"""
import numpy as np
import pytest

from oxr import abi, csg, gensteps as G, scenes
from oxr.csg import Foundry, box3, sphere, cylinder, halfspace

TWOPI = 2.0 * np.pi


def rotate_uz(v, u):
    """CLHEP Hep3Vector::rotateUz: the rotation taking the z axis onto the unit vector u
    (identity, or a flip of x and z, when u is on the z axis)."""
    v = np.asarray(v, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    up = u[0] ** 2 + u[1] ** 2
    if up > 0:
        up = np.sqrt(up)
        return np.array([(u[0] * u[2] * v[0] - u[1] * v[1]) / up + u[0] * v[2],
                         (u[1] * u[2] * v[0] + u[0] * v[1]) / up + u[1] * v[2],
                         -up * v[0] + u[2] * v[2]])
    return np.array([-v[0], v[1], -v[2]]) if u[2] < 0 else v.copy()


def frame(u):
    """(e1, e2, u): the images of x, y, z under rotate_uz, checked to be a right-handed
    orthonormal frame, so every test below uses a rotation and not an arbitrary map."""
    e1, e2, e3 = (rotate_uz(v, u) for v in np.eye(3))
    assert np.allclose(e3, u / np.linalg.norm(u), atol=1e-12)
    assert np.allclose([e1 @ e1, e2 @ e2, e1 @ e2, e1 @ e3], [1, 1, 0, 0], atol=1e-12)
    assert np.allclose(np.cross(e1, e2), e3, atol=1e-12)
    return e1, e2, e3


def generated(oracle_lib, gs, slot_offset=0, event_index=0):
    s = scenes.nested_boxes()
    c = abi.default_config(**s.config)
    c.max_bounce = 0
    return oracle_lib.Oracle(s.geometry, s.physics, c).simulate(gs, seq=False, slot_offset=slot_offset,
                                                                   event_index=event_index)['photon']


def stream(oracle_lib, photon, n=8, event=0):
    return oracle_lib.uniforms(n, skip=100000, event=event, photon=photon).astype(np.float64)


def chi2_per_dof(x, lo, hi, pdf, bins=40):
    h, e = np.histogram(x, bins=bins, range=(lo, hi))
    c = 0.5 * (e[1:] + e[:-1])
    w = pdf(c)
    exp = w / w.sum() * len(x)
    return np.sum((h - exp) ** 2 / exp) / (bins - 1)


def torch_flags_ok(p, gs):
    u = p.view(np.uint32)
    assert np.all(u[:, 3, 0] & 0xffff == abi.FLAGS['TORCH']) and np.all(u[:, 3, 3] == abi.FLAGS['TORCH'])
    assert np.all(p[:, 2, 3] == gs[0, 3, 3]) and np.all(p[:, 0, 3] == gs[0, 1, 3])


# ---------------------------------------------------------------- torch types
def test_torch_disc(oracle_lib):
    """T_DISC (storch.h:204-244): radius R*u_zenith, phi = 2 pi u_azimuth over the disc
    through gs.pos perpendicular to mom; pol = radial x mom (the fixed phase choice)."""
    mom = np.array([1.0, -2.0, 2.0]) / 3.0
    R, zen, azi, pos = 80.0, (0.25, 1.0), (0.1, 0.85), np.array([5.0, -7.0, 11.0])
    n = 3000
    gs = G.torch(n, pos=pos, mom=mom, type=G.T_DISC, radius=R, zenith=zen, azimuth=azi, wavelength=430.0, time=2.5)
    p = generated(oracle_lib, gs).astype(np.float64)
    torch_flags_ok(p.astype(np.float32), gs)
    e1, e2, m = frame(mom)
    rel = p[:, 0, :3] - pos
    assert np.allclose(p[:, 1, :3], mom, atol=1e-7)
    assert np.all(np.abs(rel @ m) < 1e-4)                                   # in the disc plane
    r = np.linalg.norm(rel, axis=1)
    U = np.stack([stream(oracle_lib, i, 2) for i in range(n)])
    uz = zen[0] + U[:, 0] * (zen[1] - zen[0])                              # first draw: zenith
    ua = azi[0] + U[:, 1] * (azi[1] - azi[0])                              # second draw: azimuth
    assert np.allclose(r, R * uz, rtol=2e-6, atol=1e-4)
    want = np.cos(TWOPI * ua)[:, None] * e1 + np.sin(TWOPI * ua)[:, None] * e2
    assert np.allclose(rel / r[:, None], want, atol=2e-5)
    pol = p[:, 2, :3]
    assert np.allclose(pol, np.cross(rel / r[:, None], m), atol=2e-5)
    assert chi2_per_dof(r / R, zen[0], zen[1], np.ones_like) < 2.0           # radius uniform (not by area)


@pytest.mark.parametrize('radius', [150.0, -150.0])
def test_torch_sphere(oracle_lib, radius):
    """T_SPHERE (storch.h:246-296): points on the sphere |R| about the origin (gs.pos is not
    added), mom radial, outward for R > 0 and inward for R < 0; cos(theta) = 1 - 2 u_zenith,
    phi = 2 pi u_azimuth; pol is the tangent at phase 2 pi gs.distance in mom's frame."""
    n, dist = 4000, 0.3
    gs = G.torch(n, pos=(100.0, 200.0, 300.0), type=G.T_SPHERE, radius=radius, distance=dist)
    p = generated(oracle_lib, gs).astype(np.float64)
    torch_flags_ok(p.astype(np.float32), gs)
    pos, mom, pol = p[:, 0, :3], p[:, 1, :3], p[:, 2, :3]
    assert np.allclose(np.linalg.norm(pos, axis=1), abs(radius), rtol=2e-6)
    assert np.allclose(mom, np.sign(radius) * pos / abs(radius), atol=2e-6)
    U = np.stack([stream(oracle_lib, i, 2) for i in range(n)])
    assert np.allclose(pos[:, 2] / abs(radius), 1.0 - 2.0 * U[:, 0], atol=2e-6)
    phi = np.arctan2(pos[:, 1], pos[:, 0]) % TWOPI
    dphi = np.angle(np.exp(1j * (phi - TWOPI * U[:, 1])))
    big = np.hypot(pos[:, 0], pos[:, 1]) > 1.0
    assert np.all(np.abs(dphi[big]) < 1e-4)
    for k in range(0, n, 97):
        e1, e2, _ = frame(mom[k])
        assert np.allclose(pol[k], np.cos(TWOPI * dist) * e1 + np.sin(TWOPI * dist) * e2, atol=2e-5)
    assert chi2_per_dof(pos[:, 2] / abs(radius), -1, 1, np.ones_like) < 2.0   # isotropic


def test_torch_sphere_marsaglia(oracle_lib):
    """T_SPHERE_MARSAGLIA (storch.h:297-350): Marsaglia (1972) rejection on pairs (u, v) =
    2 u_i - 1 until u^2 + v^2 <= 1, point (2u sqrt(1-b), 2v sqrt(1-b), 2b - 1) times |R|;
    the accepted pair is the first of the photon's stream that lands in the unit disc."""
    n, R = 3000, -90.0
    gs = G.torch(n, type=G.T_SPHERE_MARSAGLIA, radius=R, distance=0.0)
    p = generated(oracle_lib, gs).astype(np.float64)
    torch_flags_ok(p.astype(np.float32), gs)
    pos, mom, pol = p[:, 0, :3], p[:, 1, :3], p[:, 2, :3]
    for i in range(n):
        u = 2.0 * stream(oracle_lib, i, 40) - 1.0
        pairs = u.reshape(-1, 2)
        b2 = np.sum(pairs ** 2, axis=1)
        k = int(np.argmax(b2 <= 1.0))
        uu, vv = pairs[k]
        b = b2[k]
        a = 2.0 * np.sqrt(1.0 - b)
        want = np.array([a * uu, a * vv, 2.0 * b - 1.0])
        assert np.allclose(pos[i], want * abs(R), atol=2e-4), i
        assert np.allclose(mom[i], -want, atol=2e-6), i                     # R < 0: inward
    assert np.allclose(np.linalg.norm(pos, axis=1), abs(R), rtol=5e-6)
    assert np.all(np.abs(np.sum(mom * pol, axis=1)) < 1e-5)
    e1, _, _ = frame(mom[5])
    assert np.allclose(pol[5], e1, atol=2e-5)                                # phase 0: e1
    assert chi2_per_dof(pos[:, 2] / abs(R), -1, 1, np.ones_like) < 2.0


def test_torch_line_and_point(oracle_lib):
    """T_LINE (storch.h:351-386): photon k of n starts at gs.pos + R (2k/n - 1) e1(mom),
    k the absolute photon index, with mom = gs.mom and pol = -e2; T_POINT (storch.h:387-410):
    every photon at gs.pos with the same mom and pol. No uniforms are drawn."""
    mom = np.array([0.0, 0.6, 0.8])
    e1, e2, _ = frame(mom)
    n, R, pos = 1000, 40.0, np.array([1.0, 2.0, 3.0])
    gs = G.torch(n, pos=pos, mom=mom, type=G.T_LINE, radius=R)
    for off in (0, 250):
        p = generated(oracle_lib, gs, slot_offset=off).astype(np.float64)
        torch_flags_ok(p.astype(np.float32), gs)
        k = off + np.arange(n)
        want = pos + (R * (2.0 * k / n - 1.0))[:, None] * e1
        assert np.allclose(p[:, 0, :3], want, atol=2e-5)
        assert np.allclose(p[:, 1, :3], mom) and np.allclose(p[:, 2, :3], -e2, atol=1e-6)
    q = generated(oracle_lib, G.torch(50, pos=pos, mom=mom, type=G.T_POINT)).astype(np.float64)
    assert np.all(q[:, 0, :3] == pos.astype(np.float32))
    assert np.allclose(q[:, 1, :3], mom) and np.allclose(q[:, 2, :3], -e2, atol=1e-6)
    # the photon stream is not consumed: event index does not change a thing
    assert np.array_equal(generated(oracle_lib, gs, event_index=3), generated(oracle_lib, gs))


@pytest.mark.parametrize('radius', [60.0, -60.0])
def test_torch_circle(oracle_lib, radius):
    """T_CIRCLE (storch.h:411-456): photon k of n on the circle of radius |R| about gs.pos
    in the XZ plane at phi = 2 pi (ax (1-f) + ay f), f = k/n; mom radial (outwards for
    R >= 0, inwards for R < 0), pol = rotateUz(-y, mom)."""
    n, azi, pos = 720, (0.25, 0.75), np.array([10.0, 0.0, -5.0])
    gs = G.torch(n, pos=pos, type=G.T_CIRCLE, radius=radius, azimuth=azi)
    p = generated(oracle_lib, gs).astype(np.float64)
    torch_flags_ok(p.astype(np.float32), gs)
    f = np.arange(n) / n
    phi = TWOPI * (azi[0] * (1 - f) + azi[1] * f)
    radial = np.c_[np.cos(phi), np.zeros(n), np.sin(phi)]
    assert np.allclose(p[:, 0, :3], pos + abs(radius) * radial, atol=1e-4)
    assert np.allclose(p[:, 1, :3], np.sign(radius) * radial, atol=2e-6)
    for k in range(0, n, 37):
        _, e2, _ = frame(p[k, 1, :3])
        assert np.allclose(p[k, 2, :3], -e2, atol=2e-6)


def test_torch_rectangle(oracle_lib):
    """T_RECTANGLE (storch.h:457-514): n/4 photons per side in order left, right, bottom,
    top of the rectangle x in azimuth, z in zenith (XZ plane, offset by gs.pos), each aimed
    inwards along the axis normal to its side, spaced by 1/(n/4) of the side."""
    n, zen, azi, pos = 400, (-30.0, 50.0), (-20.0, 70.0), np.array([0.0, 5.0, 0.0])
    gs = G.torch(n, pos=pos, type=G.T_RECTANGLE, zenith=zen, azimuth=azi)
    p = generated(oracle_lib, gs).astype(np.float64)
    torch_flags_ok(p.astype(np.float32), gs)
    m = n // 4
    fr = np.arange(m) / m
    sides = [
        (np.c_[np.full(m, azi[0]), np.zeros(m), zen[0] + fr * (zen[1] - zen[0])], [1, 0, 0]),
        (np.c_[np.full(m, azi[1]), np.zeros(m), zen[0] + fr * (zen[1] - zen[0])], [-1, 0, 0]),
        (np.c_[azi[0] + fr * (azi[1] - azi[0]), np.zeros(m), np.full(m, zen[0])], [0, 0, 1]),
        (np.c_[azi[0] + fr * (azi[1] - azi[0]), np.zeros(m), np.full(m, zen[1])], [0, 0, -1]),
    ]
    for s, (want_pos, want_mom) in enumerate(sides):
        sl = slice(s * m, (s + 1) * m)
        assert np.allclose(p[sl, 0, :3], pos + want_pos, atol=2e-5), s
        assert np.allclose(p[sl, 1, :3], want_mom), s
        _, e2, _ = frame(np.asarray(want_mom, dtype=np.float64))
        assert np.allclose(p[sl, 2, :3], -e2, atol=1e-6), s


# ---------------------------------------------------------------- re-emission
def test_reemission_draws_and_state(oracle_lib):
    """qsim::propagate_to_boundary re-emission (qsim.h:804-825) with precooked streams that
    force it: after generation (3 draws) the first bounce draws u_scattering, u_absorption,
    then u_reemit < REEMISSIONPROB, then wavelength, mom phi, mom cos(theta), pol phi, pol
    cos(theta). The RE point must sit absorption_length * -log(u_absorption) along the
    original direction, keep the time of flight at the group velocity (no delay), take the
    ICDF wavelength of its draw and uniform_sphere (qsim.h:177-183) directions."""
    s = scenes.scintillator_sphere()
    cfg = abi.default_config(**s.config, seq_enabled=1, max_record=4)
    n = 400
    rng = np.random.default_rng(21)
    pre = rng.uniform(1e-6, 1.0, size=(n, 64)).astype(np.float32)
    pre[:, 3] = 1e-6                         # u_scattering: scattering distance of metres
    pre[:, 4] = rng.uniform(0.3, 0.9, n)     # u_absorption: absorbed tens of mm out
    pre[:, 5] = rng.uniform(0.0, 0.75, n)    # u_reemit below the 0.8 re-emission probability
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(s.gensteps(scale_photons=n / 20000), precooked=pre,
                                                              max_record=4)
    rec = r['record'].astype(np.float64)
    u = pre.astype(np.float64)
    flag1 = r['record'].view(np.uint32)[:, 1, 3, 0] & 0xffff
    assert np.all(flag1 == abi.FLAGS['BULK_REEMIT'])
    assert np.all((r['seq'][:, 0, 0] >> 4) & 0xf == 5)                          # ffs(RE) in slot 1
    p0, p1 = rec[:, 0], rec[:, 1]
    wl0 = 350.0
    abs_len = ls_table_value(s, 'abs', wl0)
    d = -abs_len * np.log(u[:, 4])
    assert np.allclose(p1[:, 0, :3], p0[:, 0, :3] + d[:, None] * p0[:, 1, :3], rtol=1e-5, atol=1e-3)
    vg = csg.C_LIGHT / scenes.LS_N_A                  # n = a + b wl gives n - wl dn/dwl = a
    assert np.allclose(p1[:, 0, 3] - p0[:, 0, 3], d / vg, rtol=1e-5, atol=1e-6)
    icdf = s.physics['icdf'].astype(np.float64)
    assert np.allclose(p1[:, 2, 3], icdf_textbook(icdf, u[:, 6]), rtol=1e-5)
    ct = 2.0 * u[:, 8] - 1.0
    st_ = np.sqrt(1.0 - ct * ct)
    mom = np.c_[np.cos(TWOPI * u[:, 7]) * st_, np.sin(TWOPI * u[:, 7]) * st_, ct]
    assert np.allclose(p1[:, 1, :3], mom, atol=2e-6)
    ct2 = 2.0 * u[:, 10] - 1.0
    st2 = np.sqrt(1.0 - ct2 * ct2)
    v = np.c_[np.cos(TWOPI * u[:, 9]) * st2, np.sin(TWOPI * u[:, 9]) * st2, ct2]
    pol = np.cross(v, mom)
    pol /= np.linalg.norm(pol, axis=1)[:, None]
    assert np.allclose(p1[:, 2, :3], pol, atol=2e-5)
    # a reemit draw above the probability absorbs instead
    pre[:, 5] = 0.95
    r2 = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(s.gensteps(scale_photons=n / 20000), precooked=pre)
    assert np.all(r2['photon'].view(np.uint32)[:, 3, 0] & 0xffff == abi.FLAGS['BULK_ABSORB'])


def ls_table_value(s, which, wl):
    wl = np.asarray(wl, dtype=np.float64)
    if which == 'abs':
        return scenes.ls_absorption_length(wl)
    raise KeyError(which)


def icdf_textbook(icdf, u):
    """Wavelength of the hd20 ICDF rows at u: linear interpolation between texel centres
    (texel j at (j + 0.5)/width) of the row that u selects (qscint.h:207-228)."""
    w = icdf.shape[1]
    out = np.empty_like(u)
    for k, uu in enumerate(u):
        if uu < 0.05:
            row, x = 1, uu * 20.0
        elif uu > 0.95:
            row, x = 2, (uu - 0.95) * 20.0
        else:
            row, x = 0, uu
        out[k] = np.interp(x, (np.arange(w) + 0.5) / w, icdf[row])
    return out


def test_reemission_statistics(oracle_lib):
    """Philox run of the scintillator sphere: of the photons whose first bulk process is
    absorption at 350 nm, the re-emitted fraction is REEMISSIONPROB(350) = 0.8; re-emitted
    directions are isotropic and independent of the incoming one, pol is a unit vector
    transverse to mom, and the re-emitted wavelengths follow the 430 nm spectrum."""
    s = scenes.scintillator_sphere()
    cfg = abi.default_config(**s.config, seq_enabled=1, max_record=3)
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(s.gensteps(scale_photons=2.0), max_record=3)
    f1 = r['record'].view(np.uint32)[:, 1, 3, 0] & 0xffff
    re = f1 == abi.FLAGS['BULK_REEMIT']
    ab = f1 == abi.FLAGS['BULK_ABSORB']
    k, m = re.sum(), (re | ab).sum()
    assert m > 30000
    assert abs(k / m - 0.8) < 4 * np.sqrt(0.8 * 0.2 / m)
    p0, p1 = r['record'][re, 0].astype(np.float64), r['record'][re, 1].astype(np.float64)
    mom, pol = p1[:, 1, :3], p1[:, 2, :3]
    assert chi2_per_dof(mom[:, 2], -1, 1, np.ones_like) < 2.0
    assert abs(np.mean(np.sum(mom * p0[:, 1, :3], axis=1))) < 4 / np.sqrt(3 * k)
    assert np.allclose(np.linalg.norm(pol, axis=1), 1.0, atol=2e-6)
    dp = np.abs(np.sum(mom * pol, axis=1))
    # pol = normalize(v x mom) loses relative precision when v is nearly parallel to mom
    assert np.quantile(dp, 0.999) < 2e-6 and dp.max() < 1e-3
    pdf = lambda w: np.exp(-0.5 * ((w - 430.0) / 15.0) ** 2) / w ** 2
    assert chi2_per_dof(p1[:, 2, 3], 380.0, 480.0, pdf) < 2.5


# ---------------------------------------------------------------- boundary table
def test_boundary_lookup_interpolation(oracle_lib):
    """qbnd::boundary_lookup (qbnd.h:103-125) on the linear-filtered, wrap-addressed texture
    (QTex.cc:260-272): between texel centres wl_min + i*wl_step it is np.interp of the
    rows; a half-texel shift or swapped weights would fail. Beyond the last centre (and
    before the first) it blends the last texel with the first (wrap)."""
    s = scenes.scintillator_sphere()
    O = oracle_lib.Oracle(s.geometry, s.physics, abi.default_config(**s.config))
    bnd = s.physics['bnd'].astype(np.float64)
    nb, nw = bnd.shape[0], bnd.shape[3]
    line = s.meta['matline_ls']
    b, j = divmod(line, 4)
    centres = csg.WL_MIN + csg.WL_STEP * np.arange(nw)
    rng = np.random.default_rng(3)
    nm = np.concatenate([rng.uniform(csg.WL_MIN, csg.WL_MAX, 4000), centres[::7], [350.0, 430.0, 612.25]])
    for k in (0, 1):
        got = O.boundary_lookup(nm.astype(np.float32), line, k).astype(np.float64)
        for c in range(4):
            row = bnd[b, j, k, :, c]
            want = np.interp(nm, centres, row)
            scale = np.abs(np.diff(row)).max() + np.abs(row).max() * 1e-6 + 1e-12
            assert np.all(np.abs(got[:, c] - want) <= 2e-4 * scale + 1e-6 * np.abs(want)), (k, c)
    # the rindex ramp is exact at texel centres and linear between them
    got = O.boundary_lookup(np.float32([350.0, 350.5, 351.0]), line, 0)[:, 0].astype(np.float64)
    want = scenes.LS_N_A + scenes.LS_N_B * np.array([350.0, 350.5, 351.0])
    assert np.allclose(got, want, atol=3e-7)
    # wrap: half a step past the last centre and before the first blend the end texels
    ends = O.boundary_lookup(np.float32([csg.WL_MAX + 0.5 * csg.WL_STEP, csg.WL_MIN - 0.5 * csg.WL_STEP]), line, 0)
    half = 0.5 * (bnd[b, j, 0, -1, :] + bnd[b, j, 0, 0, :])
    assert np.allclose(ends[0], half, rtol=1e-5) and np.allclose(ends[1], half, rtol=1e-5)


# ---------------------------------------------------------------- halfspace
def _prim_oracle(oracle_lib, tree):
    f = Foundry()
    f.add_material('M', 1.0)
    b = f.boundary('M', None, None, 'M')
    f.begin_solid('s')
    f.add_prim(tree, b)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    return oracle_lib.Oracle(f.arrays(), f.physics(), abi.default_config())


def _rays(o, d, tmin=0.0):
    o = np.atleast_2d(np.asarray(o, dtype=np.float64))
    d = np.atleast_2d(np.asarray(d, dtype=np.float64))
    r = np.zeros((len(o), 2, 4), dtype=np.float32)
    r[:, 0, :3] = o
    r[:, 0, 3] = tmin
    r[:, 1, :3] = d
    r[:, 1, 3] = 1e6
    return r


def test_halfspace_plane_crossing(oracle_lib):
    """Halfspace n.p < w cut from a box: rays from the outside side of the plane that hit
    it within the box report t = (w - o.n)/(d.n) and the plane normal n."""
    n = np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
    w = 20.0
    O = _prim_oracle(oracle_lib, halfspace(1.0, 1.0, 0.0, w) & box3(200.0))
    rng = np.random.default_rng(7)
    tgt = rng.uniform(-60, 60, size=(3000, 3))
    tgt -= ((tgt @ n) - w)[:, None] * n                                       # points on the plane
    o = tgt + rng.uniform(20, 80, size=(3000, 1)) * n + rng.normal(size=(3000, 3)) * 10.0
    o = o[(o @ n) > w + 1.0]
    tgt = tgt[:len(o)]
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    for prec in (0, 1):
        isect, valid = O.intersect_prim(0, _rays(o, d), prec)
        t = (w - o @ n) / (d @ n)
        assert valid.all()
        assert np.allclose(isect[:, 3], t, rtol=1e-5, atol=1e-4)
        assert np.allclose(isect[:, :3], n, atol=1e-6)


def _inside_cut(p, n, w, L, keep_below):
    inb = np.all(np.abs(p) < L / 2, axis=1)
    below = p @ n < w
    return inb & (below if keep_below else ~below)


@pytest.mark.parametrize('keep_below', [True, False])
def test_halfspace_tree_matches_brute_force(oracle_lib, keep_below):
    """box & halfspace and box - halfspace against a ray march of the inside test, with
    origins inside and outside both: exercises the leaf's hits, its misses and its
    exit-at-infinity signal through the CSG tree."""
    nrm = np.array([0.3, -0.5, 0.8])
    nrm /= np.linalg.norm(nrm)
    w, L = 15.0, 120.0
    hs = halfspace(*nrm, w)
    tree = (box3(L) & hs) if keep_below else (box3(L) - hs)
    O = _prim_oracle(oracle_lib, tree)
    rng = np.random.default_rng(8 + keep_below)
    m = 300
    o = rng.uniform(-1.4 * L, 1.4 * L, size=(m, 3))
    o[: m // 3] = rng.uniform(-0.45 * L, 0.45 * L, size=(m // 3, 3))          # a third start inside the box
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    from test_oracle_pins import _first_crossing
    want = _first_crossing(lambda p: _inside_cut(p, nrm, w, L, keep_below), o, d, 4 * L, step=0.05)
    isect, valid = O.intersect_prim(0, _rays(o, d), 1)
    got = np.where(valid.astype(bool), isect[:, 3], np.nan)
    agree = (np.isnan(want) & np.isnan(got)) | (np.abs(got - want) < 2e-3)
    # A complemented leaf's miss is promoted to an exit at infinity whatever side the ray
    # is on (csg_intersect_tree.h:454-462): for box - halfspace that is wrong for rays
    # inside the halfspace heading away from its plane, the reference's own limitation
    # (pinned in test_halfspace_unbounded_exit_rule); the march is compared on the rest.
    ok = np.ones(m, bool) if keep_below else ~(((o @ nrm) < w) & ((d @ nrm) < 0))
    assert ok.sum() > 200
    assert agree[ok].mean() >= 0.99, (agree[ok].mean(), np.c_[want, got][ok & ~agree][:5])


def test_halfspace_unbounded_exit_rule(oracle_lib):
    """The leaf's miss from inside (o.n - w < -1e-9) is an exit at infinity: inside box &
    halfspace, a ray heading away from the plane leaves through the box, and so does a ray
    from just below the plane (within tmin) heading up through it, which the reference's
    rule also classifies as an exit at infinity (its inside test ignores the direction)."""
    O = _prim_oracle(oracle_lib, halfspace(0.0, 0.0, 1.0, 10.0) & box3(100.0))
    r = _rays([[0, 0, 0], [0, 0, 9.99]], [[0, 0, -1], [0, 0, 1]], tmin=0.05)
    isect, valid = O.intersect_prim(0, r, 0)
    assert valid.all()
    assert np.allclose(isect[0], [0, 0, -1, 50.0], atol=1e-5)                # box bottom face
    assert np.allclose(isect[1], [0, 0, 1, 40.01], atol=1e-4)                # box top face, not the plane
    # from above the plane, heading up: no part of the solid ahead
    isect, valid = O.intersect_prim(0, _rays([[0, 0, 20.0]], [[0, 0, 1]], tmin=0.05), 0)
    assert not valid[0]
    # box - halfspace, from inside the halfspace heading away from the plane: the
    # complemented miss is promoted to an exit at infinity, so the box exit is reported
    C = _prim_oracle(oracle_lib, box3(100.0) - halfspace(0.0, 0.0, 1.0, 10.0))
    isect, valid = C.intersect_prim(0, _rays([[0, 0, 0]], [[0, 0, -1]], tmin=0.05), 0)
    assert valid[0] and np.allclose(isect[0], [0, 0, -1, 50.0], atol=1e-5)
    # and from above the plane it enters nothing heading up, the box exit bounds it
    isect, valid = C.intersect_prim(0, _rays([[0, 0, 20.0]], [[0, 0, 1]], tmin=0.05), 0)
    assert valid[0] and np.allclose(isect[0], [0, 0, 1, 30.0], atol=1e-5)
