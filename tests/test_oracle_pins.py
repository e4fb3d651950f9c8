"""Pins of the CPU oracle, independent of the GPU engine.

Each check compares the oracle with something the oracle does not itself compute:
published known-answer vectors, closed forms, textbook tables, brute force or the
reference's own test expectations.
"""
import numpy as np
import pytest

from oxr import abi, csg, gensteps as G, scenes
from oxr.csg import Foundry, box3, sphere, zsphere, cylinder, cone, convexpolyhedron, phicut, scale, translate


# ---------------------------------------------------------------- Philox / streams
def test_philox_known_answer_vectors(oracle_lib):
    """Random123 kat_vectors for philox4x32 with 10 rounds."""
    O = oracle_lib
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert list(O.philox(ctr, key)) == want


def test_stream_layout_and_conversion(oracle_lib):
    """Uniform j of photon p = ((x>>8)+1)/2^24 with x = philox((start+j)//4, p)[(start+j)%4]."""
    O = oracle_lib
    seed, skip, event, photon = 0x0000000100000007, 100000, 3, 12345
    u = O.uniforms(20, seed=seed, offset=2, skip=skip, event=event, photon=photon)
    start = 2 + skip * event
    key = [seed & 0xffffffff, seed >> 32]
    for j in range(20):
        b = (start + j) // 4
        x = O.philox([b & 0xffffffff, b >> 32, photon & 0xffffffff, photon >> 32], key)[(start + j) % 4]
        assert u[j] == np.float32(((int(x) >> 8) + 1) / 2.0 ** 24)
    v = O.uniforms(200000, photon=7)
    assert v.min() > 0.0 and v.max() <= 1.0
    assert np.all(v * 2 ** 24 == np.round(v * 2 ** 24))
    assert abs(v.mean() - 0.5) < 4 * np.sqrt(1 / 12 / len(v))


def test_seed_expansion_closed_form(oracle_lib):
    rng = np.random.default_rng(1)
    counts = rng.integers(0, 50, size=300)
    gs = G.torch(counts)
    assert np.array_equal(oracle_lib.fill_seed(gs), np.repeat(np.arange(300), counts))


# ---------------------------------------------------------------- CSG decision table
# Kensler, "Ray Tracing CSG Objects Using Single Hit Intersections" (2006): action when A is
# closer / when B is closer, per (state A, state B). 0 miss, 1 return A, 2 return B,
# 3 return flipped B, 4 loop A, 5 loop B.
E, X, M = 0, 1, 2
KENSLER = {
    1: {(E, E): (1, 2), (E, X): (4, 2), (E, M): (1, 1), (X, E): (1, 5), (X, X): (2, 1), (X, M): (1, 1),
        (M, E): (2, 2), (M, X): (2, 2), (M, M): (0, 0)},
    2: {(E, E): (4, 5), (E, X): (1, 5), (E, M): (0, 0), (X, E): (4, 2), (X, X): (1, 2), (X, M): (0, 0),
        (M, E): (0, 0), (M, X): (0, 0), (M, M): (0, 0)},
    3: {(E, E): (1, 5), (E, X): (4, 5), (E, M): (1, 1), (X, E): (1, 3), (X, X): (4, 3), (X, M): (1, 1),
        (M, E): (0, 0), (M, X): (0, 0), (M, M): (0, 0)},
}


def test_csg_lut_matches_kensler_table(oracle_lib):
    for op, table in KENSLER.items():
        for (a, b), (acl, bcl) in table.items():
            assert oracle_lib.lut(op, a, b, True) == acl, (op, a, b, 'A closer')
            assert oracle_lib.lut(op, a, b, False) == bcl, (op, a, b, 'B closer')


# ---------------------------------------------------------------- leaves: closed forms
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


@pytest.mark.parametrize('precision', [0, 1])
def test_sphere_closed_form(oracle_lib, precision):
    O = _prim_oracle(oracle_lib, sphere(100.0))
    rng = np.random.default_rng(2)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    o = -d * 400.0                                   # aimed through the centre
    isect, valid = O.intersect_prim(0, _rays(o, d), precision)
    assert valid.all()
    assert np.allclose(isect[:, 3], 300.0, atol=1e-3)
    assert np.allclose(isect[:, :3], -d, atol=1e-5)   # outward normal at entry
    isect, valid = O.intersect_prim(0, _rays(np.zeros((500, 3)), d), precision)
    assert np.allclose(isect[:, 3], 100.0, atol=1e-4)
    # impact parameter b: t = 400 - sqrt(r^2 - b^2) for a ray offset sideways by b
    b = np.linspace(0, 99, 50)
    o = np.c_[b, np.zeros(50), np.full(50, -400.0)]
    isect, valid = O.intersect_prim(0, _rays(o, np.tile([0, 0, 1.0], (50, 1))), precision)
    assert np.allclose(isect[:, 3], 400.0 - np.sqrt(100.0 ** 2 - b ** 2), atol=2e-3)


def test_box3_and_convexpolyhedron_closed_form(oracle_lib):
    O = _prim_oracle(oracle_lib, box3(100.0, 60.0, 40.0))
    planes = [(1, 0, 0, 50), (-1, 0, 0, 50), (0, 1, 0, 30), (0, -1, 0, 30), (0, 0, 1, 20), (0, 0, -1, 20)]
    P = _prim_oracle(oracle_lib, convexpolyhedron(planes, (-50, -30, -20, 50, 30, 20)))
    rng = np.random.default_rng(3)
    o = rng.uniform(-300, 300, size=(2000, 3))
    tgt = rng.uniform([-50, -30, -20], [50, 30, 20], size=(2000, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    a, va = O.intersect_prim(0, _rays(o, d), 1)
    b, vb = P.intersect_prim(0, _rays(o, d), 1)
    # slab entry/exit distance computed here independently
    lo, hi = np.array([-50, -30, -20.0]), np.array([50, 30, 20.0])
    t0, t1 = (lo - o) / d, (hi - o) / d
    tn = np.max(np.minimum(t0, t1), axis=1)
    tf = np.min(np.maximum(t0, t1), axis=1)
    want = np.where(tn > 0, tn, tf)
    ok = va.astype(bool) & vb.astype(bool)
    assert ok.mean() > 0.999
    assert np.allclose(a[ok, 3], want[ok], atol=1e-3)
    assert np.allclose(b[ok, 3], want[ok], atol=1e-3)
    assert np.allclose(a[ok, :3], b[ok, :3])


def test_cylinder_and_zsphere_caps(oracle_lib):
    C = _prim_oracle(oracle_lib, cylinder(50.0, -100.0, 100.0))
    # radial ray at z=10 from x=-300: t=250, normal -x ; axial ray from below: cap at z=-100
    r = _rays([[-300, 0, 10], [5, 5, -400]], [[1, 0, 0], [0, 0, 1]])
    i, v = C.intersect_prim(0, r, 1)
    assert v.all()
    assert np.allclose(i[0], [-1, 0, 0, 250], atol=1e-4)
    assert np.allclose(i[1], [0, 0, -1, 300], atol=1e-4)
    Z = _prim_oracle(oracle_lib, zsphere(100.0, -50.0, 60.0))
    r = _rays([[0, 0, -400], [0, 0, 400], [-400, 0, 0]], [[0, 0, 1], [0, 0, -1], [1, 0, 0]])
    i, v = Z.intersect_prim(0, r, 1)
    assert v.all()
    assert np.allclose(i[0], [0, 0, -1, 350], atol=1e-4)    # lower cap at z=-50
    assert np.allclose(i[1], [0, 0, 1, 340], atol=1e-4)     # upper cap at z=60
    assert np.allclose(i[2], [-1, 0, 0, 300], atol=1e-4)    # sphere at equator


def test_cone_and_phicut(oracle_lib):
    K = _prim_oracle(oracle_lib, cone(100.0, -50.0, 50.0, 50.0))
    # radius at height z is 75 - z/2 ; a radial ray at z=0 enters at x=-75
    i, v = K.intersect_prim(0, _rays([[-300, 0, 0]], [[1, 0, 0]]), 1)
    assert v[0] and abs(i[0, 3] - 225.0) < 1e-3
    nrm = np.array([-1.0, 0.0, 0.5]) / np.linalg.norm([-1.0, 0.0, 0.5])
    assert np.allclose(i[0, :3], nrm, atol=1e-5)
    # phicut wedge [0, 0.5 pi]: a ray along -y at x=10 from y=+300 exits through the phi=0 plane at y=0
    Pc = _prim_oracle(oracle_lib, phicut(0.0, 0.5) & box3(1000.0))
    i, v = Pc.intersect_prim(0, _rays([[10, 300, 0]], [[0, -1, 0]]), 1)
    assert v[0] and abs(i[0, 3] - 300.0) < 1e-3


# ---------------------------------------------------------------- CSG trees: brute force
def _inside_pmt(p):
    """Inside test of scenes.pmt_like_solid written from its definition."""
    def zs(p, r, z1, z2, sz):
        q = p.copy()
        q[:, 2] = q[:, 2] / sz
        return (np.sum(q * q, axis=1) < r * r) & (q[:, 2] > z1) & (q[:, 2] < z2)
    ell = zs(p, 254.0, -200.0, 254.0, 184.0 / 254.0)
    neck = (p[:, 0] ** 2 + p[:, 1] ** 2 < 50.0 ** 2) & (p[:, 2] > -330.0) & (p[:, 2] < -130.0)
    inner = zs(p, 249.0, -190.0, 249.0, 179.0 / 249.0)
    return (ell | neck) & ~inner


def _inside_combo(p):
    s = np.sum((p - np.array([30.0, 0, 0])) ** 2, axis=1) < 80.0 ** 2
    b = np.all(np.abs(p) < np.array([60.0, 60.0, 60.0]), axis=1)
    c = (p[:, 0] ** 2 + p[:, 1] ** 2 < 20.0 ** 2) & (np.abs(p[:, 2]) < 200.0)
    return (s & b) & ~c


def _first_crossing(inside, o, d, tmax, step=0.02):
    t = np.arange(step, tmax, step)
    out = np.full(len(o), np.nan)
    for k in range(len(o)):
        pts = o[k] + t[:, None] * d[k]
        ins = inside(pts)
        ch = np.nonzero(ins != ins[0])[0]
        if len(ch) == 0:
            continue
        lo, hi = t[ch[0] - 1] if ch[0] > 0 else 0.0, t[ch[0]]
        s0 = inside((o[k] + lo * d[k])[None])[0]
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if inside((o[k] + mid * d[k])[None])[0] == s0:
                lo = mid
            else:
                hi = mid
        out[k] = 0.5 * (lo + hi)
    return out


@pytest.mark.parametrize('which', ['pmt', 'combo'])
def test_csg_tree_matches_brute_force(oracle_lib, which):
    if which == 'pmt':
        tree, inside, ext = scenes.pmt_like_solid(), _inside_pmt, 350.0
    else:
        tree = (sphere(80.0, 30.0, 0.0, 0.0) & box3(120.0)) - cylinder(20.0, -200.0, 200.0)
        inside, ext = _inside_combo, 150.0
    O = _prim_oracle(oracle_lib, tree)
    rng = np.random.default_rng(4)
    n = 150
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1)[:, None] * (2.5 * ext)
    tgt = rng.uniform(-0.6 * ext, 0.6 * ext, size=(n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    want = _first_crossing(inside, o, d, 6 * ext)
    isect, valid = O.intersect_prim(0, _rays(o, d), 1)
    got = np.where(valid.astype(bool), isect[:, 3], np.nan)
    both_miss = np.isnan(want) & np.isnan(got)
    close = np.abs(got - want) < 2e-3
    agree = both_miss | close
    assert agree.mean() >= 0.99, (agree.mean(), np.c_[want, got][~agree][:5])
    isect32, valid32 = O.intersect_prim(0, _rays(o, d), 0)
    ok = valid32.astype(bool) & valid.astype(bool)
    assert np.allclose(isect32[ok, 3], isect[ok, 3], atol=5e-3)


# ---------------------------------------------------------------- boundary physics
def _fresnel(n1, n2, c1, s_frac):
    s1 = np.sqrt(1 - c1 ** 2)
    s2 = n1 / n2 * s1
    c2 = np.sqrt(1 - s2 ** 2)
    rs = ((n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)) ** 2
    rp = ((n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2)) ** 2
    return 1.0 - (rs * s_frac + rp * (1 - s_frac)), c2


def test_fresnel_snell_and_polarisation(oracle_lib):
    rng = np.random.default_rng(5)
    n = 400
    n1, n2 = 1.0, 1.5
    th = rng.uniform(0.05, 1.4, size=n)
    psi = rng.uniform(0, 2 * np.pi, size=n)
    mom = np.c_[np.sin(th), np.zeros(n), -np.cos(th)]            # travelling down onto z=0
    normal = np.tile([0.0, 0.0, 1.0], (n, 1))
    s_dir = np.tile([0.0, 1.0, 0.0], (n, 1))                     # perpendicular to plane of incidence
    p_dir = np.cross(s_dir, mom)
    pol = np.cos(psi)[:, None] * s_dir + np.sin(psi)[:, None] * p_dir
    T, c2 = _fresnel(n1, n2, np.cos(th), np.cos(psi) ** 2)
    for side, frac in (('transmit', 0.999), ('reflect', 1.001)):
        u = np.minimum(T * frac, 0.999999)[:, None].astype(np.float32)
        m2, p2, flag = oracle_lib.propagate_at_boundary(mom, pol, normal, n1, n2, u)
        want = abi.FLAGS['BOUNDARY_TRANSMIT'] if side == 'transmit' else abi.FLAGS['BOUNDARY_REFLECT']
        assert np.all(flag == want)
        assert np.allclose(np.linalg.norm(m2, axis=1), 1.0, atol=2e-6)
        assert np.allclose(np.linalg.norm(p2, axis=1), 1.0, atol=2e-6)
        assert np.all(np.abs(np.sum(m2 * p2, axis=1)) < 2e-6)
        if side == 'transmit':
            assert np.allclose(n2 * np.linalg.norm(np.cross(m2, normal), axis=1), n1 * np.sin(th), atol=5e-6)
            assert np.allclose(-m2[:, 2], c2, atol=5e-6)
        else:
            assert np.allclose(m2, mom * [1, 1, -1], atol=5e-6)
    # total internal reflection beyond the critical angle: always reflects
    th = np.linspace(np.arcsin(1 / 1.5) + 0.01, 1.5, 50)
    mom = np.c_[np.sin(th), np.zeros(50), -np.cos(th)]
    u = np.full((50, 1), 1e-6, dtype=np.float32)
    m2, p2, flag = oracle_lib.propagate_at_boundary(mom, np.tile([0, 1.0, 0], (50, 1)), np.tile([0, 0, 1.0], (50, 1)),
                                                    1.5, 1.0, u)
    assert np.all(flag == abi.FLAGS['BOUNDARY_REFLECT'])
    # normal incidence transmission 4 n1 n2 / (n1+n2)^2
    Tn = 4 * 1.0 * 1.5 / 2.5 ** 2
    for frac, want in ((0.999, 'BOUNDARY_TRANSMIT'), (1.001, 'BOUNDARY_REFLECT')):
        u = np.array([[Tn * frac]], dtype=np.float32)
        _, _, flag = oracle_lib.propagate_at_boundary([[0, 0, -1.0]], [[1.0, 0, 0]], [[0, 0, 1.0]], 1.0, 1.5, u)
        assert flag[0] == abi.FLAGS[want]


def _chi2_uniform(samples, lo, hi, pdf, bins=40):
    h, e = np.histogram(samples, bins=bins, range=(lo, hi))
    c = 0.5 * (e[1:] + e[:-1])
    w = pdf(c)
    exp = w / w.sum() * len(samples)
    return np.sum((h - exp) ** 2 / exp) / (bins - 1)


def test_rayleigh_angular_law(oracle_lib):
    mom, pol = np.array([0, 0, 1.0]), np.array([1.0, 0, 0])
    m, p = oracle_lib.sample('rayleigh', mom, pol, np.zeros(3), 200000)
    cos = m[:, 2]
    assert _chi2_uniform(cos, -1, 1, lambda c: 1 + c * c) < 2.0
    assert np.all(np.abs(np.sum(m * p, axis=1)) < 1e-5)
    # dipole: intensity ~ 1 - (k'.e)^2, so <(k'.e)^2> = 1/5 for scattered k'
    assert abs(np.mean(m[:, 0] ** 2) - 0.2) < 0.005


def test_lambertian_and_marsaglia(oracle_lib):
    m, p = oracle_lib.sample('diffuse', np.array([0, 0, -1.0]), np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), 100000)
    assert np.all(m[:, 2] >= 0)
    assert _chi2_uniform(m[:, 2], 0, 1, lambda c: c) < 2.0
    assert np.allclose(np.linalg.norm(m, axis=1), 1, atol=1e-5)
    m, _ = oracle_lib.sample('marsaglia', np.zeros(3), np.zeros(3), np.zeros(3), 100000)
    assert _chi2_uniform(m[:, 2], -1, 1, lambda c: np.ones_like(c)) < 2.0
    m, p = oracle_lib.sample('specular', np.array([0.6, 0, -0.8]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]), 1)
    assert np.allclose(m[0], [0.6, 0, 0.8], atol=1e-6) and np.allclose(p[0], [0, -1, 0], atol=1e-6)


# ---------------------------------------------------------------- generation
def _gen_only(oracle_lib, scene, gs, **cfg):
    c = abi.default_config(**dict(scene.config, **cfg))
    c.max_bounce = 0
    return oracle_lib.Oracle(scene.geometry, scene.physics, c).simulate(gs, seq=False)['photon']


def test_torch_isotropic_generation(oracle_lib):
    s = scenes.nested_boxes()
    p = _gen_only(oracle_lib, s, G.torch(100000, pos=(1, 2, 3), wavelength=420.0, type=G.T_POINT_ISOTROPIC))
    mom, pol = p[:, 1, :3], p[:, 2, :3]
    assert _chi2_uniform(mom[:, 2], -1, 1, lambda c: np.ones_like(c)) < 2.0
    assert np.all(np.abs(np.sum(mom * pol, axis=1)) < 1e-5)
    assert np.allclose(p[:, 0, :3], [1, 2, 3]) and np.all(p[:, 2, 3] == 420.0)
    assert np.all(p.view(np.uint32)[:, 3, 0] & 0xffff == abi.FLAGS['TORCH'])


def test_cerenkov_cone_and_spectrum(oracle_lib):
    s = scenes.csi_calorimeter()
    n = 50000
    beta = 0.9
    gs = G.cerenkov([n], [[0, 0, 0]], [[0, 0, 2.0]], 2.0, beta, s.meta['matline_csi'], 200.0, 700.0, scenes.CSI_N, 100.0)
    p = _gen_only(oracle_lib, s, gs)
    mom = p[:, 1, :3]
    assert np.allclose(mom[:, 2], 1.0 / (beta * scenes.CSI_N), atol=2e-5)     # cos(theta) = 1/(beta n)
    assert np.all(np.abs(np.sum(mom * p[:, 2, :3], axis=1)) < 1e-5)
    inv = 1.0 / p[:, 2, 3]
    assert _chi2_uniform(inv, 1 / 700.0, 1 / 200.0, lambda c: np.ones_like(c)) < 2.0
    assert np.all((p[:, 0, 2] >= 0) & (p[:, 0, 2] <= 2.0))


def test_scintillation_spectrum_and_time(oracle_lib):
    s = scenes.csi_calorimeter()
    n = 100000
    gs = G.scintillation([n], [[0, 0, 0]], [[0, 0, 1.0]], 1.0, 200.0, 1000.0)
    p = _gen_only(oracle_lib, s, gs)
    wl = p[:, 2, 3]
    pdf = lambda w: np.exp(-0.5 * ((w - 550.0) / 40.0) ** 2) / w ** 2   # energy-domain ICDF -> S(w) hc/w^2
    assert _chi2_uniform(wl, 400.0, 700.0, pdf) < 2.5
    t = p[:, 0, 3]
    assert abs(t.mean() - 1000.0) < 4 * 1000.0 / np.sqrt(n)
    assert np.all(np.abs(np.sum(p[:, 1, :3] * p[:, 2, :3], axis=1)) < 1e-5)


# ---------------------------------------------------------------- transport
def test_beer_lambert_and_group_velocity(oracle_lib):
    f = Foundry()
    f.add_material('Abs', 1.5, absorption_length=100.0, scattering_length=1e9)
    b = f.boundary('Abs', None, None, 'Abs')
    f.begin_solid('s')
    f.add_prim(sphere(1e5), b)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    O = oracle_lib.Oracle(f.arrays(), f.physics(), abi.default_config(max_bounce=5))
    r = O.simulate(G.torch(50000, type=G.T_POINT_ISOTROPIC), precision=1)
    p = r['photon']
    dist = np.linalg.norm(p[:, 0, :3], axis=1)
    assert np.all(p.view(np.uint32)[:, 3, 0] & 0xffff == abi.FLAGS['BULK_ABSORB'])
    assert abs(dist.mean() - 100.0) < 4 * 100.0 / np.sqrt(len(dist))
    assert np.allclose(p[:, 0, 3], dist / (csg.C_LIGHT / 1.5), rtol=1e-5)


def test_raindrop_file_source_known_answer(oracle_lib):
    """tests/test_GPUPhotonFileSource.sh:41-90: 10 input photons, 10 hits, wavelengths kept."""
    s = scenes.raindrop()
    cfg = abi.default_config(**s.config, seq_enabled=1)
    r = oracle_lib.Oracle(s.geometry, s.physics, cfg).simulate(s.gensteps(), input_photons=scenes.raindrop_input_photons())
    p = r['photon']
    hit = (p.view(np.uint32)[:, 3, 3] & abi.FLAGS['SURFACE_DETECT']) != 0
    assert hit.sum() == 10
    assert sorted(p[:, 2, 3].tolist()) == [420] * 3 + [450] * 2 + [500] * 5
    # seqhis nibbles are ffs(flag) per slot (sysrap/sseq.h:174-184): TO=3, BT=12, SD=7
    assert all(int(q[0, 0]) == 0x7c3 for q in r['seq'])


def test_precooked_stream_reproduces_philox(oracle_lib):
    s = scenes.nested_boxes()
    cfg = abi.default_config(**s.config, seq_enabled=1)
    O = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    gs = s.gensteps(scale_photons=0.01)
    a = O.simulate(gs)
    pre = np.stack([oracle_lib.uniforms(256, skip=cfg.skipahead_event, photon=i) for i in range(len(a['photon']))])
    b = O.simulate(gs, precooked=pre)
    assert np.array_equal(a['photon'].view(np.uint32), b['photon'].view(np.uint32))
    assert np.array_equal(a['seq'], b['seq'])


def test_oracle_culling_is_exact(oracle_lib):
    """The box-culled trace (used for large instance counts) equals the full brute force."""
    s = scenes.instanced_csg(num_inst=2000)
    cfg = abi.default_config(**s.config)
    rng = np.random.default_rng(11)
    n = 2000
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1)[:, None] * rng.uniform(0, 19000.0, size=(n, 1))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = _rays(o, d, 0.05)
    O = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    a = O.trace(rays)
    O.set_cull(True)
    b = O.trace(rays)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert ((a.view(np.uint32)[:, 1, 3] & 0xffff) == 3).sum() > 100     # PMT hits present


def test_polyhedron_behind_ray_is_a_miss(oracle_lib):
    """A ray from outside a convex polyhedron pointing away from it cannot hit it. The leaf
    itself reports valid (its slab interval lies behind the origin) without writing the
    isect, as the reference's leaf does; the trace must drop that zero distance."""
    planes = [(1, 0, 0, 50), (-1, 0, 0, 50), (0, 1, 0, 30), (0, -1, 0, 30), (0, 0, 1, 20), (0, 0, -1, 20)]
    P = _prim_oracle(oracle_lib, convexpolyhedron(planes, (-50, -30, -20, 50, 30, 20)))
    rng = np.random.default_rng(5)
    o = rng.uniform(-300, 300, size=(2000, 3))
    o = o[np.any(np.abs(o) > [50, 30, 20], axis=1)]
    tgt = rng.uniform([-50, -30, -20], [50, 30, 20], size=(len(o), 3))
    d = o - tgt                                          # away from a point inside
    d /= np.linalg.norm(d, axis=1)[:, None]
    isect, valid = P.intersect_prim(0, _rays(o, d, 0.05), 0)
    assert valid.all() and np.all(isect == 0.0)          # the leaf's own report
    for prec in (0, 1):
        prd = P.trace(_rays(o, d, 0.05), prec)
        assert np.all(prd[:, 1, 3].view(np.uint32) == 0xffffffff) and np.all(prd[:, 0, 3] == 0.0)


def test_list_nodes_and_polyhedron_trace_is_finite(oracle_lib):
    """Scene of list nodes, a tetrahedron and a rotated box: the box-culled trace equals
    the brute-force one, and every hit has a finite unit-scale normal and t >= tmin."""
    from test_foundry_io import mixed_scene_arrays
    g, p = mixed_scene_arrays()
    O = oracle_lib.Oracle(g, p, abi.default_config())
    rng = np.random.default_rng(12)
    o = rng.uniform(-250, 250, size=(4000, 3))
    d = rng.uniform(-100, 100, size=(4000, 3)) - o       # aimed through the objects
    d /= np.linalg.norm(d, axis=1)[:, None]
    rays = _rays(o, d, 0.05)
    a = O.trace(rays)
    O.set_cull(True)
    b = O.trace(rays)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    hit = a[:, 1, 3].view(np.uint32) != 0xffffffff
    assert hit.sum() > 1000
    n = np.linalg.norm(a[hit, 0, :3], axis=1)
    assert np.all(np.isfinite(n)) and np.all(n > 0.5) and np.all(a[hit, 0, 3] >= np.float32(0.05))


def test_single_precision_log(oracle_lib):
    """The photon loop's log (FreeBSD msun e_logf.c method in IEEE single operations) is
    faithfully rounded: within one ulp of the exact value over every Philox uniform
    k/2^24 and over random floats of all exponents; special values as the C library."""
    k = np.arange(1, 1 << 24, 97, dtype=np.float64)
    u = (k / 2.0 ** 24).astype(np.float32)
    rng = np.random.default_rng(3)
    bits = rng.integers(0x00800000, 0x7f800000, size=200000, dtype=np.uint32)
    x = np.concatenate([u, bits.view(np.float32), np.float32([1.0, 2.0, 0.5, 1.4142135, 1.4142137])])
    got = oracle_lib.xlog(x).astype(np.float64)
    exact = np.log(x.astype(np.float64))
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    err = np.abs(got - exact) / np.where(exact == 0, 1.0, ulp)
    assert err.max() < 1.0, err.max()
    assert got[-5] == 0.0 and abs(got[-4] - np.log(2.0)) < 1e-7
    sp = oracle_lib.xlog(np.float32([0.0, -1.0, np.inf, np.nan, 1e-40]))
    assert sp[0] == -np.inf and np.isnan(sp[1]) and sp[2] == np.inf and np.isnan(sp[3])
    assert abs(sp[4] - np.log(1e-40)) < 1e-4
