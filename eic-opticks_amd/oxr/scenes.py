"""Synthetic scenes for the five benchmark workloads plus the reference's raindrop.

Each builder returns a ``Scene``: CSGFoundry-layout geometry arrays, SSim-layout
physics tables, stepping-config overrides and a genstep generator. Host-side
randomness uses numpy's PCG64 seeded with 42, the seed of the reference's
integration tests (tests/test_GPURaytrace.sh:7).

Workloads (sizes as named in the benchmark definitions):
  nested_boxes      240/220/200/100 mm vacuum/lead/air/water cubes, 100k torch photons
  csi_calorimeter   8x8 CsI crystals 20x20x200 mm + SiPMs, 2000 scint + 500 Cerenkov gensteps
  instanced_csg     water sphere r=19.5 m, acrylic shell, 20000 instanced CSG solids, 10M photons
  thin_shells       50 concentric 1 mm shells at 1-20 m + 5000 instanced solids, 5M photons
  raindrop          tests/geom/opticks_raindrop.gdml restated (10 input-photon hit test)
  scintillator_sphere  re-emitting liquid scintillator in water, wavelength-dependent tables
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Dict

import numpy as np

from . import csg
from . import gensteps as G
from .csg import Foundry, box3, sphere, zsphere, cylinder, translate, scale, rotate_z_to

HC = csg.HC_EVNM


@dataclass
class Scene:
    name: str
    foundry: Foundry
    geometry: Dict[str, np.ndarray]
    physics: Dict[str, object]
    config: Dict[str, object]
    make_gensteps: Callable
    meta: Dict[str, object] = field(default_factory=dict)

    def gensteps(self, seed=42, scale_photons=1.0):
        return self.make_gensteps(np.random.default_rng(seed), scale_photons)


def _finish(name, f, config, make_gensteps, icdf=None, hd_factor=0, **meta):
    phys = f.physics()
    if icdf is not None:
        phys['icdf'] = icdf
        phys['hd_factor'] = hd_factor
    return Scene(name, f, f.arrays(), phys, config, make_gensteps, meta)


def _energy_table(pairs_ev):
    """Property tabulated against photon energy in eV -> callable of wavelength (nm),
    clamped outside the table as G4MaterialPropertyVector::Value does."""
    e = np.array([p[0] for p in pairs_ev], dtype=np.float64)
    v = np.array([p[1] for p in pairs_ev], dtype=np.float64)
    return lambda wl: np.interp(HC / wl, e, v)


# ---------------------------------------------------------------------------
def nested_boxes(num_photon=100_000):
    f = Foundry()
    f.add_material('Vacuum', 1.0)
    f.add_material('Pb', 1.0)
    f.add_material('Air', 1.0003)
    f.add_material('Water', 1.333, absorption_length=20_000.0, scattering_length=50_000.0)
    f.add_implicit_absorber('Implicit_Air_Pb')
    b_world = f.boundary('Vacuum', None, None, 'Vacuum')
    b_pb = f.boundary('Vacuum', None, None, 'Pb')
    b_air = f.boundary('Pb', None, 'Implicit_Air_Pb', 'Air')
    b_water = f.boundary('Air', None, None, 'Water')
    f.begin_solid('r0')
    f.add_prim(box3(240.0), b_world)
    f.add_prim(box3(220.0), b_pb)
    f.add_prim(box3(200.0), b_air)
    f.add_prim(box3(100.0), b_water)
    f.end_solid()
    f.add_instance(np.eye(4), 0)

    def gs(rng, s):
        n = max(1, int(round(num_photon * s)))
        return G.torch(n, pos=(0, 0, 0), wavelength=420.0, type=G.T_POINT_ISOTROPIC)

    return _finish('nested_boxes', f, dict(max_bounce=32), gs,
                   water_boundary=b_water)


# ---------------------------------------------------------------------------
def raindrop():
    """tests/geom/opticks_raindrop.gdml: water 100 mm cube in 1.1-index air in lead in
    vacuum, with a border surface drop->medium carrying EFFICIENCY 1."""
    f = Foundry()
    n_low_high = _energy_table([(1.512, 1.0), (5.512, 1.1)])
    f.add_material('Vacuum', n_low_high)
    f.add_material('Pb', n_low_high)
    f.add_material('Air', 1.1)
    f.add_material('Water', 1.333)
    f.add_sensor_surface('medium_container_bs', 1.0)
    b_world = f.boundary('Vacuum', None, None, 'Vacuum')
    b_pb = f.boundary('Vacuum', None, None, 'Pb')
    b_air = f.boundary('Pb', None, None, 'Air')
    b_water = f.boundary('Air', None, 'medium_container_bs', 'Water')
    f.begin_solid('r0')
    f.add_prim(box3(240.0), b_world)
    f.add_prim(box3(220.0), b_pb)
    f.add_prim(box3(200.0), b_air)
    f.add_prim(box3(100.0), b_water)
    f.end_solid()
    f.add_instance(np.eye(4), 0)

    def gs(rng, s):
        return G.input_photon(10)

    return _finish('raindrop', f, dict(max_bounce=31), gs)


def raindrop_input_photons():
    """The 10 photons of tests/test_GPUPhotonFileSource.sh:20-31 as sphoton (10,4,4)."""
    p = np.zeros((10, 4, 4), dtype=np.float32)
    p[:, 0, :3] = (-10.0, -30.0, -90.0)
    p[:, 0, 3] = 0.0
    p[:, 1, :3] = (0.0, 0.287348, 0.957826)
    p[:, 2, :3] = (1.0, 0.0, 0.0)
    p[:, 2, 3] = [420, 420, 420, 450, 450, 500, 500, 500, 500, 500]
    return p


# ---------------------------------------------------------------------------
CSI_N = 1.79


def csi_calorimeter(n_scint=2000, n_cer=500, mean_scint=1000.0, mean_cer=150.0):
    f = Foundry()
    f.add_material('Air', 1.0003)
    f.add_material('CsI', CSI_N, absorption_length=400.0, scattering_length=1000.0)
    f.add_material('Si', 1.0)
    # ground/unified wrapping with REFLECTIVITY 0.95: diffuse 0.95, absorb 0.05
    f.add_surface('Wrap', absorb=0.05, diffuse=0.95)
    f.add_sensor_surface('SiPM', 0.3)
    b_world = f.boundary('Air', None, None, 'Air')
    b_xtal = f.boundary('Air', 'Wrap', 'Wrap', 'CsI')
    b_sipm = f.boundary('CsI', 'SiPM', None, 'Si')
    pitch, fx, fz = 20.5, 20.0, 200.0
    f.begin_solid('r0')
    f.add_prim(box3(400.0, 400.0, 600.0), b_world)
    f.end_solid()
    f.begin_solid('r1')
    f.add_prim(box3(fx, fx, fz), b_xtal)
    f.add_prim(box3(6.0, 6.0, 0.5).moved(translate(0, 0, fz / 2 - 0.25)), b_sipm)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    centers = []
    for j in range(8):
        for i in range(8):
            cx, cy = (i - 3.5) * pitch, (j - 3.5) * pitch
            centers.append((cx, cy))
            f.add_instance(translate(cx, cy, 0.0), 1, sensor_identifier=len(centers) - 1,
                           sensor_index=len(centers) - 1)
    centers = np.array(centers)
    matline_csi = 4 * b_xtal + 3

    def gs(rng, s):
        ns = max(1, int(round(n_scint * s)))
        nc = max(1, int(round(n_cer * s)))
        ci, cj = rng.integers(1, 7, size=2)            # 3x3 cluster centre
        cluster = np.array([(ci + di) + 8 * (cj + dj) for dj in (-1, 0, 1) for di in (-1, 0, 1)])

        def points(n):
            which = cluster[rng.integers(0, 9, size=n)]
            xy = centers[which] + rng.uniform(-9.0, 9.0, size=(n, 2))
            z = rng.uniform(-94.0, 94.0, size=n)
            return np.c_[xy, z]

        def dirs(n):
            v = rng.normal(size=(n, 3))
            return v / np.linalg.norm(v, axis=1)[:, None]

        step_s = rng.uniform(0.1, 5.0, size=ns)
        scint = G.scintillation(rng.poisson(mean_scint, size=ns), points(ns),
                                dirs(ns) * step_s[:, None], step_s, 0.9 * 299.792458,
                                1000.0, matline=matline_csi)
        step_c = rng.uniform(0.1, 5.0, size=nc)
        beta = rng.uniform(0.8, 1.0, size=nc)
        cer = G.cerenkov(rng.poisson(mean_cer, size=nc), points(nc), dirs(nc) * step_c[:, None],
                         step_c, beta, matline_csi, csg.WL_MIN, csg.WL_MAX, CSI_N, mean_cer)
        return np.concatenate([scint, cer])

    icdf = csg.scint_icdf(lambda wl: np.exp(-0.5 * ((wl - 550.0) / 40.0) ** 2), hd_factor=20)
    return _finish('csi_calorimeter', f, dict(max_bounce=32), gs, icdf=icdf, hd_factor=20,
                   matline_csi=matline_csi)


# ---------------------------------------------------------------------------
LS_N_A, LS_N_B = 1.55, -1.0e-4       # liquid scintillator rindex a + b*wl (wl in nm)


def ls_absorption_length(wl):
    """UV-opaque, visible-clear: about 30 mm below 380 nm rising to 4 m above 400 nm."""
    return 30.0 + 4000.0 / (1.0 + np.exp(-(np.asarray(wl, dtype=np.float64) - 390.0) / 5.0))


def scintillator_sphere(num_photon=20_000, reemission_prob=0.8, wavelength=350.0):
    """Re-emitting scintillator: a 600 mm liquid-scintillator sphere in water inside a
    1000 mm sensor shell (efficiency 0.5), lit by an isotropic torch at 350 nm from the
    centre. Every LS property varies with wavelength: rindex a linear ramp, absorption
    length UV-opaque and visible-clear, REEMISSIONPROB a ramp (0.8 at 350 nm), so the
    table lookups interpolate between distinct texels; re-emitted light follows a
    Gaussian spectrum at 430 nm (sigma 15 nm)."""
    f = Foundry()
    f.add_material('Water', 1.333, absorption_length=1e6, scattering_length=1e6)
    f.add_material('LS', lambda wl: LS_N_A + LS_N_B * wl, absorption_length=ls_absorption_length,
                   scattering_length=lambda wl: 3000.0 * (wl / 430.0) ** 4,
                   reemission_prob=lambda wl: np.clip(reemission_prob + 1e-3 * (350.0 - wl), 0.0, 1.0))
    f.add_sensor_surface('PMT', 0.5)
    b_world = f.boundary('Water', None, None, 'Water')
    b_sens = f.boundary('Water', None, 'PMT', 'Water')
    b_ls = f.boundary('Water', None, None, 'LS')
    f.begin_solid('r0')
    f.add_prim(sphere(1200.0), b_world)
    f.add_prim(sphere(1000.0), b_sens)
    f.add_prim(sphere(600.0), b_ls)
    f.end_solid()
    f.add_instance(np.eye(4), 0)

    def gs(rng, s):
        n = max(1, int(round(num_photon * s)))
        return G.torch(n, pos=(0, 0, 0), wavelength=wavelength, type=G.T_POINT_ISOTROPIC)

    icdf = csg.scint_icdf(lambda wl: np.exp(-0.5 * ((wl - 430.0) / 15.0) ** 2), hd_factor=20)
    return _finish('scintillator_sphere', f, dict(max_bounce=32), gs, icdf=icdf, hd_factor=20,
                   matline_ls=4 * b_ls + 3, b_ls=b_ls)


# ---------------------------------------------------------------------------
def pmt_like_solid():
    """union(z-clipped ellipsoid 254/254/184, cylinder r=50 L=200) minus inner zsphere,
    positivised to intersection(union(...), complement(inner)) with 7 nodes."""
    ell = zsphere(254.0, -200.0, 254.0).moved(scale(1.0, 1.0, 184.0 / 254.0))
    neck = cylinder(50.0, -330.0, -130.0)
    inner = zsphere(249.0, -190.0, 249.0).moved(scale(1.0, 1.0, 179.0 / 249.0))
    return (ell | neck) - inner


def fibonacci_sphere(n, r):
    k = np.arange(n, dtype=np.float64) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / n)
    theta = np.pi * (1.0 + 5 ** 0.5) * k
    return r * np.c_[np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)]


def instanced_csg(num_inst=20_000, num_photon=10_000_000, num_points=1000, pmt_radius=19_400.0):
    f = Foundry()
    f.add_material('Rock', 1.0)
    f.add_material('Acrylic', 1.49, absorption_length=10_000.0, scattering_length=50_000.0)
    f.add_material('Water', 1.333, absorption_length=60_000.0, scattering_length=40_000.0)
    f.add_material('Glass', 1.5)
    f.add_implicit_absorber('Implicit_Acrylic_Rock')
    f.add_sensor_surface('Photocathode', 0.25)
    b_world = f.boundary('Rock', None, None, 'Rock')
    b_acr = f.boundary('Rock', None, 'Implicit_Acrylic_Rock', 'Acrylic')
    b_wat = f.boundary('Acrylic', None, None, 'Water')
    b_pmt = f.boundary('Water', 'Photocathode', None, 'Glass')
    f.begin_solid('r0')
    f.add_prim(sphere(20_000.0), b_world)
    f.add_prim(sphere(19_620.0), b_acr)
    f.add_prim(sphere(19_500.0), b_wat)
    f.end_solid()
    f.begin_solid('r1')
    f.add_prim(pmt_like_solid(), b_pmt)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    pos = fibonacci_sphere(num_inst, pmt_radius)
    for i, p in enumerate(pos):
        m = rotate_z_to(-p) @ translate(*p)
        f.add_instance(m, 1, sensor_identifier=i, sensor_index=i)

    def gs(rng, s):
        npt = max(1, int(round(num_points * min(1.0, s))))
        per = max(1, int(round(num_photon * s / npt)))
        u = rng.uniform(size=npt) ** (1.0 / 3.0) * 17_000.0
        v = rng.normal(size=(npt, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        return G.torch(np.full(npt, per), pos=v * u[:, None], wavelength=430.0,
                       type=G.T_POINT_ISOTROPIC)

    return _finish('instanced_csg', f, dict(max_bounce=64), gs, num_inst=num_inst)


# ---------------------------------------------------------------------------
def thin_shells(num_shell=50, num_inst=5000, num_photon=5_000_000, grazing_frac=0.1):
    """50 concentric 1 mm shells at radii 1..20 m, alternating n=1.50/1.33, with 5000
    instanced zsphere+cylinder solids in the gaps between shells (up to 20 m off origin)."""
    f = Foundry()
    f.add_material('Gap', 1.0, absorption_length=1e9, scattering_length=1e9)
    f.add_material('ShellA', 1.50, absorption_length=1e7, scattering_length=1e9)
    f.add_material('ShellB', 1.33, absorption_length=1e7, scattering_length=1e9)
    f.add_material('Rock', 1.0)
    f.add_implicit_absorber('Implicit_Gap_Rock')
    f.add_implicit_absorber('Absorb')
    radii = np.linspace(1000.0, 20_000.0, num_shell)
    b_world = f.boundary('Rock', None, 'Implicit_Gap_Rock', 'Gap')
    f.begin_solid('r0')
    f.add_prim(sphere(21_000.0), b_world)
    for k in reversed(range(num_shell)):
        mat = 'ShellA' if k % 2 == 0 else 'ShellB'
        b_out = f.boundary('Gap', None, None, mat)
        b_in = f.boundary(mat, None, None, 'Gap')
        f.add_prim(sphere(radii[k] + 1.0), b_out)
        f.add_prim(sphere(radii[k]), b_in)
    f.end_solid()
    b_obj = f.boundary('Gap', 'Absorb', None, 'Rock')
    f.begin_solid('r1')
    f.add_prim(zsphere(60.0, -30.0, 60.0) | cylinder(25.0, -80.0, -20.0), b_obj)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    rng = np.random.default_rng(42)
    gaps = 0.5 * (radii[1:] + radii[:-1] + 1.0)
    for i in range(num_inst):
        r = gaps[rng.integers(0, len(gaps))]
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        m = rotate_z_to(rng.normal(size=3)) @ translate(*(v * r))
        f.add_instance(m, 1)

    def gs(rng, s):
        n = max(1, int(round(num_photon * s)))
        return G.torch(n, pos=(0.0, 0.0, 0.0), wavelength=400.0, type=G.T_POINT_ISOTROPIC)

    return _finish('thin_shells', f, dict(max_bounce=64), gs, radii=radii,
                   grazing_frac=grazing_frac)


def thin_shell_grazing_photons(scene, n, seed=42):
    """Input photons whose directions are within 1e-3 rad of tangent to one of the outer
    shells: start near the origin region offset so the ray grazes radius R."""
    rng = np.random.default_rng(seed)
    radii = scene.meta['radii']
    R = radii[rng.integers(len(radii) // 2, len(radii), size=n)] + rng.uniform(0.0, 1.0, size=n)
    # photon starts at distance d from origin along -x, travels +y-ish so closest approach ~R
    p = np.zeros((n, 4, 4), dtype=np.float32)
    ang = rng.uniform(-1e-3, 1e-3, size=n)
    start = np.c_[R * np.cos(ang), -np.sqrt(np.maximum(0, (R + 2.0) ** 2 - R ** 2)) * np.ones(n),
                  np.zeros(n)]
    d = np.c_[np.zeros(n), np.ones(n), np.zeros(n)]
    rot = rng.uniform(0, 2 * np.pi, size=n)
    c, s = np.cos(rot), np.sin(rot)
    # rotate about the x axis to spread the photons over the sphere
    start = np.c_[start[:, 0], c * start[:, 1] - s * start[:, 2], s * start[:, 1] + c * start[:, 2]]
    d = np.c_[d[:, 0] + ang, c * d[:, 1] - s * d[:, 2], s * d[:, 1] + c * d[:, 2]]
    d /= np.linalg.norm(d, axis=1)[:, None]
    pol = np.cross(d, np.array([1.0, 0.0, 0.0]))
    pol /= np.linalg.norm(pol, axis=1)[:, None]
    p[:, 0, :3] = start
    p[:, 1, :3] = d
    p[:, 2, :3] = pol
    p[:, 2, 3] = 400.0
    return p


def serving_stream(rng, n_events=256, median=200_000, sigma=1.0, lo=1_000, hi=20_000_000, ext=60.0,
                   rate=40.0, wavelength=420.0):
    """The serving workload's request stream: n_events events arriving as a Poisson
    process at `rate` per second, photons per event log-normal (median, sigma of the log)
    clipped to [lo, hi], each event a batch of isotropic torch gensteps (one per 20k
    photons, at most 64) at random origins in the cube |x|,|y|,|z| < ext mm.
    Returns (gensteps per event, photon counts, arrival times in s)."""
    counts = np.clip(np.exp(np.log(median) + sigma * rng.normal(size=n_events)), lo, hi).astype(np.int64)
    evs = []
    for c in counts:
        ngs = int(max(1, min(64, c // 20000)))
        pos = rng.uniform(-ext, ext, size=(ngs, 3))
        per = np.full(ngs, c // ngs)
        per[0] += c - per.sum()
        evs.append(G.torch(per, pos=pos, wavelength=wavelength, type=G.T_POINT_ISOTROPIC))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_events))
    return evs, counts, arrivals


SCENES = dict(nested_boxes=nested_boxes, csi_calorimeter=csi_calorimeter,
              instanced_csg=instanced_csg, thin_shells=thin_shells, raindrop=raindrop)


def hidden_instances(num_hidden=40, hide=True):
    """A water ball in a black box with small glass spheres (solid 1) and glass cubes
    (solid 2) placed around it; the cubes are the last instances. With hide=False the
    cubes are left out of the geometry altogether, so that a launch whose visibility mask
    drops solid 2's bit can be compared with a geometry that never held them."""
    f = Foundry()
    f.add_material('Rock', 1.0)
    f.add_material('Water', 1.333, absorption_length=20_000.0, scattering_length=5_000.0)
    f.add_material('Glass', 1.5, absorption_length=1_000.0)
    f.add_implicit_absorber('Implicit_Water_Rock')
    b_world = f.boundary('Rock', None, 'Implicit_Water_Rock', 'Water')
    b_glass = f.boundary('Water', None, None, 'Glass')
    f.begin_solid('r0')
    f.add_prim(box3(2000.0), b_world)
    f.end_solid()
    f.begin_solid('r1')
    f.add_prim(sphere(40.0), b_glass)
    f.end_solid()
    f.begin_solid('r2')
    f.add_prim(box3(60.0), b_glass)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    rng = np.random.default_rng(11)
    for _ in range(40):
        f.add_instance(translate(*rng.uniform(-700, 700, size=3)), 1)
    if hide:
        for _ in range(num_hidden):
            f.add_instance(translate(*rng.uniform(-700, 700, size=3)), 2)

    def gs(rng, s):
        n = max(1, int(round(50_000 * s)))
        return G.torch(n, pos=(0.0, 0.0, 0.0), wavelength=420.0, type=G.T_POINT_ISOTROPIC)

    return _finish('hidden_instances' if hide else 'without_cubes', f, dict(max_bounce=32), gs)
