"""GDML front-end (oxr.gdml) and the photon/torch sources of the reference's example
applications (oxr.sources), on the reference's own test geometries and configurations
(tests/golden/geom, tests/golden/config) and on small inline GDML documents whose
conversion the Geant4 conventions fix."""
import math
import os

import numpy as np
import pytest

from oxr import abi, gdml, gensteps as G, scenes, sources

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
F = abi.FLAGS


def geom(name):
    return os.path.join(GOLDEN, 'geom', name + '.gdml')


def config(name):
    return os.path.join(GOLDEN, 'config', name + '.json')


def flags(r):
    return r['photon'].view(np.uint32)[:, 3, 3]


# ---------------------------------------------------------------- evaluator
def test_evaluator_units_and_expressions():
    ev = gdml.Evaluator()
    assert ev('2.034*eV') == pytest.approx(2.034e-6)
    assert ev('1.5*m + 20*cm') == pytest.approx(1700.0)
    assert ev('2^3') == 8.0 and ev('-(1+2)*3') == -9.0
    assert ev('sqrt(16)+cos(0)') == 5.0
    assert ev('90*deg') == pytest.approx(math.pi / 2)
    ev.vars['W'] = 3.0
    ev.loop['i'] = 2.0
    assert ev('-8.7 + i*2.2') == pytest.approx(-4.3)
    assert ev('W*i') == 6.0
    assert ev.name('Crystal_{i}') == 'Crystal_2'
    with pytest.raises(gdml.GdmlError):
        ev('__import__("os")')
    with pytest.raises(gdml.GdmlError):
        ev('undefined_thing + 1')
    assert gdml.strip_name('G4_WATER0x560299384e40') == 'G4_WATER'


DOC = """<?xml version="1.0"?>
<gdml>
  <define>
    <matrix coldim="2" name="RI_GLASS" values="1.0*eV 1.5 4.0*eV 1.5"/>
    <matrix coldim="2" name="RI_AIR" values="1.0*eV 1.0 4.0*eV 1.0"/>
    <matrix coldim="2" name="EFF" values="1.0*eV 0.25 4.0*eV 0.25"/>
    <matrix coldim="2" name="REFL" values="1.0*eV 0.9 4.0*eV 0.9"/>
    <variable name="i" value="0"/>
    <rotation name="rz90" z="90" unit="deg"/>
  </define>
  <materials>
    <material name="Air"><property name="RINDEX" ref="RI_AIR"/></material>
    <material name="Glass"><property name="RINDEX" ref="RI_GLASS"/></material>
    <material name="Steel"/>
  </materials>
  <solids>
    <box name="World" x="600" y="600" z="600"/>
    <box name="Holder" x="300" y="300" z="150"/>
    <box name="Core" x="2" y="2" z="2"/>
    <box name="Arm" x="2" y="2" z="2"/>
    <union name="Lshape"><first ref="Core"/><second ref="Arm"/><position name="p" x="10" y="0" z="0"/></union>
    <box name="Plate" x="10" y="10" z="1"/>
    %(SOLIDS)s
    <opticalsurface name="SensorOS" finish="polished" model="glisur" type="dielectric_dielectric" value="1">
      <property name="EFFICIENCY" ref="EFF"/>
    </opticalsurface>
    <opticalsurface name="MirrorOS" finish="ground" model="unified" type="dielectric_metal" value="0.5">
      <property name="REFLECTIVITY" ref="REFL"/>
    </opticalsurface>
  </solids>
  <structure>
    <volume name="LV_L"><materialref ref="Glass"/><solidref ref="Lshape"/></volume>
    <volume name="LV_Plate"><materialref ref="Steel"/><solidref ref="Plate"/></volume>
    %(VOLUMES)s
    <volume name="LV_Holder">
      <materialref ref="Air"/><solidref ref="Holder"/>
      <physvol name="L_pv"><volumeref ref="LV_L"/><rotationref ref="rz90"/></physvol>
      <loop for="i" from="0" to="2" step="1">
        <physvol name="Plate_{i}"><volumeref ref="LV_Plate"/>
          <position name="pp" x="100" y="0" z="20*i" unit="mm"/></physvol>
      </loop>
    </volume>
    <volume name="LV_World">
      <materialref ref="Air"/><solidref ref="World"/>
      <physvol name="Holder_pv"><volumeref ref="LV_Holder"/></physvol>
      %(PLACE)s
    </volume>
    <skinsurface name="PlateSkin" surfaceproperty="MirrorOS"><volumeref ref="LV_Plate"/></skinsurface>
    <bordersurface name="LBorder" surfaceproperty="SensorOS"><physvolref ref="L_pv"/><physvolref ref="Holder_pv"/></bordersurface>
  </structure>
  <setup name="Default" version="1.0"><world ref="LV_World"/></setup>
</gdml>
"""


def write_doc(tmp_path, solids='', volumes='', place='', name='doc.gdml'):
    p = tmp_path / name
    p.write_text(DOC % dict(SOLIDS=solids, VOLUMES=volumes, PLACE=place))
    return str(p)


def prim_aabb(g, k):
    return g.geometry['prim'].reshape(-1, 16)[k, 8:14].astype(np.float64)


def test_placement_rotation_is_inverted_and_loops_expand(tmp_path):
    """GDML physvol rotations are frame rotations: G4GDMLRead places the volume with
    the inverse of Rz Ry Rx. The L shape's arm at local +x lands at -y after a 90 deg
    z rotation (it would land at +y with the rotation applied as given)."""
    g = gdml.load_gdml(write_doc(tmp_path))
    assert g.prim_volumes[:3] == ['LV_World', 'Holder_pv', 'L_pv']
    lo_hi = prim_aabb(g, 2)
    assert np.allclose(lo_hi, [-1, -11, -1, 1, 1, 1], atol=1e-5), lo_hi
    # the loop placed three plates at z = 0, 20, 40
    zs = [0.5 * (prim_aabb(g, k)[2] + prim_aabb(g, k)[5]) for k in range(3, 6)]
    assert g.prim_volumes[3:6] == ['Plate_0', 'Plate_1', 'Plate_2'] and np.allclose(zs, [0, 20, 40])


def test_boundaries_surfaces_and_implicits(tmp_path):
    """U4TreeBorder rules: the border surface L_pv -> Holder_pv is the glass volume's isur;
    the plate's skin surface is its osur, and it has no isur because steel has no RINDEX
    (isur is only looked up from a material with RINDEX)."""
    doc = write_doc(tmp_path)
    g = gdml.load_gdml(doc)
    assert g.boundaries == ['Air///Air', 'Air//LBorder/Glass', 'Air/PlateSkin//Steel']
    assert g.surfaces == ['LBorder', 'PlateSkin']
    phys = g.physics
    bnd, opt = phys['bnd'], np.asarray(phys['optical']).reshape(-1, 4, 4)
    # sensor payload (eff, 1-eff, 0, 0); ground mirror (0, 1-R, 0, R)
    assert np.allclose(bnd[1, 2, 0, :, :], [0.25, 0.75, 0.0, 0.0])
    assert np.allclose(bnd[2, 1, 0, :, :], [0.0, 0.1, 0.0, 0.9])
    assert list(opt[1, 2]) == [1, 2, 0, 100]             # 'S' -> ems Surface, polished, value 1
    assert list(opt[2, 1]) == [2, 2, 3, 50]              # ground = 3, value 0.5
    # materials in postorder first use: Glass, Steel, Air
    assert g.materials == ['Glass', 'Steel', 'Air']
    assert list(opt[0, 0]) == [3, 0, 0, 0] and list(opt[1, 3]) == [1, 0, 0, 0]
    # steel keeps the default table rows (sproplist): n 1, lengths 1e12
    assert np.allclose(bnd[2, 3, 0, :, 0], 1.0) and np.allclose(bnd[2, 3, 0, :, 1:3], 1e12)


def test_implicit_absorber_without_surface(tmp_path):
    doc = write_doc(tmp_path, volumes='<volume name="LV_Block"><materialref ref="Steel"/><solidref ref="Core"/></volume>',
                    place='<physvol name="Block_pv"><volumeref ref="LV_Block"/><position name="q" x="-100" y="0" z="0"/></physvol>')
    g = gdml.load_gdml(doc)
    assert g.boundaries[-1] == 'Air/Implicit_RINDEX_NoRINDEX_LV_World_Block_pv//Steel'
    k = g.surfaces.index('Implicit_RINDEX_NoRINDEX_LV_World_Block_pv')
    assert np.allclose(g.physics['bnd'][-1, 1, 0, :, :], [0, 1, 0, 0])
    assert list(np.asarray(g.physics['optical']).reshape(-1, 4, 4)[-1, 1]) == [k + 1, 2, 1, 100]


def test_group_velocity_from_rindex():
    """Constant n gives c/n; for n linear in ln E the midpoint rule is exact:
    vg = c / (n + dn/dlnE)."""
    e = np.array([1.0, 2.0, 3.0, 4.0]) * 1e-6
    vg = gdml._group_velocity(np.c_[e, np.full(4, 1.5)])
    assert np.allclose(vg[:, 1], gdml.C_LIGHT / 1.5)
    a, b = 1.4, 0.05
    n = a + b * np.log(e / e[0])
    vg = gdml._group_velocity(np.c_[e, n])
    mids = 0.5 * (e[1:-1] + e[:-2])
    assert np.allclose(vg[1:-1, 0], mids)
    nbar = 0.5 * (n[1:-1] + n[:-2])
    assert np.allclose(vg[1:-1, 1], gdml.C_LIGHT / (nbar + b))


def test_polyhedra_trd_trap(tmp_path, oracle_lib):
    """trd/trap (not converted by U4Solid) become convex polyhedra whose inside test
    matches the G4 definitions; intersects agree with a brute-force march."""
    solids = ('<trd name="T" x1="40" x2="20" y1="30" y2="10" z="50"/>'
              '<trap name="P" z="40" theta="10" phi="20" y1="20" x1="10" x2="14" alpha1="5" '
              'y2="16" x3="8" x4="12" alpha2="5" aunit="deg"/>')
    volumes = ('<volume name="LV_T"><materialref ref="Glass"/><solidref ref="T"/></volume>'
               '<volume name="LV_P"><materialref ref="Glass"/><solidref ref="P"/></volume>')
    place = ('<physvol name="T_pv"><volumeref ref="LV_T"/><position name="a" x="0" y="100" z="0"/></physvol>'
             '<physvol name="P_pv"><volumeref ref="LV_P"/><position name="b" x="0" y="-100" z="0"/></physvol>')
    g = gdml.load_gdml(write_doc(tmp_path, solids, volumes, place))
    assert any('trd T' in n for n in g.notes) and any('trap P' in n for n in g.notes)
    kT = g.prim_volumes.index('T_pv')
    plan = g.geometry['plan']
    tr = plan[:6].astype(np.float64)
    rng = np.random.default_rng(3)
    p = rng.uniform(-30, 30, size=(20000, 3))
    inside_planes = np.all(p @ tr[:, :3].T < tr[:, 3], axis=1)
    z = p[:, 2]
    fx = 20.0 + (10.0 - 20.0) * (z + 25.0) / 50.0
    fy = 15.0 + (5.0 - 15.0) * (z + 25.0) / 50.0
    want = (np.abs(z) < 25.0) & (np.abs(p[:, 0]) < fx) & (np.abs(p[:, 1]) < fy)
    assert np.mean(inside_planes == want) > 0.999
    # the trap's planes: every vertex on three planes, the centroid inside all six
    tp = plan[6:12].astype(np.float64)
    assert np.all(tp[:, 3] > 0)           # origin (trap centre at z=0 line) inside
    from test_oracle_pins import _first_crossing, _rays
    O = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config())
    o = np.c_[rng.uniform(-15, 15, 300), np.full(300, 200.0), rng.uniform(-20, 20, 300)]
    d = np.tile([0.0, -1.0, 0.0], (300, 1))
    isect, valid = O.intersect_prim(kT, _rays(o, d), 0)
    inside_T = lambda q: np.all((q - [0, 100, 0]) @ tr[:, :3].T < tr[:, 3], axis=1)
    want_t = _first_crossing(inside_T, o, d, 200.0)
    got = np.where(valid.astype(bool), isect[:, 3], np.nan)
    agree = (np.isnan(want_t) & np.isnan(got)) | (np.abs(got - want_t) < 2e-3)
    assert agree.mean() > 0.99


def test_deep_subtraction_chain_is_balanced(tmp_path, oracle_lib):
    """A plate minus 40 holes: the positivised intersection chain (height 40) is
    rebalanced to height 6; intersects still match a brute-force march."""
    holes = ''.join(f'<tube name="H{k}" rmax="3" z="30" deltaphi="360" aunit="deg"/>' for k in range(40))
    chain = []
    prev = 'Base'
    for k in range(40):
        x, y = -60 + 15 * (k % 8) + 7.5, -30 + 15 * (k // 8) + 7.5
        chain.append(f'<subtraction name="S{k}"><first ref="{prev}"/><second ref="H{k}"/>'
                     f'<position name="hp{k}" x="{x}" y="{y}" z="0"/></subtraction>')
        prev = f'S{k}'
    solids = '<box name="Base" x="120" y="75" z="10"/>' + holes + ''.join(chain)
    volumes = f'<volume name="LV_S"><materialref ref="Glass"/><solidref ref="{prev}"/></volume>'
    place = '<physvol name="S_pv"><volumeref ref="LV_S"/><position name="s" x="0" y="0" z="-100"/></physvol>'
    g = gdml.load_gdml(write_doc(tmp_path, solids, volumes, place))
    k = g.prim_volumes.index('S_pv')
    num_node = int(g.geometry['prim'].reshape(-1, 16).view(np.int32)[k, 0])
    assert num_node == 127                    # 41 operands -> height 6

    centres = [(-60 + 15 * (k % 8) + 7.5, -30 + 15 * (k // 8) + 7.5) for k in range(40)]

    def inside(q):
        q = q - [0, 0, -100]
        inb = np.all(np.abs(q) < [60, 37.5, 5], axis=1)
        for cx, cy in centres:
            inb &= ~(((q[:, 0] - cx) ** 2 + (q[:, 1] - cy) ** 2 < 9.0) & (np.abs(q[:, 2]) < 15))
        return inb
    from test_oracle_pins import _first_crossing, _rays
    O = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config())
    rng = np.random.default_rng(9)
    o = np.c_[rng.uniform(-70, 70, 400), rng.uniform(-45, 45, 400), np.full(400, -50.0)]
    d = np.tile([0.0, 0.0, -1.0], (400, 1))
    isect, valid = O.intersect_prim(k, _rays(o, d), 0)
    want = _first_crossing(inside, o, d, 100.0)
    got = np.where(valid.astype(bool), isect[:, 3], np.nan)
    agree = (np.isnan(want) & np.isnan(got)) | (np.abs(got - want) < 2e-3)
    assert agree.mean() > 0.99, agree.mean()


def test_unsupported_solids_raise(tmp_path):
    doc = write_doc(tmp_path, solids='<torus name="Tor" rmin="1" rmax="2" rtor="10" deltaphi="360" aunit="deg"/>',
                    volumes='<volume name="LV_Tor"><materialref ref="Glass"/><solidref ref="Tor"/></volume>',
                    place='<physvol name="Tor_pv"><volumeref ref="LV_Tor"/></physvol>')
    with pytest.raises(gdml.GdmlError, match='torus'):
        gdml.load_gdml(doc)
    doc = write_doc(tmp_path, solids='<sphere name="Half" rmax="5" deltaphi="180" deltatheta="180" aunit="deg"/>',
                    volumes='<volume name="LV_H"><materialref ref="Glass"/><solidref ref="Half"/></volume>',
                    place='<physvol name="H_pv"><volumeref ref="LV_H"/></physvol>', name='b.gdml')
    with pytest.raises(gdml.GdmlError, match='phi'):
        gdml.load_gdml(doc)


# ---------------------------------------------------------------- reference geometries
def test_raindrop_gdml_matches_hand_built_scene():
    """tests/geom/opticks_raindrop.gdml imports to the arrays of the hand-built raindrop
    scene: same solid, nodes, instance and prims (the mesh index is the postorder lvid)."""
    g = gdml.load_gdml(geom('opticks_raindrop'))
    s = scenes.raindrop()
    for k in ('solid', 'node', 'inst'):
        assert np.array_equal(g.geometry[k].view(np.uint32), s.geometry[k].view(np.uint32)), k
    a = g.geometry['prim'].reshape(-1, 16).view(np.uint32).copy()
    b = s.geometry['prim'].reshape(-1, 16).view(np.uint32).copy()
    assert list(a[:, 5]) == [3, 2, 1, 0]
    a[:, 5] = b[:, 5] = 0
    assert np.array_equal(a, b)
    assert g.boundaries == ['VACUUM///VACUUM', 'VACUUM///G4_Pb', 'G4_Pb///G4_AIR',
                            'G4_AIR//medium_container_bs/G4_WATER']
    # GROUPVEL comes from the file when present: water 224.901 mm/ns
    assert np.allclose(g.physics['bnd'][3, 3, 1, :, 0], 224.901)


def test_photon_file_source_kat(oracle_lib):
    """tests/test_GPUPhotonFileSource.sh: 10 photons from the text file through the
    raindrop GDML give 10 hits with wavelengths 3 x 420, 2 x 450, 5 x 500 nm, history
    TO BT SD; the commented file loads 2 photons."""
    ip = sources.load_photons_txt(os.path.join(GOLDEN, 'raindrop_photons.txt'))
    assert ip.shape == (10, 4, 4)
    assert len(sources.load_photons_txt(os.path.join(GOLDEN, 'raindrop_photons_comments.txt'))) == 2
    g = gdml.load_gdml(geom('opticks_raindrop'))
    o = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config(max_bounce=31))
    r = o.simulate(G.input_photon(10), input_photons=ip, seq=True)
    fm = flags(r)
    hits = r['photon'][(fm & F['SURFACE_DETECT']) != 0]
    assert len(hits) == 10
    wl = np.round(hits[:, 2, 3]).astype(int)
    assert (wl == 420).sum() == 3 and (wl == 450).sum() == 2 and (wl == 500).sum() == 5
    assert np.all(fm == (F['TORCH'] | F['BOUNDARY_TRANSMIT'] | F['SURFACE_DETECT']))


def test_photon_file_malformed_lines(tmp_path, capsys):
    p = tmp_path / 'ph.txt'
    p.write_text('# header\n1 2 3 0 0 0 1 1 0 0 420\n1 2 three 0 0 0 1 1 0 0 420\n1 2 3\n\n4 5 6 1 0 1 0 0 1 0 500 extra\n')
    ph = sources.load_photons_txt(str(p))
    assert len(ph) == 2 and ph[1, 0, 0] == 4.0 and ph[1, 2, 3] == 500.0 and ph[1, 0, 3] == 1.0
    assert 'malformed line 3' in capsys.readouterr().err
    out = tmp_path / 'round.txt'
    sources.write_photons_txt(str(out), ph)
    assert np.array_equal(sources.load_photons_txt(str(out)), ph)


def test_torch_config_genstep():
    """src/config.cpp ReadConfig: storch fields from JSON, mom normalised."""
    tc = sources.torch_config(config('8x8SiPM_crystal'))
    gs = tc.genstep
    u = gs.view(np.uint32)
    assert tc.numphoton == 1000 and tc.event_mode == 'Minimal' and tc.max_slot == 1000000
    assert u[0, 0, 0] == G.GS_TORCH and u[0, 0, 3] == 1000
    assert np.allclose(gs[0, 1], [-8.7, -8.7, 4.0, 0.0])
    assert np.allclose(gs[0, 2], [0.0, 0.0, 1.0, 0.0])         # weight 0 from the file
    assert np.allclose(gs[0, 3], [1.0, 0.0, 0.0, 420.0])
    assert gs[0, 5, 0] == np.float32(0.5) and u[0, 5, 2] == 255 and u[0, 5, 3] == G.T_DISC
    dev = sources.torch_config(config('dev')).genstep
    m = np.array([0.0, 0.3, 1.0], dtype=np.float32)
    assert np.allclose(dev[0, 2, :3], m / np.linalg.norm(m), atol=1e-7)
    with pytest.raises(ValueError):
        sources.torch_config({'torch': {'type': 'bogus', 'gentype': 'TORCH'}})


def test_csi_sipm_array(oracle_lib):
    """tests/geom/8x8SiPM_w_CSI_optial_grease.gdml with config/8x8SiPM_crystal.json
    (tests/test_GPUPhotonSource_8x8SiPM.sh): hits are produced, all on SiPM active areas
    (EFFICIENCY 1 skin), and the scintillator supplies an HD20 ICDF."""
    g = gdml.load_gdml(geom('8x8SiPM_w_CSI_optial_grease'))
    assert len(g.prim_volumes) == 194
    assert sum(v.startswith('Crystal_') for v in g.prim_volumes) == 64
    assert sum(v.startswith('SiPM_') for v in g.prim_volumes) == 64
    assert 'Air/CrystalSkin/CrystalSkin/Crystal' in g.boundaries
    assert g.physics['icdf'].shape == (3, 4096) and g.physics['hd_factor'] == 20
    tc = sources.torch_config(config('8x8SiPM_crystal'))
    o = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config(max_bounce=32))
    r = o.simulate(tc.genstep)
    fm = flags(r)
    hit = (fm & F['SURFACE_DETECT']) != 0
    assert 100 < hit.sum() < 1000
    pos = r['photon'][hit, 0, :3]
    assert np.all(np.abs(pos[:, 2] - 8.3) <= 0.1 + 1e-4)
    sipm = g.boundaries.index('Air/SiPMActiveSkin/SiPMActiveSkin/EntranceWindow')
    u = r['photon'].view(np.uint32)
    assert np.all((u[hit, 3, 0] >> 16) & 0xffff == sipm)


def test_pfrich_loads_and_simulates(oracle_lib):
    g = gdml.load_gdml(geom('pfrich_min_FINAL'))
    assert len(g.prim_volumes) == 103 and len(g.geometry['plan']) > 0
    assert any(n.startswith('trap') for n in g.notes) and any(n.startswith('trd') for n in g.notes)
    o = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config(max_bounce=32))
    r = o.simulate(G.torch(4000, pos=(0, 0, 0), type=G.T_POINT_ISOTROPIC, wavelength=400.0))
    fm = flags(r)
    assert ((fm & F['SURFACE_DETECT']) != 0).sum() > 0 and ((fm & F['BOUNDARY_TRANSMIT']) != 0).mean() > 0.9


def test_raindrop_simg4ox_geometry(oracle_lib):
    """tests/geom/raindrop.gdml (tests/test_simg4ox.sh): a water drop in air inside a lead
    container with a border surface from the air to the lead. Lead has no RINDEX, so the
    surface lands on the inner side only; photons from the drop's centre leave the water
    and are absorbed at the container."""
    g = gdml.load_gdml(geom('raindrop'))
    assert g.prim_volumes == ['VACUUM_lv', 'container_pv', 'medium_pv', 'drop_pv']
    assert g.boundaries == ['VACUUM///VACUUM', 'VACUUM///G4_Pb', 'G4_Pb//medium_container_bs/G4_AIR',
                            'G4_AIR///G4_WATER']
    o = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config(max_bounce=32))
    r = o.simulate(G.torch([5000] * 2, pos=(0.0, 0.0, 0.0), type=G.T_POINT_ISOTROPIC, wavelength=420.0))
    fm = flags(r)
    assert ((fm & F['BOUNDARY_TRANSMIT']) != 0).mean() > 0.9
    assert ((fm & F['SURFACE_ABSORB']) != 0).mean() > 0.8


@pytest.mark.parametrize('name', ['sphere_leak', 'basic_detector', 'opticks_raindrop_with_scintillation'])
def test_other_reference_geometries(oracle_lib, name):
    g = gdml.load_gdml(geom(name))
    conf = 'sphere_leak' if name == 'sphere_leak' else 'dev'
    gs = sources.torch_config(config(conf)).genstep.copy()
    gs.view(np.uint32)[0, 0, 3] = min(int(gs.view(np.uint32)[0, 0, 3]), 5000)
    o = oracle_lib.Oracle(g.geometry, g.physics, abi.default_config(max_bounce=32))
    r = o.simulate(gs)
    assert len(r['photon']) == gs.view(np.uint32)[0, 0, 3]
    assert ((flags(r) & F['SURFACE_DETECT']) != 0).sum() > 0
