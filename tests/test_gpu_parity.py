"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on the same
seeded inputs. Integer/flag/index outputs and the FP32 photon arrays must agree bit for bit."""
import numpy as np
import pytest

from oxr import abi, gensteps as G, scenes
from oxr.csg import (Foundry, box3, sphere, zsphere, cylinder, cone, convexpolyhedron, phicut, scale, translate,
                     rotate_z_to, contiguous, discontiguous, overlap, hyperboloid, halfspace)

pytestmark = pytest.mark.gpu

SCENES = {
    'nested_boxes': (lambda: scenes.nested_boxes(), 0.2),
    'csi_calorimeter': (lambda: scenes.csi_calorimeter(), 0.01),
    'instanced_csg': (lambda: scenes.instanced_csg(num_inst=2000), 0.002),
    'thin_shells': (lambda: scenes.thin_shells(num_inst=300), 0.002),
}
_cache = {}


def scene(name):
    if name not in _cache:
        _cache[name] = SCENES[name][0]()
    return _cache[name]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def run_engine(engine_mod, s, gs, cfg, form=abi.OXR_FORM_PLAIN, inp=None, pre=None, event_index=0, slot=0):
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        if inp is not None:
            e.set_input_photons(inp)
        if pre is not None:
            e.set_precooked(pre)
        n = e.set_gensteps(gs, event_index=event_index, slot_offset=slot)
        st = e.launch(form)
        out = dict(photon=e.photons(), hit=e.hits(), seed=e.seed(), stats=st, n=n)
        if cfg.seq_enabled:
            out['seq'] = e.seq()
        if cfg.max_record:
            out['record'] = e.get(abi.OXR_GET_RECORD)
    return out


def run_oracle(oracle_lib, s, gs, cfg, inp=None, pre=None, event_index=0, slot=0, cull=False):
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    if cull:
        o.set_cull(True)
    return o.simulate(gs, event_index=event_index, slot_offset=slot, precooked=pre, input_photons=inp,
                      max_record=cfg.max_record)


@pytest.mark.parametrize('name', list(SCENES))
@pytest.mark.parametrize('schedule', [abi.OXR_RNG_PRODUCTION, abi.OXR_RNG_ALIGNED])
def test_simulate_bitexact(engine_mod, oracle_lib, name, schedule):
    s = scene(name)
    cfg = abi.default_config(**s.config, seq_enabled=1, rng_schedule=schedule)
    gs = s.gensteps(scale_photons=SCENES[name][1])
    g = run_engine(engine_mod, s, gs, cfg)
    r = run_oracle(oracle_lib, s, gs, cfg, cull=True)
    print(name, schedule, g['n'], 'photons', g['stats']['bounce_steps'], 'steps', g['stats']['num_hit'], 'hits')
    assert g['n'] == r['total'] > 0
    assert np.array_equal(bits(g['photon']), bits(r['photon']))
    assert np.array_equal(g['seq'], r['seq'])
    assert g['stats']['bounce_steps'] == int(r['bounce'].sum())
    # hits are the SD photons in photon order (SU::copy_if semantics)
    sel = (bits(r['photon'])[:, 3, 3] & cfg.hitmask) == cfg.hitmask
    assert np.array_equal(bits(g['hit']), bits(r['photon'][sel]))


@pytest.mark.parametrize('name', list(SCENES))
@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_production_kernels_bitexact(engine_mod, oracle_lib, name, form):
    """Without seqhis/records/precooked streams the engine runs its lean kernel variants;
    their photons and hits must equal the oracle's too."""
    s = scene(name)
    cfg = abi.default_config(**s.config)
    gs = s.gensteps(scale_photons=SCENES[name][1])
    g = run_engine(engine_mod, s, gs, cfg, form)
    r = run_oracle(oracle_lib, s, gs, cfg, cull=True)
    assert np.array_equal(bits(g['photon']), bits(r['photon']))
    sel = (bits(r['photon'])[:, 3, 3] & cfg.hitmask) == cfg.hitmask
    assert np.array_equal(bits(g['hit']), bits(r['photon'][sel]))
    assert g['stats']['bounce_steps'] == int(r['bounce'].sum())


@pytest.mark.parametrize('name', list(SCENES))
def test_persistent_equals_plain(engine_mod, name):
    s = scene(name)
    cfg = abi.default_config(**s.config, seq_enabled=1)
    gs = s.gensteps(scale_photons=SCENES[name][1])
    a = run_engine(engine_mod, s, gs, cfg, abi.OXR_FORM_PLAIN)
    b = run_engine(engine_mod, s, gs, cfg, abi.OXR_FORM_PERSISTENT)
    assert np.array_equal(bits(a['photon']), bits(b['photon']))
    assert np.array_equal(a['seq'], b['seq'])
    assert np.array_equal(bits(a['hit']), bits(b['hit']))
    assert a['stats']['bounce_steps'] == b['stats']['bounce_steps']


def test_raindrop_known_answer(engine_mod, oracle_lib):
    """tests/test_GPUPhotonFileSource.sh: 10 input photons -> 10 hits, wavelengths preserved."""
    s = scenes.raindrop()
    cfg = abi.default_config(**s.config, seq_enabled=1)
    inp = scenes.raindrop_input_photons()
    g = run_engine(engine_mod, s, s.gensteps(), cfg, inp=inp)
    assert len(g['hit']) == 10
    assert sorted(g['hit'][:, 2, 3].tolist()) == [420] * 3 + [450] * 2 + [500] * 5
    r = run_oracle(oracle_lib, s, s.gensteps(), cfg, inp=inp)
    assert np.array_equal(bits(g['photon']), bits(r['photon']))


def test_precooked_streams(engine_mod, oracle_lib):
    s = scene('nested_boxes')
    cfg = abi.default_config(**s.config, seq_enabled=1)
    gs = s.gensteps(scale_photons=0.02)
    n = G.num_photons(gs)
    rng = np.random.default_rng(7)
    for pre in (rng.uniform(1e-7, 1.0, size=(n, 256)).astype(np.float32),     # (N,16,16) precooked
                np.full((n, 16), 0.5, dtype=np.float32),                        # degenerate, wraps
                np.tile(np.float32([0.3, 0.7, 0.9]), (n, 1)),                   # short cycle
                np.full((n, 8), 1.0, dtype=np.float32)):                        # u=1: zero distances
        g = run_engine(engine_mod, s, gs, cfg, pre=pre)
        r = run_oracle(oracle_lib, s, gs, cfg, pre=pre)
        assert np.array_equal(bits(g['photon']), bits(r['photon']))
        assert np.array_equal(g['seq'], r['seq'])


def test_seed_expansion(engine_mod, oracle_lib):
    s = scene('nested_boxes')
    counts = np.random.default_rng(8).integers(0, 40, size=500)
    counts[[0, 7, 499]] = 0
    gs = G.torch(counts)
    g = run_engine(engine_mod, s, gs, abi.default_config(**s.config))
    assert np.array_equal(g['seed'], oracle_lib.fill_seed(gs))
    assert np.array_equal(g['seed'], np.repeat(np.arange(500), counts))


def test_event_index_and_slot_offset(engine_mod, oracle_lib):
    s = scene('csi_calorimeter')
    cfg = abi.default_config(**s.config, seq_enabled=1)
    gs = s.gensteps(scale_photons=0.003)
    g = run_engine(engine_mod, s, gs, cfg, event_index=3, slot=1000)
    r = run_oracle(oracle_lib, s, gs, cfg, event_index=3, slot=1000)
    assert np.array_equal(bits(g['photon']), bits(r['photon']))
    assert np.all(bits(g['photon'])[:, 3, 2] == 1000 + np.arange(g['n']))
    g0 = run_engine(engine_mod, s, gs, cfg, event_index=0, slot=1000)
    assert not np.array_equal(bits(g['photon']), bits(g0['photon']))


def test_slicing_matches_single_launch(engine_mod):
    """QSim::simulate genstep slicing (QSim.cc:455-526): multi-launch == single launch."""
    s = scene('csi_calorimeter')
    gs = s.gensteps(scale_photons=0.005)
    cfg = abi.default_config(**s.config)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        h1, st1 = e.simulate(gs)
        e.set_config(max_photon=1500)
        h2, st2 = e.simulate(gs)
    assert st1['num_photon'] == st2['num_photon']
    assert np.array_equal(bits(h1), bits(h2))


def test_records_match(engine_mod, oracle_lib):
    s = scene('nested_boxes')
    cfg = abi.default_config(**s.config, seq_enabled=1, max_record=10)
    gs = s.gensteps(scale_photons=0.01)
    g = run_engine(engine_mod, s, gs, cfg)
    r = run_oracle(oracle_lib, s, gs, cfg)
    assert np.array_equal(bits(g['record']), bits(r['record']))


def test_fp64_roots_mode(engine_mod, oracle_lib):
    s = scene('thin_shells')
    cfg = abi.default_config(**s.config, seq_enabled=1, fp64_roots=1)
    gs = s.gensteps(scale_photons=0.002)
    g = run_engine(engine_mod, s, gs, cfg)
    r = run_oracle(oracle_lib, s, gs, cfg, cull=True)
    assert np.array_equal(bits(g['photon']), bits(r['photon']))
    assert np.array_equal(g['seq'], r['seq'])


def test_empty_event_and_errors(engine_mod):
    s = scene('nested_boxes')
    cfg = abi.default_config(**s.config)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        assert e.set_gensteps(np.zeros((0, 6, 4), np.float32)) == 0
        st = e.launch()
        assert st['num_photon'] == 0 and len(e.hits()) == 0
        bad = G.cerenkov([10], [[0, 0, 0]], [[0, 0, 1.0]], 1.0, 0.9, 10_000, 200.0, 700.0, 1.5, 10.0)
        with pytest.raises(engine_mod.OxrError) as ex:
            e.set_gensteps(bad)
        assert ex.value.code == -1
        e.set_gensteps(G.input_photon(5))
        with pytest.raises(engine_mod.OxrError) as ex:
            e.launch()
        assert ex.value.code == -5
        # the engine keeps working after rejected calls
        e.set_gensteps(s.gensteps(scale_photons=0.001))
        assert e.launch()['num_photon'] == 100


# ---------------------------------------------------------------- geometry queries
def zoo():
    f = Foundry()
    f.add_material('M', 1.0)
    b = f.boundary('M', None, None, 'M')
    f.begin_solid('zoo')
    tet = [(1, 1, 1, 60), (-1, -1, 1, 60), (-1, 1, -1, 60), (1, -1, -1, 60)]
    tet = [tuple(list(np.array(p[:3]) / np.linalg.norm(p[:3])) + [p[3]]) for p in tet]
    prims = [
        sphere(100.0),
        sphere(50.0, 10.0, -20.0, 5.0),
        zsphere(100.0, -40.0, 70.0),
        cylinder(40.0, -80.0, 90.0),
        box3(100.0, 60.0, 30.0),
        box3(80.0).moved(rotate_z_to([1.0, 2.0, 3.0]) @ translate(5.0, -7.0, 9.0)),
        cone(80.0, -60.0, 20.0, 70.0),
        convexpolyhedron(tet, (-110, -110, -110, 110, 110, 110)),
        phicut(0.25, 1.75) & sphere(90.0),
        hyperboloid(50.0, 100.0, -100.0, 120.0),
        scenes.pmt_like_solid(),
        (sphere(80.0, 30.0, 0.0, 0.0) & box3(120.0)) - cylinder(20.0, -200.0, 200.0),
        (box3(100.0) | sphere(70.0, 0.0, 0.0, 60.0)) - (cylinder(30.0, -100.0, 100.0) | box3(20.0, 200.0, 20.0)),
        halfspace(1.0, 1.0, 0.0, 20.0) & box3(150.0),
        box3(150.0) - halfspace(0.3, -0.5, 0.8, 15.0),
        halfspace(0.0, 1.0, 0.0, 10.0) & cylinder(60.0, -80.0, 80.0),
    ]
    for t in prims:
        f.add_prim(t, b)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    return f, len(prims)


def random_rays(rng, n, ext):
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1)[:, None] * rng.uniform(0.0, 3.0 * ext, size=(n, 1))
    tgt = rng.uniform(-ext, ext, size=(n, 3))
    d = tgt - o
    k = n // 8
    d[:k] = np.eye(3)[rng.integers(0, 3, size=k)] * rng.choice([-1, 1], size=(k, 1))   # axis-aligned
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = np.zeros((n, 2, 4), dtype=np.float32)
    r[:, 0, :3] = o
    r[:, 0, 3] = rng.choice([0.0, 0.05, 1.0], size=n)
    r[:, 1, :3] = d
    r[:, 1, 3] = 1e6
    return r


@pytest.mark.parametrize('fp64', [0, 1])
def test_intersect_prim_bitexact(engine_mod, oracle_lib, fp64):
    f, nprim = zoo()
    cfg = abi.default_config(fp64_roots=fp64)
    rng = np.random.default_rng(9)
    o = oracle_lib.Oracle(f.arrays(), f.physics(), cfg)
    with engine_mod.Engine(f.arrays(), f.physics(), cfg) as e:
        for p in range(nprim):
            rays = random_rays(rng, 4000, 150.0)
            gi, gv = e.intersect_prim(p, rays)
            oi, ov = o.intersect_prim(p, rays, 0)
            assert np.array_equal(gv, ov), p
            assert np.array_equal(bits(gi), bits(oi)), p
            assert gv.mean() > 0.05, (p, gv.mean())


def list_zoo():
    f = Foundry()
    f.add_material('M', 1.0)
    b = f.boundary('M', None, None, 'M')
    f.begin_solid('lists')
    prims = [
        contiguous(box3(100.0), sphere(40.0, 60.0, 0.0, 0.0), cylinder(20.0, -90.0, 90.0)),
        discontiguous(sphere(30.0, -80.0, 0, 0), box3(40.0), sphere(25.0, 90.0, 10.0, 0)),
        overlap(box3(100.0), sphere(60.0), cylinder(35.0, -80.0, 80.0)),
        sphere(70.0) & contiguous(box3(100.0), sphere(40.0, 60.0, 0.0, 0.0)),
        (box3(120.0) - cylinder(15.0, -100.0, 100.0)) | overlap(sphere(50.0, 0, 0, 70.0), box3(60.0).moved(translate(0, 0, 60.0))),
    ]
    for t in prims:
        f.add_prim(t, b)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    return f, len(prims)


@pytest.mark.parametrize('fp64', [0, 1])
def test_list_node_prims_bitexact(engine_mod, oracle_lib, fp64):
    """List-node prims (CSG_CONTIGUOUS / DISCONTIGUOUS / OVERLAP, alone and as tree
    leaves) run the LIST kernel variants; intersects equal the oracle's."""
    f, nprim = list_zoo()
    cfg = abi.default_config(fp64_roots=fp64)
    rng = np.random.default_rng(19)
    o = oracle_lib.Oracle(f.arrays(), f.physics(), cfg)
    with engine_mod.Engine(f.arrays(), f.physics(), cfg) as e:
        for p in range(nprim):
            rays = random_rays(rng, 4000, 150.0)
            gi, gv = e.intersect_prim(p, rays)
            oi, ov = o.intersect_prim(p, rays, 0)
            assert np.array_equal(gv, ov), p
            assert np.array_equal(bits(gi), bits(oi)), p
            assert gv.mean() > 0.05, (p, gv.mean())


def list_scene():
    """Water shapes made of list nodes in air inside an absorbing box, one of them a
    sensor, lit by an isotropic torch."""
    f = Foundry()
    f.add_material('Air', 1.0003, absorption_length=1e7, scattering_length=1e7)
    f.add_material('Water', 1.333, absorption_length=2e4, scattering_length=5e3)
    f.add_implicit_absorber('Implicit_Air_Rock')
    f.add_sensor_surface('sensor', 0.5)
    b_world = f.boundary('Air', None, 'Implicit_Air_Rock', 'Air')
    b_w = f.boundary('Air', None, None, 'Water')
    b_s = f.boundary('Air', None, 'sensor', 'Water')
    f.begin_solid('r0')
    f.add_prim(box3(1000.0), b_world)
    f.add_prim(contiguous(box3(100.0).moved(translate(200, 0, 0)), sphere(60.0, 260.0, 0, 0)), b_w)
    f.add_prim(discontiguous(sphere(40.0, -200, 100, 0), sphere(40.0, -200, -100, 0)), b_s)
    f.add_prim(overlap(box3(150.0), sphere(90.0)).moved(translate(0, 250, 0)), b_w)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    gs = G.torch([20000], pos=[(0.0, 0.0, 0.0)], wavelength=420.0, type=G.T_POINT_ISOTROPIC)
    return f, gs


@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_list_scene_simulate_bitexact(engine_mod, oracle_lib, form):
    f, gs = list_scene()
    cfg = abi.default_config(max_bounce=31, seq_enabled=1)
    with engine_mod.Engine(f.arrays(), f.physics(), cfg) as e:
        e.set_gensteps(gs)
        st = e.launch(form)
        ph, seq = e.photons(), e.seq()
    r = oracle_lib.Oracle(f.arrays(), f.physics(), cfg).simulate(gs)
    print('list scene', st['num_photon'], 'photons', st['bounce_steps'], 'steps', st['num_hit'], 'hits', flush=True)
    assert np.array_equal(bits(ph), bits(r['photon']))
    assert np.array_equal(seq, r['seq'])
    assert st['num_hit'] > 0


@pytest.mark.parametrize('name', ['nested_boxes', 'csi_calorimeter', 'instanced_csg', 'thin_shells'])
def test_trace_bitexact(engine_mod, oracle_lib, name):
    s = scene(name)
    cfg = abi.default_config(**s.config)
    rng = np.random.default_rng(10)
    ext = {'nested_boxes': 120.0, 'csi_calorimeter': 200.0, 'instanced_csg': 19500.0, 'thin_shells': 20000.0}[name]
    rays = random_rays(rng, 20000, ext)
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    o.set_cull(True)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        g = e.trace(rays)
    r = o.trace(rays)
    assert np.array_equal(bits(g), bits(r))
    assert (bits(g)[:, 1, 3] & 0xffff != 0xffff).mean() > 0.5


def test_multibit_hitmask_requires_all_bits(engine_mod, oracle_lib):
    """Hit selection requires every bit of the hit mask (sphoton_selector,
    sysrap/sphoton.h:1219-1225, used by QEvt::gatherHit, qudarap/QEvt.cc:889): with
    SD|DR only detected photons that were also diffusely reflected are hits."""
    s = scene('csi_calorimeter')
    F = abi.FLAGS
    m = F['SURFACE_DETECT'] | F['SURFACE_DREFLECT']
    cfg = abi.default_config(**s.config, hitmask=m)
    gs = s.gensteps(scale_photons=SCENES['csi_calorimeter'][1])
    g = run_engine(engine_mod, s, gs, cfg)
    r = run_oracle(oracle_lib, s, gs, cfg, cull=True)
    u = bits(r['photon'])[:, 3, 3]
    sel = (u & m) == m
    assert 0 < sel.sum() < ((u & F['SURFACE_DETECT']) != 0).sum()
    assert np.array_equal(bits(g['hit']), bits(r['photon'][sel]))


def test_list_node_validation(engine_mod):
    """Engine creation rejects list nodes the kernels cannot evaluate: a contiguous list
    of more than 8 leaves (the reference's fixed enter arrays) and subs outside the prim."""
    def make(tree):
        f = Foundry()
        f.add_material('M', 1.0)
        b = f.boundary('M', None, None, 'M')
        f.begin_solid('s')
        f.add_prim(tree, b)
        f.end_solid()
        f.add_instance(np.eye(4), 0)
        return f
    nine = contiguous(*[sphere(10.0, 15.0 * k, 0, 0) for k in range(9)])
    with pytest.raises(engine_mod.OxrError, match='more than 8'):
        f = make(nine)
        engine_mod.Engine(f.arrays(), f.physics(), abi.default_config())
    f = make(overlap(box3(10.0), sphere(7.0)))
    a = f.arrays()
    a['node'].view(np.uint32).reshape(-1, 16)[0, 1] = 7        # subOffset past the prim's nodes
    with pytest.raises(engine_mod.OxrError, match='subs outside'):
        engine_mod.Engine(a, f.physics(), abi.default_config())


def test_debug_arrays_enabled_after_gensteps(engine_mod, oracle_lib):
    """seqhis and records switched on (and enlarged) by set_config after the gensteps of a
    larger event were uploaded: the launch sizes the event buffers for the config in force,
    and the arrays equal the oracle's."""
    s = scene('nested_boxes')
    small = s.gensteps(scale_photons=0.001)
    big = s.gensteps(scale_photons=0.3)
    with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config, seq_enabled=1, max_record=2)) as e:
        e.set_gensteps(small)
        e.launch()
        e.set_config(seq_enabled=0, max_record=0)
        n = e.set_gensteps(big)
        e.set_config(seq_enabled=1, max_record=6)
        print('debug arrays after gensteps:', n, 'photons', flush=True)
        e.launch()
        ph, seq, rec = e.photons(), e.seq(), e.get(abi.OXR_GET_RECORD)
    cfg = abi.default_config(**s.config, seq_enabled=1, max_record=6)
    r = run_oracle(oracle_lib, s, big, cfg)
    assert np.array_equal(bits(ph), bits(r['photon']))
    assert np.array_equal(seq, r['seq'])
    assert np.array_equal(bits(rec), bits(r['record']))


def test_csg_stacks_in_registers_for_shallow_trees(engine_mod):
    """Tuning csg_stack 1 keeps the CSG evaluation stacks in registers when every tree is
    one or two levels deep (the instanced-CSG solid, height 2; the thin-shell solid, height
    1); leaf-only scenes (height 0) and a tree of height 3 keep the private-memory stacks,
    and so does csg_stack 0. Photons of both settings equal the oracle's in
    test_gpu_tuning_forms."""
    want = {'instanced_csg': (2, 1), 'thin_shells': (1, 1), 'csi_calorimeter': (0, 0), 'nested_boxes': (0, 0)}
    for name, (h, reg) in want.items():
        s = scene(name)
        with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config)) as e:
            assert e.info()['csg_in_registers'] == 0 and e.tuning()['csg_stack'] == 0
            e.set_tuning(csg_stack=1)
            i = e.info()
            assert e.tuning()['csg_stack'] == reg, name
        assert (i['csg_height'], i['csg_in_registers']) == (h, reg), name
    f = Foundry()
    f.add_material('Air', 1.0)
    f.add_material('Glass', 1.5)
    b0 = f.boundary('Air', None, None, 'Air')
    b1 = f.boundary('Air', None, None, 'Glass')
    f.begin_solid('r0')
    f.add_prim(box3(1000.0), b0)
    deep = ((sphere(50.0) | sphere(40.0, 30.0)) | (sphere(30.0, -30.0) | sphere(20.0, 0.0, 40.0))) | box3(20.0)
    f.add_prim(deep, b1)
    f.end_solid()
    f.add_instance(np.eye(4), 0)
    with engine_mod.Engine(f.arrays(), f.physics(), abi.default_config()) as e:
        e.set_tuning(csg_stack=1)
        i = e.info()
    assert (i['csg_height'], i['csg_in_registers']) == (3, 0)


def test_simple_leaf_kernels_chosen_by_the_leaves(engine_mod, oracle_lib):
    """A geometry whose leaves are all spheres, zspheres, cylinders or boxes runs the
    kernels compiled with those leaves only (every measured event does); one cone leaf
    anywhere selects the general kernels. Photons of both equal the oracle's."""
    for name in ('instanced_csg', 'thin_shells', 'csi_calorimeter', 'nested_boxes'):
        s = scene(name)
        with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config)) as e:
            assert e.info()['simple_leaves'] == 1, name
    for with_cone in (False, True):
        f = Foundry()
        f.add_material('Air', 1.0)
        f.add_material('Glass', 1.5)
        b0 = f.boundary('Air', None, None, 'Air')
        b1 = f.boundary('Air', None, None, 'Glass')
        f.begin_solid('r0')
        f.add_prim(box3(1000.0), b0)
        f.add_prim(sphere(60.0) | cylinder(30.0, -90.0, 90.0), b1)
        if with_cone:
            f.add_prim(cone(40.0, -20.0, 10.0, 20.0).moved(translate(0.0, 0.0, 200.0)), b1)
        f.end_solid()
        f.add_instance(np.eye(4), 0)
        g, p = f.arrays(), f.physics()
        gs = G.torch([4000], pos=(0.0, 0.0, -300.0), wavelength=440.0, type=G.T_POINT_ISOTROPIC)
        cfg = abi.default_config()
        with engine_mod.Engine(g, p, cfg) as e:
            assert e.info()['simple_leaves'] == (0 if with_cone else 1)
            e.set_gensteps(gs)
            e.launch(abi.OXR_FORM_PERSISTENT)
            got = e.photons()
        o = oracle_lib.Oracle(g, p, cfg)
        o.set_cull(True)
        want = o.simulate(gs, seq=False)['photon']
        assert np.array_equal(bits(got), bits(want)), with_cone



def test_csg_arrays_read_from_lds_when_they_fit(engine_mod, oracle_lib):
    """The simple-leaf kernels copy the CSG node and transform arrays and the prim table
    into LDS when they take at most 16 KB and the copy costs no resident workgroup: 60
    placed boxes (8.2 KB with their transforms and prim entries) are copied, 300 (41 KB) are
    not, and the instanced-CSG event at its tuned launch is. Photons equal the oracle's
    with and without the copy."""
    import json
    from oxr import autotune
    rng = np.random.default_rng(5)
    for nbox, want in ((60, 1), (300, 0)):
        f = Foundry()
        f.add_material('Air', 1.0)
        f.add_material('Glass', 1.5)
        b0 = f.boundary('Air', None, None, 'Air')
        b1 = f.boundary('Air', None, None, 'Glass')
        f.begin_solid('r0')
        f.add_prim(box3(3000.0), b0)
        for x, y, z in rng.uniform(-1200.0, 1200.0, size=(nbox, 3)):
            f.add_prim(box3(40.0).moved(translate(float(x), float(y), float(z))), b1)
        f.end_solid()
        f.add_instance(np.eye(4), 0)
        g, p = f.arrays(), f.physics()
        gs = G.torch([20000], pos=(0.0, 0.0, 0.0), wavelength=440.0, type=G.T_POINT_ISOTROPIC)
        cfg = abi.default_config()
        print('boxes', nbox, flush=True)
        with engine_mod.Engine(g, p, cfg) as e:
            i = e.info()
            assert (i['simple_leaves'], i['csg_in_lds']) == (1, want), nbox
            e.set_gensteps(gs)
            e.launch(abi.OXR_FORM_PERSISTENT)
            got = e.photons()
        o = oracle_lib.Oracle(g, p, cfg)
        o.set_cull(True)
        assert np.array_equal(bits(got), bits(o.simulate(gs, seq=False)['photon'])), nbox
    s = scene('instanced_csg')
    cfg = abi.default_config(**s.config)
    gs = s.gensteps(scale_photons=SCENES['instanced_csg'][1])
    with open(autotune.TABLE) as fh:
        tune = json.load(fh)['instanced_csg']
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_tuning(**tune)
        assert e.info()['csg_in_lds'] == 1
        e.set_gensteps(gs)
        e.launch(abi.OXR_FORM_PERSISTENT)
        got = e.photons()
    want = run_oracle(oracle_lib, s, gs, cfg, cull=True)['photon']
    assert np.array_equal(bits(got), bits(want))

def test_max_photon_from_device_memory(engine_mod):
    """max_photon 0 (the default, as the reference's MaxSlot 0) sizes a launch slice from
    the device memory: 87% of it over the bytes per photon, rounded down to a million
    (SEventConfig::HeuristicMaxSlot); debug arrays shrink it; an explicit value is kept."""
    s = scene('nested_boxes')
    with engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config)) as e:
        total = e.info()['total_bytes']
        got = e.config()['max_photon']
        e.set_config(seq_enabled=1, max_record=10, max_photon=0)
        dbg = e.config()['max_photon']
        e.set_config(max_photon=1500)
        fixed = e.config()['max_photon']
    # the engine sizes from the device properties' memory; the query reports hipMemGetInfo's
    # total, which may differ by a little
    assert got % 1000000 == 0 and abs(got - 0.87 * total / 149) < 2e6 and got > 1_000_000_000
    assert dbg % 1000000 == 0 and abs(dbg - 0.87 * total / (149 + 32 + 640)) < 2e6 and dbg < got
    assert fixed == 1500
