"""Benchmark: photons/s and bounce-steps/s on one MI355X.

Headline step: one simulate launch (generate + propagate + hit compaction) of the
instanced-CSG workload (water sphere r=19.5 m in an acrylic shell, 20,000 instances of a
7-node CSG solid, 10M torch photons from 1,000 interior points, max_bounce 64) with
gensteps resident in HBM, persistent lane-refill form with the committed autotuned
launch settings, timed over --steps launches. Also reported: the plain form, the
roofline of that launch, a CPU baseline (the CPU restatement on every host core), the
calorimeter, nested-box and thin-shell workloads, the serving form on its 256-event
stream (p50/p99 event latency against sequential per-event launches) and served vs
per-event photons/s on the instanced-CSG geometry.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
python bench.py --workload {instanced_csg,csi_calorimeter,thin_shells,nested_boxes} --only
    (the timed launches of one workload alone, for kernel traces and counters)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, 'eic-opticks_amd'), os.path.join(ROOT, 'oracle')]

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ensure_built():
    from oxr import engine
    if not os.path.exists(engine.LIB_PATH):
        from oxr import build
        build.build()


def load_tuning(name):
    path = os.path.join(ROOT, 'eic-opticks_amd', 'oxr', 'tuning_gfx950.json')
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get(name)
    return None


def read_profile(workload):
    """Figures measured by rocprofv3 for the simulate launch of this workload, from the
    committed summary profiles/pmc_traffic.txt: 'traffic' = HBM bytes per launch
    (2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE, copy-kernel calibration), 'valu_busy' and
    'valu_util' (percent) of the propagate kernel."""
    path = os.path.join(ROOT, 'profiles', 'pmc_traffic.txt')
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        parts = line.split()
        if not parts or parts[0] != workload:
            continue
        for kv in parts[1:]:
            k, _, v = kv.partition('=')
            try:
                out[k] = float(v)
            except ValueError:
                pass
    return out


def algorithmic_bytes(stats, n_photon, n_gs, n_hit):
    """Bytes the hot path must move per launch, counted at visit granularity by the
    counting launch: per photon 4 (seed) + 96/ppg (genstep) + 64 (photon) + 1 (hit flag);
    per BVH node visit 64; per instance entered 64; per prim test 64 per CSG node plus 64
    per transformed leaf; per bounce step 4 table lookups x 2 texels x 16 B + 16 B optical;
    per hit 64 read + 64 write (compaction)."""
    ppg = n_photon / max(n_gs, 1)
    per_photon = 4 + 96.0 / ppg + 64 + 1
    per_step = 4 * 2 * 16 + 16
    return (n_photon * per_photon + stats['node_visits'] * 64 + stats['instance_visits'] * 64
            + stats['csg_node_bytes'] + stats['bounce_steps'] * per_step + n_hit * 128)


def timed_launches(e, form, steps, warmup):
    """Wall time of `steps` synchronous launches; the returned stats carry the mean
    kernel_ms and generate_ms (HIP events on the engine stream) over those launches."""
    for _ in range(warmup):
        e.launch(form)
    ks, gs = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = e.launch(form)
        ks.append(st['kernel_ms'])
        gs.append(st['generate_ms'])
    dt = time.perf_counter() - t0
    st = dict(st, generate_ms=float(np.mean(gs)))
    return dt, st, float(np.mean(ks))


# The GPU host allots 16 cores to a one-GPU job, whatever os.cpu_count() shows there.
HOST_CORE_SHARE = 16


def host_threads():
    """The host cores this process may use: its affinity mask, within the host's share."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, HOST_CORE_SHARE))


def cpu_baseline(scene, gs, cfg, threads, budget_s=12.0, inp=None):
    """The CPU restatement (oracle) of the same workload on `threads` host threads: a
    bounded sample of the event's first photons, photons/s over the sample."""
    import oracle as O
    o = O.Oracle(scene.geometry, scene.physics, cfg)
    o.set_cull(True)
    n_total = int(np.asarray(gs).view(np.uint32)[:, 0, 3].astype(np.int64).sum())
    probe = min(max(20000, 1000 * threads), n_total)
    t0 = time.perf_counter()
    o.simulate(gs, nthreads=threads, begin=0, end=probe, seq=False, input_photons=inp)
    rate = probe / max(time.perf_counter() - t0, 1e-6)
    sample = int(min(n_total, max(probe, rate * budget_s)))
    t0 = time.perf_counter()
    r = o.simulate(gs, nthreads=threads, begin=0, end=sample, seq=False, input_photons=inp)
    dt = time.perf_counter() - t0
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return dict(value=round(sample / dt, 1), unit='photons/s', cores=threads, nproc=os.cpu_count(),
                affinity_cores=affinity, core_share=HOST_CORE_SHARE, kind='port',
                sample=f'first {sample} of {n_total} photons of the same event, {threads} std::threads, '
                       f'{dt:.1f} s', bounce_steps_per_s=round(float(r['bounce'].sum()) / dt, 1),
                photons_per_s_per_core=round(sample / dt / threads, 1))


def history_chi2_leg(engine_mod, scene, gs, cfg, sample, threads, tune=None, inp=None, begin=0):
    """seqhis chi2/ndf of the GPU FP32 run (event 1) against the FP64 oracle (event 2) over
    the same photon index range [begin, begin+sample): independent random streams, so
    chi2/ndf near 1 means the two histories are samples of one distribution."""
    import oracle as O
    from oxr import abi, qcf
    c = abi.default_config(**{k: getattr(cfg, k) for k, _ in abi.Config._fields_})
    c.seq_enabled = 1
    e = engine_mod.Engine(scene.geometry, scene.physics, c)
    if tune:
        e.set_tuning(**tune)
    if inp is not None:
        e.set_input_photons(inp)
    e.set_gensteps(gs, event_index=1)
    e.launch(abi.OXR_FORM_PERSISTENT)
    ga = e.seq()[begin:begin + sample]
    e.close()
    o = O.Oracle(scene.geometry, scene.physics, c)
    o.set_cull(True)
    t0 = time.perf_counter()
    r = o.simulate(gs, event_index=2, precision=1, nthreads=threads, begin=begin, end=begin + sample,
                   input_photons=inp)
    dt = time.perf_counter() - t0
    res = qcf.history_chi2(qcf.seqhis(ga), qcf.seqhis(r['seq']))
    res.update(sample=int(len(ga)), oracle_s=round(dt, 2))
    return res


def precision_leg(engine_mod, threads, sample=5000):
    """Thin shells, 5M photons (10% grazing input photons): FP32, FP32 with FP64 roots and
    FP32 with the refine trace, throughput on the GPU, and each against the FP64 oracle on
    the same streams over a sample of grazing and isotropic photons."""
    import oracle as O
    from oxr import abi, qcf, scenes, gensteps as G
    s = scenes.thin_shells()
    n_total = 5_000_000
    ng = n_total // 10
    inp = scenes.thin_shell_grazing_photons(s, ng)
    gs = np.concatenate([G.input_photon(ng), s.gensteps(scale_photons=(n_total - ng) / 5e6)])
    tune = load_tuning('thin_shells')
    rows = {}
    o = O.Oracle(s.geometry, s.physics, abi.default_config(**s.config, seq_enabled=1))
    o.set_cull(True)
    t0 = time.perf_counter()
    rg = o.simulate(gs, precision=1, nthreads=threads, begin=0, end=sample, input_photons=inp)
    ri = o.simulate(gs, precision=1, nthreads=threads, begin=ng, end=ng + sample, input_photons=inp)
    oracle_s = time.perf_counter() - t0
    for f64, refine in ((0, 0), (1, 0), (0, 1)):
        cfg = abi.default_config(**s.config, fp64_roots=f64, propagate_refine=refine)
        e = engine_mod.Engine(s.geometry, s.physics, cfg)
        if tune:
            e.set_tuning(**tune)
        e.set_input_photons(inp)
        n = e.set_gensteps(gs)
        dt, st, kms = timed_launches(e, abi.OXR_FORM_PERSISTENT, 3, 1)
        e.set_config(seq_enabled=1)
        e.set_gensteps(gs)
        e.launch(abi.OXR_FORM_PERSISTENT)
        ph, sq = e.photons(), e.seq()
        e.close()
        gra = qcf.endpoint_errors(ph[:sample], sq[:sample], rg['photon'], rg['seq'])
        iso = qcf.endpoint_errors(ph[ng:ng + sample], sq[ng:ng + sample], ri['photon'], ri['seq'])
        rows['fp64_roots' if f64 else ('fp32_refine' if refine else 'fp32')] = {
            'photons_per_s': round(3 * n / dt, 1), 'kernel_ms': round(kms, 3),
            'bounce_steps': int(st['bounce_steps']),
            'grazing_vs_fp64': {k: (round(v, 6) if isinstance(v, float) else v) for k, v in gra.items()},
            'isotropic_vs_fp64': {k: (round(v, 6) if isinstance(v, float) else v) for k, v in iso.items()}}
    rows['fp64_roots_cost'] = round(rows['fp32']['photons_per_s'] / rows['fp64_roots']['photons_per_s'], 4)
    rows['refine_cost'] = round(rows['fp32']['photons_per_s'] / rows['fp32_refine']['photons_per_s'], 4)
    rows['refine'] = 'fp32_refine: PropagateRefine on, refine distance 5000 mm (traces longer than 5.05 m re-traced from 0.99 t)' 
    rows['workload'] = ('thin_shells: 50 x 1 mm shells at 1-20 m (n 1.50/1.33) + 5000 instanced zsphere|cylinder; '
                        f'{n_total} photons at 400 nm, {ng} grazing (within 1e-3 rad of tangent) input photons '
                        'then isotropic torch from the origin; max_bounce 64')
    rows['oracle_sample'] = f'{sample} grazing + {sample} isotropic photons, FP64 oracle, {threads} threads, {oracle_s:.1f} s'
    return rows, (s, gs, inp, ng)


def run_server(engine_mod, e, evs, arrivals, step_photons, out):
    """Serve the events; the consumer takes each request's hits into the reused host
    array out, as run_sequential does."""
    srv = engine_mod.Server(e, max_step_photons=step_photons)
    res = {}

    def feeder():
        t0 = time.perf_counter()
        for k, (gs, t) in enumerate(zip(evs, arrivals)):
            d = t - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            srv.submit(k, gs, 1000 + k)

    th = threading.Thread(target=feeder)
    t_start = time.perf_counter()
    th.start()
    photons = 0
    for _ in evs:
        r = srv.wait(600000)
        srv.take_hits(r['request_id'], r['num_hit'], out=out)
        res[r['request_id']] = r['t_done'] - r['t_submit']
        photons += r['num_photon']
    wall = time.perf_counter() - t_start
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array([res[k] for k in range(len(evs))])
    return lat, wall, photons, st


def run_sequential(e, evs, arrivals, out):
    """Per-event plain launches, one event at a time in arrival order (the reference's
    one-request-at-a-time service, CSGOptiXService.h:76-84); hits into the reused host
    array out."""
    from oxr import abi
    lat = np.zeros(len(evs))
    t0 = time.perf_counter()
    busy = 0.0
    photons = 0
    for k, (gs, t) in enumerate(zip(evs, arrivals)):
        now = time.perf_counter() - t0
        if now < t:
            time.sleep(t - now)
        a = time.perf_counter()
        h, st = e.simulate(gs, event_index=1000 + k, form=abi.OXR_FORM_PLAIN, out=out)
        b = time.perf_counter()
        busy += b - a
        photons += st['num_photon']
        lat[k] = (b - t0) - t
    return lat, time.perf_counter() - t0, photons, busy


def served_instanced_csg(engine_mod, abi, scenes, s3, cfg3, t3, per_event=True):
    """128 events submitted at once on the instanced-CSG geometry: served (8M-photon steps)
    against one persistent launch of all their gensteps (the device rate of this photon
    mix) and against per-event plain launches one at a time."""
    e3 = engine_mod.Engine(s3.geometry, s3.physics, cfg3)
    if t3:
        e3.set_tuning(**t3)
    evs, counts, _ = scenes.serving_stream(np.random.default_rng(43), n_events=128, ext=16000.0,
                                           wavelength=430.0)
    zero = np.zeros(len(evs))
    e3.set_gensteps(np.concatenate(evs), event_index=7)
    e3.launch(abi.OXR_FORM_PERSISTENT)
    t_one = time.perf_counter()
    e3.launch(abi.OXR_FORM_PERSISTENT)
    one_s = time.perf_counter() - t_one
    hitbuf = np.empty((int(counts.max()), 4, 4), dtype=np.float32)
    hitbuf.fill(0.0)
    lat_s, wall_s, ph_s, sst = run_server(engine_mod, e3, evs, zero, 8 << 20, hitbuf)
    r = {'workload': '128 events submitted at once, log-normal photons (median 200k, sigma 1, clipped 1k-20M), '
                     'torch gensteps at random interior origins, instanced_csg geometry resident',
         'events': len(evs), 'photons': int(counts.sum()),
         'served_photons_per_s': round(ph_s / wall_s, 1),
         'single_launch_photons_per_s': round(float(counts.sum()) / one_s, 1),
         'served_vs_single_launch': round((ph_s / wall_s) / (float(counts.sum()) / one_s), 3),
         'served_p50_ms': round(1e3 * float(np.percentile(lat_s, 50)), 1),
         'served_p99_ms': round(1e3 * float(np.percentile(lat_s, 99)), 1),
         'served_steps': int(sst['steps']),
         'served_kernel_ms': round(float(sst['kernel_ms']), 2), 'served_wall_ms': round(1e3 * wall_s, 2),
         'served_device_busy_frac': round(float(sst['busy_ms']) / (1e3 * wall_s), 3)}
    if per_event:
        lat_b, wall_b, ph_b, busy_b = run_sequential(e3, evs, zero, hitbuf)
        r.update({'per_event_plain_photons_per_s': round(ph_b / wall_b, 1),
                  'ratio': round((ph_s / wall_s) / (ph_b / wall_b), 3),
                  'per_event_p50_ms': round(1e3 * float(np.percentile(lat_b, 50)), 1),
                  'per_event_p99_ms': round(1e3 * float(np.percentile(lat_b, 99)), 1)})
    e3.close()
    return r


WORKLOADS = {
    'instanced_csg': 'instanced_csg: water sphere r=19.5 m in a 120 mm acrylic shell, 20000 instanced 7-node CSG '
                     'solids (zsphere|cylinder minus inner zsphere) at r=19.4 m, 10M torch photons at 430 nm from '
                     '1000 interior points, max_bounce 64',
    'csi_calorimeter': 'csi_calorimeter: 8x8 CsI 20x20x200 mm crystals + 6x6 mm SiPMs, 2000 scintillation + 500 '
                       'Cerenkov gensteps, Poisson 1000/150 photons, max_bounce 32',
    'nested_boxes': 'nested_boxes: 240/220/200/100 mm vacuum/lead/air/water cubes, 100k torch photons at 420 nm '
                    'from the water centre, max_bounce 32',
    'thin_shells': 'thin_shells: 50 x 1 mm shells at 1-20 m (n 1.50/1.33) + 5000 instanced zsphere|cylinder, '
                   '5M torch photons at 400 nm from the origin, max_bounce 64',
}


def roofline_block(e, name, n, n_gs, n_hit, kms, generate_ms):
    """Algorithmic bytes of one launch (counting launch over the same photons) over the
    measured mean launch time, against the 8 TB/s HBM peak; beside it the HBM traffic and
    the vector-issue counters rocprofv3 measured for this workload (committed summary)."""
    cnt = e.launch(counting=True)
    ab = algorithmic_bytes(cnt, n, n_gs, n_hit)
    achieved = ab / (kms / 1e3) / 1e9
    prof = read_profile(name)
    traffic = prof.get('traffic')
    return dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                kernel=('k_simulate_persist generating in place (one simulate launch)' if e.tuning().get('generate_inplace')
                        else 'k_generate + k_simulate_persist (one simulate launch)'),
                generate_ms=round(generate_ms, 4), propagate_ms=round(kms - generate_ms, 4),
                algorithmic_bytes_per_launch=int(ab),
                node_visits=int(cnt['node_visits']), prim_tests=int(cnt['prim_tests']),
                csg_node_bytes=int(cnt['csg_node_bytes']), instance_visits=int(cnt['instance_visits']),
                measured_limiter='vector ALU issue under lane divergence' if prof else None,
                hbm_measured_frac=(round(traffic / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None),
                valu_busy_pct=prof.get('valu_busy'), valu_lane_util_pct=prof.get('valu_util'),
                counters_source=(f'profiles/pmc_traffic.txt ({name}): rocprofv3 --pmc on bench.py --workload {name} '
                                 '--only, read from the committed summary, not measured in this run') if prof else None,
                note=('the algorithmic bytes price every BVH node, CSG node, transform and table read as if it came '
                      'from HBM; they exceed what HBM can deliver in the measured time, so most are served from LDS, '
                      'L1, L2 and the Infinity Cache: hbm_measured_frac is the HBM share, and the kernel is bound by '
                      'vector issue under lane divergence (valu_lane_util_pct)') if achieved > HBM_PEAK_GBS else None)


def setup(engine_mod, abi, scenes, name, overrides=None):
    from oxr import gensteps as G
    s = scenes.SCENES[name]()
    cfg = abi.default_config(**s.config)
    e = engine_mod.Engine(s.geometry, s.physics, cfg)
    tune = dict(load_tuning(name) or {}, **(overrides or {}))
    if tune:
        e.set_tuning(**tune)
    inp = None
    if name == 'thin_shells':
        n_total = 5_000_000
        ng = n_total // 10
        inp = scenes.thin_shell_grazing_photons(s, ng)
        gs = np.concatenate([G.input_photon(ng), s.gensteps(scale_photons=(n_total - ng) / 5e6)])
        e.set_input_photons(inp)
    else:
        gs = s.gensteps()
    n = e.set_gensteps(gs)
    return s, cfg, e, gs, n, tune, inp


def measure(engine_mod, abi, scenes, name, steps, warmup, roofline=True, overrides=None):
    """Persistent and plain forms of one workload, timed over `steps` launches each."""
    s, cfg, e, gs, n, tune, inp = setup(engine_mod, abi, scenes, name, overrides)
    dt, st, kms = timed_launches(e, abi.OXR_FORM_PERSISTENT, steps, warmup)
    dtp, stp, kmsp = timed_launches(e, abi.OXR_FORM_PLAIN, steps, warmup)
    out = dict(workload=WORKLOADS[name], photons=n, gensteps=len(gs),
               photons_per_s=round(steps * n / dt, 1),
               bounce_steps_per_s=round(steps * st['bounce_steps'] / dt, 1),
               ms_per_step=round(1e3 * dt / steps, 3), kernel_ms=round(kms, 3),
               hits=int(st['num_hit']), bounce_steps=int(st['bounce_steps']),
               plain_form=dict(photons_per_s=round(steps * n / dtp, 1), kernel_ms=round(kmsp, 3),
                               bounce_steps_per_s=round(steps * stp['bounce_steps'] / dtp, 1)),
               persistent_vs_plain=round(dtp / dt, 3), tuning=e.tuning(), max_bounce=cfg.max_bounce)
    log(f"[bench] {name}: persistent {out['photons_per_s'] / 1e6:.1f} M photons/s "
        f"({kms:.2f} ms), plain {out['plain_form']['photons_per_s'] / 1e6:.1f} M photons/s ({kmsp:.2f} ms)")
    if roofline:
        e.set_gensteps(gs)
        out['roofline'] = roofline_block(e, name, n, len(gs), st['num_hit'], kms, st['generate_ms'])
    e.close()
    return out, (s, cfg, gs, tune, inp)


def only(engine_mod, abi, scenes, args):
    """The timed launches of one workload alone (persistent then plain), so that a kernel
    trace or counter run of this command holds only that workload's launches."""
    overrides = {k: int(v) for k, v in (kv.split('=', 1) for kv in args.tune)}
    out, _ = measure(engine_mod, abi, scenes, args.workload, args.steps, args.warmup, roofline=False,
                     overrides=overrides)
    print(json.dumps(dict(metric=f'photons/s on 1 MI355X ({args.workload})', value=out['photons_per_s'],
                          unit='photons/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
                          ms_per_step=out['ms_per_step'], higher_is_better=True, scaling='weak',
                          vs_baseline=None, dtype='f32', data='synthetic',
                          config={'workload': out['workload'], 'photons': out['photons'], 'form': 'persistent',
                                  'tuning': out['tuning']}, detail=out)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--quick', action='store_true', help='headline only: no CPU, serving, secondary or precision legs')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-serving', action='store_true')
    ap.add_argument('--no-precision', action='store_true', help='skip the thin-shell precision and history-chi2 legs')
    ap.add_argument('--workload', choices=sorted(WORKLOADS), default='instanced_csg')
    ap.add_argument('--only', action='store_true', help='time only the --workload launches (for profiling)')
    ap.add_argument('--served-only', action='store_true',
                    help='run only the served instanced-CSG stream and its single launch (for profiling)')
    ap.add_argument('--tune', action='append', default=[], metavar='KEY=VALUE',
                    help='with --only: override a field of the committed tuning (e.g. csg_stack=1)')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one MI355X: this engine runs one process on one GPU and the '
              'multi-GPU form is a design note only; run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    ensure_built()
    from oxr import abi, scenes
    from oxr import engine as engine_mod
    if args.only:
        only(engine_mod, abi, scenes, args)
        return
    if args.served_only:
        s3 = scenes.instanced_csg()
        r = served_instanced_csg(engine_mod, abi, scenes, s3, abi.default_config(**s3.config),
                                 load_tuning('instanced_csg'), per_event=False)
        print(json.dumps(r), flush=True)
        return
    threads = host_threads()

    # ---------------------------------------------------------------- headline: instanced CSG
    name = args.workload
    head, (s3, cfg3, gs3, t3, inp3) = measure(engine_mod, abi, scenes, name, args.steps, args.warmup)
    out = {
        'metric': 'photons/s & bounce-steps/s on 1 MI355X; p99 event latency vs per-event launch',
        'value': head['photons_per_s'], 'unit': 'photons/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': head['ms_per_step'], 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': head['workload'], 'photons': head['photons'], 'gensteps': head['gensteps'],
                   'form': 'persistent', 'tuning': head['tuning'], 'max_bounce': head['max_bounce']},
        'bounce_steps_per_s': head['bounce_steps_per_s'], 'kernel_ms': head['kernel_ms'],
        'plain_form': head['plain_form'], 'persistent_vs_plain': head['persistent_vs_plain'],
        'roofline': head['roofline'],
    }
    out['cpu_baseline'] = None if (args.quick or args.no_cpu) else cpu_baseline(s3, gs3, cfg3, threads, inp=inp3)
    if out['cpu_baseline']:
        log(f"[bench] cpu baseline: {out['cpu_baseline']['value']:.0f} photons/s on {threads} threads")
    if args.quick:
        print(json.dumps(out), flush=True)
        return

    # ---------------------------------------------------------------- other workloads
    work = {}
    for w in ('csi_calorimeter', 'nested_boxes'):
        if w == name:
            continue
        r, (s, cfg, gs, tune, inp) = measure(engine_mod, abi, scenes, w, args.steps, args.warmup)
        if not args.no_cpu:
            r['cpu_baseline'] = cpu_baseline(s, gs, cfg, threads, budget_s=5.0)
        work[w] = r
    out['workloads'] = work

    if not args.no_serving:
        # ------------------------------------------------------------ serving stream (C2 geometry)
        s = scenes.csi_calorimeter()
        e = engine_mod.Engine(s.geometry, s.physics, abi.default_config(**s.config))
        tune = load_tuning('csi_calorimeter')
        if tune:
            e.set_tuning(**tune)
        evs, counts, arrivals = scenes.serving_stream(np.random.default_rng(42))
        hitbuf = np.empty((int(counts.max()), 4, 4), dtype=np.float32)   # one host hit array, reused
        hitbuf.fill(0.0)                                                   # pages touched before timing
        lat_s, wall_s, ph_s, sst = run_server(engine_mod, e, evs, arrivals, 4 << 20, hitbuf)
        lat_b, wall_b, ph_b, busy_b = run_sequential(e, evs, arrivals, hitbuf)
        e.close()
        out['serving'] = {
            'workload': '256 events, Poisson arrivals 40/s, log-normal photons (median 200k, sigma 1, '
                        'clipped 1k-20M), torch gensteps, csi_calorimeter geometry resident',
            'photons': int(counts.sum()),
            'served': {'p50_ms': round(1e3 * float(np.percentile(lat_s, 50)), 2),
                       'p99_ms': round(1e3 * float(np.percentile(lat_s, 99)), 2),
                       'photons_per_s_wall': round(ph_s / wall_s, 1),
                       'photons_per_s_busy': round(float(sst['photons']) / (float(sst['busy_ms']) / 1e3), 1),
                       'steps': int(sst['steps'])},
            'per_event_plain': {'p50_ms': round(1e3 * float(np.percentile(lat_b, 50)), 2),
                                'p99_ms': round(1e3 * float(np.percentile(lat_b, 99)), 2),
                                'photons_per_s_wall': round(ph_b / wall_b, 1),
                                'photons_per_s_busy': round(ph_b / busy_b, 1)},
        }
        log(f"[bench] serving: {out['serving']}")
        # ------------------------------------------------------------ served vs per-event, instanced CSG
        out['served_instanced_csg'] = served_instanced_csg(engine_mod, abi, scenes, s3, cfg3, t3)
        log(f"[bench] served instanced_csg: {out['served_instanced_csg']}")

    if not args.no_precision:
        chi2 = {}
        prec, (s5, gs5, inp5, ng5) = precision_leg(engine_mod, threads)
        out['thin_shells_precision'] = prec
        log(f"[bench] thin shells precision: {prec}")
        chi2['instanced_csg'] = history_chi2_leg(engine_mod, s3, gs3, cfg3, 20_000, threads, tune=t3)
        s2 = scenes.csi_calorimeter()
        chi2['csi_calorimeter'] = history_chi2_leg(engine_mod, s2, s2.gensteps(), abi.default_config(**s2.config),
                                                   50_000, threads, tune=load_tuning('csi_calorimeter'))
        s1 = scenes.nested_boxes()
        chi2['nested_boxes'] = history_chi2_leg(engine_mod, s1, s1.gensteps(), abi.default_config(**s1.config),
                                                100_000, threads)
        chi2['thin_shells_isotropic'] = history_chi2_leg(engine_mod, s5, gs5, abi.default_config(**s5.config),
                                                         20_000, threads, tune=load_tuning('thin_shells'),
                                                         inp=inp5, begin=ng5)
        out['history_chi2'] = {'definition': 'seqhis chi2/ndf, GPU FP32 (event 1) vs FP64 oracle (event 2), '
                                              'histories with a+b > 30 (c2per = c2sum/c2n)', **chi2}
        log(f"[bench] history chi2: {out['history_chi2']}")
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
