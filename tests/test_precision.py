"""CPU checks of the history-comparison helpers (oxr.qcf) and of the FP32-vs-FP64
precision bounds on the thin-shell workload.

The GPU kernels equal the oracle's FP32 arithmetic bit for bit (tests/test_gpu_parity.py),
so the FP32 (and FP32 + FP64-roots) versus FP64 comparison here bounds the GPU's error
as well; tests/test_gpu_precision.py repeats it through the C ABI."""
import numpy as np
import pytest

from oxr import abi, qcf, scenes, gensteps as G


def pack(slots):
    """seqhis (2,) u64 from a list of nibbles, slot 0 first."""
    lo = hi = 0
    for k, v in enumerate(slots):
        if k < 16:
            lo |= v << (4 * k)
        else:
            hi |= v << (4 * (k - 16))
    return [lo, hi]


def test_history_chi2_closed_form():
    h1, h2, h3 = pack([1, 3, 4]), pack([1, 3, 12, 4]), pack([1, 5])
    a = np.array([h1] * 50 + [h2] * 50 + [h3] * 5, dtype=np.uint64)
    b = np.array([h1] * 60 + [h2] * 40 + [h3] * 5, dtype=np.uint64)
    r = qcf.history_chi2(a, b, cut=30)
    # h3 has a+b = 10 <= 30 and is excluded (ana/nbase.py chi2 cut)
    want = 10.0 ** 2 / 110 + 10.0 ** 2 / 90
    assert r['c2n'] == 2 and r['c2c'] == 1 and r['histories'] == 3
    assert abs(r['c2sum'] - want) < 1e-12 and abs(r['c2per'] - want / 2) < 1e-12
    assert qcf.history_chi2(a, a)['c2sum'] == 0.0


def test_history_chi2_distinguishes_high_word():
    # histories equal in the first 16 slots and different beyond must count separately
    a = np.array([pack([1] * 17)] * 40, dtype=np.uint64)
    b = np.array([pack([1] * 16 + [2])] * 40, dtype=np.uint64)
    r = qcf.history_chi2(a, b, cut=30)
    assert r['histories'] == 2 and r['c2n'] == 2 and abs(r['c2sum'] - 80.0) < 1e-12


def test_first_divergence_and_points():
    a = np.array([pack([1, 3, 4]), pack([1, 3, 4]), pack([1] * 20), pack([])], dtype=np.uint64)
    b = np.array([pack([1, 3, 4]), pack([1, 5, 4]), pack([1] * 19 + [2]), pack([])], dtype=np.uint64)
    assert list(qcf.first_divergence(a, b)) == [-1, 1, 19, -1]
    assert list(qcf.num_points(a)) == [3, 3, 20, 0]


def test_chi2_of_independent_streams_is_statistical(oracle_lib):
    """Two runs of the nested-box event on independent streams (different event index)
    are samples of one distribution: chi2/ndf stays near 1."""
    s = scenes.nested_boxes(num_photon=20_000)
    cfg = abi.default_config(**s.config)
    cfg.seq_enabled = 1
    o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
    gs = s.gensteps()
    r1 = o.simulate(gs, event_index=1, nthreads=8)
    r2 = o.simulate(gs, event_index=2, nthreads=8, precision=1)
    c = qcf.history_chi2(qcf.seqhis(r1['seq']), qcf.seqhis(r2['seq']))
    assert c['c2n'] >= 3
    assert c['c2per'] < 3.0, c
    # and a planted difference (every 10th photon of B absorbed at once) is caught
    h2 = qcf.seqhis(r2['seq']).copy()
    h2[::10] = pack([15, 4])
    assert qcf.history_chi2(qcf.seqhis(r1['seq']), h2)['c2per'] > 10.0


@pytest.fixture(scope='module')
def thin_runs(oracle_lib):
    s = scenes.thin_shells()
    ng = 1000
    inp = scenes.thin_shell_grazing_photons(s, ng)
    gs = np.concatenate([G.input_photon(ng), s.gensteps(scale_photons=1000 / 5e6)])
    out = {}
    for name, prec, f64 in (('fp32', 0, 0), ('fp64roots', 0, 1), ('fp64', 1, 0)):
        cfg = abi.default_config(**s.config)
        cfg.seq_enabled = 1
        cfg.fp64_roots = f64
        o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
        o.set_cull(True)
        out[name] = o.simulate(gs, precision=prec, nthreads=8, input_photons=inp)
    return ng, out


@pytest.mark.parametrize('variant', ['fp32', 'fp64roots'])
def test_thin_shells_isotropic_within_tolerance(thin_runs, variant):
    """Photons from the origin cross the shells near normal incidence: no history
    diverges from FP64 and end points agree within 0.05 mm and 1e-3 ns."""
    ng, R = thin_runs
    a, b = R[variant], R['fp64']
    e = qcf.endpoint_errors(a['photon'][ng:], a['seq'][ng:], b['photon'][ng:], b['seq'][ng:])
    assert e['fully_recorded'] > 100
    assert e['divergent_fraction'] == 0.0
    assert e['pos_mm_max'] < 0.05 and e['time_ns_max'] < 1e-3, e


def test_thin_shells_grazing_divergence_is_counted(thin_runs):
    """Within 1e-3 rad of grazing the Fresnel/TIR decisions flip with the last bits of
    the incidence angle: a large fraction of histories diverge from FP64 in either
    variant; FP64 roots shrink the median end-point error of the photons that agree."""
    ng, R = thin_runs
    b = R['fp64']
    e32 = qcf.endpoint_errors(R['fp32']['photon'][:ng], R['fp32']['seq'][:ng], b['photon'][:ng], b['seq'][:ng])
    e64 = qcf.endpoint_errors(R['fp64roots']['photon'][:ng], R['fp64roots']['seq'][:ng], b['photon'][:ng], b['seq'][:ng])
    for e in (e32, e64):
        assert e['fully_recorded'] > 100
        assert 0.0 < e['divergent_fraction'] < 0.8, e
    assert e64['pos_mm_p50'] < e32['pos_mm_p50'], (e32, e64)
