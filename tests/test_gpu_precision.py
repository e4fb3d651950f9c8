"""GPU precision on the thin-shell workload (50 concentric 1 mm shells at 1-20 m plus
instanced zsphere+cylinder solids), through the C ABI:

* with grazing input photons (within 1e-3 rad of tangent to the outer shells) and
  isotropic torch photons, the FP32 kernel and the FP64-roots kernel each equal the
  oracle's same-precision arithmetic bit for bit;
* against the FP64 oracle on the same random streams, isotropic photons keep every
  history and end within 0.05 mm / 1e-3 ns; grazing histories diverge in a counted
  fraction, and FP64 roots reduce the median end-point error of those that agree."""
import numpy as np
import pytest

from oxr import abi, gensteps as G, qcf, scenes

pytestmark = pytest.mark.gpu

NG = 1000   # grazing input photons, then 1000 isotropic torch photons


@pytest.fixture(scope='module')
def thin():
    s = scenes.thin_shells()
    inp = scenes.thin_shell_grazing_photons(s, NG)
    gs = np.concatenate([G.input_photon(NG), s.gensteps(scale_photons=1000 / 5e6)])
    return s, inp, gs


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def gpu_run(engine_mod, s, inp, gs, fp64_roots, form):
    cfg = abi.default_config(**s.config, seq_enabled=1, fp64_roots=fp64_roots)
    with engine_mod.Engine(s.geometry, s.physics, cfg) as e:
        e.set_input_photons(inp)
        e.set_gensteps(gs)
        e.launch(form)
        return dict(photon=e.photons(), seq=e.seq())


@pytest.fixture(scope='module')
def oracle_runs(oracle_lib, thin):
    s, inp, gs = thin
    out = {}
    for name, prec, f64 in (('fp32', 0, 0), ('fp64roots', 0, 1), ('fp64', 1, 0)):
        cfg = abi.default_config(**s.config, seq_enabled=1, fp64_roots=f64)
        o = oracle_lib.Oracle(s.geometry, s.physics, cfg)
        o.set_cull(True)
        out[name] = o.simulate(gs, precision=prec, nthreads=8, input_photons=inp)
    return out


@pytest.mark.parametrize('fp64_roots', [0, 1])
@pytest.mark.parametrize('form', [abi.OXR_FORM_PLAIN, abi.OXR_FORM_PERSISTENT])
def test_thin_shells_grazing_bitexact(engine_mod, thin, oracle_runs, fp64_roots, form):
    s, inp, gs = thin
    g = gpu_run(engine_mod, s, inp, gs, fp64_roots, form)
    r = oracle_runs['fp64roots' if fp64_roots else 'fp32']
    print('thin_shells grazing', 'fp64_roots', fp64_roots, 'form', form, len(g['photon']), 'photons', flush=True)
    assert np.array_equal(bits(g['photon']), bits(r['photon']))
    assert np.array_equal(g['seq'], r['seq'])


def test_thin_shells_against_fp64(engine_mod, thin, oracle_runs):
    s, inp, gs = thin
    b = oracle_runs['fp64']
    rep = {}
    for f64 in (0, 1):
        g = gpu_run(engine_mod, s, inp, gs, f64, abi.OXR_FORM_PERSISTENT)
        iso = qcf.endpoint_errors(g['photon'][NG:], g['seq'][NG:], b['photon'][NG:], b['seq'][NG:])
        gra = qcf.endpoint_errors(g['photon'][:NG], g['seq'][:NG], b['photon'][:NG], b['seq'][:NG])
        rep[f64] = (iso, gra)
        print('fp64_roots', f64, 'isotropic', iso, 'grazing', gra, flush=True)
        assert iso['fully_recorded'] > 100 and iso['divergent_fraction'] == 0.0
        assert iso['pos_mm_max'] < 0.05 and iso['time_ns_max'] < 1e-3
        assert gra['fully_recorded'] > 100 and 0.0 < gra['divergent_fraction'] < 0.8
    assert rep[1][1]['pos_mm_p50'] < rep[0][1]['pos_mm_p50']
