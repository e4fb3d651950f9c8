"""The C++ facade (include/oxr_g4cx.hpp) on the CPU: the photon-file reader, the
genstep makers against the Python builders, OPTICKS_MAX_SLOT-style integer specs, and
the example application's argument errors (tests/test_GPUPhotonFileSource.sh Test 3)."""
import os
import subprocess

import numpy as np
import pytest

from oxr import gensteps as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
EXE = os.path.join(ROOT, 'eic-opticks_amd', 'oxr', 'photon_file_source')

PROBE = r'''
#include "oxr_g4cx.hpp"
#include <cstdio>
int main(int argc, char** argv) {
    const char* specs[] = {"M1", "K100", "123", "h3", "X4", "G2"};
    for (const char* s : specs) std::printf("%lld\n", (long long)oxr::ParseIntSpec(s));
    std::vector<oxr::sphoton> p = oxr::load_photons_txt(argv[1]);
    std::printf("%zu\n", p.size());
    for (const oxr::sphoton& q : p) std::printf("%.9g %.9g %.9g %.9g\n", q.pos[0], q.mom[2], q.pol[0], q.wavelength);
    oxr::CerenkovStep c;
    c.trackid = 7; c.matline = 3; c.numphoton = 120; c.pos[0] = 1; c.pos[1] = 2; c.pos[2] = 3; c.time = 0.5;
    c.delta_position[2] = 4; c.step_length = 4; c.charge = -1; c.weight = 1; c.pre_velocity = 280; c.post_velocity = 279;
    c.beta_inverse = 1.05; c.wmin_nm = 200; c.wmax_nm = 700; c.max_cos = 0.7; c.max_sin2 = 0.51;
    c.mean_photons1 = 30; c.mean_photons2 = 31;
    oxr::quad6 g = oxr::MakeGenstep_G4Cerenkov_modified(c);
    const unsigned* u = reinterpret_cast<const unsigned*>(g.f);
    for (int k = 0; k < 24; k++) std::printf("%u ", u[k]);
    std::printf("\n");
    oxr::ScintillationStep s;
    s.numphoton = 1000; s.pos[0] = -1; s.time = 2; s.step_length = 0.3; s.mean_velocity = 150; s.scintillation_time = 1000;
    oxr::quad6 h = oxr::MakeGenstep_DsG4Scintillation_r4695(s);
    u = reinterpret_cast<const unsigned*>(h.f);
    for (int k = 0; k < 24; k++) std::printf("%u ", u[k]);
    std::printf("\n");
    oxr::SEvt* ev = oxr::SEvt::Get_EGPU();
    oxr::SEvt::AddGenstep(g); oxr::SEvt::AddGenstep(h);
    ev->SetInputPhoton(p);
    std::printf("%lld %lld\n", (long long)ev->GetNumGenstepFromGenstep(0), (long long)ev->GetNumPhotonCollected(0));
    return 0;
}
'''


@pytest.fixture(scope='module')
def probe(tmp_path_factory, engine_mod):
    d = tmp_path_factory.mktemp('probe')
    src = d / 'probe.cpp'
    src.write_text(PROBE)
    exe = d / 'probe'
    subprocess.check_call(['g++', '-std=c++17', '-Wall', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe),
                           '-L', os.path.dirname(EXE), '-loxr', f'-Wl,-rpath,{os.path.dirname(EXE)}'])
    out = subprocess.run([str(exe), os.path.join(GOLDEN, 'raindrop_photons_comments.txt')], capture_output=True,
                         text=True, timeout=60, check=True)
    return out.stdout.split('\n')


def test_int_specs(probe):
    """sstr::ParseIntSpec: M1 -> 1e6, K100 -> 1e5, X4 -> 2^4-1."""
    assert [int(x) for x in probe[:6]] == [1000000, 100000, 123, 300, 15, 2000000000]


def test_photon_file_reader(probe):
    assert int(probe[6]) == 2
    a = [float(x) for x in probe[7].split()]
    b = [float(x) for x in probe[8].split()]
    assert a == [-10.0, pytest.approx(0.957826), 1.0, 420.0] and b[3] == 450.0


def test_genstep_makers_match_python_builders(probe):
    """MakeGenstep_G4Cerenkov_modified / _DsG4Scintillation_r4695 write the scerenkov /
    sscint fields (sysrap/scerenkov.h:29-58, sysrap/sscint.h:34-58) at the offsets the
    Python builders use; gencodes 18 and 5 as U4 sets them."""
    c = np.array([int(x) for x in probe[9].split()], dtype=np.uint32).reshape(6, 4)
    f = c.view(np.float32)
    assert c[0].tolist() == [18, 7, 3, 120]
    assert np.allclose(f[1], [1, 2, 3, 0.5]) and np.allclose(f[2], [0, 0, 4, 4])
    assert c[3, 0] == 11 and np.allclose(f[3, 1:], [-1, 1, 280])
    assert np.allclose(f[4], [1.05, 200, 700, 0.7]) and np.allclose(f[5], [0.51, 30, 31, 279])
    py = G.cerenkov(120, (1, 2, 3), (0, 0, 4), 4.0, 1 / 1.05, 3, 200.0, 700.0, 1.5, 30.0, time=0.5)
    # the fields both write with the same meaning
    assert np.allclose(py[0, 1], f[1]) and np.allclose(py[0, 2], f[2]) and py[0, 4, 0] == np.float32(1.05)
    s = np.array([int(x) for x in probe[10].split()], dtype=np.uint32).reshape(6, 4)
    sf = s.view(np.float32)
    assert s[0].tolist() == [5, 0, 0, 1000] and sf[1, 0] == -1 and sf[1, 3] == 2
    assert np.allclose(sf[3, 1:], [-1, 1, 150]) and s[4, 0] == 1 and sf[5, 0] == 1000
    py = G.scintillation(1000, (-1, 0, 0), (0, 0, 0), 0.3, 150.0, 1000.0, time=2.0)
    assert np.allclose(py[0, 1], sf[1]) and np.allclose(py[0, 3, 1:], sf[3, 1:]) and py[0, 5, 0] == sf[5, 0]
    assert probe[11].split() == ['3', '1122']


def test_example_argument_errors(engine_mod):
    r = subprocess.run([EXE, '-g', '/nonexistent'], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and 'Required argument not used: -p' in r.stderr
    r = subprocess.run([EXE, '-p', os.path.join(GOLDEN, 'raindrop_photons.txt'), '-g', '/nonexistent/cf'],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and 'no CSGFoundry folder' in r.stderr


# The pure virtual members of sysrap/SSimulator.h:16-37, as a host built against that
# interface declares them.
SSIMULATOR_LIKE = r'''
struct SSimulatorLike {
    virtual ~SSimulatorLike() = default;
    virtual double render_launch() = 0;
    virtual double simtrace_launch() = 0;
    virtual double simulate_launch() = 0;
    virtual double launch() = 0;
    virtual const char* desc() const = 0;
    virtual double simulate(int eventID, bool reset = false) = 0;
    virtual double simtrace(int eventID) = 0;
    virtual double render(const char* stem = nullptr) = 0;
    virtual void reset(int eventID) = 0;
};
'''


def test_simulator_implements_the_launcher_protocol(tmp_path):
    """oxr::Simulator<Base> overrides every pure virtual of the launcher interface, so a
    host can hold it through that base (compiled only: no engine is created here)."""
    src = tmp_path / 'sim_iface.cpp'
    src.write_text('#include "oxr_g4cx.hpp"\n#include <type_traits>\n' + SSIMULATOR_LIKE + r'''
static_assert(!std::is_abstract<oxr::Simulator<SSimulatorLike>>::value, "all launcher members implemented");
static_assert(std::is_base_of<SSimulatorLike, oxr::Simulator<SSimulatorLike>>::value, "held through the interface");
static_assert(!std::is_abstract<oxr::Simulator<>>::value, "usable on its own");
int main() { return 0; }
''')
    subprocess.check_call(['g++', '-std=c++17', '-Wall', '-Werror', '-fsyntax-only', '-I', os.path.join(ROOT, 'include'),
                           str(src)])
