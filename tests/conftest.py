import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ('eic-opticks_amd', 'oracle'):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs through the HIP library)')


@pytest.fixture(scope='session')
def oracle_lib():
    so = os.path.join(ROOT, 'oracle', 'liboxr_oracle.so')
    if not os.path.exists(so):
        subprocess.check_call(['make', '-s', '-C', os.path.join(ROOT, 'oracle')])
    import oracle
    return oracle


@pytest.fixture(scope='session')
def engine_mod():
    from oxr import engine
    if not os.path.exists(engine.LIB_PATH):
        from oxr import build
        build.build()
    engine.lib()
    return engine
