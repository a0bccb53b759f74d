import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))

HOST_EXE = os.path.join(REPO, "axisem_amd", "fortran", "_build", "axisem_gpu")
LIB = os.path.join(REPO, "axisem_amd", "libaxihip.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long CPU reference comparison")


@pytest.fixture(scope="session")
def host_exe():
    if not os.path.exists(HOST_EXE):
        pytest.skip("drop-in host not built (needs the reference tree at build time)")
    return HOST_EXE


@pytest.fixture(scope="session")
def state_cache(tmp_path_factory, host_exe):
    """Record (once per session) the state of a golden case or batch member
    with the drop-in host in record mode; returns a function name -> (rundir,
    state path)."""
    import cases
    from axisem_amd import inputs

    base = tmp_path_factory.mktemp("states")
    done = {}

    def get(name):
        if name not in done:
            rundir = str(base / name)
            inputs.write_run_dir(rundir, cases.mesh_of(name), cases.ALL_CONFIGS[name])
            st = os.path.join(rundir, "state.bin")
            inputs.record_state(host_exe, rundir, st)
            done[name] = (rundir, st)
        return done[name]

    return get
