import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gpu-cpm_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(__file__)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # The product library is built by __graft_entry__.build() (or
    # gpu-cpm_amd/build.py); build it here only when nothing is built and the
    # compiler is present, so that a fresh checkout's CPU suite can import it.
    lib = os.path.join(ROOT, "gpu-cpm_amd", "gpucpm", "libgpucpm.so")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(lib) and os.path.exists(hipcc):
        import subprocess
        subprocess.run([sys.executable, os.path.join(ROOT, "gpu-cpm_amd", "build.py"), "-q"], check=True)


@pytest.fixture
def engine_options():
    """Set kernel options (small, compact, full, attempts, narrow, box3) and
    the launch form (mode = phase | global | lds, tile) for the engines this
    test makes through helpers.make_cpm; cleared afterwards."""
    import helpers

    def set_options(**kw):
        helpers.ENGINE_OPTIONS.update(kw)

    yield set_options
    helpers.ENGINE_OPTIONS.clear()
