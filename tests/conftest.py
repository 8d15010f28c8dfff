import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _ensure_cpu_libs():
    """Build the host-only libraries if absent (fast, CPU only).  The HIP
    library is never built from a test: build() does that beforehand."""
    from hicops_amd import build
    if not (ROOT / "oracle" / "liboracle.so").exists():
        build.build_oracle()
    if not (ROOT / "hicops_amd" / "libhcx_host.so").exists():
        build.build_host()
    # the C++ entry-point library links the HIP library; g++ only
    if (ROOT / "hicops_amd" / "libhcx.so").exists() and not (ROOT / "hicops_amd" / "libhcx_dslim.so").exists():
        build.build_mirror()


_ensure_cpu_libs()


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (plumbing only; counting devices does not initialise HIP)
        return torch.cuda.device_count() > 0
    except Exception:  # noqa: BLE001
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def small_db():
    from hicops_amd import workload as W
    p = W.scaled("small", npep=3000, nspec=200, dM=100.0)
    db = W.Database(p)
    q = W.make_queries(db, p)
    return p, db, q
