import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


def _ensure_built():
    import subprocess

    if not (ROOT / "oracle" / "libtit_oracle.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    if not (ROOT / "titsolver_amd" / "libtit_hip.so").exists():
        subprocess.run(["make", "-s", "-j8", "-C",
                        str(ROOT / "titsolver_amd" / "csrc")], check=True)


_ensure_built()


@pytest.fixture(autouse=True)
def _reset_oracle_mutation():
    import oracle

    oracle.set_mutation(0)
    yield
    oracle.set_mutation(0)
