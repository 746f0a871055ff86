import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the whole GPU session (fails loudly without a GPU)."""
    from turbogp_amd import Context
    c = Context(0)
    print(f"[ctx] {c.info.name.decode()} {c.info.arch.decode()} CUs={c.info.cu_count}", flush=True)
    yield c
    c.close()
