import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "atom-3d-engine_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) visible to HIP")


@pytest.fixture(scope="session")
def renderer():
    """One HIP renderer for the whole GPU session (the library has no CPU fallback).  Its scenes
    render with the table-driven kernels (code generation off); the generated kernels are
    checked against them in test_codegen.py."""
    from mm3e import hip
    r = hip.Renderer()
    r.set_codegen(hip.CODEGEN_OFF)
    yield r
    r.close()
