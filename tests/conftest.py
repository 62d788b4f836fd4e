import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; parity tests through the C ABI")


def _ensure_built():
    """Tests load what is built; only if nothing is built do they build it."""
    from yaoray_amd import native

    if not (os.path.exists(native.lib_path("libyrh.so")) and os.path.exists(native.lib_path("libyaoray_host.so"))):
        from yaoray_amd import build

        build.build_all()
    from oracle import oracle_py

    if not os.path.exists(os.path.join(oracle_py.HERE, "liboracle.so")):
        oracle_py.build(with_reference=False)


_ensure_built()


@pytest.fixture(scope="session")
def device():
    from yaoray_amd import Device

    dev = Device(0)
    yield dev
    dev.synchronize()
    dev.close()


@pytest.fixture(scope="session")
def ref_available():
    from oracle import oracle_py

    return oracle_py.ref_available()
