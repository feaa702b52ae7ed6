import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI of libswe.so)")


def _ensure_built():
    """Build what is missing (only when nothing is built yet)."""
    import oracle as O

    if not os.path.exists(O.LIB_PATH):
        O.build()
    lib = os.path.join(ROOT, "simplewinderosion_amd", "libswe.so")
    if not os.path.exists(lib):
        from simplewinderosion_amd import build

        build.build()


_ensure_built()


@pytest.fixture
def statics():
    """Restore the class-level parameter statics after a test changes them."""
    from simplewinderosion_amd.world import Wind, World

    saved = ({k: getattr(Wind, k) for k in ("maxAge", "boundarylayer", "suspension", "gravity", "pspeed")},
             {k: getattr(World, k) for k in ("SEED", "lrate", "maxdiff", "settling")})
    yield
    for k, v in saved[0].items():
        setattr(Wind, k, v)
    for k, v in saved[1].items():
        setattr(World, k, v)
