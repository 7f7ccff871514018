"""Shared pytest configuration: the ``gpu`` marker, import path, library build."""

import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))

_LIB = ROOT / "realspace_exatdiagonalization_py_amd" / "libedx.so"
if not _LIB.exists():  # build once if nothing is built yet (builder loaded by path)
    import importlib.util

    _spec = importlib.util.spec_from_file_location("_edx_build", _LIB.parent / "csrc" / "build.py")
    _mod = importlib.util.module_from_spec(_spec)
    _spec.loader.exec_module(_mod)
    _mod.build()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
