"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and refuses to run without a gfx950 device (no CPU fallback)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tenet_hip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tenet_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_and_exports_every_declared_symbol():
    from tenet_hip import _lib

    declared = _declared()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(_lib.lib, s)]
    assert not missing, missing
    # the ctypes table covers exactly the declared entry points
    assert sorted(_lib.EXPORTED) == declared
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(r" T (tenet_[a-z0-9_]+)", out)))
    assert exported == declared


def test_version_string():
    import tenet_hip

    assert "gfx950" in tenet_hip.version()


def test_no_cpu_fallback_without_device():
    import tenet_hip
    from tenet_hip import _lib

    import ctypes

    n = ctypes.c_int(-1)
    assert _lib.lib.tenet_device_count(ctypes.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a device is visible: the no-device path is not reachable here")
    with pytest.raises(tenet_hip.TenetError) as e:
        tenet_hip.Context(0)
    assert e.value.status == _lib.TENET_ERR_NODEVICE


def test_error_paths_without_device():
    """Argument validation happens before any device work."""
    from tenet_hip import _lib

    import ctypes

    assert _lib.lib.tenet_create(0, None) == _lib.TENET_ERR_ARG
    assert _lib.lib.tenet_malloc(None, 16, None) == _lib.TENET_ERR_ARG
    assert b"null" in _lib.lib.tenet_last_error()
    assert _lib.lib.tenet_plan_destroy(None) == 0
    assert _lib.lib.tenet_graph_destroy(None) == 0
    assert _lib.lib.tenet_destroy(None) == 0


def test_brickwork_schedule_counts():
    """n = 64, depth 20: 32 odd-layer + 31 even-layer gates per layer pair."""
    from tenet_hip.tebd import brickwork_schedule
    from oracle import tenet_oracle as O

    s = brickwork_schedule(64, 20)
    assert len(s) == 630
    assert s == list(O.brickwork_schedule(64, 20))
    assert s[:3] == [0, 2, 4] and s[32:35] == [61, 59, 57]


def test_set_log_writes_errors_to_the_file(tmp_path):
    """tenet_set_log(1, path): every error message set afterwards is appended
    to the file; level 0 stops it; a bad level is refused."""
    from tenet_hip import _lib

    path = tmp_path / "tenet.log"
    assert _lib.lib.tenet_set_log(3, None) == _lib.TENET_ERR_ARG
    assert _lib.lib.tenet_set_log(1, str(path).encode()) == 0
    try:
        assert _lib.lib.tenet_malloc(None, 16, None) == _lib.TENET_ERR_ARG
    finally:
        assert _lib.lib.tenet_set_log(0, None) == 0
    assert _lib.lib.tenet_malloc(None, 16, None) == _lib.TENET_ERR_ARG
    lines = path.read_text().splitlines()
    assert len(lines) == 1 and "error: null context" in lines[0]
