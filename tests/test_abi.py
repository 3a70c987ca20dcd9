"""The C-ABI library loads and exports every entry point include/*.h declares
(no device calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

from mpx_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(mpxkkt_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 8
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTED)


def test_default_params_and_version():
    p = _lib.default_params()
    assert p.max_trials == 10 and abs(p.mu_b - 1e-2) < 1e-15 and p.segments == 1
    assert _lib.load().mpxkkt_version() >= 100


def test_argument_errors_without_device():
    """Argument checks run before any device call and report through last_error."""
    lib = _lib.load()
    pr = _lib.Problem()
    pr.dtype, pr.model, pr.batch = 0, 1, 0
    nbytes = ctypes.c_size_t(0)
    st = lib.mpxkkt_workspace_size(ctypes.byref(pr), ctypes.byref(_lib.default_params()),
                                   ctypes.byref(nbytes))
    assert st == _lib.ERR_ARG
    assert b"batch" in lib.mpxkkt_last_error()


def test_phase_count_matches_header():
    src = open(os.path.join(ROOT, "include", "mpx_kkt.h")).read()
    m = re.search(r"#define\s+MPXKKT_PHASES\s+(\d+)", src)
    assert m and int(m.group(1)) == _lib.PHASES
