"""The engine library loads and exports every entry point include/*.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from squishy_volumes_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(mpm_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_symbols()
    for required in ("mpm_create", "mpm_step_to", "mpm_read_particles", "mpm_read_grid", "mpm_destroy", "mpm_error_string"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not abi.LIB_PATH.exists():
        pytest.fail(f"{abi.LIB_PATH} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(str(abi.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_covers_the_header():
    assert set(declared_symbols()) == set(abi.SIGNATURES)


def test_struct_layouts_match_header_sizes():
    # Sizes of the C structs as laid out by the C compiler of the platform.
    assert ctypes.sizeof(abi.MpmParticleParameters) == 36
    assert ctypes.sizeof(abi.MpmParticlesView) == 8 * 12
    assert ctypes.sizeof(abi.MpmFrameInput) == 8 + 12 + 12 + 4 + 4 + 32 + 48
    assert ctypes.sizeof(abi.MpmColliderMesh) == 8 + 16
    assert ctypes.sizeof(abi.MpmStatus) == 72


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(abi.MissingEngineLibrary):
        abi.load_library(tmp_path / "nope.so")


def test_python_mirror_holds_the_header_constants():
    """Return codes, option bits and the per-call substep cap in abi.py match
    include/mpm_engine.h."""
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "mpm_engine.h").read_text(), flags=re.S)
    codes = dict(re.findall(r"\bMPM_((?:OK|ERR_\w+))\s*=\s*(-?\d+)", text))
    assert len(codes) >= 12
    for name, value in codes.items():
        assert getattr(abi, name) == int(value), name
    cap = re.search(r"#define MPM_MAX_SUBSTEPS_PER_CALL \(1u << (\d+)\)", text)
    assert cap and abi.MAX_SUBSTEPS_PER_CALL == 1 << int(cap.group(1))
