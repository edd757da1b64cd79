"""The C-ABI library loads without a GPU and exports every entry point include/palmtracer.h declares."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "palmtracer-cpp_amd", "libpalmtracer_amd.so")
HEADER = os.path.join(ROOT, "include", "palmtracer.h")
REFERENCE_HOT_PATH = {"Localization", "Fitting", "Wavelett", "AutoThreshold"}


def declared():
    text = open(HEADER).read()
    return set(re.findall(r"PT_API\s+[\w\s\*]+?PT_CALL\s+(\w+)\s*\(", text))


def test_header_declares_reference_hot_path():
    names = declared()
    assert {"Localization", "Fitting", "Wavelett"} <= names
    assert len(names) >= 8


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert declared() <= exported
    # Nothing but the declared C ABI is exported (internal C++ symbols stay hidden).
    extra = {s for s in exported - declared() if not s.startswith("_")}
    assert not extra, extra


def test_package_fails_loudly_without_library(tmp_path, monkeypatch):
    import importlib
    import palmtracer_amd
    mod = importlib.reload(palmtracer_amd)
    monkeypatch.setattr(mod, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(mod, "_lib", None)
    with pytest.raises(RuntimeError):
        mod.lib()
