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


FORWARD_PROBE = r'''
import ctypes, sys
lib = ctypes.CDLL(sys.argv[1])
lib.BlinkingReconnection.restype = ctypes.c_uint64
lib.BlinkingReconnection.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64,
                                     ctypes.c_uint64, ctypes.c_double]
lib.TrackAnalysis.restype = ctypes.c_bool
print(lib.BlinkingReconnection(None, None, 7, 0.1, 1, 3, 2.0))
'''


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
@pytest.mark.parametrize("with_host", [False, True])
def test_host_exports_forward(tmp_path, with_host):
    """Host-side exports go to the library named by PALMTRACER_HOST_LIBRARY, else return 0."""
    env = dict(os.environ)
    env.pop("PALMTRACER_HOST_LIBRARY", None)
    if with_host:
        src = tmp_path / "host.c"
        src.write_text("#include <stdint.h>\n"
                       "uint64_t BlinkingReconnection(const double* i, double* o, uint64_t n, double p, uint64_t m,"
                       " uint64_t d, double s) { return n * 6 + m + d; }\n")
        so = tmp_path / "libhost.so"
        subprocess.run(["gcc", "-shared", "-fPIC", str(src), "-o", str(so)], check=True)
        env["PALMTRACER_HOST_LIBRARY"] = str(so)
    probe = tmp_path / "probe.py"
    probe.write_text(FORWARD_PROBE)
    r = subprocess.run([os.sys.executable, str(probe), LIB], capture_output=True, text=True, env=env, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == ("46" if with_host else "0")
    if not with_host:
        assert "PALMTRACER_HOST_LIBRARY" in r.stderr
