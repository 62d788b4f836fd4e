"""The C-ABI library loads and exports every symbol include/yrh.h declares; the
ctypes bindings cover exactly that set. No GPU calls."""
import ctypes
import os
import re

from yaoray_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(yrh\w*|yrhost\w*)\s*\(", text, flags=re.M))


def test_yrh_exports_every_declared_symbol():
    names = declared(os.path.join(ROOT, "include", "yrh.h"))
    assert "yrh_trace_closest" in names and "yrh_upload_geometry" in names
    lib = ctypes.CDLL(native.lib_path("libyrh.so"))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(native.YRH_SIGNATURES), names ^ set(native.YRH_SIGNATURES)


def test_host_library_exports_its_c_api():
    names = declared(os.path.join(ROOT, "yaoray_amd", "host", "yaoray_host.h"))
    native.yrh()
    lib = ctypes.CDLL(native.lib_path("libyaoray_host.so"))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(native.HOST_SIGNATURES), names ^ set(native.HOST_SIGNATURES)


def test_only_c_abi_symbols_are_exported():
    # The host library's yr:: symbols share their mangled names with the
    # reference's; they must stay hidden so neither interposes the other.
    import subprocess

    for name, prefix in (("libyrh.so", "yrh_"), ("libyaoray_host.so", "yrhost_")):
        out = subprocess.run(["nm", "-D", "--defined-only", native.lib_path(name)], capture_output=True,
                             text=True, check=True).stdout
        funcs = [line.split()[-1] for line in out.splitlines() if " T " in line]
        assert funcs and all(f.startswith(prefix) for f in funcs), [f for f in funcs if not f.startswith(prefix)]
