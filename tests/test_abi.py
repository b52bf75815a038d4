"""The C ABI library loads without a GPU and exports every function include/rgkc.h declares; with
no device, context creation fails with RGKC_E_NODEV instead of falling back to any CPU path."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "rgkc.h")
LIB = os.path.join(ROOT, "rapidgkc_amd", "librgkc.so")


def declared():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rgkc_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    if not os.path.exists(LIB):
        pytest.skip("librgkc.so not built")
    names = declared()
    assert len(names) >= 14
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = ctypes.CDLL(LIB)
    for n in names:
        getattr(L, n)


def test_python_binding_matches_header():
    from rapidgkc_amd import native
    assert sorted(native.EXPORTS) == declared()


def test_no_device_fails_loudly():
    if not os.path.exists(LIB):
        pytest.skip("librgkc.so not built")
    import torch  # noqa: F401  (only to know whether this host has a device)
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from rapidgkc_amd import native
    with pytest.raises(native.RgkcError) as e:
        native.Counter(0)
    assert e.value.code == -7
