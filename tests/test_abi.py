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


@pytest.mark.parametrize("k,p,parts", [(31, 11, 64), (25, 9, 31), (55, 13, 128), (21, 7, 16)])
def test_store_minimizer_length(k, p, parts):
    """The minimizer length is recovered from a store's records alone (host code of the library,
    no device): the oracle's store of seeded reads built with (k, p)."""
    if not os.path.exists(LIB):
        pytest.skip("librgkc.so not built")
    import numpy as np
    import oracle_api as oa
    from rapidgkc_amd import native, synth
    g = synth.genome(91, 60_000, gc=0.45)
    chars, offs = synth.short_reads(g, 92, 2000, 150, 0.01)
    st, po, _, _ = oa.store(chars, offs, k, p, parts)
    assert native.store_minimizer_length(st, po, k) == p
    # too few records decide nothing
    with pytest.raises(native.RgkcError):
        native.store_minimizer_length(st[:0], np.zeros(2, dtype=np.uint64), k)
