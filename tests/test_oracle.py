"""Pins the oracle (oracle/oracle.cpp) before anything is checked against it.

Each oracle function is checked against (a) an independent closed form written here from the
reference's definitions, (b) the reference's own cpu_funcs.cpp built into oracle/_ref (skipped
when that build is absent), (c) brute force over the raw reads, and (d) invariants. The checks
are chosen so that a dropped term, a wrong shift or a swapped strand fails at least one.
"""
import itertools
import os

import numpy as np
import pytest

import oracle_api as oa
from rapidgkc_amd import synth

CODE = {"A": 0, "C": 1, "G": 2, "T": 3, "N": 0}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A", "N": "T"}


def enc(s):
    v = 0
    for c in s:
        v = (v << 2) | CODE[c]
    return v


def rcs(s):
    return "".join(COMP[c] for c in reversed(s))


def valid_str(s):
    # minimizer_filter.h:19-20 read as strings: not [AC][AC]A at the start, not AAA at the end
    bad_start = s[0] in "AC" and s[1] in "AC" and s[2] == "A"
    return not bad_start and s[-3:] != "AAA"


def closed_minimizers(read, k, p):
    mask = (1 << (2 * p)) - 1
    out = []
    for i in range(len(read) - k + 1):
        m = mask
        for j in range(i, i + k - p + 1):
            f = read[j:j + p]
            r = rcs(f)
            for s in (f, r):
                if valid_str(s):
                    m = min(m, enc(s))
        out.append(m)
    return out


def rand_read(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(list(alphabet), size=n))


# ---------------------------------------------------------------- a4 filter truth table
@pytest.mark.parametrize("p", [5, 7, 9])
def test_filter_truth_table(p):
    O = oa.oracle()
    for t in itertools.product("ACGT", repeat=p):
        s = "".join(t)
        assert O.or_pmer_valid(enc(s), p) == int(valid_str(s)), s


# ---------------------------------------------------------------- a3 packing
def closed_pack(read):
    out = []
    i = 0
    while i + 3 <= len(read):
        out.append(0xC0 | (CODE[read[i]] << 4) | (CODE[read[i + 1]] << 2) | CODE[read[i + 2]])
        i += 3
    rem = len(read) - i
    t = rem << 6
    for j in range(rem):
        t |= CODE[read[i + j]] << (4 - 2 * j)
    out.append(t)
    return out


def test_pack_closed_form_and_reference():
    rng = np.random.default_rng(7)
    R = oa.ref(8)
    for n in list(range(1, 40)) + [150, 151, 152]:
        r = rand_read(rng, n, "ACGTN")
        got = oa.pack(r.encode())
        assert got.tolist() == closed_pack(r)
        if R is not None:
            a = np.frombuffer(r.encode() + b"\0", dtype=np.uint8).copy()
            out = np.zeros(n // 3 + 1, dtype=np.uint8)
            R.ref_read_compression(a.ctypes.data, n, out.ctypes.data)
            assert out.tolist() == got.tolist()


# ---------------------------------------------------------------- a5/a6 minimizers and cuts
@pytest.mark.parametrize("k,p", [(10, 5), (15, 7), (31, 11), (25, 9)])
def test_minimizers_closed_form(k, p):
    rng = np.random.default_rng(k * 100 + p)
    for trial in range(30):
        n = int(rng.integers(k, k + 60))
        alph = "ACGT" if trial % 3 else "AC"  # low-complexity reads hit the filter and the mask
        r = rand_read(rng, n, alph)
        assert oa.minimizers(r.encode(), k, p).tolist() == closed_minimizers(r, k, p)


def test_reference_test_vector():
    # cpu_funcs.cpp:656-667 runs gen_skm_offs on this read with k=10, p=5 (it prints, no expected
    # values). Expected values from the closed form: cut wherever the minimizer value changes.
    read = "AAAAAAAAACCCCCCCAAAAAAATGA"
    k, p = 10, 5
    mm = closed_minimizers(read, k, p)
    exp_offs = [0] + [i for i in range(1, len(mm)) if mm[i] != mm[i - 1]] + [len(mm)]
    exp_mm = [mm[o] for o in exp_offs[:-1]]
    offs, mms = oa.skm_offsets(read.encode(), k, p)
    assert offs.tolist() == exp_offs and mms.tolist() == exp_mm
    # Hand check of the first window AAAAAAAAAC: forward p-mers AAAAA and AAAAC both start with
    # [AC][AC]A and are filtered; their reverse complements TTTTT and GTTTT pass, so the
    # minimizer is GTTTT = 0b1011111111.
    assert mm[0] == enc("GTTTT") == 0b1011111111
    R = oa.ref(8)
    if R is not None:
        o2 = np.zeros(64, dtype=np.int32); m2 = np.zeros(64, dtype=np.uint32)
        n = R.ref_gen_skm_offs(read.encode(), len(read), k, p, o2.ctypes.data, m2.ctypes.data)
        assert o2[: n + 1].tolist() == exp_offs and m2[:n].tolist() == exp_mm


@pytest.mark.parametrize("k,p", [(31, 11), (25, 9), (21, 7), (55, 13)])
def test_skm_offsets_vs_reference(k, p):
    R = oa.ref(8)
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(k + p)
    for trial in range(200):
        n = int(rng.integers(k, 400))
        r = rand_read(rng, n, "ACGT" if trial % 4 else "ACA")
        offs, mms = oa.skm_offsets(r.encode(), k, p)
        o2 = np.zeros(n + 2, dtype=np.int32); m2 = np.zeros(n + 1, dtype=np.uint32)
        c = R.ref_gen_skm_offs(r.encode(), n, k, p, o2.ctypes.data, m2.ctypes.data)
        assert o2[: c + 1].tolist() == offs.tolist()
        assert m2[:c].tolist() == mms.tolist()


# ---------------------------------------------------------------- a7/a8/a13 closed forms
def test_partition_bytes_revcomp_closed_forms():
    O = oa.oracle()
    for mm in [0, 1, 5, 511, 512, 0x3FFFFF, 123456]:
        for parts in [31, 512, 1]:
            assert O.or_partition(mm, parts) == ((~mm) & 0xFFFFFFFF) % parts
    for k in [10, 31, 55]:
        for beg in range(0, 7):
            for end in range(beg + 1, beg + 9):
                L = end - beg + k - 1
                x = beg % 3 + L
                exp = (x + 2) // 3 + (1 if x % 3 == 0 else 0)
                assert O.or_skm_bytes(beg, end, k) == exp
    rng = np.random.default_rng(3)
    for k in [1, 5, 21, 31, 32]:
        for _ in range(50):
            s = rand_read(rng, k)
            assert O.or_revcomp64(enc(s), k) == enc(rcs(s))


# ---------------------------------------------------------------- a9/a10 store vs reference
@pytest.mark.parametrize("k,p,parts", [(31, 11, 512), (21, 7, 31), (25, 9, 64)])
def test_store_vs_reference(k, p, parts):
    R = oa.ref(8)
    if R is None:
        pytest.skip("oracle/_ref not built")
    g = synth.genome(11, 20000, gc=0.45)
    chars, offs = synth.short_reads(g, 12, 3000, 150, 0.01)
    st, po, ps, pk = oa.store(chars, offs, k, p, parts)
    ptr = oa.P()
    po2 = np.zeros(parts + 1, dtype=np.uint64); s2 = np.zeros(parts, dtype=np.uint64); k2 = np.zeros(parts, dtype=np.uint64)
    assert R.ref_gen_superkmers(chars.ctypes.data, offs.ctypes.data, len(offs) - 1, k, p, parts, oa.ctypes.byref(ptr),
                                po2.ctypes.data, s2.ctypes.data, k2.ctypes.data) == 0
    ref_bytes = np.ctypeslib.as_array((oa.ctypes.c_uint8 * int(po2[-1])).from_address(ptr.value)).copy()
    R.ref_free(ptr)
    assert po.tolist() == po2.tolist()
    assert ps.tolist() == s2.tolist() and pk.tolist() == k2.tolist()
    assert np.array_equal(st, ref_bytes)  # byte-identical, garbage bits included
    assert int(pk.sum()) == int(sum(max(0, int(l) - k + 1) for l in np.diff(offs)))


# ---------------------------------------------------------------- a12..a17 counting
@pytest.mark.parametrize("k,p", [(31, 11), (21, 7), (55, 13), (33, 11)])
def test_count_vs_reference_and_bruteforce(k, p):
    parts = 64
    g = synth.genome(21, 30000, gc=0.5, tandem_frac=0.0)
    chars, offs = synth.short_reads(g, 22, 4000, 150, 0.01)
    st, po, ps, pk = oa.store(chars, offs, k, p, parts)
    keys, cnts, ro = oa.count_store(st, po, k)
    # union of partition tables == brute force over the reads (pins extraction, canonical form,
    # sort and run-length independently of minimizers and partitioning)
    bk, bc, bd = oa.bruteforce(chars, offs, k)
    sk, sc = oa.sort_table(keys, cnts)
    assert len(sk) == len(bk) == bd
    assert np.array_equal(sk, bk) and np.array_equal(sc, bc)
    assert int(cnts.sum()) == int(pk.sum())
    # partitions are key-disjoint
    assert len(np.unique(sk)) == len(sk)
    R = oa.ref(oa.key_width(k))
    if R is None:
        return
    for i in range(parts):
        seg = np.ascontiguousarray(st[po[i]:po[i + 1]])
        if len(seg) == 0:
            continue
        n = int(pk[i])
        rk = np.zeros(n, dtype=oa.key_dtype(k))
        assert R.ref_extract_sorted(seg.ctypes.data, len(seg), int(ps[i]), n, k, rk.ctypes.data) == n
        # run-length of the reference's sorted keys
        chg = np.ones(n, dtype=bool)
        chg[1:] = rk[1:] != rk[:-1]
        starts = np.flatnonzero(chg)
        rc = np.diff(np.append(starts, n)).astype(np.uint32)
        lo, hi = ro[i], ro[i + 1]
        assert np.array_equal(keys[lo:hi], rk[starts])
        assert np.array_equal(cnts[lo:hi], rc)
        assert R.ref_count_partition(seg.ctypes.data, len(seg), int(ps[i]), n, k) == hi - lo


def test_rc_invariance_and_filter():
    k, p, parts = 31, 11, 128
    g = synth.genome(31, 20000)
    chars, offs = synth.short_reads(g, 32, 2000, 120, 0.02)
    comp = np.zeros(256, dtype=np.uint8)
    for a, b in zip(b"ACGT", b"TGCA"):
        comp[a] = b
    rc_chars = chars.copy()
    for i in range(len(offs) - 1):
        s, e = int(offs[i]), int(offs[i + 1])
        rc_chars[s:e] = comp[chars[s:e][::-1]]
    t1 = oa.count_store(*oa.store(chars, offs, k, p, parts)[:2], k)
    t2 = oa.count_store(*oa.store(rc_chars, offs, k, p, parts)[:2], k)
    for a, b in zip(t1, t2):
        assert np.array_equal(a, b)  # same partitions, same rows: the path is strand-symmetric
    keys, cnts, _ = t1
    fk, fc, _ = oa.count_store(*oa.store(chars, offs, k, p, parts)[:2], k, 2, 5)
    sel = (cnts >= 2) & (cnts <= 5)
    assert np.array_equal(fk, keys[sel]) and np.array_equal(fc, cnts[sel])


def test_golden_fixtures():
    """Committed vectors written by tests/golden/make_golden.py from the reference's own code."""
    path = os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz")
    if not os.path.exists(path):
        pytest.skip("golden fixtures not generated")
    z = np.load(path)
    for tag in ("k31", "k55"):
        k, p, parts = (int(x) for x in z[f"{tag}_kpp"])
        chars, offs = z["chars"], z["offs"]
        st, po, ps, pk = oa.store(chars, offs, k, p, parts)
        assert np.array_equal(po, z[f"{tag}_part_offs"])
        assert np.array_equal(st, z[f"{tag}_store"])
        keys, cnts, ro = oa.count_store(st, po, k)
        kk = z[f"{tag}_keys"]
        if k > 32:
            kk = kk.view(oa.key_dtype(k)).reshape(-1)
        assert np.array_equal(keys, kk) and np.array_equal(cnts, z[f"{tag}_counts"])
        assert np.array_equal(ro, z[f"{tag}_row_offs"])
