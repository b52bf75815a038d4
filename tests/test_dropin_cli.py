"""The reference's own host program (gkc_main.cpp + cpu_funcs.cpp) linked with the drop-in GPU
layer (rapidgkc_amd/host/dropin_gpu.cpp over librgkc.so) in place of its .cu files.

Its GPU path (phase 1 = GenSuperkmerGPU, phase 2 = kmc_counting_GPU_streams) must report the
same "Total number of distinct kmers" (gkc_main.cpp:385) as its own -cpuonly path and as the
oracle's brute-force count, for 64-bit and 128-bit keys, in memory and disk store modes.
The binaries are built from the reference tree in the development container; tests skip where
they were not built.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_api as oa
from rapidgkc_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "rapidgkc_amd", "host", "_dropin")


def write_fasta(path, chars, offs):
    with open(path, "w") as f:
        for i in range(len(offs) - 1):
            f.write(f">r{i}\n{chars[offs[i]:offs[i + 1]].tobytes().decode()}\n")


def run(binary, args, cwd):
    print(f"[dropin] {binary} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(BIN, binary)] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-2000:]
    m = re.search(r"Total number of distinct kmers: (\d+)", out)
    assert m, out[-2000:]
    return int(m.group(1))


@pytest.mark.gpu
@pytest.mark.parametrize("k,p,binary,mode", [(31, 11, "rapidgkc", "-im"), (25, 9, "rapidgkc", "disk"),
                                              (55, 13, "l_rapidgkc", "-im")])
def test_dropin_gpu_matches_cpu(tmp_path, k, p, binary, mode):
    if not os.path.exists(os.path.join(BIN, binary)):
        pytest.skip("drop-in binaries not built (needs the reference tree at build time)")
    g = synth.genome(13, 150_000, gc=0.44)
    chars, offs = synth.short_reads(g, 14, 30_000, 150, 0.01)
    fa = tmp_path / "reads.fa"
    write_fasta(fa, chars, offs)
    _, _, want = oa.bruteforce(chars, offs, k)
    store = ["-im"] if mode == "-im" else ["-tmp", str(tmp_path)]
    common = ["-t", "4", "-t2", "4", "-k", str(k), "-p", str(p), "-skm", "128", "-rb", "5000"] + store
    cpu = run(binary, common + ["-cpuonly", "-ngpu", "0", "-read", str(fa)], tmp_path)
    gpu = run(binary, common + ["-ngpu", "1", "-tgpu", "4", "-gpuonly", "-read", str(fa)], tmp_path)
    assert cpu == want
    assert gpu == want


RGKC = os.path.join(ROOT, "rapidgkc_amd", "host", "rgkc")


def read_table(path):
    """Table file written by rgkc -o: header, per-partition row counts, keys, counts."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"RGKCTAB1"
    k, W, parts, cmin, cmax, _ = np.frombuffer(raw[8:32], dtype=np.uint32)
    total = int(np.frombuffer(raw[32:40], dtype=np.uint64)[0])
    rows = np.frombuffer(raw[40:40 + 8 * parts], dtype=np.uint64)
    o = 40 + 8 * int(parts)
    kd = np.uint64 if W == 8 else np.dtype([("lo", "<u8"), ("hi", "<u8")])
    keys = np.frombuffer(raw[o:o + total * int(W)], dtype=kd)
    cnts = np.frombuffer(raw[o + total * int(W):o + total * int(W) + 4 * total], dtype=np.uint32)
    return int(k), rows, keys, cnts


@pytest.mark.gpu
@pytest.mark.parametrize("k,p,fmt,store", [(31, 11, "fa", "-im"), (25, 9, "fq.gz", "disk"), (55, 13, "fa", "-im"),
                                           (31, 11, "fq.gz", "-im-host")])
def test_rgkc_cli_table(tmp_path, k, p, fmt, store):
    """The standalone CLI: loader thread -> GPU phase 1 -> partition stores -> GPU count -> table
    file; rows per partition equal the oracle's. -im keeps the store on the device (batches appended
    by the batched path, rgkc_stream_add); -im -ds 0 and disk mode keep host stores (memory or
    <tmp>/<id>.skm in the reference format) that phase 2 reads back group by group."""
    import gzip
    if not os.path.exists(RGKC):
        pytest.skip("rgkc not built")
    parts, cmin, cmax = 64, 2, 500
    g = synth.genome(23, 120_000, gc=0.46)
    chars, offs = synth.short_reads(g, 24, 25_000, 150, 0.01)
    path = tmp_path / f"reads.{fmt}"
    if fmt == "fa":
        write_fasta(path, chars, offs)
    else:
        with gzip.open(path, "wt") as f:
            for i in range(len(offs) - 1):
                s = chars[offs[i]:offs[i + 1]].tobytes().decode()
                f.write(f"@r{i}\n{s}\n+\n{'I' * len(s)}\n")
    out = tmp_path / "table.bin"
    st_args = {"-im": ["-im"], "-im-host": ["-im", "-ds", "0"]}.get(store, ["-tmp", str(tmp_path)])
    args = ["-t", "4", "-k", str(k), "-p", str(p), "-skm", str(parts), "-ci", str(cmin), "-cx", str(cmax),
            "-o", str(out), "-gb", "1000000"] + st_args + ["-read", str(path)]
    print(f"[rgkc] {' '.join(args)}", flush=True)
    r = subprocess.run([RGKC] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    _, _, want_distinct = oa.bruteforce(chars, offs, k)
    m = re.search(r"Total number of distinct kmers: (\d+)", r.stderr)
    assert m and int(m.group(1)) == want_distinct
    assert ("(device store)" in r.stderr) == (store == "-im"), r.stderr[-2000:]
    kk, rows, keys, cnts = read_table(out)
    st, opo, _, _ = oa.store(chars, offs, k, p, parts)
    ok_, oc_, oro = oa.count_store(st, opo, k, cmin, cmax)
    assert np.array_equal(rows, np.diff(oro))
    ro = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
    for i in range(parts):
        a = oa.sort_table(keys[ro[i]:ro[i + 1]], cnts[ro[i]:ro[i + 1]])
        assert np.array_equal(a[0], ok_[oro[i]:oro[i + 1]]) and np.array_equal(a[1], oc_[oro[i]:oro[i + 1]])


def test_rgkc_cli_arguments(tmp_path):
    """Argument handling needs no device: usage without -read, refusals of -cpuonly / -hpc 1,
    unknown flags, like the reference's ArgParser (utilities.hpp:222-265)."""
    if not os.path.exists(RGKC):
        pytest.skip("rgkc not built")
    r = subprocess.run([RGKC, "-k", "31"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([RGKC, "-cpuonly", "-read", "x.fa"], capture_output=True, text=True)
    assert r.returncode == 1 and "GPU only" in r.stderr
    r = subprocess.run([RGKC, "-hpc", "1", "-read", "x.fa"], capture_output=True, text=True)
    assert r.returncode == 1 and "homopolymer" in r.stderr
    r = subprocess.run([RGKC, "-bogus", "1", "-read", "x.fa"], capture_output=True, text=True)
    assert r.returncode == 1 and "Wrong argument: -bogus" in r.stderr
