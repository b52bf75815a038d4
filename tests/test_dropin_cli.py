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
