"""Builds the native libraries in-tree: the gfx950 HIP library (librgkc.so), the synthetic data
library (librgkc_synth.so) and the host drop-in layer. Run as ``python -m rapidgkc_amd.build``."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = ["k_util.hip", "k_phase1.hip", "k_count.hip", "k_sort.hip", "k_index.hip", "k_mlp.hip", "rgkc_api.hip"]
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    out = os.path.join(HERE, "librgkc.so")
    _run([hipcc, *HIPFLAGS, "-shared", "-o", out + ".tmp", *srcs])
    os.replace(out + ".tmp", out)
    return out


def build_synth():
    out = os.path.join(HERE, "librgkc_synth.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", out + ".tmp",
          os.path.join(CSRC, "synth.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_host():
    """The C++ drop-in layer (reference-shaped API + CLI) over the C ABI."""
    mk = os.path.join(HERE, "host", "Makefile")
    if os.path.exists(mk):
        _run(["make", "-s", "-C", os.path.join(HERE, "host"), "-j8"])
    # the reference's own host program linked with the drop-in layer, where the reference tree is
    # present (this container); the built binaries travel with the repository like the libraries
    ref = os.environ.get("RGKC_REFERENCE", "/root/reference")
    if os.path.exists(os.path.join(ref, "gkc_main.cpp")):
        _run(["make", "-s", "-f", os.path.join(HERE, "host", "Makefile.dropin"), f"REF={ref}", "-j8"])


def build_all():
    build_synth()
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
    print("built", file=sys.stderr)
