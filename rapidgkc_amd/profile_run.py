"""One pass of the counting path on a seeded workload, for profilers (rocprofv3 counter runs).

    python3 -m rapidgkc_amd.profile_run --workload genome_100m_30x --scale 1.0 --copy

runs phase 1 and the count stage once on the resident batch, then (with --copy) one 1 GiB
device-to-device copy: HIP's blit copy kernel of known size, whose FETCH_SIZE / WRITE_SIZE readings
calibrate the same counters for the count kernel.
"""
import argparse
import ctypes
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from rapidgkc_amd import native, synth  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="genome_100m_30x")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--copy", action="store_true", help="also run one 1 GiB device-to-device copy")
    args = ap.parse_args()
    wl = synth.WORKLOADS[args.workload]
    chars, offs = wl.make_reads(args.scale, threads=min(16, os.cpu_count() or 8))
    c = native.Counter(0)
    c.set_reads(chars, offs)
    c.run(wl.k, wl.p, wl.parts, wl.min_count, wl.max_count)
    st = c.stats()
    print(f"{wl.name} scale {args.scale}: k-mers {st['n_kmers']} rows {st['rows']} expand {st['expand_ms']:.2f} "
          f"count {st['count_ms']:.2f} ms "
          f"phase 1 {st['p1_count_ms'] + st['p1_emit_ms']:.2f} ms", flush=True)
    c.close()
    if args.copy:
        hip = ctypes.CDLL("libamdhip64.so")
        n = 1 << 30
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
            raise RuntimeError("hipMalloc failed")
        if hip.hipMemset(a, 1, ctypes.c_size_t(n)) or hip.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) or hip.hipDeviceSynchronize():  # device to device
            raise RuntimeError("hipMemcpy failed")
        hip.hipFree(a)
        hip.hipFree(b)
        print(f"copied {n} bytes device to device", flush=True)


if __name__ == "__main__":
    main()
