"""GPU, large inputs: properties that need no oracle at this size.

Every tuning mode of the path must give the same store (per-partition bytes, records) and the same
table (distinct keys, rows per partition): superkmer lists on / off / overflowing (tiles rescanned),
heavy sub-buckets expanded into keys or read as records. The workloads are the k=55 tandem-repeat
set (128-bit keys, heavy repeats, many sweep splits) and the mixed short + long read set, at half
size.
"""
import numpy as np
import pytest

from rapidgkc_amd import synth

pytestmark = pytest.mark.gpu

CASES = [("tandem_200m_k55", 0.5), ("mixed_50m", 0.5)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_modes_agree_at_size(case):
    from rapidgkc_amd import native
    name, scale = case
    wl = synth.WORKLOADS[name]
    print(f"[gpu] full-size modes: {name} scale {scale}", flush=True)
    chars, offs = wl.make_reads(scale, threads=16)
    c = native.Counter(0)
    c.set_reads(chars, offs)
    ref = None
    for lists, heavy_off in ((0, 0), (0, 1), (0xFFFFFFFF, 0), (64, 0)):
        c.set_p1_lists(lists)
        c.set_heavy_expand(heavy_off)
        po, sc, kc = c.superkmers(wl.k, wl.p, wl.parts)
        rows, distinct = c.count(wl.min_count, wl.max_count)
        st = c.stats()
        if lists == 64:
            assert st["p1_rescans"] > 0
        if heavy_off == 0:
            assert st["heavy_groups"] > 0
        sig = (int(distinct), rows.tobytes(), np.asarray(sc).tobytes(), np.asarray(kc).tobytes(), np.asarray(po).tobytes())
        print(f"[gpu]   lists {lists:#x} heavy_off {heavy_off}: distinct {distinct}", flush=True)
        ref = ref or sig
        assert sig == ref, f"lists {lists:#x} heavy_off {heavy_off}"
    assert int(rows.sum()) == int(ref[0]) or wl.min_count > 1
    c.close()
