"""rs_sieve_submit_device_noalloc (the entry point that allocates nothing:
record scratch and work counter in the library's static device array, hits
in the caller's buffer) against rs_sieve_submit_device on the same offsets:
bit-exact, on two calls in a row, the second with a smaller input, for a
per-wavefront layout and a shared-table layout."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def noalloc_hits(rs, sv, low, dev, count, hits_buf, nhits_buf):
    sv.submit_device_noalloc(low, dev, hits_buf, nhits_buf, count=count)
    sv.sync()
    n = int(nhits_buf.download()[0])
    assert n <= hits_buf.count // 6, "hit buffer overflowed"
    raw = hits_buf.download(np.uint64).reshape(-1, 3)[:n]  # (p, k | n << 32, job | pad)
    out = sorted((int(r[0]), int(r[1] & 0xffffffff), int(r[1] >> 32)) for r in raw)
    assert all(int(r[2] & 0xffffffff) == 0 for r in raw)
    return out


@pytest.mark.parametrize("layout", ["wave", "group"])
def test_noalloc_equals_submit_device(rs, gpu, layout):
    if layout == "wave":
        plan = rs.Plan(rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2))
        m, tile, low, width = 512, 4, 10**9 + 1, 1 << 24
    else:
        plan = rs.Plan(rs.Abcd.synth(2, 60, 3, 10**9, 1000000, 1200000, 10**4, seed=41))
        m, tile, low, width = 2048, 4 | rs.TILE_GROUP8, 10**9 + 1, 1 << 22
    offs = rs.PrimeGen(low + width).range(low, low + width)
    dev = rs.DeviceBuffer(0, offs)
    sv = rs.Sieve(plan, m=m, tile=tile)
    hits_buf = rs.DeviceBuffer.empty(0, 6 * (1 << 16))
    nhits_buf = rs.DeviceBuffer.empty(0, 1)
    for count in (len(offs), len(offs) // 3):
        sv.submit_device(low, dev, count=count)
        want = sv.collect()
        got = noalloc_hits(rs, sv, low, dev, count, hits_buf, nhits_buf)
        assert got == want, (layout, count)
        assert len(want) > 0
    sv.close()
    for b in (dev, hits_buf, nhits_buf):
        b.free()
