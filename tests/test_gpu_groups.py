"""Shared-table layouts (RS_TILE_GROUP4 / RS_TILE_GROUP8): the 4 or 8
wavefronts of a workgroup sieve one prime with one baby table in LDS.

Bar: bit-exact against the oracle and, on tiny primes, the exhaustive
search; both arithmetic paths (double precision below 2^50, 64-bit integers
above), every tile, orders of b^Q below m (repeated baby values, overflowing
buckets), and the batched and serving front ends.
"""
import numpy as np
import pytest

from test_gpu_parity import is_prime, run_oracle, tiny_job

pytestmark = pytest.mark.gpu


def sieve(rs, plan, m, tile, low, offs):
    sv = rs.Sieve(plan, m=m, tile=tile, max_batch=1 << 16)
    assert sv.tile == tile and sv.kernel.startswith("rs_sieve_group")
    sv.submit(low, offs)
    got = sv.collect()
    sv.close()
    return got


def test_group_tiny_primes_vs_exhaustive(rs, orc, gpu):
    """Small p: orders of g below m, p | k, p | b; each group size and tile."""
    plan = tiny_job(rs, 6, {5: list(range(100, 700, 3)), 11: list(range(101, 650, 2)),
                            12: [100, 150, 200, 640], 35: list(range(300, 700))})
    primes = [q for q in range(5, 40000, 2) if is_prime(q)]
    offs = np.array([q - 3 for q in primes], np.uint32)
    rk, rr, _ = plan.rows()
    for m, tile in ((512, 4 | rs.TILE_GROUP8), (512, 1 | rs.TILE_GROUP4),
                    (1024, 2 | rs.TILE_GROUP8), (768, 4 | rs.TILE_GROUP4)):
        job = orc.Job(plan.base, plan.Q, m, plan.minN, plan.maxN, rk, rr, plan.c)
        brute = []
        for q in primes:
            brute += orc.brute_prime(job, q)
        assert sieve(rs, plan, m, tile, 3, offs) == brute, (m, tile)


@pytest.mark.parametrize("c", [-1, 1])
def test_group_windows_both_paths(rs, orc, gpu, c):
    """A wide job (60 k, n over 2e5) on windows below and above 2^50 and
    near 2^62, every tile and both group sizes, against the oracle."""
    plan = rs.Plan(rs.Abcd.synth(2, 60, 3, 10**9, 1000000, 1200000, 10**4, seed=41, c=c))
    cases = [(10**9 + 1, 1 << 18), ((1 << 50) - (1 << 17), 1 << 17),
             ((1 << 56) + 1, 1 << 16), ((1 << 62) - (1 << 16), 1 << 16)]
    total = 0
    for i, (low, width) in enumerate(cases):
        offs = rs.PrimeGen(low + width, threads=16).range(low, low + width)
        for m in (2048, 4096):
            want = run_oracle(orc, plan, m, low, offs)
            total += len(want)
            tiles = (4 | rs.TILE_GROUP8, 2 | rs.TILE_GROUP4) if (i + m // 2048) % 2 else (
                1 | rs.TILE_GROUP8, 4 | rs.TILE_GROUP4)
            for tile in tiles:
                assert sieve(rs, plan, m, tile, low, offs) == want, (low, m, tile)
    assert total > 0


def test_group_edge_shape_near_2_62(rs, orc, gpu):
    """The many-k edge shape (base 2, 500 odd k in [3, 1e9], n in [1e6, 2e6]
    presieved by primes < 1e6) with tables of 4096 and 8192 shared by 8
    wavefronts, p just below 2^62, against the oracle; device-resident
    submits and the device prime generator give the same."""
    plan = rs.Plan(rs.Abcd.synth(2, 500, 3, 10**9, 1000000, 2000000, 10**6, seed=3))
    low = (1 << 62) - (1 << 17)
    offs = rs.PrimeGen(1 << 62, threads=16).range(low, 1 << 62)
    want = run_oracle(orc, plan, 4096, low, offs)
    for m in (4096, 8192):
        assert sieve(rs, plan, m, 4 | rs.TILE_GROUP8, low, offs) == want, m
    sv = rs.Sieve(plan, m=4096, tile=4 | rs.TILE_GROUP8)
    sv.submit_range(low, 1 << 62)
    assert sv.collect() == want
    assert sv.timing()["primes"] == len(offs)
    sv.close()


def test_group_job_in_batch_and_server(rs, orc, gpu):
    """A shared-table job next to per-wavefront jobs in one batch (its own
    launch in the graph) and in the server: each job equals the oracle."""
    wide = rs.Plan(rs.Abcd.synth(2, 60, 3, 10**9, 1000000, 1200000, 10**4, seed=41))
    small = rs.Plan(rs.Abcd.synth(745, 6, 3, 10**6, 20000, 30000, 1000, seed=7))
    low = 10**9 + 1
    offs = rs.PrimeGen(low + (1 << 18)).range(low, low + (1 << 18))
    want = [run_oracle(orc, wide, 2048, low, offs), run_oracle(orc, small, 256, low, offs)]
    bat = rs.Batch([wide, small], [2048, 256], tiles=[4 | rs.TILE_GROUP8, 2])
    dev = rs.DeviceBuffer(0, offs)
    bat.add_chunk(0, low, dev)
    bat.add_chunk(1, low, dev)
    for _ in range(2):  # capture, then replay
        bat.run()
        res = bat.collect()
        for j in range(2):
            assert [(p, k, n) for (jj, p, k, n) in res if jj == j] == want[j], j
    bat.close()
    dev.free()
    srv = rs.Server(hbm_budget=1 << 33, max_active=4, chunk_width=1 << 17, lds_cap=65536)
    ids = [srv.submit(wide, low, low + (1 << 18), m=2048, tile=4 | rs.TILE_GROUP8),
           srv.submit(small, low, low + (1 << 18), m=256)]
    for _ in range(50):
        srv.step()
        if all(srv.poll(j)[0] == srv.DONE for j in ids):
            break
    assert [srv.results(j) for j in ids] == want
    srv.close()


def test_group_layout_rules(rs, gpu):
    """m must be a multiple of 64 x the group size; group layouts do not take
    row-number lists."""
    plan = rs.Plan(rs.Abcd.synth(2, 60, 3, 10**9, 1000000, 1200000, 10**4, seed=41))
    for m, tile in ((2048 + 256, 4 | rs.TILE_GROUP8), (2048 + 64, 4 | rs.TILE_GROUP4),
                    (2048, 4 | rs.TILE_GROUP4 | rs.TILE_GROUP8),
                    (2048, 4 | rs.TILE_GROUP8 | rs.TILE_ROW_INDEX)):
        with pytest.raises(rs.RsError) as e:
            rs.Sieve(plan, m=m, tile=tile)
        assert e.value.status == rs.RS_ERR_ARG, (m, tile)
