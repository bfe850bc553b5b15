"""The cubic-residue row test (RS_TILE_CUBIC) through the whole sieve.

For p = 1 (mod 3) below 2^50 and Q a multiple of 3, the set-up drops the
rows (k, r) whose k b^r is not a cube mod p.  The factor set must not
change: every case here runs the sieve with the flag against the exhaustive
search or the oracle (and the sieve without it), bit-exact.  The masks
themselves are checked row by row against Python's pow in
test_gpu_filters.py.
"""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def is_prime(n):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def oracle_hits(orc, plan, m, low, offs):
    rk, rr, _ = plan.rows()
    job = orc.Job(plan.base, plan.Q, m, plan.minN, plan.maxN, rk, rr, plan.c)
    hits, _ = orc.bsgs_many(job, low, offs, threads=16)
    return hits


def sieve(rs, plan, m, tile, low, offs):
    sv = rs.Sieve(plan, m=m, tile=tile, max_batch=1 << 16)
    applied = bool(sv.tile & rs.TILE_CUBIC)
    sv.submit(low, offs)
    hits = sv.collect()
    sv.close()
    return hits, applied


@pytest.mark.parametrize("c", [-1, 1])
def test_cubic_tiny_primes_vs_exhaustive(rs, orc, gpu, c):
    """Small primes (p = 7, 13, ... with orders of b^Q below m, p | k, p | b,
    p <= k) with Q = 18 and 6: the sieve with the cubic test equals the
    exhaustive search."""
    lists = {5: list(range(100, 700, 3)), 11: list(range(101, 650, 2)),
             12: [100, 150, 200, 640], 35: list(range(300, 700)), 91: list(range(120, 500, 7))}
    primes = [q for q in range(5, 30000, 2) if is_prime(q)]
    low = 3
    offs = np.array([q - low for q in primes], np.uint32)
    for Q in (18, 6):
        a = rs.Abcd.from_lists(6, lists)
        a.set_form(c)
        plan = rs.Plan(a, Q)
        rk, rr, _ = plan.rows()
        job = orc.Job(plan.base, plan.Q, 1024, plan.minN, plan.maxN, rk, rr, plan.c)
        brute = []
        for q in primes:
            brute += orc.brute_prime(job, q)
        assert len(brute) > 100
        for tile in (2, 4 | rs.TILE_ROW_INDEX):
            got, applied = sieve(rs, plan, 1024, tile | rs.TILE_CUBIC, low, offs)
            assert applied, (Q, tile)
            assert got == brute, (Q, tile, c)


def test_cubic_r745_window(rs, orc, gpu):
    """The R745-shaped job on a 2^22 window at 1e12 with the cubic test, in
    both per-wavefront builds: the oracle's factors, and the same as without
    the test."""
    plan = rs.Plan(rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2))
    assert plan.Q % 3 == 0
    low = 10**12 + 1
    offs = rs.PrimeGen(low + 2**22).range(low, low + 2**22)
    want = oracle_hits(orc, plan, 512, low, offs)
    plain, applied = sieve(rs, plan, 512, 4, low, offs)
    assert not applied and plain == want
    for tile in (4, 4 | rs.TILE_ROW_INDEX):
        got, applied = sieve(rs, plan, 512, tile | rs.TILE_CUBIC, low, offs)
        assert applied and got == want, tile


def test_cubic_random_jobs_vs_oracle(rs, orc, gpu):
    """Seeded random jobs with Q a multiple of 3 (6, 12 or 18), both forms,
    every tile, with the cubic test, over 2^16-wide windows of p from 2^20
    to just above 2^50 (where the test stops applying): the oracle's
    factors."""
    rng = random.Random(3)
    lows = [(1 << 20) + 1, (1 << 20) + 1, 10**9 + 7, 10**12 + 39, (1 << 49) + 1,
            (1 << 50) - (1 << 15) + 1]
    offs_of = {}
    total = applied_n = 0
    for case in range(16):
        base = rng.randrange(3, 1031)
        nk = rng.choice([1, 3, 8, 25])
        kmax = rng.choice([10**3, 10**6, 10**9])
        nmin = rng.randrange(1000, 200000)
        width = rng.choice([600, 3000, 40000])
        c = rng.choice([-1, 1])
        Q = rng.choice([6, 12, 18])
        m = 64 * rng.randrange(8, 33)
        tile = rng.choice([1, 2, 4]) | (rs.TILE_ROW_INDEX if rng.random() < 0.3 else 0)
        low = rng.choice(lows)
        if low not in offs_of:
            offs_of[low] = rs.PrimeGen(low + (1 << 16), threads=16).range(low, low + (1 << 16))
        offs = offs_of[low]
        plan = rs.Plan(rs.Abcd.synth(base, nk, 3, kmax, nmin, nmin + width, 1000, seed=case,
                                     c=c), Q)
        got, applied = sieve(rs, plan, m, tile | rs.TILE_CUBIC, low, offs)
        applied_n += applied
        want = oracle_hits(orc, plan, m, low, offs)
        assert got == want, (case, base, nk, kmax, nmin, width, c, Q, m, tile, low)
        total += len(got)
    assert total > 100 and applied_n >= 12


def test_cubic_in_a_batch_equals_alone(rs, orc, gpu):
    """Jobs with and without the cubic test batched together: each job's
    hits equal its run alone and the oracle."""
    plans = [rs.Plan(rs.Abcd.synth(b, nk, 3, 10**6, 100000, 100000 + w, 10**4, seed=s), Q)
             for b, nk, w, s, Q in ((31, 4, 30000, 1, 18), (1030, 9, 8000, 2, 12),
                                    (12, 2, 50000, 3, 6))]
    low = 10**9 + 7
    offs = rs.PrimeGen(low + 2**20).range(low, low + 2**20)
    ms = [640, 512, 704]
    tiles = [4 | rs.TILE_CUBIC, 2, 2 | rs.TILE_CUBIC]
    bat = rs.Batch(plans, ms, tiles=tiles)
    dev = rs.DeviceBuffer(0, offs)
    for j in range(len(plans)):
        bat.add_chunk(j, low, dev)
    bat.run()
    res = bat.collect()
    bat.close()
    for j, plan in enumerate(plans):
        mine = [(p, k, n) for jj, p, k, n in res if jj == j]
        alone, _ = sieve(rs, plan, ms[j], tiles[j], low, offs)
        assert mine == alone == oracle_hits(orc, plan, ms[j], low, offs), j
    dev.free()
