"""The sieve kernel's own row filter and the batched / serving distributions
against the oracle.

Row filter: the set-up phase folds the Legendre test of the reference
(kernel_v0.24.6.cu:491-521) into per-row pass masks from per-k Jacobi
constants, quadratic-character tables, (2/p)^z, reciprocity and, for
k*b^n+1, (-1/p).  Each mask bit is compared with Euler's criterion on
c*k*b^n mod p, n = r + Q, row by row, over k on every path (no table, table with p
above and below the odd part, even k of odd and even valuation, k = 1, a
table budget that runs out), bases of every 2-adic shape, p in every class
mod 8 on both sides of 2^50, p dividing k or b, and both forms.

Where Q is a multiple of 3 and p = 1 (mod 3) lies below 2^50, the masks
also carry the cubic-residue test (c k b^n a cube mod p), checked against
Python's pow.

Distributions: the 64-job batched set and the 24-job serving set drawn from
the heterogeneous job distribution, each job against the oracle on an
oracle-sized window.
"""
import random

import numpy as np
import pytest


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


def odd_part(k):
    while k % 2 == 0:
        k //= 2
    return k


# k on every path of the set-up's Jacobi code: odd part 1; small and large
# odd parts with a character table (odd part <= 2^28); odd parts above 2^28,
# factored: primes with a table of their own (3, 100000007), a prime above
# 2^28 (binary Jacobi loop), a square factor (3^2 in 2^30 - 1 and
# 9 * 100000007), a cube (3^3); 2-adic valuations 1, 2, 3, 5, 30; the
# reference's own k values
K_PATHS = [1, 3, 5, 7, 2, 4, 8, 1 << 30, 8 * 3**5 * 7, 1000003, 2 * 1000003, 4 * 1000003,
           8 * 1000003, 32 * 12345679, 4 * 268435399, 268435459, 999999937, 2 * 999999937,
           2147483647, 2147483646, 18300, 29586, 3 * 100000007, 9 * 100000007,
           2 * 3 * 100000007, 3 * 268435459, 3**3 * 5 * 7 * 11 * 13 * 17 * 19 * 23,
           (1 << 30) - 1]
# nine odd parts just below 2^28: their tables exceed the per-job table
# budget, so the last ones fall back to the loop
K_BUDGET = [(1 << 28) - 2 * j - 1 for j in range(9)]
# three times seventeen primes just below 2^28: their factor tables exceed
# that budget, so the largest primes go through the loop
Q_BUDGET = [q for q in range((1 << 28) - 1, (1 << 28) - 4000, -2) if is_prime(q)][:17]


def test_k_cases_cover_the_paths():
    """The k and prime samples reach every branch they are meant to."""
    assert any(odd_part(k) == 1 for k in K_PATHS)
    assert any(3 <= odd_part(k) <= 1 << 28 for k in K_PATHS)
    assert any(odd_part(k) > 1 << 28 for k in K_PATHS)
    vals = {(k & -k).bit_length() - 1 for k in K_PATHS}
    assert {0, 1, 2, 3, 5, 30} <= vals
    # words of the character tables of K_BUDGET exceed 2^26, of the factor
    # tables of 3 Q_BUDGET 2^27
    assert sum((odd_part(k) + 31) // 32 for k in K_BUDGET) > 1 << 26
    assert len(Q_BUDGET) == 17 and sum((q + 31) // 32 for q in Q_BUDGET) > 1 << 27
    assert all(3 * q < 1 << 31 for q in Q_BUDGET)
    # factored odd parts: with a prime of a table, a prime above 2^28, a
    # square, a cube
    big = [odd_part(k) for k in K_PATHS if odd_part(k) > 1 << 28]
    assert any(n % 100000007 == 0 for n in big) and any(n % 268435459 == 0 for n in big)
    assert any(n % 9 == 0 and n % 27 for n in big) and any(n % 27 == 0 for n in big)
    ps = sample_primes()
    assert {p % 8 for p in ps if p > 1 << 50} == {1, 3, 5, 7}
    assert any(p < 1 << 50 < q for p, q in zip(ps, ps[1:]))
    assert 1000003 in ps and 5 in ps and 149 in ps  # p | k, p | 745
    assert 100000007 in ps and 268435459 in ps  # the prime factor of k itself


def sample_primes():
    rng = random.Random(99)
    ps = [q for q in range(3, 3000, 2) if is_prime(q)]
    # around the odd parts: p = n (p | k), just below and above
    for n in (1000003, 12345679, 100000007, 268435399, 268435459, 999999937):
        ps += [q for q in range(n - 200, n + 200) if q % 2 and is_prime(q)]
    ps.append(1000003)
    for bits in (20, 29, 31, 33, 40, 49, 50, 51, 56, 62, 63):
        got = {1: 0, 3: 0, 5: 0, 7: 0}
        while min(got.values()) < 10:
            q = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
            if q < 1 << 63 and got[q % 8] < 10 and is_prime(q):
                got[q % 8] += 1
                ps.append(q)
    # both sides of 2^50 (the set-up's double-precision / integer switch)
    ps += [q for q in range((1 << 50) - 300, (1 << 50) + 300) if q % 2 and is_prime(q)]
    return sorted(set(ps))


def cubic_applies(plan, p):
    """The set-up's cubic-residue test: Q a multiple of 3 (so n = r mod 3),
    at most 1024 distinct k, and p = 1 (mod 3) below 2^50."""
    nk = len(set(plan.rows()[0].tolist()))
    return plan.Q % 3 == 0 and nk <= 1024 and p % 3 == 1 and p < 1 << 50


def expected_masks(plan, primes, cubic=True):
    rk, rr, _ = plan.rows()
    nw = (plan.nrows + 63) // 64
    out = np.zeros((len(primes), nw), np.uint64)
    # a row (k, r) can hold a factor of p only if c k b^n is a square mod p
    # for its n = Q x + r >= 1 (c = 1 for k*b^n-1, -1 for k*b^n+1); b^Q is a
    # square, so n = r + Q decides (and p | b fails every row); where the
    # cubic test applies, c k b^n must be a cube too (b^Q is one, 3 | Q)
    c = 1 if plan.c == -1 else -1
    for i, p in enumerate(primes):
        br = [pow(plan.base, r + plan.Q, p) for r in range(plan.Q)]
        cub = cubic and cubic_applies(plan, p)
        words = [0] * nw
        for j, (k, r) in enumerate(zip(rk.tolist(), rr.tolist())):
            x = c * k * br[r] % p
            if x and pow(x, (p - 1) // 2, p) == 1 and (not cub or pow(x, (p - 1) // 3, p) == 1):
                words[j // 64] |= 1 << (j % 64)
        out[i] = words
    return out


def plan_of(rs, base, ks, c):
    Q = 18
    lists = {k: list(range(200000, 200000 + 2 * Q)) for k in ks}
    a = rs.Abcd.from_lists(base, lists)
    a.set_form(c)
    plan = rs.Plan(a, Q)
    assert plan.Q == Q and plan.nrows == Q * len(ks)
    return plan


@pytest.mark.gpu
@pytest.mark.parametrize("c", [-1, 1])
@pytest.mark.parametrize("base", [745, 2, 1030, 12])
def test_row_masks_vs_euler(rs, gpu, base, c):
    """Every row bit of the set-up's masks = Euler's criterion."""
    primes = sample_primes()
    plan = plan_of(rs, base, K_PATHS, c)
    got = rs.dev_row_masks(plan, primes)
    want = expected_masks(plan, primes)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, [primes[i] for i in bad[:10]]
    # the quadratic test passes about half of the rows of a prime not
    # dividing k or b, and the cubic test a third of those where it applies
    quad = expected_masks(plan, primes, cubic=False)
    frac = np.unpackbits(quad.view(np.uint8)).sum() / (len(primes) * plan.nrows)
    assert 0.3 < frac < 0.6
    cub = [i for i, p in enumerate(primes) if cubic_applies(plan, p) and p > 1 << 20]
    assert len(cub) > 100
    kept = (np.unpackbits(want[cub].view(np.uint8)).sum() /
            np.unpackbits(quad[cub].view(np.uint8)).sum())
    assert 0.25 < kept < 0.42


@pytest.mark.gpu
@pytest.mark.parametrize("c", [-1, 1])
def test_row_masks_table_budget_fallback(rs, gpu, c):
    """Character tables for the first k until the job's budget runs out, the
    loop for the rest: the masks equal Euler's criterion either way."""
    primes = [q for q in sample_primes() if q > 1 << 28][:200] + [3, 5, 7, 11, 13, 1000003]
    plan = plan_of(rs, 745, K_BUDGET, c)
    assert np.array_equal(rs.dev_row_masks(plan, primes), expected_masks(plan, primes))


@pytest.mark.gpu
@pytest.mark.parametrize("c", [-1, 1])
def test_row_masks_factor_table_budget(rs, gpu, c):
    """Factor tables for the smallest primes of the factored odd parts until
    their budget runs out, the loop for the rest: the masks equal Euler's
    criterion either way, for p on both sides of the largest factor."""
    primes = ([q for q in sample_primes() if q > 1 << 28][:200] + [3, 5, 7, 11, 13]
              + Q_BUDGET[::4] + [q for q in range(Q_BUDGET[-1] - 100, Q_BUDGET[0] + 100)
                                 if q % 2 and is_prime(q)][:40])
    plan = plan_of(rs, 745, [3 * q for q in Q_BUDGET], c)
    assert np.array_equal(rs.dev_row_masks(plan, sorted(set(primes))),
                          expected_masks(plan, sorted(set(primes))))


def run_oracle(orc, plan, m, low, offs):
    rk, rr, _ = plan.rows()
    job = orc.Job(plan.base, plan.Q, m, plan.minN, plan.maxN, rk, rr, plan.c)
    hits, _ = orc.bsgs_many(job, low, offs, threads=16)
    return hits


WINDOW_LOW = 10**9 + 1
WINDOW = 1 << 18


@pytest.mark.gpu
def test_batched_distribution_vs_oracle(rs, orc, gpu):
    """The 64-job batched set (base U[3,1030], k-count log-U[1,200], n-width
    log-U[1e4,5e5], presieve < 1e5; seed 4) in one batch, table sizes from
    the cost model and tiles cycling through every layout: each job's hits
    equal the oracle on a 2^18-wide window at 1e9, on the first run and on
    the graph replay."""
    import workloads

    plans = [workloads.job_plan(sp) for sp in workloads.job_specs(64, seed=4)]
    offs = rs.PrimeGen(WINDOW_LOW + WINDOW).range(WINDOW_LOW, WINDOW_LOW + WINDOW)
    ms = [pl.auto_m() for pl in plans]
    tiles = [(1, 2, 4)[i % 3] | (rs.TILE_ROW_INDEX if i % 2 and pl.nrows <= 512 else 0)
             for i, pl in enumerate(plans)]
    bat = rs.Batch(plans, ms, hit_capacity=1 << 22, tiles=tiles)
    dev = rs.DeviceBuffer(0, offs)
    for j in range(len(plans)):
        bat.add_chunk(j, WINDOW_LOW, dev)
    bat.run()
    first = bat.collect()
    bat.run()
    assert bat.collect() == first
    bat.close()
    dev.free()
    total = 0
    for j, pl in enumerate(plans):
        mine = [(p, k, n) for (jj, p, k, n) in first if jj == j]
        assert mine == run_oracle(orc, pl, ms[j], WINDOW_LOW, offs), j
        total += len(mine)
    assert total > 100


@pytest.mark.gpu
def test_serving_distribution_vs_oracle(rs, orc, gpu):
    """The 24-job serving set (seed 6) through the continuous-batching
    server, all arriving at once with room for 8 at a time, each over two
    windows of 2^17: every job's hits equal the oracle."""
    import workloads

    plans = [workloads.job_plan(sp) for sp in workloads.job_specs(24, seed=6)]
    half = WINDOW // 2
    srv = rs.Server(hbm_budget=1 << 34, max_active=8, chunk_width=half, lds_cap=65536,
                    hit_capacity=1 << 22)
    ids = [srv.submit(pl, WINDOW_LOW, WINDOW_LOW + WINDOW) for pl in plans]
    steps = 0
    while any(srv.poll(j)[0] in (srv.QUEUED, srv.RUNNING) for j in ids):
        srv.step()
        steps += 1
        assert steps < 200
    offs = rs.PrimeGen(WINDOW_LOW + WINDOW).range(WINDOW_LOW, WINDOW_LOW + WINDOW)
    total = 0
    for j, pl in zip(ids, plans):
        assert srv.poll(j)[0] == srv.DONE
        got = srv.results(j)
        assert got == run_oracle(orc, pl, pl.auto_m(), WINDOW_LOW, offs), j
        total += len(got)
    srv.close()
    assert total > 50
