"""Parity of the HIP sieve (through the C ABI) with the CPU oracle.

Bar: bit-exact.  Every case compares the sorted (p, k, n) hit list of the
device with the oracle's BSGS restatement and, for tiny inputs, with an
exhaustive search; every device factor is re-verified by modpow.
"""
import os
import random
import subprocess

import numpy as np
import pytest

from conftest import R745, ROOT

pytestmark = pytest.mark.gpu

KNOWN_P = 102297149781479  # divides 18300*745^242670-1 (datless: n not in sr_745.abcd)


def py_montmul(a, b, p):
    return a * b * pow(2, -64, p) % p


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


def run_device(rs, plan, low, offs, m=0, max_batch=1 << 16, device=0):
    sv = rs.Sieve(plan, m=m, max_batch=max_batch, device=device)
    sv.submit(low, offs)
    hits = sv.collect()
    sv.close()
    return hits


def run_oracle(orc, plan, m, low, offs):
    rk, rr, _ = plan.rows()
    job = orc.Job(plan.base, plan.Q, m, plan.minN, plan.maxN, rk, rr, plan.c)
    hits, _ = orc.bsgs_many(job, low, offs, threads=16)
    return hits


# ------------------------------------------------------------ arithmetic


def test_dev_montmul_random_and_edges(rs, gpu):
    rng = random.Random(1)
    a, b, p = [], [], []
    for bits in (3, 20, 31, 32, 33, 40, 50, 61, 62, 63):
        for _ in range(300):
            q = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
            q = max(q, 3)
            if q >= 1 << 63:
                q = (1 << 63) - 25
            a.append(rng.randrange(q))
            b.append(rng.randrange(q))
            p.append(q)
    # special operands: 0, 1, p-1, a >= p (unreduced), p near 2^63
    for q in ((1 << 63) - 25, (1 << 62) + 135, 4294967311, 3, 5):
        for x, y in ((0, 1), (1, 1), (q - 1, q - 1), (1 << 32, q - 1), ((1 << 64) - 1, q - 1)):
            a.append(x % (1 << 64))
            b.append(y % q)
            p.append(q)
    # the reference's lost-carry case: bits 32..63 of a*b all ones
    for _ in range(500):
        q = (1 << 63) - 2 * rng.randrange(1, 1 << 40) - 1
        bb = rng.randrange(2, 1 << 31)
        h = rng.randrange(1, bb)
        aa = ((h << 64) + (1 << 64) - 1) // bb
        assert aa < 1 << 64
        assert ((aa * bb) >> 32) & 0xFFFFFFFF == 0xFFFFFFFF
        a.append(aa)
        b.append(bb)
        p.append(q)
    got = rs.dev_montmul(a, b, p)
    want = [py_montmul(x, y, q) for x, y, q in zip(a, b, p)]
    assert [int(v) for v in got] == want


def test_dev_halfmul(rs, gpu):
    rng = random.Random(2)
    k, y, p = [], [], []
    for bits in (5, 32, 33, 45, 62, 63):
        for _ in range(500):
            q = max(3, rng.getrandbits(bits) | 1 | (1 << (bits - 1)))
            q = min(q, (1 << 63) - 25)
            k.append(rng.getrandbits(32))
            y.append(rng.randrange(q))
            p.append(q)
    for q in ((1 << 63) - 25, 3):
        k += [0, 1, (1 << 32) - 1]
        y += [q - 1, q - 1, q - 1]
        p += [q, q, q]
    got = rs.dev_halfmul(k, y, p)
    want = [kk * yy * pow(2, -32, q) % q for kk, yy, q in zip(k, y, p)]
    assert [int(v) for v in got] == want


def test_dev_jacobi_vs_euler(rs, gpu):
    rng = random.Random(3)
    ps = [3, 5, 7, 11, 13, 101, 65537, 4294967311, 1000000007]
    while len(ps) < 60:
        q = rng.getrandbits(rng.choice((33, 40, 50, 62))) | 1
        if is_prime(q):
            ps.append(q)
    a, p = [], []
    for q in ps:
        for x in [0, 1, 2, 4, 8, 3, 745, 18300, 29586, (1 << 32) - 1] + [
                rng.getrandbits(32) for _ in range(40)]:
            a.append(x)
            p.append(q)
    got = rs.dev_jacobi(a, p)
    for x, q, j in zip(a, p, got):
        e = pow(x, (q - 1) // 2, q)
        want = 0 if x % q == 0 else (1 if e == 1 else -1)
        assert j == want, (x, q)


# ------------------------------------------------------------ BSGS parity


def tiny_job(rs, base, kn):
    return rs.Plan(rs.Abcd.from_lists(base, kn))


def test_single_window_job_two_giant_steps(rs, orc, gpu):
    """A job whose x-span fits one baby window still gets two giant steps
    per row at tile 1 (the second searches past maxN, whose n are dropped):
    it equals the exhaustive search on small primes."""
    plan = tiny_job(rs, 6, {5: list(range(100, 400, 3)), 11: [101, 131, 377], 35: [200, 399]})
    primes = [q for q in range(5, 40000, 2) if is_prime(q)]
    offs = np.array([q - 3 for q in primes], np.uint32)
    rk, rr, _ = plan.rows()
    job = orc.Job(plan.base, plan.Q, 256, plan.minN, plan.maxN, rk, rr, plan.c)
    brute = []
    for q in primes:
        brute += orc.brute_prime(job, q)
    assert len(brute) > 0
    sv = rs.Sieve(plan, m=256, tile=1)
    assert sv.T == 2
    sv.submit(3, offs)
    assert sv.collect() == brute
    sv.close()


def test_tiny_primes_vs_exhaustive(rs, orc, gpu):
    """Small p (orders of b^Q below m, p | k, p | b) against brute force."""
    plan = tiny_job(rs, 6, {5: list(range(100, 700, 3)), 11: list(range(101, 650, 2)),
                            12: [100, 150, 200, 640], 35: list(range(300, 700))})
    primes = [q for q in range(5, 60000, 2) if is_prime(q)]
    low = 3
    offs = np.array([q - low for q in primes], np.uint32)
    rk, rr, _ = plan.rows()
    for m in (64, 192, 1024):
        got = run_device(rs, plan, low, offs, m=m)
        job = orc.Job(plan.base, plan.Q, m, plan.minN, plan.maxN, rk, rr, plan.c)
        brute = []
        for q in primes:
            brute += orc.brute_prime(job, q)
        assert got == brute, f"m={m}"
    # independent Python check on the smallest primes
    small = [h for h in got if h[0] < 400]
    rows = set(zip(rk.tolist(), rr.tolist()))
    want = []
    for q in primes:
        if q >= 400:
            break
        for k, r in sorted(rows):
            for n in range(plan.minN, plan.maxN + 1):
                if n % plan.Q == r and k * pow(plan.base, n, q) % q == 1:
                    want.append((q, k, n))
    assert small == sorted(want)


def test_sierpinski_tiny_primes_vs_exhaustive(rs, orc, gpu):
    """k*b^n+1 on small primes (both residues mod 4, orders below m, p | k):
    device = exhaustive search = oracle BSGS."""
    a = rs.Abcd.from_lists(6, {5: list(range(100, 700, 3)), 11: list(range(101, 650, 2)),
                               12: [100, 150, 200, 640], 35: list(range(300, 700))})
    a.set_form(1)
    plan = rs.Plan(a)
    assert plan.c == 1
    primes = [q for q in range(5, 40000, 2) if is_prime(q)]
    low = 3
    offs = np.array([q - low for q in primes], np.uint32)
    rk, rr, _ = plan.rows()
    for m in (64, 192):
        got = run_device(rs, plan, low, offs, m=m)
        job = orc.Job(plan.base, plan.Q, m, plan.minN, plan.maxN, rk, rr, 1)
        brute = []
        for q in primes:
            brute += orc.brute_prime(job, q)
        assert got == brute, f"m={m}"
        assert got == run_oracle(orc, plan, m, low, offs)
    assert got and all(rs.verify_factor(p, k, 6, n, 1) for p, k, n in got)


def test_sierpinski_window_and_mixed_batch(rs, orc, gpu):
    """A synthetic k*b^n+1 set over a window of p near 1e7 against the
    oracle, and Riesel and Sierpinski jobs batched together = alone."""
    sp = rs.Plan(rs.Abcd.synth(2, 40, 3, 10**6, 100000, 130000, 10**4, seed=31, c=1))
    ri = rs.Plan(rs.Abcd.synth(2, 40, 3, 10**6, 100000, 130000, 10**4, seed=31, c=-1))
    assert sp.c == 1 and ri.c == -1
    low = 10**7 + 1
    offs = rs.PrimeGen(low + 2**20).range(low, low + 2**20)
    want = run_oracle(orc, sp, 512, low, offs)
    got = run_device(rs, sp, low, offs, m=512)
    assert got == want and len(got) > 0
    assert all(rs.verify_factor(p, k, 2, n, 1) for p, k, n in got)
    bat = rs.Batch([sp, ri], [512, 384])
    dev = rs.DeviceBuffer(0, offs)
    bat.add_chunk(0, low, dev)
    bat.add_chunk(1, low, dev)
    bat.run()
    res = bat.collect()
    assert [(p, k, n) for j, p, k, n in res if j == 0] == got
    assert [(p, k, n) for j, p, k, n in res if j == 1] == run_oracle(orc, ri, 384, low, offs)
    bat.close()


def test_cli_sierpinski_file(tmp_path, gpu):
    import pyrsieve as rsmod
    a = rsmod.Abcd.synth(2, 12, 3, 10**5, 2000, 6000, 300, seed=8, c=1)
    path = tmp_path / "s.abcd"
    a.write(path)
    assert "^$a+1 [" in path.read_text().splitlines()[0]
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    out = tmp_path / "f.txt"
    r = subprocess.run([exe, "-i", str(path), "-p", "1000001", "-P", "1100000", "-O", str(out),
                        "-m", "7"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().splitlines()
    assert lines and all(l.endswith("+1") for l in lines)
    for l in lines:
        p, rest = l.split(" | ")
        k, rest = rest.split("*")
        b, rest = rest.split("^")
        n = rest[:-2]
        assert (int(k) * pow(int(b), int(n), int(p)) + 1) % int(p) == 0


def test_small_base745_job_full_range(rs, orc, gpu):
    """The CPU-sized base-745 workload in full: 4 odd k in [3, 1e6], n in
    [10000, 20000] presieved by odd primes < 1000, p = all primes in
    [1e9, 1e9 + 1e7]."""
    abcd = rs.Abcd.synth(745, 4, 3, 10**6, 10000, 20000, 1000, seed=1)
    plan = rs.Plan(abcd)
    low = 10**9
    offs = rs.PrimeGen(low + 10**7).range(low, low + 10**7 + 1)
    assert len(offs) == 482449
    for m in (0, 512):
        got = run_device(rs, plan, low, offs, m=m, max_batch=1 << 17)
        mm = m or plan.auto_m()
        want = run_oracle(orc, plan, mm, low, offs)
        assert got == want
    for p, k, n in got:
        assert orc.verify(p, k, plan.base, n)
    assert len(got) > 0


def test_r745_known_answer_and_window(rs, orc, gpu):
    """sr_745.abcd on the reference's default test range (:837-839)."""
    plan = rs.Plan(rs.Abcd.read(R745))
    low = 102297149770000 + 1
    offs = rs.PrimeGen(low + 2500000).range(low, low + 2500000)
    for m in (512, 0, 2048):
        got = run_device(rs, plan, low, offs, m=m)
        assert (KNOWN_P, 18300, 242670) in got
        want = run_oracle(orc, plan, m or plan.auto_m(), low, offs)
        assert got == want
    # with verification (dat mode) the datless hit is dropped
    assert not plan.contains(18300, 242670)


def test_r745_shaped_window(rs, orc, gpu):
    """The R745-shaped benchmark job: 22 odd k in [3, 1e7], n in [180000,
    250000] presieved by odd primes < 1e5, a 2^22 window of p at 1e12."""
    abcd = rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2)
    plan = rs.Plan(abcd)
    low = 10**12 + 1
    offs = rs.PrimeGen(low + 2**22).range(low, low + 2**22)
    got = run_device(rs, plan, low, offs, m=0, max_batch=1 << 15)
    want = run_oracle(orc, plan, plan.auto_m(), low, offs)
    assert got == want


def test_full_size_r745_layouts_agree(rs, gpu):
    """The whole R745-shaped benchmark range, p in [1e12, 1e12 + 2^34) (about
    6.2e8 primes, generated on the device), sieved twice with different
    layouts: m = 512, tile 4, full row entries (the 4-wave build) and m = 768,
    tile 2 with row-number lists (a 5-wave build).  The factor set is fixed by
    the mathematics, so the two must be equal at full size, and every factor
    must verify by modpow."""
    plan = rs.Plan(rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2))
    lo, hi = 10**12, 10**12 + (1 << 34)
    res = []
    for m, tile in ((512, 4), (768, 2 | rs.TILE_ROW_INDEX)):
        sv = rs.Sieve(plan, m=m, tile=tile)
        sv.submit_range(lo, hi)
        res.append(sv.collect())
        sv.close()
    assert res[0] == res[1]
    assert len(res[0]) > 500
    assert all(rs.verify_factor(p, k, 745, n) for p, k, n in res[0])


def test_large_p_near_2_62_and_2_63(rs, orc, gpu):
    plan = rs.Plan(rs.Abcd.synth(2, 12, 3, 10**9, 1000000, 1003000, 1000, seed=3))
    for low in ((1 << 62) - (1 << 18), (1 << 63) - (1 << 19)):
        offs = rs.PrimeGen(low + (1 << 18), threads=16).range(low, low + (1 << 18))
        assert all(is_prime(low + int(o)) for o in offs[:50])
        got = run_device(rs, plan, low, offs, m=1024)
        want = run_oracle(orc, plan, 1024, low, offs)
        assert got == want


def test_submit_rejects_p_past_2_63(rs, orc, gpu):
    """rs_sieve_submit refuses a call with any p = low + offset at or past
    2^63 (it reads the offsets for their largest only where low + 2^32 - 1
    could reach 2^63), or with low below 3, and sieves nothing for it: the
    sieve then still gives the oracle's hits for valid primes."""
    plan = rs.Plan(rs.Abcd.synth(2, 12, 3, 10**9, 1000000, 1003000, 1000, seed=3))
    sv = rs.Sieve(plan, m=1024)
    low = (1 << 63) - (1 << 19)
    offs = rs.PrimeGen(low + (1 << 18), threads=16).range(low, low + (1 << 18))
    for bad_low, bad in ((low, np.append(offs[:5], np.uint32((1 << 19) + 1))),
                         (2, np.array([1], np.uint32))):
        with pytest.raises(rs.RsError) as e:
            sv.submit(bad_low, bad)
        assert e.value.status == rs.RS_ERR_ARG
    sv.submit(low, offs)
    assert sv.collect() == run_oracle(orc, plan, 1024, low, offs)
    sv.close()


def test_zero_count_submit_prepares_staging(rs, orc, gpu):
    """rs_sieve_submit with no primes only makes the pinned staging ready:
    it sieves nothing, and later submits (twice, so both buffers are used)
    give the oracle's hits."""
    plan = rs.Plan(rs.Abcd.synth(745, 6, 3, 10**6, 20000, 30000, 1000, seed=5))
    low = 10**9 + 1
    offs = rs.PrimeGen(low + (1 << 20)).range(low, low + (1 << 20))
    sv = rs.Sieve(plan, m=512, max_batch=1 << 15)
    sv.submit(low, np.empty(0, np.uint32))
    assert sv.collect() == []
    sv.submit(low, offs)
    sv.submit(low, offs[: len(offs) // 3])
    got = sv.collect()
    sv.close()
    want = run_oracle(orc, plan, 512, low, offs)
    assert got == sorted(want + [h for h in want if h[0] - low <= int(offs[len(offs) // 3 - 1])])


def planted_pairs(b, p, nmin, nmax):
    """(k, n) with 1 < k < 2^31 and p | k*b^n - 1, n in [nmin, nmax]: k is
    b^-n mod p whenever that is small (about one n in p/2^31)."""
    inv = pow(b, -1, p)
    v = pow(inv, nmin, p)
    out = []
    for n in range(nmin, nmax + 1):
        if 1 < v < 1 << 31:
            out.append((v, n))
        v = v * inv % p
    return out


def test_p_around_2_50_both_baby_chains(rs, orc, gpu):
    """Primes on both sides of 2^50, where the baby-step chain switches from
    double precision to 64-bit integers: factors planted for the nearest
    primes below and above 2^50 must be found, and the whole window's hits
    equal the oracle's."""
    b = 5
    lo_p = max(q for q in range((1 << 50) - 400, 1 << 50) if is_prime(q))
    hi_p = min(q for q in range((1 << 50) + 1, (1 << 50) + 400) if is_prime(q))
    kn = {}
    planted = []
    for q in (lo_p, hi_p):
        pairs = planted_pairs(b, q, 1 << 20, (1 << 20) + (1 << 21))
        assert pairs, q
        for k, n in pairs[:3]:
            kn.setdefault(k, set()).update(range(n - 30, n + 31))
            planted.append((q, k, n))
    plan = rs.Plan(rs.Abcd.from_lists(b, {k: sorted(v) for k, v in kn.items()}))
    low = lo_p - (1 << 12)
    offs = rs.PrimeGen(hi_p + (1 << 12), threads=16).range(low, hi_p + (1 << 12))
    m = plan.auto_m()
    got = run_device(rs, plan, low, offs, m=m)
    want = run_oracle(orc, plan, m, low, offs)
    assert got == want
    for h in planted:
        assert h in got, h
    assert all(rs.verify_factor(p, k, b, n) for p, k, n in got)


def test_random_jobs_vs_oracle(rs, orc, gpu):
    """Seeded random jobs (base, k count and size, n window, sign, m, tile
    with and without row-number lists) over 2^16-wide windows of p from 2^20
    to 2^62 against the oracle: layouts, tiles and both arithmetic paths on
    shapes no fixed case has.  Small windows give hits; at large p both
    sides must agree on (almost always) none."""
    rng = random.Random(20261204)
    lows = [(1 << 20) + 1] * 4 + [10**9 + 7] * 2 + [
        10**12 + 39, (1 << 50) - (1 << 16) + 1, (1 << 50) + 1, (1 << 56) + 1,
        (1 << 62) - (1 << 17)]
    offs_of = {}
    total = 0
    for case in range(30):
        base = rng.randrange(3, 1031)
        nk = rng.choice([1, 3, 8, 25, 60])
        kmax = rng.choice([10**3, 10**6, 10**9, (1 << 31) - 1])
        nmin = rng.randrange(1000, 200000)
        width = rng.choice([200, 3000, 40000])
        c = rng.choice([-1, 1])
        m = 64 * rng.randrange(1, 33)
        tile = rng.choice([1, 2, 4]) | (rs.TILE_ROW_INDEX if rng.random() < 0.3 else 0)
        low = rng.choice(lows)
        if low not in offs_of:
            offs_of[low] = rs.PrimeGen(low + (1 << 16), threads=16).range(low, low + (1 << 16))
        offs = offs_of[low]
        plan = rs.Plan(rs.Abcd.synth(base, nk, 3, kmax, nmin, nmin + width, 1000, seed=case,
                                     c=c))
        sv = rs.Sieve(plan, m=m, tile=tile, max_batch=1 << 16)
        sv.submit(low, offs)
        got = sv.collect()
        sv.close()
        want = run_oracle(orc, plan, m, low, offs)
        assert got == want, (case, base, nk, kmax, nmin, width, c, m, tile, low)
        assert all(rs.verify_factor(p, k, base, n, c) for p, k, n in got[:200]), case
        total += len(got)
    assert total > 200


def test_row_index_layout_all_tiles(rs, orc, gpu):
    """Active rows listed by row number into the shared row table
    (TILE_ROW_INDEX): every tile with and without it equals the oracle on a
    window, and the exhaustive search on tiny primes."""
    plan = rs.Plan(rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2))
    assert plan.nrows <= 512
    low = 10**9 + 1
    offs = rs.PrimeGen(low + 2**20).range(low, low + 2**20)
    want = run_oracle(orc, plan, 512, low, offs)
    assert len(want) > 0
    for tile in (1, 2, 4):
        for flag in (0, rs.TILE_ROW_INDEX):
            sv = rs.Sieve(plan, m=512, tile=tile | flag)
            assert sv.tile == tile | flag
            sv.submit(low, offs)
            assert sv.collect() == want, (tile, flag)
            sv.close()
    tiny = tiny_job(rs, 6, {5: list(range(100, 700, 3)), 11: list(range(101, 650, 2)),
                            12: [100, 150, 200, 640], 35: list(range(300, 700))})
    primes = [q for q in range(5, 30000, 2) if is_prime(q)]
    toffs = np.array([q - 3 for q in primes], np.uint32)
    rk, rr, _ = tiny.rows()
    job = orc.Job(tiny.base, tiny.Q, 128, tiny.minN, tiny.maxN, rk, rr, tiny.c)
    brute = []
    for q in primes:
        brute += orc.brute_prime(job, q)
    sv = rs.Sieve(tiny, m=128, tile=4 | rs.TILE_ROW_INDEX)
    sv.submit(3, toffs)
    assert sv.collect() == brute
    sv.close()


def test_many_rows_sliced_all_tiles(rs, orc, gpu):
    """200 k of base 1030 give ~1000+ active (k, r) rows, more than one LDS
    row slice (512): every slice and every giant-step tile width must agree
    with the oracle."""
    plan = rs.Plan(rs.Abcd.synth(1030, 200, 3, 10**7, 100000, 120000, 10**5, seed=21))
    assert plan.nrows > 1024
    low = 10**9 + 1
    offs = rs.PrimeGen(low + 2**19).range(low, low + 2**19)
    m = plan.auto_m()
    want = run_oracle(orc, plan, m, low, offs)
    assert len(want) > 0
    for tile in (1, 2, 4):
        sv = rs.Sieve(plan, m=m, max_batch=1 << 15, tile=tile)
        sv.submit(low, offs)
        got = sv.collect()
        sv.close()
        assert got == want, tile


def test_many_k_base2_near_2_62(rs, orc, gpu):
    """Many-k / large-p edge: base 2, 500 odd k in [3, 1e9], n in [1e6, 2e6]
    presieved by primes < 1e6, p in a window just below 2^62."""
    plan = rs.Plan(rs.Abcd.synth(2, 500, 3, 10**9, 1000000, 2000000, 10**6, seed=3))
    low = (1 << 62) - (1 << 18)
    offs = rs.PrimeGen(1 << 62, threads=16).range(low, 1 << 62)
    m = plan.auto_m()
    got = run_device(rs, plan, low, offs, m=m, max_batch=1 << 14)
    want = run_oracle(orc, plan, m, low, offs)
    assert got == want


def test_k_near_2_31_second_baby_entries(rs, orc, gpu):
    """k just below 2^31 (the largest k of the format): giant-step products
    left in [0, 2p) exceed p for many items and half the baby values are in
    the table twice; small and large p, two tiles, against the oracle."""
    plan = rs.Plan(rs.Abcd.synth(6, 40, (1 << 31) - (1 << 24), (1 << 31) - 1, 30000, 90000,
                                 3000, seed=17, odd_k=False))
    assert max(plan.rows()[0]) > (1 << 31) - (1 << 24)
    for low, width in ((10**7 + 1, 1 << 20), ((1 << 62) + 1, 1 << 16)):
        offs = rs.PrimeGen(low + width, threads=16).range(low, low + width)
        m = plan.auto_m()
        want = run_oracle(orc, plan, m, low, offs)
        for tile in (1, 4):
            sv = rs.Sieve(plan, m=m, max_batch=1 << 15, tile=tile)
            sv.submit(low, offs)
            got = sv.collect()
            sv.close()
            assert got == want, (low, tile)
        if low < 1 << 32:
            assert len(want) > 0


def test_many_hits_small_p_datless(rs, orc, gpu):
    """p near 1e6: many hits per prime, exercises the append buffer."""
    plan = rs.Plan(rs.Abcd.synth(5, 10, 3, 10**5, 2000, 60000, 200, seed=4, odd_k=False))
    low = 10**6 + 1
    offs = rs.PrimeGen(low + 10**5).range(low, low + 10**5)
    got = run_device(rs, plan, low, offs, m=256)
    want = run_oracle(orc, plan, 256, low, offs)
    assert got == want and len(got) > 1000


def test_device_resident_and_repeat(rs, gpu):
    plan = rs.Plan(rs.Abcd.synth(745, 8, 3, 10**6, 50000, 80000, 1000, seed=5))
    low = 10**11 + 1
    offs = rs.PrimeGen(low + 2**21).range(low, low + 2**21)
    a = run_device(rs, plan, low, offs, m=384, max_batch=5000)
    sv = rs.Sieve(plan, m=384, max_batch=1 << 14)
    buf = rs.DeviceBuffer(0, offs)
    sv.submit_device(low, buf)
    b = sv.collect()
    sv.submit_device(low, buf, start=0, count=len(offs) // 2)
    sv.submit_device(low, buf, start=len(offs) // 2)
    c = sv.collect()
    t = sv.timing()
    sv.close()
    buf.free()
    assert a == b == c
    assert t["primes"] == 2 * len(offs) and t["kernel_ms"] > 0


def test_layout_beyond_lds_is_rejected(rs, gpu):
    """A baby table that cannot fit the 160 KB of LDS of a compute unit (m =
    32768: 16384 buckets, 256 KB) is refused at creation with RS_ERR_ARG, not
    launched; the next sieve on the device works."""
    plan = rs.Plan(rs.Abcd.synth(745, 4, 3, 10**6, 10000, 20000, 1000, seed=1))
    with pytest.raises(rs.RsError) as e:
        rs.Sieve(plan, m=32768)
    assert e.value.status == rs.RS_ERR_ARG
    low = 10**9 + 1
    offs = rs.PrimeGen(low + (1 << 16)).range(low, low + (1 << 16))
    assert run_device(rs, plan, low, offs, m=512) == run_device(rs, plan, low, offs, m=256)


def test_autotune_grid_and_descent(rs, orc, gpu):
    """The autotuner's full grid times each (m, tile) pair once (with
    row-number tiles only for plans of at most 512 rows), its descent returns
    a pair it timed, and the tuned sieve equals the oracle."""
    low = 10**9 + 7
    offs = rs.PrimeGen(low + (1 << 18)).range(low, low + (1 << 18))
    g = rs.GpuPrimeGen(low + (1 << 18) + 1)
    buf = g.device_range(low, low + (1 << 18))
    for nk, width in ((6, 20000), (120, 60000)):  # few rows; more than 512 rows
        plan = rs.Plan(rs.Abcd.synth(1030, nk, 3, 10**6, 100000, 100000 + width, 1000, seed=7))
        for effort in (0, 1):
            m, tile, cands = rs.autotune(plan, low, buf, effort=effort)
            pairs = [(cm, ct) for cm, ct, _ in cands]
            assert len(pairs) == len(set(pairs)), (nk, effort)
            assert (m, tile) in pairs
            assert min(cands, key=lambda c: c[2])[:2] == (m, tile)
            if plan.nrows > 512:
                assert all(ct & rs.TILE_ROW_INDEX == 0 for _, ct in pairs)
        sv = rs.Sieve(plan, m=m, tile=tile)
        sv.submit(low, offs)
        assert sv.collect() == run_oracle(orc, plan, m, low, offs), nk
        sv.close()
    buf.free()


def shape_ok(plan, m, tw):
    """The job-shape limits of rsieve.h (rs_sieve_create refuses the rest)."""
    X = plan.maxN // plan.Q - plan.minN // plan.Q + 1
    T = max(2, -(-(-(-X // m)) // tw) * tw)
    ys = T + (1 if T % 2 == 0 else 0)
    return (m % 64 == 0 and 64 <= m <= 1 << 16 and T <= 65535 and plan.Q * ys < 65536
            and min(plan.nrows, 512) * T + 4096 < 65536)


def test_autotune_wide_plan_below_the_shape_limit(rs, gpu):
    """A plan over 512 rows whose grid starts below the job-shape limit
    (base 2, 500 k, n in [1e6, 3e6]: at half the cost model's m, T breaks
    min(rows, 512) T + 4096 < 65536): both efforts skip the shapes outside
    the limits, time the rest and return a pair that builds."""
    plan = rs.Plan(rs.Abcd.synth(2, 500, 3, 10**6, 10**6, 3 * 10**6, 1000, seed=3,
                                 odd_k=True))
    assert plan.nrows > 512
    m0 = plan.auto_m()
    assert not shape_ok(plan, max(64, m0 // 2 // 64 * 64), 1)
    low = 10**9 + 7
    g = rs.GpuPrimeGen(low + (1 << 16) + 1)
    buf = g.device_range(low, low + (1 << 16))
    for effort in (0, 1):
        m, tile, cands = rs.autotune(plan, low, buf, effort=effort)
        assert m >= 64 and (m, tile) in [(cm, ct) for cm, ct, _ in cands], effort
        assert all(shape_ok(plan, cm, max(1, ct & 7)) for cm, ct, _ in cands)
        sv = rs.Sieve(plan, m=m, tile=tile)
        sv.close()
    buf.free()


def test_hit_buffer_overflow_is_reported(rs, gpu):
    plan = rs.Plan(rs.Abcd.synth(5, 10, 3, 10**5, 2000, 60000, 200, seed=4, odd_k=False))
    low = 10**6 + 1
    offs = rs.PrimeGen(low + 10**5).range(low, low + 10**5)
    sv = rs.Sieve(plan, m=256, hit_capacity=16)
    sv.submit(low, offs)
    with pytest.raises(RuntimeError, match="overflowed"):
        sv.collect()
    sv.close()


def test_cli_default_range_known_answer(tmp_path, gpu):
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    out = tmp_path / "f.txt"
    r = subprocess.run([exe, "-i", R745, "-O", str(out), "-v", "-m", "8"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = f"{KNOWN_P} | 18300*745^242670-1"
    assert line in r.stdout and line in out.read_text()
    out2 = tmp_path / "g.txt"
    r = subprocess.run([exe, "-i", R745, "-O", str(out2), "-m", "8"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert line not in r.stdout


def test_cli_readme_small_m_lines(tmp_path, gpu):
    """The README's lines with -m 4 (m = 32, below this build's smallest
    table) and -m 1 (automatic size) run, print the same factors as -m 8 and
    find the reference's debug prime on the default range."""
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    outs = {}
    for mflag in ("8", "4", "1"):
        out = tmp_path / f"f{mflag}.txt"
        r = subprocess.run([exe, "-i", R745, "-b", "5", "-m", mflag, "-s", "256", "-Q", "18",
                            "-v", "-O", str(out)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (mflag, r.stdout[-2000:], r.stderr)
        outs[mflag] = sorted(out.read_text().splitlines())
        if mflag == "4":
            assert "smallest table, 64, is used" in r.stdout
    assert f"{KNOWN_P} | 18300*745^242670-1" in outs["8"]
    assert outs["4"] == outs["8"] == outs["1"]


def test_factor_removal_after_sieving(rs, orc, tmp_path, gpu):
    """Dat mode: verified factors' candidates leave the ABCD; sieving the
    rewritten file over the same primes finds none of its candidates, and
    the CLI's -R writes the same file."""
    abcd = rs.Abcd.synth(5, 10, 3, 10**5, 2000, 9000, 200, seed=4, odd_k=2)
    before = abcd.info()["nn"]
    path = tmp_path / "in.abcd"
    abcd.write(path)
    plan = rs.Plan(abcd)
    low = 10**6 + 1
    offs = rs.PrimeGen(low + 2 * 10**5).range(low, low + 2 * 10**5)
    hits = run_device(rs, plan, low, offs, m=256)
    assert hits == run_oracle(orc, plan, 256, low, offs)
    found = [h for h in hits if plan.contains(h[1], h[2])]
    assert found and all(rs.verify_factor(p, k, 5, n) for p, k, n in found)
    gone = abcd.remove(found)
    assert gone == len({(k, n) for _, k, n in found})
    assert abcd.info()["nn"] == before - gone
    plan2 = rs.Plan(abcd)
    again = run_device(rs, plan2, low, offs, m=256)
    assert not [h for h in again if plan2.contains(h[1], h[2])]
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    out = tmp_path / "left.abcd"
    r = subprocess.run([exe, "-i", str(path), "-p", str(low), "-P", str(low + 2 * 10**5 - 1),
                        "-O", str(tmp_path / "f.txt"), "-m", "7", "-R", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"Removed {gone} candidates" in r.stdout
    for x, y in zip(abcd.lists(), rs.Abcd.read(out).lists()):
        assert np.array_equal(x, y)


def test_batched_jobs_equal_alone_and_oracle(rs, orc, gpu):
    """Heterogeneous jobs (bases, k counts, n windows, m) in one launch: each
    job's hits equal its plain run and the oracle; rerun replays the graph."""
    specs = [(745, 6, 10**6, 20000, 30000, 1000, 0), (2, 9, 10**8, 100000, 140000, 2000, 256),
             (5, 3, 10**5, 5000, 7000, 500, 128), (1030, 12, 10**7, 40000, 90000, 3000, 0),
             (3, 1, 10**4, 1000, 100000, 200, 1024)]
    plans, ms, bufs, lows = [], [], [], []
    for i, (b, nk, kmax, nlo, nhi, ps, m) in enumerate(specs):
        plans.append(rs.Plan(rs.Abcd.synth(b, nk, 3, kmax, nlo, nhi, ps, seed=10 + i,
                                           odd_k=(b % 2 == 0))))
        ms.append(m)
        low = 10**9 * (i + 1) + 1
        lows.append(low)
        bufs.append(rs.PrimeGen(low + 2**21).range(low, low + 2**21))
    bat = rs.Batch(plans, ms)
    dev = [rs.DeviceBuffer(0, o) for o in bufs]
    for j in range(len(plans)):
        half = len(bufs[j]) // 2
        bat.add_chunk(j, lows[j], dev[j], 0, half)
        bat.add_chunk(j, lows[j], dev[j], half)
    bat.run()
    first = bat.collect()
    bat.run()  # graph replay
    second = bat.collect()
    assert first == second
    for j, pl in enumerate(plans):
        mine = [(p, k, n) for (jj, p, k, n) in first if jj == j]
        m = bat.info[j]["m"]
        alone = run_device(rs, pl, lows[j], bufs[j], m=m)
        assert mine == alone, j
        assert mine == run_oracle(orc, pl, m, lows[j], bufs[j]), j
    assert len(first) > 0
    bat.close()


def test_batch_large_chunk_single_launch_equals_plain(rs, gpu):
    """A chunk large enough for its own single-job launch inside the batch
    graph (row table in block-shared LDS), next to small chunks of other jobs
    in class launches: every job's hits equal its plain run, on replay too."""
    big = rs.Plan(rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2))
    small = rs.Plan(rs.Abcd.synth(5, 6, 3, 10**6, 20000, 60000, 1000, seed=8, odd_k=False))
    low = 10**12 + 1
    gpg = rs.GpuPrimeGen(low + (1 << 28))
    buf = gpg.device_range(low, low + (1 << 28))
    assert buf.count > 9_000_000
    sbuf = gpg.device_range(low, low + (1 << 21))
    bat = rs.Batch([big, small], [512, 256])
    bat.add_chunk(0, low, buf)
    bat.add_chunk(1, low, sbuf)
    bat.run()
    first = bat.collect()
    bat.run()
    assert bat.collect() == first
    bat.close()
    for j, (pl, m, b) in enumerate(((big, 512, buf), (small, 256, sbuf))):
        sv = rs.Sieve(pl, m=m)
        sv.submit_device(low, b)
        want = sv.collect()
        sv.close()
        assert [(p, k, n) for (jj, p, k, n) in first if jj == j] == want, j
    assert any(jj == 0 for (jj, _, _, _) in first)


def test_gpu_primes_equal_cpu_generator(rs, gpu):
    """GPU segmented sieve (+ Miller-Rabin above the sieving bound) vs the
    host generator, including tiny ranges, 1e12, 1e14 (sieving primes to
    1e7, ragged ends; the second range is wide enough for the wide segment
    shape, 2^20 odd numbers per workgroup) and ranges near 2^62/2^63."""
    cases = [(3, 1 << 20, 1 << 22), (10**9, 10**9 + (1 << 24), 10**12),
             (10**12, 10**12 + (1 << 26), 10**12 + (1 << 27)),
             (10**14 + 12345, 10**14 + 12345 + (1 << 28) + 777, 10**14 + (1 << 29)),
             (10**14 + 999, 10**14 + 999 + (1 << 30) + 12345, 10**14 + (1 << 31)),
             ((1 << 62) - (1 << 20), 1 << 62, 1 << 62),
             ((1 << 63) - (1 << 20), 1 << 63, 1 << 63)]
    for lo, hi, hmax in cases:
        g = rs.GpuPrimeGen(hmax)
        got = g.range(lo, hi)
        want = rs.PrimeGen(hi, threads=16).range(lo, hi)
        assert np.array_equal(got, want), (lo, hi)


def test_submit_range_equals_host_primes(rs, orc, gpu):
    plan = rs.Plan(rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2))
    low, width = 10**12 + 1, 1 << 24
    offs = rs.PrimeGen(low + width).range(low, low + width)
    a = run_device(rs, plan, low, offs, m=576)
    sv = rs.Sieve(plan, m=576)
    sv.submit_range(low, low + width // 2)
    sv.submit_range(low + width // 2, low + width)
    b = sv.collect()
    t = sv.timing()
    sv.close()
    assert a == b
    assert t["primes"] == len(offs)
    assert a == run_oracle(orc, plan, 576, low, offs)


def test_cli_gpu_generated_primes(tmp_path, gpu):
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    outs = []
    for extra in ([], ["-g"]):
        out = tmp_path / f"f{len(extra)}.txt"
        r = subprocess.run([exe, "-i", R745, "-O", str(out), "-v", "-m", "8"] + extra,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(sorted(out.read_text().splitlines()))
    assert outs[0] == outs[1] and f"{KNOWN_P} | 18300*745^242670-1" in outs[1]


def test_cli_batches_across_windows(rs, tmp_path, gpu):
    """The CLI's CPU-generated batches (16384 or 65536 primes; launches of
    whole batches sliced from prime windows of 2^24, 2^25, then 2^26
    numbers, so batches span window ends, and collects paced by the hits
    seen)
    cover every prime of [P0, P1] once: a job with thousands of factors
    gives the same lines as the GPU-generated run and as the library's own
    sieve, verified and in the n-lists by default, every datless hit with -v
    (the reference's "don't verify")."""
    abcd = rs.Abcd.synth(5, 10, 3, 10**5, 2000, 9000, 200, seed=4, odd_k=2)
    path = tmp_path / "in.abcd"
    abcd.write(path)
    lo, hi = 10**6 + 1, 8 * 10**7
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    outs = []
    runs = (["-b", "1"], ["-b", "3"], ["-g"], ["-b", "1", "-v"])
    for extra in runs:
        out = tmp_path / f"f{len(outs)}.txt"
        r = subprocess.run([exe, "-i", str(path), "-p", str(lo), "-P", str(hi), "-m", "8",
                            "-O", str(out)] + extra, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (extra, r.stderr)
        outs.append(sorted(out.read_text().splitlines()))
    plan = rs.Plan(abcd)
    offs = rs.PrimeGen(hi + 1).range(lo, hi + 1)
    hits = run_device(rs, plan, lo, offs, m=512)
    every = sorted(rs.format_factor(p, k, 5, n).strip() for p, k, n in hits)
    listed = sorted(rs.format_factor(p, k, 5, n).strip() for p, k, n in hits
                    if plan.contains(k, n))
    assert len(listed) > 1000 and len(every) > len(listed)
    for extra, got in zip(runs, outs):
        want = every if "-v" in extra else listed
        assert got == want, (extra, len(got), len(want), sorted(set(got) ^ set(want))[:5])


def test_cli_from_small_p(rs, orc, tmp_path, gpu):
    """The CLI from p = 211 (just past the input's presieve bound, where a
    2^30-wide first window holds 5% of its numbers as primes) to 3e7: every
    listed candidate's factor, as the library's own sieve and the oracle
    find them, with the small primes' dense hits collected without loss."""
    abcd = rs.Abcd.synth(5, 10, 3, 10**5, 2000, 9000, 200, seed=4, odd_k=2)
    path = tmp_path / "in.abcd"
    abcd.write(path)
    lo, hi = 211, 3 * 10**7
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    out = tmp_path / "f.txt"
    r = subprocess.run([exe, "-i", str(path), "-p", str(lo), "-P", str(hi), "-m", "8", "-b", "2",
                        "-O", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "dropped" not in r.stderr
    got = sorted(out.read_text().splitlines())
    plan = rs.Plan(abcd)
    offs = rs.PrimeGen(hi + 1).range(lo, hi + 1)
    hits = run_device(rs, plan, lo, offs, m=512)
    small = np.array([o for o in offs if lo + int(o) < 10**6], np.uint32)
    assert [h for h in hits if h[0] < 10**6] == run_oracle(orc, plan, 512, lo, small)
    want = sorted(rs.format_factor(p, k, 5, n).strip() for p, k, n in hits
                  if plan.contains(k, n))
    assert len(want) > 1000
    assert got == want, (len(got), len(want), sorted(set(got) ^ set(want))[:5])


def test_cli_edge_ranges(rs, tmp_path, gpu):
    """The CLI on ranges at the edges of its batching: a prime gap (no prime
    in [1328, 1360]: no launch, no factor, a clean summary), one prime (the
    reference's debug prime alone, datless), and 2^22 numbers below 2^62
    (host primes proved by Miller-Rabin above the sieving bound) with CPU
    and with GPU primes, against the library's own sieve."""
    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")

    def run(*flags):
        out = tmp_path / f"f{len(list(tmp_path.iterdir()))}.txt"
        r = subprocess.run([exe, "-i", R745, "-m", "8", "-O", str(out)] + [str(f) for f in flags],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (flags, r.stderr)
        return r.stdout, (sorted(out.read_text().splitlines()) if out.exists() else [])

    stdout, lines = run("-p", 1328, "-P", 1360)
    assert "Factors found: 0" in stdout and not lines
    stdout, lines = run("-p", KNOWN_P, "-P", KNOWN_P, "-v")
    assert lines == [f"{KNOWN_P} | 18300*745^242670-1"]
    lo, hi = (1 << 62) - (1 << 22), (1 << 62) - 1
    got_cpu = run("-p", lo, "-P", hi, "-v")[1]
    got_gpu = run("-p", lo, "-P", hi, "-v", "-g")[1]
    plan = rs.Plan(rs.Abcd.read(R745))
    offs = rs.PrimeGen(hi + 1).range(lo, hi + 1)
    hits = run_device(rs, plan, lo, offs, m=512)
    want = sorted(rs.format_factor(p, k, 745, n).strip() for p, k, n in hits)
    assert got_cpu == got_gpu == want


@pytest.mark.parametrize("quantum_ms", [0.0, 0.05])
def test_server_results_alone_and_in_the_mix(rs, orc, gpu, quantum_ms):
    """Jobs arriving between steps are admitted as slots and budget free up;
    each job's hits equal the job sieved alone (and the oracle), datless and
    with device-resident n-lists (dat); with a time quantum the windows of
    each job are sized by its cost (2^20 up to 2^22 here)."""
    specs = [(745, 6, 10**6, 20000, 30000, 1000), (2, 9, 10**8, 100000, 140000, 2000),
             (5, 3, 10**5, 5000, 7000, 500), (1030, 12, 10**7, 40000, 90000, 3000),
             (3, 1, 10**4, 1000, 100000, 200), (31, 4, 10**6, 30000, 60000, 1000)]
    plans = [rs.Plan(rs.Abcd.synth(b, nk, 3, kmax, nlo, nhi, ps, seed=20 + i,
                                   odd_k=(b % 2 == 0)))
             for i, (b, nk, kmax, nlo, nhi, ps) in enumerate(specs)]
    width = 1 << 22
    ranges = [(10**9 + 1 + i * width, 10**9 + 1 + i * width + (2 + i % 3) * width)
              for i in range(len(plans))]
    ranges[1] = ranges[0]  # two jobs on the same windows share their primes
    # a budget for about two jobs at a time (the largest footprint x 2.5)
    probe = rs.Server(hbm_budget=1 << 40, max_active=1, chunk_width=width, lds_cap=16384)
    per_job = max(probe.job_bytes(pl, dat=(i == 3)) for i, pl in enumerate(plans))
    probe.close()
    srv = rs.Server(hbm_budget=int(2.5 * per_job), max_active=3, chunk_width=width,
                    lds_cap=16384, quantum_ms=quantum_ms)
    ids, pending = [], list(range(len(plans)))
    steps = 0
    while pending or any(srv.poll(j)[0] in (0, 1) for j in ids):
        if pending and steps % 2 == 0:  # arrivals while others run
            i = pending.pop(0)
            ids.append(srv.submit(plans[i], *ranges[i], dat=(i == 3)))
        srv.step()
        steps += 1
        assert steps < 400
    for i, j in enumerate(ids):
        st, done, nh = srv.poll(j)
        assert st == srv.DONE and done == ranges[i][1] - ranges[i][0]
        got = srv.results(j)
        sv = rs.Sieve(plans[i], m=256 if i % 2 else 0)
        sv.submit_range(*ranges[i])
        alone = sv.collect()
        sv.close()
        if i == 3:
            alone = [h for h in alone if plans[i].contains(h[1], h[2])]
        assert got == alone, i
    offs = rs.PrimeGen(ranges[2][1]).range(ranges[2][0], ranges[2][1])
    assert srv.results(ids[2]) == run_oracle(orc, plans[2], 256, ranges[2][0], offs)
    st = srv.stats()
    # primes are counted per job: jobs 0 and 1 share their windows and each
    # sieves them
    gen = rs.PrimeGen(max(hi for _, hi in ranges))
    assert st["primes"] == sum(len(gen.range(lo, hi)) for lo, hi in ranges)
    assert st["used"] == 0
    srv.close()


def test_server_rejects_job_over_budget(rs, gpu):
    plan = rs.Plan(rs.Abcd.synth(745, 4, 3, 10**6, 10000, 20000, 1000, seed=1))
    srv = rs.Server(hbm_budget=1 << 16, max_active=2, chunk_width=1 << 24)
    j = srv.submit(plan, 10**9, 10**9 + (1 << 24))
    assert srv.step() == (0, 0, 0)
    assert srv.poll(j)[0] == srv.REJECTED
    srv.release(j)
    srv.close()
