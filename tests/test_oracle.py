"""Pins the CPU oracle (oracle/oracle.cpp) before its results are trusted.

Each oracle function is checked against something it does not compute
itself: Python big-integer arithmetic, Euler's criterion, exhaustive search,
closed forms, and the golden facts of the reference's only fixtures
(sr_745.abcd and the datless debug prime of kernel_v0.24.6.cu:244).
"""
import random

import numpy as np
import pytest

from conftest import R745

KNOWN_P = 102297149781479


def small_primes(n):
    s = np.ones(n + 1, bool)
    s[:2] = False
    for i in range(2, int(n**0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return np.nonzero(s)[0]


# ---------------------------------------------------------------- arithmetic


def test_modul64_closed_form(orc):
    rng = random.Random(1)
    for _ in range(2000):
        z = rng.getrandbits(rng.randint(2, 63)) | 1
        z = max(z, 3)
        x, y = rng.randrange(z), rng.getrandbits(64)
        assert orc.modul64(x, y, z) == ((x << 64) | y) % z


def test_mprime_identity(orc):
    rng = random.Random(2)
    for _ in range(2000):
        p = rng.getrandbits(63) | 1
        assert p * orc.mprime(p) % (1 << 64) == (1 << 64) - 1


def test_montmul_vs_bigint(orc):
    rng = random.Random(3)
    cases = []
    for _ in range(3000):
        p = max(3, rng.getrandbits(rng.randint(2, 63)) | 1)
        cases.append((rng.getrandbits(64), rng.randrange(p), p))
    # lost-carry edge of kernel_v0.24.6.cu:226: bits 32..63 of a*b all ones
    for _ in range(300):
        p = (1 << 63) - 2 * rng.randrange(1, 1 << 40) - 1
        b = rng.randrange(2, 1 << 31)
        a = ((rng.randrange(1, b) << 64) + (1 << 64) - 1) // b
        assert a < 1 << 64 and ((a * b) >> 32) & 0xFFFFFFFF == 0xFFFFFFFF
        cases.append((a, b, p))
    for a, b, p in cases:
        assert orc.montmul(a, b, p, orc.mprime(p)) == a * b * pow(2, -64, p) % p


def test_jacobi_vs_euler_and_bruteforce(orc):
    for p in small_primes(2000)[1:]:
        p = int(p)
        squares = {x * x % p for x in range(1, p)}
        for a in range(0, 3 * p, 7):
            want = 0 if a % p == 0 else (1 if a % p in squares else -1)
            assert orc.jacobi(a, p) == want
    rng = random.Random(4)
    for _ in range(500):
        p = rng.getrandbits(40) | 1
        if pow(2, p - 1, p) != 1:
            continue
        a = rng.getrandbits(32)
        e = pow(a, (p - 1) // 2, p)
        assert orc.jacobi(a, p) == (0 if a % p == 0 else (1 if e == 1 else -1))


def test_powmod_verify(orc):
    rng = random.Random(5)
    for _ in range(1000):
        p = max(3, rng.getrandbits(62))
        b, e = rng.getrandbits(32), rng.getrandbits(32)
        assert orc.powmod(b, e, p) == pow(b, e, p)
    assert orc.verify(KNOWN_P, 18300, 745, 242670) == 1
    assert orc.verify(KNOWN_P, 18300, 745, 242671) == 0


# ---------------------------------------------------------------- reader/planner


def test_abcd_golden_stats(orc):
    d = orc.read_abcd(R745)
    assert d["base"] == 745
    assert d["ks"].tolist() == [1160, 1416, 3662, 6366, 6894, 9162, 9650, 11138, 11964, 15864,
                                18300, 19770, 20136, 21290, 22034, 22500, 22598, 24950, 25736,
                                26036, 28080, 29586]
    counts = np.diff(d["kstart"]).tolist()
    assert counts == [2100, 1524, 4419, 3070, 2869, 3698, 3394, 3438, 3743, 3809, 2787, 1712,
                      4585, 1991, 1951, 1732, 3104, 2959, 2346, 3124, 4206, 2550]
    assert len(d["ns"]) == 65111 and d["minN"] == 182644 and d["maxN"] == 250000
    # every list is strictly increasing; first n of 1160 from the header
    for i in range(22):
        ns = d["ns"][d["kstart"][i]:d["kstart"][i + 1]]
        assert np.all(np.diff(ns) > 0)
    assert d["ns"][0] == 182711


def test_planner_golden(orc):
    d = orc.read_abcd(R745)
    args = (d["ks"], d["kstart"], d["ns"], d["minN"], d["maxN"])
    pl = orc.plan(*args)
    assert pl["Q"] == 18 and len(pl["rows_k"]) == 198
    assert orc.plan(*args, Qin=18)["Q"] == 18
    # Q scan table: subs / R for Q = 2, 6, 12, 18 (independent recount)
    for Q, subs, R in ((2, 40, 33679), (6, 72, 11227), (12, 144, 5614), (18, 198, 3743)):
        rows = {(int(k), int(n) % Q) for i, k in enumerate(d["ks"])
                for n in d["ns"][d["kstart"][i]:d["kstart"][i + 1]]}
        assert len(rows) == subs
        assert d["maxN"] // Q - d["minN"] // Q + 1 == R
        assert orc.plan(*args, Qin=Q)["Q"] == Q
    # tMin/tMax by -m (m = 2 << flag), kernel_v0.24.6.cu:1148-1158
    for m, tmin, tmax in ((32, 317, 435), (512, 19, 28), (1024, 9, 14), (2048, 4, 7),
                          (8192, 1, 2), (16384, 0, 1)):
        assert orc.tminmax(d["minN"], d["maxN"], 18, m) == (tmin, tmax)
    # rows sorted by (k, r); bitmaps hold exactly the file's n (MSB first)
    rk, rr, bm = pl["rows_k"], pl["rows_r"], pl["bitmaps"]
    assert list(zip(rk, rr)) == sorted(zip(rk, rr))
    x0 = d["minN"] // 18
    got = set()
    for row in range(len(rk)):
        for w in range(pl["words"]):
            v = int(bm[row, w])
            for bit in range(32):
                if v >> (31 - bit) & 1:
                    got.add((int(rk[row]), 18 * (x0 + 32 * w + bit) + int(rr[row])))
    want = {(int(k), int(n)) for i, k in enumerate(d["ks"])
            for n in d["ns"][d["kstart"][i]:d["kstart"][i + 1]]}
    assert got == want


def test_planner_never_picks_odd_q(orc):
    # one k with a single n: no even Q beats K*(maxN-minN) = 0 work
    ks = np.array([3], np.int32)
    kstart = np.array([0, 1], np.int32)
    ns = np.array([100], np.int32)
    assert orc.plan(ks, kstart, ns, 100, 100)["Q"] % 2 == 0


# ---------------------------------------------------------------- BSGS


def test_bsgs_equals_bruteforce_small(orc):
    rng = random.Random(6)
    for trial in range(6):
        base = rng.choice([2, 3, 5, 6, 10, 745])
        ks = sorted(rng.sample(range(1, 400), 4))
        rows_k, rows_r = [], []
        Q = rng.choice([2, 4, 6])
        for k in ks:
            for r in sorted(rng.sample(range(Q), rng.randint(1, Q))):
                rows_k.append(k)
                rows_r.append(r)
        minN, maxN = rng.randint(1, 300), rng.randint(400, 3000)
        for m in (1, 7, 64, 256):
            job = orc.Job(base, Q, m, minN, maxN, rows_k, rows_r)
            for p in small_primes(3000)[1:]:
                p = int(p)
                if base % p == 0:
                    continue
                got, _ = orc.bsgs_prime(job, p)
                assert got == orc.brute_prime(job, p), (trial, m, p)


def test_bsgs_small_vs_python_pow(orc):
    job = orc.Job(7, 4, 32, 5, 700, [3, 3, 3, 10, 10], [0, 1, 3, 1, 2])
    for p in small_primes(500)[2:]:
        p = int(p)
        if p == 7:
            continue
        want = sorted((p, k, n) for k, r in [(3, 0), (3, 1), (3, 3), (10, 1), (10, 2)]
                      for n in range(5, 701) if n % 4 == r and k * pow(7, n, p) % p == 1)
        assert orc.bsgs_prime(job, p)[0] == want


def test_bsgs_sierpinski_equals_bruteforce_and_pow(orc):
    """k*b^n+1: the oracle's BSGS (negated k, Jacobi symbols against (-1/p))
    = exhaustive search = Python pow, over primes 1 and 3 mod 4, and the
    Riesel and Sierpinski factor sets of the same rows differ."""
    rng = random.Random(16)
    for trial in range(4):
        base = rng.choice([2, 3, 5, 10])
        ks = sorted(rng.sample(range(1, 300), 3))
        Q = rng.choice([2, 4, 6])
        rows = [(k, r) for k in ks for r in sorted(rng.sample(range(Q), rng.randint(1, Q)))]
        rows_k, rows_r = [k for k, _ in rows], [r for _, r in rows]
        minN, maxN = rng.randint(1, 100), rng.randint(200, 900)
        for m in (1, 32, 128):
            job = orc.Job(base, Q, m, minN, maxN, rows_k, rows_r, 1)
            riesel = orc.Job(base, Q, m, minN, maxN, rows_k, rows_r, -1)
            differ = False
            for p in small_primes(1500)[1:]:
                p = int(p)
                if base % p == 0:
                    continue
                got, _ = orc.bsgs_prime(job, p)
                assert got == orc.brute_prime(job, p), (trial, m, p)
                if m == 32 and p < 400:
                    want = sorted((p, k, n) for k, r in rows for n in range(minN, maxN + 1)
                                  if n % Q == r and (k * pow(base, n, p) + 1) % p == 0)
                    assert got == want
                differ |= got != orc.bsgs_prime(riesel, p)[0]
            assert differ


def test_datless_known_answer(orc):
    d = orc.read_abcd(R745)
    pl = orc.plan(d["ks"], d["kstart"], d["ns"], d["minN"], d["maxN"])
    job = orc.Job(745, 18, 512, d["minN"], d["maxN"], pl["rows_k"], pl["rows_r"])
    hits, _ = orc.bsgs_prime(job, KNOWN_P)
    assert hits == [(KNOWN_P, 18300, 242670)]
    assert pow(745, 242670, KNOWN_P) * 18300 % KNOWN_P == 1
    # n = 242670 is not in the file's list for k = 18300
    i = d["ks"].tolist().index(18300)
    assert 242670 not in d["ns"][d["kstart"][i]:d["kstart"][i + 1]]


def test_bsgs_many_matches_per_prime(orc):
    job = orc.Job(745, 6, 128, 10000, 13000, [4, 4, 22, 22, 22], [0, 3, 1, 2, 4])
    offs = orc.true_primes(10**7, 10**7 + 200000)
    hits, mm = orc.bsgs_many(job, 10**7, offs, threads=4)
    want = []
    for o in offs:
        want += orc.bsgs_prime(job, 10**7 + int(o))[0]
    assert hits == want and mm > 0


# ---------------------------------------------------------------- primes


def test_true_primes_vs_numpy_sieve(orc):
    ps = small_primes(2_000_000)
    got = orc.true_primes(1_000_001, 2_000_000) + 1_000_001
    assert got.tolist() == ps[ps >= 1_000_001].tolist()


def test_mask_multipliers_closed_form(orc):
    # the reference's magicMultipliers (kernel_v0.24.6.cu:1843) equal
    # 128^-1 mod the group products (:1842, :1329-1336)
    mods = [19 * 13 * 67 * 11, 103 * 23 * 17 * 5, 107 * 97 * 3 * 7, 43 * 41 * 101, 83 * 73 * 29,
            61 * 47 * 59, 31 * 71 * 79, 37 * 89 * 53]
    assert mods == [182039, 201365, 217959, 178063, 175711, 169153, 173879, 174529]
    mult = [126574, 105402, 69815, 23649, 133156, 83255, 164370, 85901]
    for M, c in zip(mods, mult):
        assert orc.inv128(M) == c == pow(128, -1, M)


def test_reference_generator_is_partial_sieve(orc):
    assert orc.small_primes(30).tolist()[:5] == [3, 5, 7, 11, 13]
    assert int(orc.small_primes(27)[-1]) == 107          # mask primes end at 107
    s = 4                                                 # 16384 small primes
    qmax = int(orc.small_primes(4096 * s)[-1])
    assert qmax == 180511
    # below qmax^2 the partial sieve gives exactly the true primes
    low = 10**9 + 1
    ref = orc.ref_primes(low, s, 4096, 4096 * orc.test_array_factor(s))
    tru = orc.true_primes(low, low + 2 * 4096 * orc.test_array_factor(s))
    assert ref.tolist() == tru[:4096].tolist()
    # far above qmax^2 the survivors include composites
    low = (1 << 62) + 1
    ref = orc.ref_primes(low, s, 2048, 2048 * 24)
    comp = [o for o in ref.tolist() if pow(2, low + o - 1, low + o) != 1]
    assert len(comp) > 0
