"""Host side of librsieve (reader, planner, prime generator, synthetic
input, verification, formats, error paths) against the oracle, through the
C ABI.  CPU only."""
import os
import random
import re

import numpy as np
import pytest

from conftest import R745

KNOWN_P = 102297149781479


def test_reader_matches_oracle(rs, orc):
    a = rs.Abcd.read(R745)
    d = orc.read_abcd(R745)
    info = a.info()
    assert (info["base"], info["nk"], info["nn"], info["minN"], info["maxN"]) == (
        745, 22, 65111, 182644, 250000)
    ks, kstart, ns = a.lists()
    assert np.array_equal(ks, d["ks"]) and np.array_equal(kstart, d["kstart"])
    assert np.array_equal(ns, d["ns"])


def test_planner_matches_oracle(rs, orc):
    a = rs.Abcd.read(R745)
    d = orc.read_abcd(R745)
    for Qin in (0, 2, 6, 18, 22):
        pl = rs.Plan(a, Qin)
        op = orc.plan(d["ks"], d["kstart"], d["ns"], d["minN"], d["maxN"], Qin)
        rk, rr, bm = pl.rows()
        assert pl.Q == op["Q"]
        assert np.array_equal(rk, op["rows_k"]) and np.array_equal(rr, op["rows_r"])
        assert np.array_equal(bm, op["bitmaps"])
        for m in (32, 512, 2048, 16384):
            assert pl.tminmax(m) == orc.tminmax(d["minN"], d["maxN"], pl.Q, m)


def test_plan_contains_is_the_file(rs, orc):
    a = rs.Abcd.read(R745)
    pl = rs.Plan(a)
    ks, kstart, ns = a.lists()
    rng = random.Random(1)
    for i in range(len(ks)):
        lst = set(ns[kstart[i]:kstart[i + 1]].tolist())
        for n in rng.sample(sorted(lst), 50):
            assert pl.contains(int(ks[i]), n)
        for n in rng.sample(range(182644, 250001), 200):
            assert pl.contains(int(ks[i]), n) == (n in lst)
    assert not pl.contains(18300, 242670)
    assert not pl.contains(1161, 182711)
    assert not pl.contains(1160, 182643) and not pl.contains(1160, 250001)


def test_writer_round_trip(rs, tmp_path):
    a = rs.Abcd.read(R745)
    out = tmp_path / "r.abcd"
    a.write(out)
    b = rs.Abcd.read(out)
    for x, y in zip(a.lists(), b.lists()):
        assert np.array_equal(x, y)
    first = out.read_text().splitlines()[0]
    assert first == "ABCD 1160*745^$a-1 [182711]"


def test_reader_errors(rs, tmp_path):
    with pytest.raises(rs.RsError) as e:
        rs.Abcd.read(tmp_path / "missing.abcd")
    assert e.value.status == rs.RS_ERR_IO
    bad = tmp_path / "bad.abcd"
    for text in ("12\n", "ABCD 5*2^$a-1 [10]\nxyz\n", "ABCD 5*2^$a-1\n",
                 "ABCD 5*2^$a-1 [10]\n0\n", "ABCD 5*2^$a-1 [10]\nABCD 7*2^$a+1 [10]\n",
                 "ABCD 5*2^$a*1 [10]\n"):
        bad.write_text(text)
        with pytest.raises(rs.RsError) as e:
            rs.Abcd.read(bad)
        assert e.value.status == rs.RS_ERR_FORMAT, text


def test_sierpinski_form_read_write(rs, tmp_path):
    src = tmp_path / "s.abcd"
    src.write_text("ABCD 27*2^$a+1 [100]\n4\n10\nABCD 33*2^$a+1 [101]\n6\n")
    a = rs.Abcd.read(src)
    assert a.c == 1
    ks, kstart, ns = a.lists()
    assert ks.tolist() == [27, 33] and ns.tolist() == [100, 104, 114, 101, 107]
    out = tmp_path / "o.abcd"
    a.write(out)
    assert out.read_text() == src.read_text()
    assert rs.Plan(a).c == 1 and rs.Abcd.read(R745).c == -1
    b = rs.Abcd.from_lists(2, {27: [100, 104]})
    assert b.c == -1
    b.set_form(1)
    assert b.c == 1 and rs.Plan(b).c == 1


def test_sierpinski_verify_and_format(rs, orc):
    # 3*2^5+1 = 97 is prime; 5 | 3*2^3+1 = 25
    assert rs.verify_factor(97, 3, 2, 5, 1) and not rs.verify_factor(97, 3, 2, 5, -1)
    assert rs.verify_factor(5, 3, 2, 3, 1) and orc.verify(5, 3, 2, 3, 1)
    assert rs.format_factor(5, 3, 2, 3, 1) == "5 | 3*2^3+1\n"
    rng = random.Random(9)
    for _ in range(500):
        p = rng.getrandbits(40) | 1
        k, b, n = rng.getrandbits(31), rng.randint(2, 2000), rng.getrandbits(16)
        want = (k * pow(b, n, p) + 1) % p == 0
        assert rs.verify_factor(p, k, b, n, 1) == want == bool(orc.verify(p, k, b, n, 1))


def test_synth_presieve_exact_sierpinski(rs):
    a = rs.Abcd.synth(2, 6, 3, 10**6, 3000, 4500, 2000, seed=7, c=1)
    assert a.c == 1
    ks, kstart, ns = a.lists()
    qs = [q for q in range(3, 2000, 2) if all(q % d for d in range(3, int(q**0.5) + 1, 2))]
    for i, k in enumerate(ks.tolist()):
        kept = set(ns[kstart[i]:kstart[i + 1]].tolist())
        for n in range(3000, 4501, 7):  # a sample of n
            survives = all((k * pow(2, n, q) + 1) % q for q in qs)
            assert (n in kept) == survives, (k, n)


def test_plan_rejects_odd_q(rs):
    a = rs.Abcd.read(R745)
    with pytest.raises(rs.RsError):
        rs.Plan(a, 3)


def test_primes_range_matches_oracle(rs, orc):
    g = rs.PrimeGen(10**12 + 2**24, threads=4)
    for lo, hi in ((0, 100000), (1, 3), (2, 3), (999, 1000), (10**9 + 7, 10**9 + 2**20),
                   (10**12, 10**12 + 2**22)):
        got = g.range(lo, hi)
        want = orc.true_primes(lo, hi)
        assert np.array_equal(got, want), (lo, hi)


@pytest.mark.parametrize("threads", [1, 4, 16])
def test_primes_range_across_segments_and_blocks(rs, orc, threads):
    """A window of 4.3 segments of 2 MiB wheel bytes (63M numbers each; one
    block on one thread, four blocks of one segment and a piece of the next
    on four, sixteen blocks of about a quarter segment on sixteen, each
    prime's next multiple carried from segment to segment), with sieving
    primes to 1.01e7 and an odd, ragged end, equals the oracle; so do
    pieces of it cut at arbitrary points."""
    lo = 102297149770000 + 1
    hi = lo + (1 << 28) - 77
    g = rs.PrimeGen(hi, threads=threads)
    want = orc.true_primes(lo, hi)
    got = g.range(lo, hi)
    assert np.array_equal(got, want)
    cuts = [lo, lo + 12345, lo + (1 << 22) + 1, lo + (5 << 22) - 3, hi]
    parts = [c0 - lo + g.range(c0, c1).astype(np.int64) for c0, c1 in zip(cuts, cuts[1:])]
    assert np.array_equal(np.concatenate(parts), want.astype(np.int64))


def test_primes_range_concurrent_calls(rs, orc):
    """Calls on one generator from several threads at once (the ctypes calls
    release the GIL): one call keeps the generator's per-block buffers, the
    others use their own; every result equals the oracle."""
    import threading
    lo = 10**12 + 1
    g = rs.PrimeGen(lo + 2**26, threads=4)
    wins = [(lo + i * 2**23 + 17 * i, lo + (i + 1) * 2**23) for i in range(8)]
    got = [None] * len(wins)

    def work(j):
        for _ in range(3):
            got[j] = g.range(*wins[j])

    th = [threading.Thread(target=work, args=(j,)) for j in range(len(wins))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for (a, b), offs in zip(wins, got):
        assert np.array_equal(offs, orc.true_primes(a, b)), (a, b)


def test_primes_batch_is_reference_layout(rs, orc):
    g = rs.PrimeGen(10**12 + 2**24, threads=4)
    low = 10**12 + 1
    offs, nxt = g.batch(low, 50000)
    want = orc.true_primes(low, low + 4 * 10**6)[:50000]
    assert np.array_equal(offs, want)
    assert nxt == low + int(offs[-1]) + 2
    # even low is moved down to the odd below it (kernel_v0.24.6.cu:1515-1519)
    offs2, _ = g.batch(low + 1, 10)
    assert np.array_equal(offs2, offs[:10])


def test_primes_equal_reference_generator_below_qmax_squared(rs, orc):
    s = 4
    low = 10**9 + 1
    g = rs.PrimeGen(low + 10**6, small_prime_scale=s)
    ours, _ = g.batch(low, 4096)
    ref = orc.ref_primes(low, s, 4096, 4096 * orc.test_array_factor(s))
    assert np.array_equal(ours, ref)


def test_primes_true_above_qmax_squared(rs):
    low = (1 << 62) + 1
    g = rs.PrimeGen(low + 2**16, small_prime_scale=4, threads=4)
    offs = g.range(low, low + 2**16)
    vals = [low + int(o) for o in offs]
    for v in vals:
        assert all(pow(a, v - 1, v) == 1 for a in (2, 3, 5, 7))
    # and nothing prime is missing: check every odd number in a slice
    for v in range(low, low + 4000, 2):
        isp = all(pow(a, v - 1, v) == 1 for a in (2, 3, 5, 7, 11, 13))
        assert (v in set(vals)) == isp


def test_factor_removal(rs, tmp_path):
    a = rs.Abcd.from_lists(745, {18: [100, 118, 136], 40: [101, 119], 66: [104]})
    # (18, 118) and (66, 104) go, (40, 7) is not a candidate, (18, 118) twice
    gone = a.remove([(1009, 18, 118), (1013, 66, 104), (1019, 40, 7), (1021, 18, 118)])
    assert gone == 2
    ks, kstart, ns = a.lists()
    assert ks.tolist() == [18, 40] and kstart.tolist() == [0, 2, 4]
    assert ns.tolist() == [100, 136, 101, 119]
    inf = a.info()
    assert (inf["nk"], inf["nn"], inf["minN"], inf["maxN"]) == (2, 4, 100, 136)
    path = tmp_path / "left.abcd"
    a.write(path)
    b = rs.Abcd.read(path)
    for x, y in zip(a.lists(), b.lists()):
        assert np.array_equal(x, y)
    assert a.remove([]) == 0
    assert a.remove([(3, 18, 100), (3, 18, 136), (3, 40, 101), (3, 40, 119)]) == 4
    assert a.info()["nk"] == 0
    with pytest.raises(rs.RsError):
        rs.Plan(a)


def test_synth_presieve_exact(rs):
    a = rs.Abcd.synth(745, 5, 3, 10**6, 10000, 12000, 1000, seed=11)
    ks, kstart, ns = a.lists()
    qs = [q for q in range(3, 1000, 2) if all(q % d for d in range(3, int(q**0.5) + 1, 2))
          and 745 % q]
    assert len(ks) == 5 and np.all(ks % 2 == 1) and np.all(np.diff(ks) > 0)
    for i, k in enumerate(ks.tolist()):
        kept = set(ns[kstart[i]:kstart[i + 1]].tolist())
        assert kept
        for n in range(10000, 12001):
            survives = all((k * pow(745, n, q) - 1) % q for q in qs)
            assert (n in kept) == survives, (k, n)


def test_synth_presieve_exact_many_k_base2(rs):
    # more k and larger q than above: the per-q discrete logs use baby steps
    # t_k*b^j with m > 1 and orders up to q-1
    a = rs.Abcd.synth(2, 40, 3, 10**9, 1000, 2500, 20000, seed=5)
    ks, kstart, ns = a.lists()
    sieve = np.ones(20000, dtype=bool)
    sieve[:3] = False
    for i in range(2, 142):
        if sieve[i]:
            sieve[i * i::i] = False
    qs = [int(q) for q in np.nonzero(sieve)[0] if q > 2]
    width = 1501
    dead = np.zeros((len(ks), width), dtype=bool)
    kv = ks.astype(object)
    for q in qs:
        pw = np.empty(width, dtype=np.int64)
        v = pow(2, 1000, q)
        for j in range(width):
            pw[j] = v
            v = v * 2 % q
        for i, k in enumerate(ks.tolist()):
            dead[i] |= (k % q) * pw % q == 1
    for i in range(len(ks)):
        kept = np.zeros(width, dtype=bool)
        kept[ns[kstart[i]:kstart[i + 1]] - 1000] = True
        assert np.array_equal(kept, ~dead[i]), int(kv[i])


def test_synth_is_seeded(rs):
    a = rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2)
    b = rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=2)
    c = rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=3)
    for x, y in zip(a.lists(), b.lists()):
        assert np.array_equal(x, y)
    assert not np.array_equal(a.lists()[0], c.lists()[0])


def test_verify_and_format(rs, orc):
    assert rs.verify_factor(KNOWN_P, 18300, 745, 242670)
    assert not rs.verify_factor(KNOWN_P, 18300, 745, 242671)
    assert not rs.verify_factor(KNOWN_P + 2, 18300, 745, 242670)
    line = rs.format_factor(KNOWN_P, 18300, 745, 242670)
    assert line == "102297149781479 | 18300*745^242670-1\n"
    assert re.fullmatch(r"\d+ \| \d+\*\d+\^\d+-1\n", line)
    rng = random.Random(3)
    for _ in range(500):
        p = rng.getrandbits(62) | 1
        k, b, n = rng.getrandbits(31), rng.randint(2, 2000), rng.getrandbits(20)
        assert rs.verify_factor(p, k, b, n) == bool(orc.verify(p, k, b, n))


def test_auto_m_is_multiple_of_64(rs):
    pl = rs.Plan(rs.Abcd.read(R745))
    m = pl.auto_m()
    assert m % 64 == 0 and 256 <= m <= 2048
