"""Seeded synthetic job sets for the batched and serving paths.

Heterogeneous jobs (no reference counterpart: the reference sieves one
ABCD file per process, kernel_v0.24.6.cu:1537-1611):
  base     uniform in [3, 1030],
  k-count  log-uniform in [1, 200], k drawn in [3, 10^7] with the parity a
           Riesel k*b^n-1 needs (odd k for even b, even k for odd b),
  n-range  width log-uniform in [10^4, 5*10^5] starting at 10^5,
  presieve by odd primes < 10^5,
so job shapes (Q, rows, T, m) differ by orders of magnitude.
"""
import math
import random

import pyrsieve as rs

NMIN = 100000
PRESIEVE = 100000
KMAX = 10**7


def r745_plan(seed=2):
    """R745-shaped job: base 745, 22 odd k in [3, 1e7], n in [180000,
    250000] presieved by odd primes < 1e5 (k parity as the ABCD in the
    reference's tests, sr_745.abcd)."""
    return rs.Plan(rs.Abcd.synth(745, 22, 3, KMAX, 180000, 250000, PRESIEVE, seed=seed))


def job_specs(njobs, seed=4):
    """[(base, nk, nmin, nmax, synth_seed)] drawn from the distribution above."""
    rng = random.Random(seed)
    out = []
    for j in range(njobs):
        base = rng.randint(3, 1030)
        nk = max(1, min(200, int(round(math.exp(rng.uniform(0.0, math.log(200)))))))
        width = int(round(math.exp(rng.uniform(math.log(1e4), math.log(5e5)))))
        out.append((base, nk, NMIN, NMIN + width, seed * 1000 + j))
    return out


def job_plan(spec):
    base, nk, nmin, nmax, sseed = spec
    parity = 1 if base % 2 == 0 else 2
    return rs.Plan(rs.Abcd.synth(base, nk, 3, KMAX, nmin, nmax, PRESIEVE, seed=sseed,
                                 odd_k=parity))


def poisson_arrivals(njobs, rate_per_s, seed=5):
    """Arrival times (s) of a Poisson process with the given rate."""
    rng = random.Random(seed)
    t, out = 0.0, []
    for _ in range(njobs):
        t += rng.expovariate(rate_per_s)
        out.append(t)
    return out
