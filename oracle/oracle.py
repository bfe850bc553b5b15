"""ctypes wrapper for liboracle.so, the CPU restatement of the reference sieve.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
compare results and by bench.py's cpu_baseline leg; never by the product path.
Each function mirrors one in oracle.cpp, which cites kernel_v0.24.6.cu.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
i32 = ctypes.c_int
i64 = ctypes.c_longlong
P = ctypes.POINTER


class OrHit(ctypes.Structure):
    _fields_ = [("p", u64), ("k", u32), ("n", u32)]


class OrJob(ctypes.Structure):
    _fields_ = [("base", u32), ("Q", i32), ("m", i32), ("minN", i32),
                ("maxN", i32), ("nrows", i32), ("rows_k", P(i32)),
                ("rows_r", P(i32)), ("c", i32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "or_modul64": (u64, [u64, u64, u64]),
            "or_mprime": (u64, [u64]),
            "or_montmul": (u64, [u64, u64, u64, u64]),
            "or_jacobi": (i32, [u64, u64]),
            "or_powmod": (u64, [u64, u64, u64]),
            "or_verify": (i32, [u64, u64, u64, u64, i32]),
            "or_abcd_count": (i32, [ctypes.c_char_p, P(i32), P(i32)]),
            "or_abcd_read": (i32, [ctypes.c_char_p, P(u32), P(i32), P(i32), P(i32),
                                   P(i32), P(i32)]),
            "or_plan": (i32, [i32, P(i32), P(i32), P(i32), i32, i32, i32, P(i32),
                              P(i32), P(i32), i32, P(i32), P(i32), P(u32)]),
            "or_tminmax": (None, [i32, i32, i32, i32, P(i32), P(i32)]),
            "or_bsgs_prime": (i32, [P(OrJob), u64, P(OrHit), i32, P(i64)]),
            "or_brute_prime": (i32, [P(OrJob), u64, P(OrHit), i32]),
            "or_bsgs_many": (i64, [P(OrJob), u64, P(u32), i64, P(OrHit), i64, i32,
                                   P(i64)]),
            "or_small_primes": (i32, [i32, P(u32)]),
            "or_inv128": (i64, [i64]),
            "or_ref_primes": (i32, [u64, i32, i32, i32, P(u32)]),
            "or_test_array_factor": (i32, [i32]),
            "or_true_primes": (i64, [u64, u64, P(u32), i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(P(ct))


modul64 = lambda x, y, z: lib().or_modul64(x, y, z)
mprime = lambda p: lib().or_mprime(p)
montmul = lambda a, b, p, mp: lib().or_montmul(a, b, p, mp)
jacobi = lambda a, n: lib().or_jacobi(a, n)
powmod = lambda b, e, p: lib().or_powmod(b, e, p)
def verify(p, k, b, n, c=-1):
    """1 if p | k*b^n + c (c = -1 Riesel, +1 Sierpinski)."""
    return lib().or_verify(p, k, b, n, c)
inv128 = lambda M: lib().or_inv128(M)
test_array_factor = lambda s: lib().or_test_array_factor(s)


def read_abcd(path):
    """Returns dict(base, ks, kstart, ns, minN, maxN) (kernel_v0.24.6.cu:938-1083)."""
    nk, nl = i32(), i32()
    rc = lib().or_abcd_count(path.encode(), ctypes.byref(nk), ctypes.byref(nl))
    if rc != 0:
        raise OSError(f"or_abcd_count({path}) = {rc}")
    ks = np.zeros(nk.value, np.int32)
    kstart = np.zeros(nk.value + 1, np.int32)
    ns = np.zeros(nl.value, np.int32)
    base = u32()
    mn, mx = i32(), i32()
    rc = lib().or_abcd_read(path.encode(), ctypes.byref(base), _ptr(ks, i32),
                            _ptr(kstart, i32), _ptr(ns, i32), ctypes.byref(mn),
                            ctypes.byref(mx))
    if rc != 0:
        raise ValueError(f"or_abcd_read({path}) = {rc}")
    return dict(base=base.value, ks=ks, kstart=kstart, ns=ns, minN=mn.value,
                maxN=mx.value)


def plan(ks, kstart, ns, minN, maxN, Qin=0):
    """Planner (kernel_v0.24.6.cu:1086-1255): dict(Q, rows_k, rows_r, bitmaps)."""
    ks = np.ascontiguousarray(ks, np.int32)
    kstart = np.ascontiguousarray(kstart, np.int32)
    ns = np.ascontiguousarray(ns, np.int32)
    Q, nr, w = i32(), i32(), i32()
    args = (len(ks), _ptr(ks, i32), _ptr(kstart, i32), _ptr(ns, i32), minN, maxN,
            Qin, ctypes.byref(Q), ctypes.byref(nr), ctypes.byref(w))
    rc = lib().or_plan(*args, 0, None, None, None)
    assert rc == 0
    rk = np.zeros(nr.value, np.int32)
    rr = np.zeros(nr.value, np.int32)
    bm = np.zeros(nr.value * w.value, np.uint32)
    rc = lib().or_plan(*args, nr.value, _ptr(rk, i32), _ptr(rr, i32), _ptr(bm, u32))
    assert rc == 0
    return dict(Q=Q.value, rows_k=rk, rows_r=rr, words=w.value,
                bitmaps=bm.reshape(nr.value, w.value), minN=minN, maxN=maxN)


def tminmax(minN, maxN, Q, m):
    a, b = i32(), i32()
    lib().or_tminmax(minN, maxN, Q, m, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


class Job:
    """Keeps the numpy row arrays alive behind an OrJob struct."""

    def __init__(self, base, Q, m, minN, maxN, rows_k, rows_r, c=-1):
        self.rows_k = np.ascontiguousarray(rows_k, np.int32)
        self.rows_r = np.ascontiguousarray(rows_r, np.int32)
        self.s = OrJob(base, Q, m, minN, maxN, len(self.rows_k),
                       _ptr(self.rows_k, i32), _ptr(self.rows_r, i32), c)


def _hits(buf, n):
    return [(int(h.p), int(h.k), int(h.n)) for h in buf[:n]]


def bsgs_prime(job, p, cap=4096):
    buf = (OrHit * cap)()
    mm = i64()
    n = lib().or_bsgs_prime(ctypes.byref(job.s), p, buf, cap, ctypes.byref(mm))
    if n > cap:
        return bsgs_prime(job, p, n)
    return _hits(buf, n), mm.value


def brute_prime(job, p, cap=4096):
    buf = (OrHit * cap)()
    n = lib().or_brute_prime(ctypes.byref(job.s), p, buf, cap)
    if n > cap:
        return brute_prime(job, p, n)
    return _hits(buf, n)


def bsgs_many(job, low, offs, threads=0, cap=1 << 16):
    """All hits for p = low + offs[i], sorted by (p, k, n); plus montmul count."""
    offs = np.ascontiguousarray(offs, np.uint32)
    buf = (OrHit * cap)()
    mm = i64()
    n = lib().or_bsgs_many(ctypes.byref(job.s), low, _ptr(offs, u32), len(offs), buf,
                           cap, threads, ctypes.byref(mm))
    if n > cap:
        return bsgs_many(job, low, offs, threads, n)
    return _hits(buf, n), mm.value


def small_primes(count):
    out = np.zeros(count, np.uint32)
    lib().or_small_primes(count, _ptr(out, u32))
    return out


def ref_primes(low, s, array_size, test_array_size):
    """Reference partial-sieve prime window (kernel_v0.24.6.cu:1838-1987)."""
    out = np.zeros(array_size, np.uint32)
    c = lib().or_ref_primes(low, s, array_size, test_array_size, _ptr(out, u32))
    return out[:c]


def true_primes(lo, hi, cap=None):
    if cap is None:
        cap = max(16, int((hi - lo) / max(1.0, np.log(max(hi, 3)) - 1.2)) + 1024)
    out = np.zeros(cap, np.uint32)
    c = lib().or_true_primes(lo, hi, _ptr(out, u32), cap)
    if c > cap:
        return true_primes(lo, hi, c)
    return out[:c]
