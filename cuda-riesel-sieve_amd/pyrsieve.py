"""ctypes binding of librsieve.so (include/rsieve.h) for tests and benchmarks.

Every call goes through the C ABI; a missing library or a failing call raises
(there is no fallback path).  The objects mirror the reference program's
pieces: Abcd (the ABCD reader, kernel_v0.24.6.cu:113-137, :938-1083), Plan
(the subsequence planner, :1086-1255), PrimeGen (the prime-array generator,
:1838-1987) and Sieve (device set-up, launch and result collection,
:1367-1611).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librsieve.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "rsieve.h")

u64, u32, i32, i64 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int64
vp = ctypes.c_void_p
P = ctypes.POINTER

RS_OK = 0
RS_ERR_ARG = -1
RS_ERR_HIP = -2
RS_ERR_NOMEM = -3
RS_ERR_IO = -4
RS_ERR_FORMAT = -5
RS_ERR_OVERFLOW = -6
RS_ERR_BUDGET = -7
RS_ERR_STATE = -8
RS_ERR_RANGE = -9


class Factor(ctypes.Structure):
    _fields_ = [("p", u64), ("k", u32), ("n", u32)]


class JobFactor(ctypes.Structure):
    _fields_ = [("p", u64), ("k", u32), ("n", u32), ("job", u32), ("pad", u32)]

_SIGS = {
    "rs_strerror": (ctypes.c_char_p, [i32]),
    "rs_abcd_read": (i32, [ctypes.c_char_p, P(vp)]),
    "rs_abcd_from_lists": (i32, [u32, i32, P(i32), P(i32), P(i32), P(vp)]),
    "rs_abcd_info": (i32, [vp, P(u32), P(i32), P(i32), P(i32), P(i32)]),
    "rs_abcd_lists": (i32, [vp, P(i32), P(i32), P(i32)]),
    "rs_abcd_write": (i32, [vp, ctypes.c_char_p]),
    "rs_abcd_remove": (i32, [vp, vp, i64, P(i64)]),
    "rs_abcd_free": (None, [vp]),
    "rs_abcd_synth": (i32, [u32, i32, u32, u32, i32, i32, i32, i32, u32, u64, P(vp)]),
    "rs_abcd_form": (i32, [vp, P(i32)]),
    "rs_abcd_set_form": (i32, [vp, i32]),
    "rs_plan_create": (i32, [vp, i32, P(vp)]),
    "rs_plan_form": (i32, [vp, P(i32)]),
    "rs_plan_info": (i32, [vp, P(u32), P(i32), P(i32), P(i32), P(i32), P(i32)]),
    "rs_plan_rows": (i32, [vp, P(i32), P(i32), P(u32)]),
    "rs_plan_contains": (i32, [vp, u32, u32]),
    "rs_plan_tminmax": (i32, [vp, i32, P(i32), P(i32)]),
    "rs_plan_auto_m": (i32, [vp]),
    "rs_plan_free": (None, [vp]),
    "rs_primegen_create": (i32, [i32, u64, i32, P(vp)]),
    "rs_primes_range": (i32, [vp, u64, u64, P(u32), i64, P(i64)]),
    "rs_primes_batch": (i32, [vp, u64, P(u32), i64, P(u64)]),
    "rs_primegen_free": (None, [vp]),
    "rs_gpu_primegen_create": (i32, [i32, u64, P(vp)]),
    "rs_gpu_primes_range": (i32, [vp, u64, u64, vp, i64, P(i64)]),
    "rs_gpu_primegen_free": (None, [vp]),
    "rs_verify_factor": (i32, [u64, u32, u32, u32, i32]),
    "rs_format_factor": (i32, [P(Factor), u32, i32, ctypes.c_char_p, i32]),
    "rs_sieve_create": (i32, [i32, vp, i32, i32, i64, i64, P(vp)]),
    "rs_sieve_info": (i32, [vp, P(i32), P(i32), P(i32), P(i32), P(i32)]),
    "rs_sieve_kernel_name": (ctypes.c_char_p, [vp]),
    "rs_sieve_submit": (i32, [vp, u64, P(u32), i64]),
    "rs_sieve_submit_device": (i32, [vp, u64, vp, i64]),
    "rs_sieve_submit_range": (i32, [vp, u64, u64]),
    "rs_sieve_submit_device_noalloc": (i32, [vp, u64, vp, i64, vp, i64, vp]),
    "rs_sieve_collect": (i32, [vp, P(Factor), i64, P(i64), P(i64)]),
    "rs_sieve_sync": (i32, [vp]),
    "rs_sieve_timing": (i32, [vp, P(ctypes.c_double), P(i64)]),
    "rs_sieve_destroy": (None, [vp]),
    "rs_autotune": (i32, [i32, vp, u64, vp, i64, i32, P(i32), P(i32), P(i32), P(i32),
                          P(ctypes.c_double), i32, P(i32)]),
    "rs_batch_create": (i32, [i32, i32, P(vp), P(i32), P(i32), i64, P(vp)]),
    "rs_batch_job_info": (i32, [vp, i32, P(i32), P(i32), P(i32), P(i32)]),
    "rs_batch_add_chunk": (i32, [vp, i32, u64, vp, i64]),
    "rs_batch_clear": (i32, [vp]),
    "rs_batch_run": (i32, [vp]),
    "rs_batch_collect": (i32, [vp, P(JobFactor), i64, P(i64), P(i64)]),
    "rs_batch_timing": (i32, [vp, P(ctypes.c_double), P(i64)]),
    "rs_batch_destroy": (None, [vp]),
    "rs_server_create": (i32, [i32, i64, i32, u64, i32, i64, ctypes.c_double, P(vp)]),
    "rs_server_job_bytes": (i32, [vp, vp, i32, P(i64)]),
    "rs_server_submit": (i32, [vp, vp, i32, i32, i32, u64, u64, ctypes.c_double, P(i64)]),
    "rs_server_step": (i32, [vp, P(i32), P(i32), P(i64)]),
    "rs_server_poll": (i32, [vp, i64, P(i32), P(u64), P(i64)]),
    "rs_server_results": (i32, [vp, i64, P(Factor), i64, P(i64)]),
    "rs_server_release": (i32, [vp, i64]),
    "rs_server_stats": (i32, [vp, P(i64), P(i64), P(ctypes.c_double), P(i64)]),
    "rs_server_destroy": (None, [vp]),
    "rs_dev_montmul": (i32, [i32, P(u64), P(u64), P(u64), P(u64), i64]),
    "rs_dev_halfmul": (i32, [i32, P(u32), P(u64), P(u64), P(u64), i64]),
    "rs_dev_jacobi": (i32, [i32, P(u32), P(u64), P(i32), i64]),
    "rs_dev_row_masks": (i32, [i32, vp, P(u64), i64, P(u64)]),
    "rs_dev_alloc": (i32, [i32, i64, P(vp)]),
    "rs_dev_upload": (i32, [i32, vp, vp, i64]),
    "rs_dev_download": (i32, [i32, vp, vp, i64]),
    "rs_dev_free": (i32, [i32, vp]),
    "rs_dev_mad_rate": (i32, [i32, i32, i32, P(ctypes.c_double)]),
    "rs_dev_copy_rate": (i32, [i32, i64, P(ctypes.c_double)]),
    "rs_device_count": (i32, []),
}

_lib = None



# tile flags: active rows listed by row number (rsieve.h RS_TILE_ROW_INDEX);
# one baby table shared by the 4 or 8 wavefronts of a workgroup
# (RS_TILE_GROUP4, RS_TILE_GROUP8)
TILE_ROW_INDEX = 16
TILE_GROUP4 = 32
TILE_GROUP8 = 64
TILE_SPARSE = 128
TILE_CUBIC = 256

class RsError(RuntimeError):
    def __init__(self, what, status):
        self.status = status
        msg = lib().rs_strerror(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} ({status})")


def lib():
    """Loads librsieve.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _chk(what, st):
    if st != RS_OK:
        raise RsError(what, st)


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


def device_count():
    return lib().rs_device_count()


def strerror(st):
    return lib().rs_strerror(st).decode()


class Abcd:
    """An ABCD sieve file image (base, k list, n lists)."""

    h = vp()

    def __init__(self, handle):
        self.h = vp(handle)

    @classmethod
    def read(cls, path):
        h = vp()
        _chk(f"rs_abcd_read({path})", lib().rs_abcd_read(str(path).encode(), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_lists(cls, base, kn):
        """kn: dict k -> increasing list of n (insertion order kept)."""
        ks = np.array(list(kn.keys()), np.int32)
        counts = [len(v) for v in kn.values()]
        kstart = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        ns = np.array([n for v in kn.values() for n in v], np.int32)
        h = vp()
        _chk("rs_abcd_from_lists", lib().rs_abcd_from_lists(
            base, len(ks), _p(ks, i32), _p(kstart, i32), _p(ns, i32), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def synth(cls, base, nk, kmin, kmax, nmin, nmax, presieve_bound, seed, odd_k=True, c=-1):
        """odd_k: True/1 odd k only, 2 even k only, False/0 any parity;
        c: -1 for k*b^n-1, +1 for k*b^n+1."""
        h = vp()
        _chk("rs_abcd_synth", lib().rs_abcd_synth(base, nk, kmin, kmax, int(odd_k), c, nmin, nmax,
                                                  presieve_bound, seed, ctypes.byref(h)))
        return cls(h.value)

    @property
    def c(self):
        """Form k*b^n + c of the candidates (-1 or +1)."""
        v = i32()
        _chk("rs_abcd_form", lib().rs_abcd_form(self.h, ctypes.byref(v)))
        return v.value

    def set_form(self, c):
        _chk("rs_abcd_set_form", lib().rs_abcd_set_form(self.h, c))

    def remove(self, factors):
        """Drops the candidates (k, n) of factors [(p, k, n), ...]; returns how many."""
        arr = (Factor * max(1, len(factors)))()
        for i, (p, k, n) in enumerate(factors):
            arr[i] = Factor(p, k, n)
        gone = i64()
        _chk("rs_abcd_remove", lib().rs_abcd_remove(self.h, arr, len(factors), ctypes.byref(gone)))
        return gone.value

    def info(self):
        b, nk, nn, lo, hi = u32(), i32(), i32(), i32(), i32()
        _chk("rs_abcd_info", lib().rs_abcd_info(self.h, ctypes.byref(b), ctypes.byref(nk),
                                                ctypes.byref(nn), ctypes.byref(lo),
                                                ctypes.byref(hi)))
        return dict(base=b.value, nk=nk.value, nn=nn.value, minN=lo.value, maxN=hi.value)

    def lists(self):
        inf = self.info()
        ks = np.zeros(inf["nk"], np.int32)
        kstart = np.zeros(inf["nk"] + 1, np.int32)
        ns = np.zeros(inf["nn"], np.int32)
        _chk("rs_abcd_lists", lib().rs_abcd_lists(self.h, _p(ks, i32), _p(kstart, i32),
                                                  _p(ns, i32)))
        return ks, kstart, ns

    def write(self, path):
        _chk("rs_abcd_write", lib().rs_abcd_write(self.h, str(path).encode()))

    def __del__(self):
        if _lib is not None and self.h:
            _lib.rs_abcd_free(self.h)
            self.h = vp()


class Plan:
    """Subsequence plan: Q, rows (k, r) sorted, bitmaps, n window."""

    h = vp()

    def __init__(self, abcd, Qin=0):
        h = vp()
        _chk("rs_plan_create", lib().rs_plan_create(abcd.h, Qin, ctypes.byref(h)))
        self.h = h
        b, Q, nr, w, lo, hi = u32(), i32(), i32(), i32(), i32(), i32()
        _chk("rs_plan_info", lib().rs_plan_info(self.h, ctypes.byref(b), ctypes.byref(Q),
                                                ctypes.byref(nr), ctypes.byref(w),
                                                ctypes.byref(lo), ctypes.byref(hi)))
        self.base, self.Q, self.nrows, self.words = b.value, Q.value, nr.value, w.value
        self.minN, self.maxN = lo.value, hi.value
        cc = i32()
        _chk("rs_plan_form", lib().rs_plan_form(self.h, ctypes.byref(cc)))
        self.c = cc.value

    def rows(self):
        rk = np.zeros(self.nrows, np.int32)
        rr = np.zeros(self.nrows, np.int32)
        bm = np.zeros(self.nrows * self.words, np.uint32)
        _chk("rs_plan_rows", lib().rs_plan_rows(self.h, _p(rk, i32), _p(rr, i32), _p(bm, u32)))
        return rk, rr, bm.reshape(self.nrows, self.words)

    def contains(self, k, n):
        return bool(lib().rs_plan_contains(self.h, k, n))

    def tminmax(self, m):
        a, b = i32(), i32()
        _chk("rs_plan_tminmax", lib().rs_plan_tminmax(self.h, m, ctypes.byref(a),
                                                      ctypes.byref(b)))
        return a.value, b.value

    def auto_m(self):
        return lib().rs_plan_auto_m(self.h)

    def __del__(self):
        if _lib is not None and self.h:
            _lib.rs_plan_free(self.h)
            self.h = vp()


class PrimeGen:
    """True-prime generator (uint32 offsets from a uint64 base)."""

    h = vp()

    def __init__(self, hi, small_prime_scale=256, threads=0):
        h = vp()
        _chk("rs_primegen_create", lib().rs_primegen_create(small_prime_scale, hi, threads,
                                                            ctypes.byref(h)))
        self.h = h

    def range(self, lo, hi):
        n = i64()
        est = int((hi - lo) / max(1.0, np.log(max(hi, 3)) - 1.2) * 1.1) + 1024
        buf = np.zeros(est, np.uint32)
        st = lib().rs_primes_range(self.h, lo, hi, _p(buf, u32), len(buf), ctypes.byref(n))
        if st == RS_ERR_OVERFLOW:
            buf = np.zeros(n.value, np.uint32)
            st = lib().rs_primes_range(self.h, lo, hi, _p(buf, u32), len(buf), ctypes.byref(n))
        _chk("rs_primes_range", st)
        return buf[: n.value]

    def batch(self, low, count):
        buf = np.zeros(count, np.uint32)
        nxt = u64()
        _chk("rs_primes_batch", lib().rs_primes_batch(self.h, low, _p(buf, u32), count,
                                                      ctypes.byref(nxt)))
        return buf, nxt.value

    def __del__(self):
        if _lib is not None and self.h:
            _lib.rs_primegen_free(self.h)
            self.h = vp()


class Server:
    """Continuous-batching sieve server (jobs admitted by HBM budget)."""

    QUEUED, RUNNING, DONE, REJECTED = 0, 1, 2, -1
    h = vp()

    def __init__(self, hbm_budget, max_active=8, chunk_width=1 << 30, lds_cap=16384,
                 hit_capacity=1 << 20, device=0, quantum_ms=0.0):
        h = vp()
        _chk("rs_server_create", lib().rs_server_create(device, hbm_budget, max_active,
                                                        chunk_width, lds_cap, hit_capacity,
                                                        quantum_ms, ctypes.byref(h)))
        self.h = h

    def job_bytes(self, plan, dat=False):
        """Device bytes admission charges for a job of this plan."""
        v = i64()
        _chk("rs_server_job_bytes", lib().rs_server_job_bytes(self.h, plan.h, int(dat),
                                                              ctypes.byref(v)))
        return v.value

    def submit(self, plan, p_lo, p_hi, m=0, tile=0, dat=False, ns_per_prime=0.0):
        jid = i64()
        _chk("rs_server_submit", lib().rs_server_submit(self.h, plan.h, m, tile, int(dat), p_lo,
                                                        p_hi, ns_per_prime, ctypes.byref(jid)))
        return jid.value

    def step(self):
        """(active, queued, primes sieved in this step)."""
        a, q, n = i32(), i32(), i64()
        _chk("rs_server_step", lib().rs_server_step(self.h, ctypes.byref(a), ctypes.byref(q),
                                                    ctypes.byref(n)))
        return a.value, q.value, n.value

    def poll(self, jid):
        st, done, nh = i32(), u64(), i64()
        _chk("rs_server_poll", lib().rs_server_poll(self.h, jid, ctypes.byref(st),
                                                    ctypes.byref(done), ctypes.byref(nh)))
        return st.value, done.value, nh.value

    def results(self, jid):
        cnt = i64()
        cap = 1024
        while True:
            buf = (Factor * cap)()
            st = lib().rs_server_results(self.h, jid, buf, cap, ctypes.byref(cnt))
            if st == RS_ERR_OVERFLOW:
                cap = cnt.value
                continue
            _chk("rs_server_results", st)
            return [(int(f.p), int(f.k), int(f.n)) for f in buf[: cnt.value]]

    def release(self, jid):
        _chk("rs_server_release", lib().rs_server_release(self.h, jid))

    def stats(self):
        u, b, ms, n = i64(), i64(), ctypes.c_double(), i64()
        _chk("rs_server_stats", lib().rs_server_stats(self.h, ctypes.byref(u), ctypes.byref(b),
                                                      ctypes.byref(ms), ctypes.byref(n)))
        return dict(used=u.value, budget=b.value, kernel_ms=ms.value, primes=n.value)

    def close(self):
        if self.h:
            lib().rs_server_destroy(self.h)
            self.h = vp()

    def __del__(self):
        if _lib is not None and self.h:
            self.close()


class GpuPrimeGen:
    """Prime offsets generated on the device (into a DeviceBuffer)."""

    h = vp()

    def __init__(self, hi_max, device=0):
        h = vp()
        _chk("rs_gpu_primegen_create", lib().rs_gpu_primegen_create(device, hi_max,
                                                                    ctypes.byref(h)))
        self.h = h
        self.device = device

    def device_range(self, lo, hi):
        """Primes of [lo, hi) as offsets from lo in a DeviceBuffer (stays in HBM)."""
        cap = int(2 * (hi - lo) / max(1.0, np.log(max(hi - lo, 3)))) + 64
        buf = DeviceBuffer.empty(self.device, cap)
        n = i64()
        _chk("rs_gpu_primes_range", lib().rs_gpu_primes_range(self.h, lo, hi, buf.ptr, cap,
                                                              ctypes.byref(n)))
        buf.count = n.value
        return buf

    def range(self, lo, hi):
        """Primes of [lo, hi) as a host uint32 array (through a device buffer)."""
        cap = int(2 * (hi - lo) / max(1.0, np.log(max(hi - lo, 3)))) + 64
        buf = DeviceBuffer(self.device, np.zeros(cap, np.uint32))
        n = i64()
        _chk("rs_gpu_primes_range", lib().rs_gpu_primes_range(self.h, lo, hi, buf.ptr, cap,
                                                              ctypes.byref(n)))
        out = np.zeros(n.value, np.uint32)
        _chk("rs_dev_download", lib().rs_dev_download(self.device, out.ctypes.data_as(vp),
                                                      buf.ptr, out.nbytes))
        buf.free()
        return out

    def __del__(self):
        if _lib is not None and self.h:
            _lib.rs_gpu_primegen_free(self.h)
            self.h = vp()


def verify_factor(p, k, b, n, c=-1):
    return bool(lib().rs_verify_factor(p, k, b, n, c))


def format_factor(p, k, b, n, c=-1):
    f = Factor(p, k, n)
    buf = ctypes.create_string_buffer(128)
    r = lib().rs_format_factor(ctypes.byref(f), b, c, buf, 128)
    if r < 0:
        raise RsError("rs_format_factor", r)
    return buf.value.decode()


class DeviceBuffer:
    """Device memory holding uint32 prime offsets (HBM-resident input)."""

    ptr = vp()

    def __init__(self, device, host_array):
        a = np.ascontiguousarray(host_array)
        self.device, self.nbytes, self.count = device, a.nbytes, a.size
        self.ptr = vp()
        _chk("rs_dev_alloc", lib().rs_dev_alloc(device, max(a.nbytes, 4), ctypes.byref(self.ptr)))
        if a.nbytes:
            _chk("rs_dev_upload", lib().rs_dev_upload(device, self.ptr, a.ctypes.data_as(vp),
                                                      a.nbytes))

    @classmethod
    def empty(cls, device, count):
        """Uninitialised device buffer for `count` uint32 offsets."""
        self = cls.__new__(cls)
        self.device, self.nbytes, self.count = device, 4 * count, count
        self.ptr = vp()
        _chk("rs_dev_alloc", lib().rs_dev_alloc(device, max(4 * count, 4), ctypes.byref(self.ptr)))
        return self

    def download(self, dtype=np.uint32):
        """The buffer's contents as a numpy array."""
        out = np.zeros(self.nbytes // np.dtype(dtype).itemsize, dtype)
        _chk("rs_dev_download", lib().rs_dev_download(self.device, out.ctypes.data_as(vp),
                                                      self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().rs_dev_free(self.device, self.ptr)
            self.ptr = vp()

    def __del__(self):
        if _lib is not None and self.ptr:
            self.free()


class Sieve:
    """A sieve job resident on one GPU."""

    h = vp()

    def __init__(self, plan, m=0, max_batch=1 << 20, hit_capacity=1 << 20, device=0, tile=0):
        h = vp()
        _chk("rs_sieve_create", lib().rs_sieve_create(device, plan.h, m, tile, max_batch,
                                                      hit_capacity, ctypes.byref(h)))
        self.h = h
        self.plan = plan
        mm, T, slots, lds, tw = i32(), i32(), i32(), i32(), i32()
        _chk("rs_sieve_info", lib().rs_sieve_info(self.h, ctypes.byref(mm), ctypes.byref(T),
                                                  ctypes.byref(slots), ctypes.byref(lds),
                                                  ctypes.byref(tw)))
        self.m, self.T, self.slots, self.lds_bytes = mm.value, T.value, slots.value, lds.value
        self.tile = tw.value
        self.kernel = lib().rs_sieve_kernel_name(self.h).decode()

    def submit(self, low, offsets):
        offs = np.ascontiguousarray(offsets, np.uint32)
        _chk("rs_sieve_submit", lib().rs_sieve_submit(self.h, low, _p(offs, u32), len(offs)))

    def submit_device(self, low, dbuf, start=0, count=None):
        count = dbuf.count - start if count is None else count
        ptr = vp(dbuf.ptr.value + 4 * start)
        _chk("rs_sieve_submit_device", lib().rs_sieve_submit_device(self.h, low, ptr, count))

    def submit_device_noalloc(self, low, dbuf, hits_buf, nhits_buf, start=0, count=None):
        """Sieves device-resident offsets without allocating: hits into
        hits_buf (a DeviceBuffer of 6 words per hit), their number into
        nhits_buf (one word).  Asynchronous; read both after sync()."""
        count = dbuf.count - start if count is None else count
        ptr = vp(dbuf.ptr.value + 4 * start)
        _chk("rs_sieve_submit_device_noalloc",
             lib().rs_sieve_submit_device_noalloc(self.h, low, ptr, count, hits_buf.ptr,
                                                  hits_buf.count // 6, nhits_buf.ptr))

    def submit_range(self, lo, hi):
        """Primes of [lo, hi) generated on the device, then sieved."""
        _chk("rs_sieve_submit_range", lib().rs_sieve_submit_range(self.h, lo, hi))

    def collect(self):
        """All hits since the last collect as a sorted list of (p, k, n)."""
        cnt, lost = i64(), i64()
        cap = 4096
        while True:
            buf = (Factor * cap)()
            st = lib().rs_sieve_collect(self.h, buf, cap, ctypes.byref(cnt), ctypes.byref(lost))
            if st == RS_ERR_OVERFLOW:
                cap = cnt.value
                continue
            _chk("rs_sieve_collect", st)
            if lost.value:
                raise RuntimeError(f"device hit buffer overflowed: {lost.value} hits lost")
            return [(int(f.p), int(f.k), int(f.n)) for f in buf[: cnt.value]]

    def sync(self):
        _chk("rs_sieve_sync", lib().rs_sieve_sync(self.h))

    def timing(self):
        """Kernel ms and primes since the last call (HIP events)."""
        a, n = ctypes.c_double(), i64()
        _chk("rs_sieve_timing", lib().rs_sieve_timing(self.h, ctypes.byref(a), ctypes.byref(n)))
        return dict(kernel_ms=a.value, primes=n.value)

    def close(self):
        if self.h:
            lib().rs_sieve_destroy(self.h)
            self.h = vp()

    def __del__(self):
        if _lib is not None and self.h:
            self.close()


class Batch:
    """Many jobs sieved in one persistent (graph-captured) launch."""

    h = vp()

    def __init__(self, plans, ms=None, hit_capacity=1 << 20, device=0, tiles=None):
        hs = (vp * len(plans))(*[pl.h.value for pl in plans])
        msa = np.ascontiguousarray(ms if ms is not None else [0] * len(plans), np.int32)
        tla = np.ascontiguousarray(tiles if tiles is not None else [0] * len(plans), np.int32)
        h = vp()
        _chk("rs_batch_create", lib().rs_batch_create(device, len(plans), hs, _p(msa, i32),
                                                      _p(tla, i32), hit_capacity,
                                                      ctypes.byref(h)))
        self.h = h
        self.plans = plans
        self.info = []
        for j in range(len(plans)):
            mm, T, lds, tw = i32(), i32(), i32(), i32()
            _chk("rs_batch_job_info", lib().rs_batch_job_info(self.h, j, ctypes.byref(mm),
                                                              ctypes.byref(T), ctypes.byref(lds),
                                                              ctypes.byref(tw)))
            self.info.append(dict(m=mm.value, T=T.value, lds_bytes=lds.value, tile=tw.value))

    def add_chunk(self, job, low, dbuf, start=0, count=None):
        count = dbuf.count - start if count is None else count
        _chk("rs_batch_add_chunk", lib().rs_batch_add_chunk(self.h, job, low,
                                                            vp(dbuf.ptr.value + 4 * start),
                                                            count))

    def clear(self):
        _chk("rs_batch_clear", lib().rs_batch_clear(self.h))

    def run(self):
        _chk("rs_batch_run", lib().rs_batch_run(self.h))

    def collect(self):
        """Hits of the last run as a sorted list of (job, p, k, n)."""
        cnt, lost = i64(), i64()
        cap = 4096
        while True:
            buf = (JobFactor * cap)()
            st = lib().rs_batch_collect(self.h, buf, cap, ctypes.byref(cnt), ctypes.byref(lost))
            if st == RS_ERR_OVERFLOW:
                cap = cnt.value
                continue
            _chk("rs_batch_collect", st)
            if lost.value:
                raise RuntimeError(f"device hit buffer overflowed: {lost.value} hits lost")
            return [(int(f.job), int(f.p), int(f.k), int(f.n)) for f in buf[: cnt.value]]

    def timing(self):
        a, n = ctypes.c_double(), i64()
        _chk("rs_batch_timing", lib().rs_batch_timing(self.h, ctypes.byref(a), ctypes.byref(n)))
        return dict(kernel_ms=a.value, primes=n.value)

    def close(self):
        if self.h:
            lib().rs_batch_destroy(self.h)
            self.h = vp()

    def __del__(self):
        if _lib is not None and self.h:
            self.close()


def autotune(plan, low, dbuf, count=None, device=0, effort=0):
    """(best_m, best_tile, [(m, tile, ns_per_prime), ...]) measured on dbuf.
    effort 0: full (m, tile) grid; 1: coordinate descent."""
    count = dbuf.count if count is None else count
    bm, bt, n = i32(), i32(), i32()
    cap = 256
    cm = np.zeros(cap, np.int32)
    ct = np.zeros(cap, np.int32)
    ns = np.zeros(cap, np.float64)
    _chk("rs_autotune", lib().rs_autotune(device, plan.h, low, dbuf.ptr, count, effort,
                                          ctypes.byref(bm),
                                          ctypes.byref(bt), _p(cm, i32), _p(ct, i32),
                                          _p(ns, ctypes.c_double), cap, ctypes.byref(n)))
    k = min(n.value, cap)
    return bm.value, bt.value, list(zip(cm[:k].tolist(), ct[:k].tolist(), ns[:k].tolist()))


def dev_montmul(a, b, p, device=0):
    a, b, p = (np.ascontiguousarray(x, np.uint64) for x in (a, b, p))
    out = np.zeros_like(a)
    _chk("rs_dev_montmul", lib().rs_dev_montmul(device, _p(a, u64), _p(b, u64), _p(p, u64),
                                                _p(out, u64), len(a)))
    return out


def dev_halfmul(k, y, p, device=0):
    k = np.ascontiguousarray(k, np.uint32)
    y, p = (np.ascontiguousarray(x, np.uint64) for x in (y, p))
    out = np.zeros_like(y)
    _chk("rs_dev_halfmul", lib().rs_dev_halfmul(device, _p(k, u32), _p(y, u64), _p(p, u64),
                                                _p(out, u64), len(k)))
    return out


def dev_jacobi(a, p, device=0):
    a = np.ascontiguousarray(a, np.uint32)
    p = np.ascontiguousarray(p, np.uint64)
    out = np.zeros(len(a), np.int32)
    _chk("rs_dev_jacobi", lib().rs_dev_jacobi(device, _p(a, u32), _p(p, u64), _p(out, i32),
                                              len(a)))
    return out


def dev_row_masks(plan, primes, device=0):
    """The sieve kernel's own row filter for each prime: a (len(primes), NW)
    uint64 array, bit j of word w set iff row 64 w + j passes."""
    p = np.ascontiguousarray(primes, np.uint64)
    nw = (plan.nrows + 63) // 64
    out = np.zeros((len(p), nw), np.uint64)
    _chk("rs_dev_row_masks", lib().rs_dev_row_masks(device, plan.h, _p(p, u64), len(p),
                                                    _p(out, u64)))
    return out


def dev_mad_rate(blocks=8192, iters=4096, device=0):
    r = ctypes.c_double()
    _chk("rs_dev_mad_rate", lib().rs_dev_mad_rate(device, blocks, iters, ctypes.byref(r)))
    return r.value


def dev_copy_rate(nbytes=1 << 30, device=0):
    r = ctypes.c_double()
    _chk("rs_dev_copy_rate", lib().rs_dev_copy_rate(device, nbytes, ctypes.byref(r)))
    return r.value
