// diag.hip -- device helpers behind the C ABI that are not the sieve
// itself: exact-arithmetic checks for tests, device buffers for
// HBM-resident inputs, and the multiply-add throughput probe.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../host/host.h"

extern "C" __global__ void rs_test_montmul(const uint64_t*, const uint64_t*, const uint64_t*,
                                           uint64_t*, long long);
extern "C" __global__ void rs_test_halfmul(const uint32_t*, const uint64_t*, const uint64_t*,
                                           uint64_t*, long long);
extern "C" __global__ void rs_test_jacobi(const uint32_t*, const uint64_t*, int32_t*, long long);
extern "C" __global__ void rs_mad_probe(uint64_t*, int);
extern "C" __global__ void rs_copy_probe(const uint4*, uint4*, long long);

#define HIPCHK(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) return RS_ERR_HIP; \
    } while (0)

// ------------------------------------------------------------ test helpers

template <typename F>
static rs_status run_elementwise(int device, int64_t count, F launch_fn) {
    if (count < 0) return RS_ERR_ARG;
    if (count == 0) return RS_OK;
    HIPCHK(hipSetDevice(device));
    return launch_fn();
}

extern "C" rs_status rs_dev_montmul(int device, const uint64_t* a, const uint64_t* b,
                                    const uint64_t* p, uint64_t* out, int64_t count) {
    if (!a || !b || !p || !out) return RS_ERR_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (p[i] % 2 == 0 || p[i] < 3 || p[i] >= (1ull << 63) || b[i] >= p[i]) return RS_ERR_ARG;
    return run_elementwise(device, count, [&]() -> rs_status {
        uint64_t *da, *db, *dp, *dout;
        const size_t sz = (size_t)count * 8;
        HIPCHK(hipMalloc(&da, sz));
        HIPCHK(hipMalloc(&db, sz));
        HIPCHK(hipMalloc(&dp, sz));
        HIPCHK(hipMalloc(&dout, sz));
        HIPCHK(hipMemcpy(da, a, sz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(db, b, sz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp, p, sz, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(rs_test_montmul, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           0, da, db, dp, dout, (long long)count);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, sz, hipMemcpyDeviceToHost));
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dp);
        (void)hipFree(dout);
        return RS_OK;
    });
}

extern "C" rs_status rs_dev_halfmul(int device, const uint32_t* k, const uint64_t* y,
                                    const uint64_t* p, uint64_t* out, int64_t count) {
    if (!k || !y || !p || !out) return RS_ERR_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (p[i] % 2 == 0 || p[i] < 3 || p[i] >= (1ull << 63) || y[i] >= p[i]) return RS_ERR_ARG;
    return run_elementwise(device, count, [&]() -> rs_status {
        uint32_t* dk;
        uint64_t *dy, *dp, *dout;
        const size_t sz = (size_t)count * 8;
        HIPCHK(hipMalloc(&dk, (size_t)count * 4));
        HIPCHK(hipMalloc(&dy, sz));
        HIPCHK(hipMalloc(&dp, sz));
        HIPCHK(hipMalloc(&dout, sz));
        HIPCHK(hipMemcpy(dk, k, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dy, y, sz, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp, p, sz, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(rs_test_halfmul, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           0, dk, dy, dp, dout, (long long)count);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, sz, hipMemcpyDeviceToHost));
        (void)hipFree(dk);
        (void)hipFree(dy);
        (void)hipFree(dp);
        (void)hipFree(dout);
        return RS_OK;
    });
}

extern "C" rs_status rs_dev_jacobi(int device, const uint32_t* a, const uint64_t* p,
                                   int32_t* out, int64_t count) {
    if (!a || !p || !out) return RS_ERR_ARG;
    for (int64_t i = 0; i < count; ++i)
        if (p[i] % 2 == 0 || p[i] < 3) return RS_ERR_ARG;
    return run_elementwise(device, count, [&]() -> rs_status {
        uint32_t* da;
        uint64_t* dp;
        int32_t* dout;
        HIPCHK(hipMalloc(&da, (size_t)count * 4));
        HIPCHK(hipMalloc(&dp, (size_t)count * 8));
        HIPCHK(hipMalloc(&dout, (size_t)count * 4));
        HIPCHK(hipMemcpy(da, a, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dp, p, (size_t)count * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(rs_test_jacobi, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, 0,
                           da, dp, dout, (long long)count);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, (size_t)count * 4, hipMemcpyDeviceToHost));
        (void)hipFree(da);
        (void)hipFree(dp);
        (void)hipFree(dout);
        return RS_OK;
    });
}

// ------------------------------------------------------------ device buffers

extern "C" rs_status rs_dev_alloc(int device, int64_t bytes, void** ptr) {
    if (!ptr || bytes <= 0) return RS_ERR_ARG;
    *ptr = nullptr;
    HIPCHK(hipSetDevice(device));
    if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) return RS_ERR_NOMEM;
    return RS_OK;
}

extern "C" rs_status rs_dev_upload(int device, void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return RS_OK;
}

extern "C" rs_status rs_dev_download(int device, void* dst, const void* src, int64_t bytes) {
    if (!dst || !src || bytes < 0) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return RS_OK;
}

extern "C" rs_status rs_dev_free(int device, void* ptr) {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipFree(ptr));
    return RS_OK;
}

// Owners of a device buffer and of two events, released on every return.
namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace

extern "C" rs_status rs_dev_mad_rate(int device, int blocks, int iters, double* ops_per_s) {
    if (!ops_per_s || blocks < 1 || iters < 1) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    DevBuf out;
    if (hipMalloc(&out.p, (size_t)blocks * 256 * 8) != hipSuccess) return RS_ERR_NOMEM;
    uint64_t* const o = (uint64_t*)out.p;
    EventPair ev;
    HIPCHK(hipEventCreate(&ev.a));
    HIPCHK(hipEventCreate(&ev.b));
    hipLaunchKernelGGL(rs_mad_probe, dim3(blocks), dim3(256), 0, 0, o, iters);  // warm-up
    HIPCHK(hipEventRecord(ev.a, 0));
    hipLaunchKernelGGL(rs_mad_probe, dim3(blocks), dim3(256), 0, 0, o, iters);
    HIPCHK(hipEventRecord(ev.b, 0));
    HIPCHK(hipEventSynchronize(ev.b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev.a, ev.b));
    *ops_per_s = (double)blocks * 256.0 * 8.0 * iters / (ms * 1e-3);
    return RS_OK;
}

extern "C" rs_status rs_dev_copy_rate(int device, int64_t bytes, double* bytes_per_s) {
    if (!bytes_per_s || bytes < 16 || bytes % 16 != 0) return RS_ERR_ARG;
    HIPCHK(hipSetDevice(device));
    DevBuf sbuf, dbuf;
    if (hipMalloc(&sbuf.p, (size_t)bytes) != hipSuccess ||
        hipMalloc(&dbuf.p, (size_t)bytes) != hipSuccess)
        return RS_ERR_NOMEM;
    const uint4* const src = (const uint4*)sbuf.p;
    uint4* const dst = (uint4*)dbuf.p;
    HIPCHK(hipMemset(sbuf.p, 1, (size_t)bytes));
    EventPair ev;
    HIPCHK(hipEventCreate(&ev.a));
    HIPCHK(hipEventCreate(&ev.b));
    const long long n = bytes / 16;
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int grid = cus * 16;
    hipLaunchKernelGGL(rs_copy_probe, dim3(grid), dim3(256), 0, 0, src, dst, n);  // warm-up
    HIPCHK(hipEventRecord(ev.a, 0));
    hipLaunchKernelGGL(rs_copy_probe, dim3(grid), dim3(256), 0, 0, src, dst, n);
    HIPCHK(hipEventRecord(ev.b, 0));
    HIPCHK(hipEventSynchronize(ev.b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev.a, ev.b));
    *bytes_per_s = 2.0 * (double)bytes / (ms * 1e-3);  // read + write
    return RS_OK;
}
