// gpuprimes.h -- internal interface of the GPU prime generator used by the
// device sieve (the public one is rs_gpu_primes_range in rsieve.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rsieve.h"

// Enqueues the generation of every prime of [lo, hi) as offsets p - lo into
// out (capacity cap) on `stream`; *count_dev receives the device address of
// the prime count, valid once the stream reaches this point.  Calls share
// the generator's work buffers, so they must be ordered (one stream).
rs_status gpu_primes_enqueue(rs_gpu_primegen* g, uint64_t lo, uint64_t hi, uint32_t* out,
                             int64_t cap, hipStream_t stream, const uint32_t** count_dev);

// Upper bound on the number of primes in a window of the given width.
int64_t gpu_primes_capacity(uint64_t width);

// The largest hi the generator accepts.
uint64_t rs_gpu_primegen_hi(const rs_gpu_primegen* g);
