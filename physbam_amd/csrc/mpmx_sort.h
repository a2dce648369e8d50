// Deterministic device primitives: exclusive scan of uint32 arrays and a
// stable LSD radix sort of (key, slot) pairs.  Both are bit-reproducible: no
// atomics decide an output position.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mpmx {

// Scratch sizes (bytes) needed by the primitives for n items.
size_t scan_scratch_bytes(uint32_t n);
size_t sort_scratch_bytes(uint32_t n, int key_bits);

// out[i] = sum_{j<i} in[i]; in == out allowed.  total (device, may be null)
// receives the sum of all elements.
void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* scratch,
                    hipStream_t stream);

// Stable sort of keys[0..n) with values = slot index 0..n-1.  Sorted keys and
// slots end in keys_out / vals_out; keys_tmp / vals_tmp are ping-pong buffers
// of n entries.  Only the low key_bits bits of each key take part, except that
// 0xffffffff (invalid) always sorts last.
void radix_sort_pairs(const uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_out, uint32_t* keys_tmp,
                      uint32_t* vals_tmp, uint32_t n, int key_bits, void* scratch, hipStream_t stream);

// The sort of radix_sort_pairs on device arrays without allocating: the
// ping-pong arrays and the histogram live in a static device array of
// sort_noalloc_scratch_bytes() bytes on each GPU (at most SORT_NOALLOC_MAX_KEYS
// keys).  `device` is the GPU the stream belongs to; its array's address is
// looked up once (hipGetSymbolAddress with that device current).  Calls that
// overlap on one GPU share the array and race.  Throws std::runtime_error on
// a count above the maximum or a failed lookup.
constexpr uint32_t SORT_NOALLOC_MAX_KEYS = 1u << 20;
size_t sort_noalloc_scratch_bytes();
void radix_sort_pairs_noalloc(const uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                              int key_bits, int device, hipStream_t stream);

}  // namespace mpmx
