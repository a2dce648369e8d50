// Deterministic scan and stable LSD radix sort for gfx950 (wave64).
//
// The radix sort replaces the x-stripe colouring of PhysBAM's
// GATHER_SCATTER::Prepare_Scatter (Hybrid_Methods/Iterators/GATHER_SCATTER.cpp:32-86):
// particles are ordered by the tile that owns their stencil base so that P2G and
// G2P work on block-owned grid tiles.  Each pass is histogram -> scan ->
// stable scatter; ranks inside a block come from wave ballots, so the output
// equals std::stable_sort of the same keys.
#include "mpmx_sort.h"

#include <mutex>
#include <stdexcept>
#include <string>

namespace mpmx {

namespace {

constexpr int WAVE = 64;
constexpr int THREADS = 256;
constexpr int WAVES = THREADS / WAVE;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = THREADS * SCAN_ITEMS;  // 2048
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = THREADS * SORT_ITEMS;  // 4096
constexpr int MAX_DIGIT_BITS = 9;

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & (WAVE - 1);
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// Exclusive scan across a 256-thread block; returns the prefix, sets total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(x, o, WAVE);
    if (lane >= o) x += y;
  }
  if (lane == WAVE - 1) s_wave[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < WAVES; i++) {
    uint32_t t = s_wave[i];
    off += i < w ? t : 0u;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

__global__ __launch_bounds__(THREADS) void scan_reduce_kernel(const uint32_t* in, uint32_t n, uint32_t* partial) {
  __shared__ uint32_t s_wave[WAVES];
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) sum += base + k < n ? in[base + k] : 0u;
  uint32_t total;
  block_exclusive_scan(sum, s_wave, total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Single block: exclusive scan of the per-tile partial sums (any count).
__global__ __launch_bounds__(THREADS) void scan_partials_kernel(uint32_t* partial, uint32_t np, uint32_t* total_out) {
  __shared__ uint32_t s_wave[WAVES];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < np; base += THREADS) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < np ? partial[i] : 0u;
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(v, s_wave, tot);
    if (i < np) partial[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(THREADS) void scan_downsweep_kernel(const uint32_t* in, uint32_t* out, uint32_t n,
                                                                 const uint32_t* partial) {
  __shared__ uint32_t s_wave[WAVES];
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < n ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, s_wave, total) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
}

// Mask of active lanes whose digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, int bits, bool active) {
  uint64_t m = __ballot(active);
  for (int b = 0; b < bits; b++) {
    const bool bit = (digit >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

__global__ __launch_bounds__(THREADS) void radix_upsweep_kernel(const uint32_t* keys, uint32_t n, int shift, int bits,
                                                                uint32_t* hist, uint32_t nblocks) {
  __shared__ uint32_t h[1 << MAX_DIGIT_BITS];
  const uint32_t nb = 1u << bits, mask = nb - 1;
  for (uint32_t i = threadIdx.x; i < nb; i += THREADS) h[i] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * SORT_TILE;
  const int lane = threadIdx.x & (WAVE - 1);
  for (int k = 0; k < SORT_ITEMS; k++) {
    const uint32_t i = base + k * THREADS + threadIdx.x;
    const bool active = i < n;
    const uint32_t d = active ? (keys[i] >> shift) & mask : 0u;
    const uint64_t m = match_digit(d, bits, active);
    if (active && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&h[d], (uint32_t)__popcll(m));
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nb; i += THREADS) hist[i * nblocks + blockIdx.x] = h[i];
}

__global__ __launch_bounds__(THREADS) void radix_downsweep_kernel(const uint32_t* keys_in, const uint32_t* vals_in,
                                                                  uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                                                                  int shift, int bits, const uint32_t* hist,
                                                                  uint32_t nblocks) {
  __shared__ uint32_t running[1 << MAX_DIGIT_BITS];
  __shared__ uint32_t wcnt[WAVES][1 << MAX_DIGIT_BITS];
  const uint32_t nb = 1u << bits, mask = nb - 1;
  for (uint32_t i = threadIdx.x; i < nb; i += THREADS) {
    running[i] = hist[i * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < WAVES; w++) wcnt[w][i] = 0;
  }
  __syncthreads();
  const uint32_t base = blockIdx.x * SORT_TILE;
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  for (int k = 0; k < SORT_ITEMS; k++) {
    const uint32_t i = base + k * THREADS + threadIdx.x;
    const bool active = i < n;
    const uint32_t key = active ? keys_in[i] : 0u;
    const uint32_t val = active ? (vals_in ? vals_in[i] : i) : 0u;
    const uint32_t d = (key >> shift) & mask;
    const uint64_t m = match_digit(d, bits, active);
    const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt());
    if (active && lane == __ffsll((unsigned long long)m) - 1) wcnt[w][d] = (uint32_t)__popcll(m);
    __syncthreads();
    if (active) {
      uint32_t pos = running[d] + rank;
      for (int q = 0; q < w; q++) pos += wcnt[q][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nb; j += THREADS) {
      uint32_t s = 0;
#pragma unroll
      for (int q = 0; q < WAVES; q++) { s += wcnt[q][j]; wcnt[q][j] = 0; }
      running[j] += s;
    }
    __syncthreads();
  }
}

inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

}  // namespace

size_t scan_scratch_bytes(uint32_t n) { return sizeof(uint32_t) * (cdiv(n, SCAN_TILE) + 1); }

void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* scratch,
                    hipStream_t stream) {
  if (n == 0) return;
  uint32_t* partial = (uint32_t*)scratch;
  const uint32_t nb = cdiv(n, SCAN_TILE);
  scan_reduce_kernel<<<nb, THREADS, 0, stream>>>(in, n, partial);
  scan_partials_kernel<<<1, THREADS, 0, stream>>>(partial, nb, total);
  scan_downsweep_kernel<<<nb, THREADS, 0, stream>>>(in, out, n, partial);
}

size_t sort_scratch_bytes(uint32_t n, int key_bits) {
  (void)key_bits;
  const uint32_t nblocks = cdiv(n > 0 ? n : 1, SORT_TILE);
  const uint32_t hist = nblocks << MAX_DIGIT_BITS;
  return sizeof(uint32_t) * hist + scan_scratch_bytes(hist) + 256;
}

void radix_sort_pairs(const uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_out, uint32_t* keys_tmp,
                      uint32_t* vals_tmp, uint32_t n, int key_bits, void* scratch, hipStream_t stream) {
  if (n == 0) return;
  const uint32_t nblocks = cdiv(n, SORT_TILE);
  uint32_t* hist = (uint32_t*)scratch;
  void* scan_scratch = (void*)(hist + ((size_t)nblocks << MAX_DIGIT_BITS));
  if (key_bits < 1) key_bits = 1;
  const int passes = (key_bits + MAX_DIGIT_BITS - 1) / MAX_DIGIT_BITS;
  const int per = (key_bits + passes - 1) / passes;
  // Ping-pong so that the last pass lands in keys_out / vals_out.
  const uint32_t* kin = keys_in;
  const uint32_t* vin = nullptr;
  int shift = 0;
  for (int p = 0; p < passes; p++) {
    const int bits = (p == passes - 1) ? key_bits - shift : per;
    const bool last = p == passes - 1;
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint32_t* kout = to_out ? keys_out : keys_tmp;
    uint32_t* vout = to_out ? vals_out : vals_tmp;
    (void)last;
    radix_upsweep_kernel<<<nblocks, THREADS, 0, stream>>>(kin, n, shift, bits, hist, nblocks);
    scan_exclusive(hist, hist, nblocks << bits, nullptr, scan_scratch, stream);
    radix_downsweep_kernel<<<nblocks, THREADS, 0, stream>>>(kin, vin, kout, vout, n, shift, bits, hist, nblocks);
    kin = kout;
    vin = vout;
    shift += bits;
  }
}

// ------------------------------------------------- sort without allocation --
namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// Layout of the static array: keys_tmp and vals_tmp (SORT_NOALLOC_MAX_KEYS
// uint32 each), then the histogram and scan partials of sort_scratch_bytes for
// that count (recomputed here as constants: nblocks << MAX_DIGIT_BITS
// histogram entries, one scan partial per SCAN_TILE of them plus one, and the
// 256-byte tail).
constexpr size_t NOALLOC_TMP_BYTES = align256((size_t)SORT_NOALLOC_MAX_KEYS * sizeof(uint32_t));
constexpr size_t NOALLOC_HIST = ((size_t)(SORT_NOALLOC_MAX_KEYS + SORT_TILE - 1) / SORT_TILE) << MAX_DIGIT_BITS;
constexpr size_t NOALLOC_SORT_BYTES =
    sizeof(uint32_t) * NOALLOC_HIST + sizeof(uint32_t) * ((NOALLOC_HIST + SCAN_TILE - 1) / SCAN_TILE + 1) + 256;
constexpr size_t NOALLOC_BYTES = 2 * NOALLOC_TMP_BYTES + align256(NOALLOC_SORT_BYTES);

// One copy per GPU, held until the process ends.  In host code this name is
// a stand-in, not a device address: only the pointer that hipGetSymbolAddress
// returns for the current device is passed to kernels.
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `device`'s array, looked up once per device with that
// device current (the caller's current device is restored).
void* scratch_on(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return g_scratch_ptr[device];
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (cur != device && (e = hipSetDevice(device)) != hipSuccess)
    throw std::runtime_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (cur != device) (void)hipSetDevice(cur);
  if (e != hipSuccess || !ptr)
    throw std::runtime_error(std::string("hipGetSymbolAddress of the sort scratch array: ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return ptr;
}

}  // namespace

size_t sort_noalloc_scratch_bytes() { return NOALLOC_BYTES; }

void radix_sort_pairs_noalloc(const uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                              int key_bits, int device, hipStream_t stream) {
  if (n > SORT_NOALLOC_MAX_KEYS)
    throw std::runtime_error("the sort without allocation takes at most " + std::to_string(SORT_NOALLOC_MAX_KEYS) +
                             " keys, got " + std::to_string(n));
  unsigned char* base = (unsigned char*)scratch_on(device);
  if (n == 0) return;
  if (sort_scratch_bytes(n, key_bits) > NOALLOC_SORT_BYTES) throw std::runtime_error("sort scratch layout mismatch");
  uint32_t* keys_tmp = (uint32_t*)base;
  uint32_t* vals_tmp = (uint32_t*)(base + NOALLOC_TMP_BYTES);
  void* sort_scratch = (void*)(base + 2 * NOALLOC_TMP_BYTES);
  radix_sort_pairs(keys_in, keys_out, vals_out, keys_tmp, vals_tmp, n, key_bits, sort_scratch, stream);
}

}  // namespace mpmx
