// Deterministic scan and stable LSD radix sort for gfx950 (wave64).
//
// The radix sort replaces the x-stripe colouring of PhysBAM's
// GATHER_SCATTER::Prepare_Scatter (Hybrid_Methods/Iterators/GATHER_SCATTER.cpp:32-86):
// particles are ordered by the tile that owns their stencil base so that P2G and
// G2P work on block-owned grid tiles.  Each pass is histogram -> scan ->
// stable scatter; ranks inside a block come from wave ballots, so the output
// equals std::stable_sort of the same keys.
#include "mpmx_sort.h"

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

}  // namespace mpmx
