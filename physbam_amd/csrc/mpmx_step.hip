// Explicit APIC MPM step on gfx950: binning, P2G with fused stress, grid
// update with gravity and walls, G2P with the F/B/X/V update.
//
// Reference step (Public_Library/Hybrid_Methods/Examples_And_Drivers/MPM_DRIVER.cpp):
//   Update_Simulated_Particles :1142-1157   -> invalid particles carry INVALID_KEY
//   Update_Particle_Weights    :233-241     -> recomputed in registers (weights())
//   Prepare_Scatter            GATHER_SCATTER.cpp:32-86 -> radix sort by tile
//   Particle_To_Grid           :245-306     -> p2g_kernel
//   Apply_Forces (symplectic)  :959-1060    -> stress folded into p2g_kernel,
//                                              gravity/walls/friction in grid_kernel
//   Grid_To_Particle           :450-515     -> g2p_kernel
//   Reflect_Or_Invalidate      :1219-1239   -> g2p_kernel
//
// Data layout in HBM
//   particles: structure of arrays, component-major (X[a*n + p], F[k*n + p]
//              with k the column-major entry), stored in tile order and
//              re-gathered by the binning permutation every few steps;
//   grid:      tile-blocked nodes (64 per tile: 4^3 in 3D, 8^2 in 2D), each
//              node {m, p_x, p_y, p_z} (momentum during P2G, velocity after the
//              grid update); only tiles touched this step are cleared and used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "mpmx.h"
#include "mpmx_adjoint.h"
#include "mpmx_device.h"
#include "mpmx_engine.h"
#include "mpmx_sort.h"

namespace mpmx {

#define HIPCHECK(x)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) throw Error{MPMX_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)

constexpr int BLOCK = 256;
constexpr int EXTRA_BLOCKS = 256;  // grid of the mover / slow-path kernels (grid-stride)

template<typename T, int D> struct Args {
  Geom<T, D> g;                // geometry of scene 0 (host-side use and the grid export)
  const Geom<T, D>* scenes;    // per-scene geometry (one entry for one scene)
  const uint16_t* tscene;      // tile -> scene (null: one scene)
  const uint16_t* pscene;      // particle slot -> scene (null: one scene)
  uint32_t ntotal;             // tiles of all scenes
  T dt;
  T grav[D];
  int model;
  T thr, root;
  int nwalls;
  WallDev<T> walls[MAX_WALLS];
  T* X; T* V; T* F; T* B;
  const T* mass; const T* vol; const T* mu; const T* la;
  uint8_t* valid;
  uint32_t* keys;
  uint32_t n;
  const uint32_t* order;  // sorted position -> slot, null for identity
  const uint32_t* tile_begin;
  const uint32_t* tile_end;
  const uint32_t* ptiles;
  const uint32_t* gtiles;
  // [0] binned tiles, [1] grid tiles, [3] slow-stress particles, [4] movers
  uint32_t* counts;
  uint32_t* slow_list;   // particles whose stress needs the FP64 SVD path
  uint32_t* mover_list;  // particles whose stencil base left their binned tile
  uint32_t* mover_tile;  // ... and that tile
  uint32_t* mstamp;      // per tile: step stamp of the movers whose base is in it
  uint32_t next_stamp;   // stamp of the next step (written by G2P)
  Node4<T>* grid;
  Node4<T>* gridf;
  Node4<T>* gridu;  // if set, the grid update also stores (m, v before the collision objects)
  // if set, the grid update also zeroes these tiles of this grid (the fused
  // step's other grid, as its P2G of two steps back left it)
  Node4<T>* clear_grid;
  const uint32_t* clear_tiles;
  const uint32_t* clear_count;
  // fused G2P -> P2G: P2G target of the next step and the next step's mover
  // list (movers are appended from the bottom by the tile kernel and from the
  // top, index mcap-1-k, by the mover kernel; counts[5] and counts[7] count
  // them; the current list counts are counts[4] and counts[6])
  Node4<T>* grid_next;
  uint32_t* mover_list2;
  uint32_t* mover_tile2;
  uint32_t mcap;
  uint32_t* err;
  // reflection boundary conditions: bit 2a (2a+1) set when particles leaving
  // through the low (high) side of axis a are reflected instead of
  // invalidated; band: the tiles along the domain sides are always active
  uint32_t reflect_sides;
  int band;
};

// The geometry always comes from the scene table (one entry for one scene),
// so only the fields a kernel uses are loaded.
template<typename T, int D> __device__ __forceinline__ Geom<T, D> tile_geom(const Args<T, D>& a, uint32_t t) {
  return a.scenes[a.tscene ? a.tscene[t] : 0];
}
template<typename T, int D> __device__ __forceinline__ Geom<T, D> particle_geom(const Args<T, D>& a, uint32_t p) {
  return a.scenes[a.pscene ? a.pscene[p] : 0];
}

// Node origin of tile t (node coordinates of its scene).
template<typename T, int D> __device__ __forceinline__ void tile_origin(uint32_t t, const Geom<T, D>& g, int* o) {
  t -= g.tile0;
#pragma unroll
  for (int a = D - 1; a >= 0; a--) {
    o[a] = (int)(t % (uint32_t)g.GT[a]) * Tile<D>::E;
    t /= (uint32_t)g.GT[a];
  }
}

// Storage slot of node nc of a scene: tile-blocked, 64 nodes per tile.
template<typename T, int D> __device__ __forceinline__ uint32_t node_slot(const int* nc, const Geom<T, D>& g) {
  uint32_t t = 0, l = 0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    t = t * (uint32_t)g.GT[a] + (uint32_t)(nc[a] / Tile<D>::E);
    l = l * Tile<D>::E + (uint32_t)(nc[a] % Tile<D>::E);
  }
  return (g.tile0 + t) * 64u + l;
}

// Region-local flat index -> offsets (row-major, last axis fastest).
template<int D> __device__ __forceinline__ void region_offsets(int k, int* off) {
#pragma unroll
  for (int a = D - 1; a >= 0; a--) {
    off[a] = k % Tile<D>::R;
    k /= Tile<D>::R;
  }
}

template<typename T> __device__ __forceinline__ void atomic_add(T* p, T v) { atomicAdd(p, v); }

// List append with one atomic per wavefront: every lane with `want` gets its
// own entry index.  Must be reached by all lanes of the wavefront (it votes).
// One returning atomic per lane on a shared counter would serialise in L2.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
  const uint64_t m = __ballot(want);
  if (m == 0) return 0;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  return base + (uint32_t)__popcll(m & below);
}

// Entry of a list of n that work item r takes.  Neighbouring list entries are
// usually neighbours in space, so consecutive lanes are sent n/64 entries
// apart: their node atomics then rarely hit the same address.  A bijection on
// [0, n): the first 64*floor(n/64) entries are transposed, the rest kept.
__device__ __forceinline__ uint32_t spread_index(uint32_t r, uint32_t n) {
  const uint32_t m = n >> 6;
  return r < (m << 6) ? (r & 63u) * m + (r >> 6) : r;
}

// Tile holding the stencil base of x (node coordinates, ghost offset included).
template<typename T, int D> __device__ __forceinline__ uint32_t base_tile(const Geom<T, D>& g, const T* x) {
  uint32_t t = 0;
#pragma unroll
  for (int a = 0; a < D; a++) t = t * (uint32_t)g.GT[a] + (uint32_t)((stencil_base<T, D>(g, x, a) + 3) / Tile<D>::E);
  return g.tile0 + t;
}

// ----------------------------------------------------------------- bins ----
// Particles are binned by the tile holding their stencil base.  Bins are
// rebuilt (key -> stable radix sort -> tile ranges) only when enough particles
// have left their binned tile; in between, such particles ("movers") are listed
// by P2G each step and transferred through HBM by p2g_extra_kernel and
// g2p_mover_kernel, and the tiles holding them are marked in mstamp so that the
// grid tiles they touch are cleared and updated.
// Binning key: the tile holding the particle's stencil base times 64 plus the
// base's cell inside that tile, so that the sort groups particles by tile
// and, inside a tile, by cell.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void key_kernel(Args<T, D> a, uint32_t* range_err) {
  const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
  if (p >= a.n) return;
  uint32_t key = INVALID_KEY;
  if (a.valid[p]) {
    T x[D];
#pragma unroll
    for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
    const Geom<T, D> g = particle_geom(a, p);
    key = tile_key<T, D>(g, x);
    if (key == INVALID_KEY) {
      atomicAdd(range_err, 1u);
    } else {
      int lc[D];
#pragma unroll
      for (int c = 0; c < D; c++) lc[c] = (stencil_base<T, D>(g, x, c) + 3) % Tile<D>::E;
      key = key * 64u + cell_label<D>(lc);
    }
  }
  a.keys[p] = key;
}

// Stratified order inside each tile.  The sort leaves a tile's particles
// grouped by cell; with K = round(n_t / 64) (the tile's particles per cell),
// position j of the tile goes to stratum j % K, and the strata are laid out
// one after another: a run of 64 consecutive positions then holds about one
// particle of every cell, so the fused kernel's 256-particle chunks hold about
// four of each (one round of its P2G).  A bijection on the tile's positions.
__global__ __launch_bounds__(BLOCK) void stratify_kernel(const uint32_t* sk, const uint32_t* order_in,
                                                         uint32_t* order_out, const uint32_t* tb,
                                                         const uint32_t* te, uint32_t n) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t key = sk[i];
  if (key == INVALID_KEY) {
    order_out[i] = order_in[i];
    return;
  }
  const uint32_t t = key >> 6, b = tb[t], nt = te[t] - b, j = i - b;
  const uint32_t K = max(1u, (nt + 32u) >> 6);
  const uint32_t Q = nt / K, Rm = nt % K, sidx = j % K, q = j / K;
  order_out[b + sidx * Q + min(sidx, Rm) + q] = order_in[i];
}

__global__ __launch_bounds__(BLOCK) void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = v;
}

// Tile ranges of the sorted keys.  A tile is binned iff its stamp equals the
// current bin stamp, so nothing has to be cleared between rebuilds.
__global__ __launch_bounds__(BLOCK) void bounds_kernel(const uint32_t* sk, uint32_t n, uint32_t* tb, uint32_t* te,
                                                       uint32_t* stamp_arr, uint32_t stamp) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  if (sk[i] == INVALID_KEY) return;
  const uint32_t t = sk[i] >> 6;  // keys are tile * 64 + cell
  if (i == 0 || (sk[i - 1] >> 6) != t) {
    tb[t] = i;
    stamp_arr[t] = stamp;
  }
  if (i == n - 1 || sk[i + 1] == INVALID_KEY || (sk[i + 1] >> 6) != t) te[t] = i + 1;
}

// Grid-tile list in two launches, no device-wide scan: gcount_kernel counts
// the touched tiles of each 256-tile block; glist_kernel sums the counts of
// the blocks before its own, ranks its tiles with wave ballots, writes their
// ids to gtiles and clears their nodes (and those of up to two more node
// arrays).  The last block writes the total to counts[1].
template<typename T, int D>
__device__ __forceinline__ bool grid_tile_flag(const Args<T, D>& a, const uint32_t* bstamp, uint32_t bs, uint32_t ms,
                                               uint32_t t) {
  const Geom<T, D> g = tile_geom(a, t);
  int c[D];
  uint32_t r = t - g.tile0;
#pragma unroll
  for (int k = D - 1; k >= 0; k--) {
    c[k] = (int)(r % (uint32_t)g.GT[k]);
    r /= (uint32_t)g.GT[k];
  }
  bool f = false;
  if (a.band) {  // tiles holding ghost nodes or their mirror images along some axis
#pragma unroll
    for (int k = 0; k < D; k++) {
      constexpr int E = Tile<D>::E;  // the tile covers ghosted nodes [E c, E c + E)
      f = f || E * c[k] < 6 || E * c[k] + E > g.G[k] - 6;
    }
  }
#pragma unroll
  for (int m = 0; m < (1 << D); m++) {
    bool ok = true;
    uint32_t q = 0;
#pragma unroll
    for (int k = 0; k < D; k++) {
      const int cc = c[k] - ((m >> k) & 1);
      ok = ok && cc >= 0;
      q = q * (uint32_t)g.GT[k] + (uint32_t)(cc < 0 ? 0 : cc);
    }
    q += g.tile0;
    if (ok && (bstamp[q] == bs || a.mstamp[q] == ms)) f = true;
  }
  return f;
}

// GRID: touched grid tiles (octant rule); otherwise binned tiles (stamp == bs).
// lists 1: the slow and mover lists of the step about to run start empty;
// lists 2: the list bookkeeping at the end of a fused step, done here to save
// a launch (this kernel runs right after the side kernel and reads no
// counter, and the tile count in counts[1] is replaced only by glist_kernel
// after it): the next mover lists become current ([4] <- [5], [6] <- [7]),
// the pending grid's tile count becomes the last grid's (last[1] <- [1]),
// [9] <- movers for the host's rebin policy, [10] <- movers of the step just
// taken, [11] <- the P2G's FP64-stress count, and the slow list is consumed.
template<typename T, int D, bool GRID>
__global__ __launch_bounds__(BLOCK) void gcount_kernel(Args<T, D> a, const uint32_t* bstamp, uint32_t bs, uint32_t ms,
                                                       uint32_t* block_cnt, int lists, uint32_t* last) {
  const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
  if (lists == 1 && t == 0) {
    a.counts[3] = 0;
    a.counts[4] = 0;
    a.counts[5] = 0;
    a.counts[6] = 0;
    a.counts[7] = 0;
  } else if (lists == 2 && t == 0) {
    uint32_t* counts = a.counts;
    counts[10] = counts[4] + counts[6];  // movers of the step just taken
    counts[4] = counts[5];
    counts[6] = counts[7];
    counts[9] = counts[4] + counts[6];
    last[1] = counts[1];
    counts[11] = counts[3];  // FP64-stress particles of the P2G just done
    counts[3] = 0;
    counts[5] = 0;
    counts[7] = 0;
  }
  const bool f = t < a.ntotal && (GRID ? grid_tile_flag<T, D>(a, bstamp, bs, ms, t) : bstamp[t] == bs);
  const int cnt = __syncthreads_count(f ? 1 : 0);
  if (threadIdx.x == 0) block_cnt[blockIdx.x] = (uint32_t)cnt;
}

template<typename T, int D, bool GRID>
__global__ __launch_bounds__(BLOCK) void glist_kernel(Args<T, D> a, const uint32_t* bstamp, uint32_t bs, uint32_t ms,
                                                      const uint32_t* block_cnt, uint32_t* gtiles, uint32_t* total,
                                                      Node4<T>* z0, Node4<T>* z1, Node4<T>* z2) {
  __shared__ uint32_t s_w[BLOCK / 64], s_tiles[BLOCK];
  __shared__ uint32_t s_base;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // offset of this block: sum of the counts of the blocks before it
  uint32_t part = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += BLOCK) part += block_cnt[b];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
  if (lane == 0) s_w[wv] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; w++) sum += s_w[w];
    s_base = sum;
    if (blockIdx.x == gridDim.x - 1) *total = sum + block_cnt[blockIdx.x];
  }
  __syncthreads();
  const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
  const bool f = t < a.ntotal && (GRID ? grid_tile_flag<T, D>(a, bstamp, bs, ms, t) : bstamp[t] == bs);
  const uint64_t bal = __ballot(f);
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  if (lane == 0) s_w[wv] = (uint32_t)__popcll(bal);
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; w++) {
    off += w < wv ? s_w[w] : 0u;
    tot += s_w[w];
  }
  if (f) {
    const uint32_t r = off + (uint32_t)__popcll(bal & lt);
    gtiles[s_base + r] = t;
    s_tiles[r] = t;
  }
  if (!z0) return;
  __syncthreads();
  Node4<T> zn;
  zn.m = zn.x = zn.y = zn.z = 0;
  for (uint32_t k = threadIdx.x; k < tot * 64u; k += BLOCK) {
    const uint32_t slot = s_tiles[k >> 6] * 64u + (k & 63u);
    z0[slot] = zn;
    if (z1) z1[slot] = zn;
    if (z2) z2[slot] = zn;
  }
}


// ------------------------------------------------------------------ P2G ----
// One workgroup per binned tile (grid-stride over the tile list).  Mass and
// momentum, including the stress impulse -dt * tau_p * grad w_ip of
// MPM_FINITE_ELEMENTS::Add_Forces (Forces/MPM_FINITE_ELEMENTS.cpp:84-99), are
// accumulated in LDS over the tile's stencil region and flushed to HBM with
// one atomic per touched node and component.
//
// The tile's particles are taken in chunks of 512.  Each chunk is grouped by
// base cell in LDS (wave-ballot ranks, stable), which also sets aside movers
// (base outside the tile) for p2g_extra_kernel.  No LDS atomics (on gfx950
// they cost ~40 plain LDS accesses each).  Round k takes particles 4k..4k+3 of
// every cell:
//   phase A: thread t = 4c + j handles particle 4k+j of cell c: it computes
//            the stress and transfer coefficients and stores them in LDS;
//   phase B: lane l of wave w owns cell l (64 cells, 64 lanes) and scatters
//            its particle 4k+w into the wave's private copy of the region
//            with ordered read-modify-writes.  The 64 lanes of one instruction
//            address 64 distinct nodes, a compiler barrier keeps each read
//            after the previous write in program order, LDS executes one
//            wave's accesses in order, and no other wave writes the copy, so
//            every update is race-free;
//   flush:   the four copies are summed per node and added to HBM.
// Particles whose F needs the FP64 SVD (inverted, nearly degenerate, or in the
// neo-Hookean clamp branch) are queued; p2g_extra_kernel adds their stress
// impulse afterwards.
template<typename T, int D> struct P2GCoef {
  // m, mv[D], mC[D*D], Q (symmetric, D(D+1)/2), xe[D]
  static constexpr bool MASS = true, SYM = true;
  static constexpr int NQ = D * (D + 1) / 2;
  static constexpr int M = 0, MV = 1, MC = 1 + D, Q = 1 + D + D * D, XE = 1 + D + D * D + NQ;
  static constexpr int N = XE + D;
};
// The scatter of the G2P adjoint in the same form (vbar_i += w (vt + Bbar dz)
// - Q grad w with Q = -Gbar, no mass): vt[D], Bbar[D*D], Q[D*D], xe[D] -- the
// order of AdjCoef, with Gbar negated.
template<typename T, int D> struct AdjScatterCoef {
  static constexpr bool MASS = false, SYM = false;
  static constexpr int M = 0, MV = 0, MC = D, Q = D + D * D, XE = D + 2 * D * D;
  static constexpr int N = XE + D;
};

template<typename T> struct Vec4Type;
template<> struct Vec4Type<float> { typedef float type __attribute__((ext_vector_type(4))); };
template<> struct Vec4Type<double> { typedef double type __attribute__((ext_vector_type(4))); };


// The P2G of one particle with transfer coefficients cf (layout C), whose
// base node is region node base_li, added into this wavefront's copy of the
// tile's stencil region.  The active lanes of a wavefront hold distinct cells,
// so the nodes of one read-modify-write step are distinct; the compiler
// barrier after each write keeps the steps in order (a later step of one lane
// may touch the node an earlier step of another lane wrote).
template<typename T, int D, class C = P2GCoef<T, D>>
__device__ __forceinline__ void p2g_scatter_vals(const Geom<T, D>& g, const T* cf, int base_li,
                                                 typename Vec4Type<T>::type* reg) {
  constexpr int R = Tile<D>::R;
  using T4 = typename Vec4Type<T>::type;
  const T m = C::MASS ? cf[C::M] : (T)0;
  T mv[D], mC[D][D], Q[D][D], w[3][D], dw[3][D], dz[3][D];
#pragma unroll
  for (int c = 0; c < D; c++) mv[c] = cf[C::MV + c];
#pragma unroll
  for (int j = 0; j < D; j++)
#pragma unroll
    for (int c = 0; c < D; c++) mC[c][j] = cf[C::MC + c + D * j];
  if constexpr (C::SYM) {
    int qq = 0;
#pragma unroll
    for (int c = 0; c < D; c++)
#pragma unroll
      for (int j = c; j < D; j++) {
        const T qv = cf[C::Q + qq++];
        Q[c][j] = qv;
        Q[j][c] = qv;
      }
  } else {
#pragma unroll
    for (int j = 0; j < D; j++)
#pragma unroll
      for (int c = 0; c < D; c++) Q[c][j] = cf[C::Q + c + D * j];
  }
#pragma unroll
  for (int c = 0; c < D; c++) {
    const T xe = cf[C::XE + c];
    const T inv = g.inv_dx[c];
    const T xx = xe * inv;
    const T q0 = xx - (T)1.5, q1 = xx - (T)1, q2 = xx - (T)0.5;
    w[0][c] = (T)0.5 * q0 * q0;
    w[1][c] = (T)0.75 - q1 * q1;
    w[2][c] = (T)0.5 * q2 * q2;
    dw[0][c] = q0 * inv;
    dw[1][c] = (T)(-2) * q1 * inv;
    dw[2][c] = q2 * inv;
#pragma unroll
    for (int o = 0; o < 3; o++) dz[o][c] = (T)o * g.dx[c] - xe;
  }
  // Separable form: with per-axis factors w, dw and dz = o dx - xe,
  //   m_i   += w_x w_y w_z m
  //   p_i,c += w_x w_y w_z (mv_c + sum_j mC_cj dz_j) - sum_j Q_cj (grad w)_j
  // is built axis by axis.  Per x plane, the z-dependent part
  //   w_x w_y (w_z dz_z mC_c2 - dw_z Q_c2) = w_y H_c(z)
  // is formed once for the three z offsets, so that a node costs one FMA for
  // its mass and two per momentum component: w_z P_c(x, y) + w_y H_c(z).
  if constexpr (D == 3) {
    T wzd[3], wydz[3];
#pragma unroll
    for (int o = 0; o < 3; o++) {
      wzd[o] = w[o][2] * dz[o][2];
      wydz[o] = w[o][1] * dz[o][1];
    }
#pragma unroll
    for (int ox = 0; ox < 3; ox++) {
      const T wx = w[ox][0], dwx = dw[ox][0], dzx = dz[ox][0];
      T ax[3], bx[3], dx_[3], H[3][3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        ax[c] = wx * (mv[c] + mC[c][0] * dzx) - Q[c][0] * dwx;  // coefficient of w_y w_z
        bx[c] = wx * mC[c][1];                                  // of w_y dz_y w_z
        dx_[c] = -wx * Q[c][1];                                 // of dw_y w_z
        const T ex = wx * mC[c][2], fx = -wx * Q[c][2];         // of w_y w_z dz_z and of w_y dw_z
#pragma unroll
        for (int oz = 0; oz < 3; oz++) H[oz][c] = wzd[oz] * ex + dw[oz][2] * fx;  // coefficient of w_y
      }
      const T mx = wx * m;
#pragma unroll
      for (int oy = 0; oy < 3; oy++) {
        const T wy = w[oy][1], dwy = dw[oy][1], wyz = wydz[oy];
        T P[3];
#pragma unroll
        for (int c = 0; c < 3; c++) P[c] = wy * ax[c] + wyz * bx[c] + dwy * dx_[c];  // coefficient of w_z
        const T mxy = wy * mx;
#pragma unroll
        for (int oz = 0; oz < 3; oz++) {
          const T wz = w[oz][2];
          const int li = base_li + (ox * R + oy) * R + oz;
          T4 cur = reg[li];
          if constexpr (C::MASS) cur.x += wz * mxy;
          cur.y += wz * P[0];
          cur.z += wz * P[1];
          cur.w += wz * P[2];
          cur.y += wy * H[oz][0];
          cur.z += wy * H[oz][1];
          cur.w += wy * H[oz][2];
          reg[li] = cur;
          __asm__ __volatile__("" ::: "memory");
        }
      }
    }
  } else {
    T wyd[3];
#pragma unroll
    for (int o = 0; o < 3; o++) wyd[o] = w[o][D - 1] * dz[o][D - 1];
#pragma unroll
    for (int ox = 0; ox < 3; ox++) {
      const T wx = w[ox][0], dwx = dw[ox][0], dzx = dz[ox][0];
      T ax[2], ex[2], fx[2];
#pragma unroll
      for (int c = 0; c < 2; c++) {
        ax[c] = wx * (mv[c] + mC[c][0] * dzx) - Q[c][0] * dwx;
        ex[c] = wx * mC[c][D - 1];
        fx[c] = -wx * Q[c][D - 1];
      }
      const T mx = wx * m;
#pragma unroll
      for (int oy = 0; oy < 3; oy++) {
        const T wy = w[oy][D - 1], dwy = dw[oy][D - 1], wyz = wyd[oy];
        const int li = base_li + ox * R + oy;
        T4 cur = reg[li];
        if constexpr (C::MASS) cur.x += wy * mx;
        cur.y += wy * ax[0];
        cur.z += wy * ax[1];
        cur.y += wyz * ex[0];
        cur.z += wyz * ex[1];
        cur.y += dwy * fx[0];
        cur.z += dwy * fx[1];
        reg[li] = cur;
        __asm__ __volatile__("" ::: "memory");
      }
    }
  }
}

// Phase B of the tile P2G: the particle in coefficient slot sl, whose base
// node is region node base_li, is added into this wavefront's copy of the
// tile's stencil region (p2g_scatter_vals on the slot's coefficients).
template<typename T, int D, class C = P2GCoef<T, D>, int ROW = BLOCK>
__device__ __forceinline__ void p2g_scatter_slot(const Geom<T, D>& g, const T (*s_c)[ROW], int sl, int base_li,
                                                 typename Vec4Type<T>::type* reg) {
  T cf[C::N];
#pragma unroll
  for (int q = 0; q < C::N; q++) cf[q] = s_c[q][sl];
  p2g_scatter_vals<T, D, C>(g, cf, base_li, reg);
}

// The unfused P2G (first step of a run, adaptive substeps, taped steps):
// one workgroup per binned tile; particles whose base is outside the tile go
// to the mover list for p2g_extra_kernel.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) void p2g_kernel(Args<T, D> a) {
  constexpr int R = Tile<D>::R, RN = Tile<D>::RN, NW = BLOCK / 64, E = Tile<D>::E;
  constexpr int JT = 2, CH = JT * BLOCK;  // particles per chunk
  using C = P2GCoef<T, D>;
  using T4 = typename Vec4Type<T>::type;
  __shared__ T4 s_reg[NW][RN];  // per-wave copies of the region: (m, p_x, p_y, p_z)
  __shared__ union {
    T c[C::N][BLOCK];            // phase A -> phase B coefficients
    uint32_t wc[JT][NW][64];     // grouping: per (pass, wave) cell counts, then offsets
  } s_u;
  __shared__ uint32_t s_idx[CH];  // chunk's particle slots grouped by cell
  __shared__ uint32_t s_cb[64], s_cn[64], s_max;
  auto& s_c = s_u.c;
  const uint32_t ntile = a.counts[0];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int a_cell = threadIdx.x >> 2, a_j = threadIdx.x & 3;  // phase A assignment
  const int a_slot = a_j * 64 + a_cell;                         // = phase B slot (wave a_j, lane a_cell)
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (uint32_t seg = blockIdx.x; seg < ntile; seg += gridDim.x) {
    const uint32_t t = a.ptiles[seg];
    const Geom<T, D> g = tile_geom(a, t);
    T dinv[D];
#pragma unroll
    for (int c = 0; c < D; c++) dinv[c] = (T)4 * g.inv_dx[c] * g.inv_dx[c];  // D_p^-1 (PARTICLE_GRID_WEIGHTS_SPLINE.cpp:152-158)
    int org[D];
    tile_origin<T, D>(t, g, org);
    for (int k = threadIdx.x; k < NW * RN; k += BLOCK) {
      T4 z = {0, 0, 0, 0};
      s_reg[k / RN][k % RN] = z;
    }
    const uint32_t tb = a.tile_begin[t], te = a.tile_end[t];
    for (uint32_t c0 = tb; c0 < te; c0 += CH) {
      for (int k = threadIdx.x; k < JT * NW * 64; k += BLOCK) (&s_u.wc[0][0][0])[k] = 0;
      __syncthreads();
      // ---- group the chunk by base cell; movers go to the mover list
      uint32_t cell_j[JT], rank_j[JT], slot_j[JT];
#pragma unroll
      for (int j = 0; j < JT; j++) {
        const uint32_t i = c0 + j * BLOCK + threadIdx.x;
        uint32_t cell = 64, p = 0;
        if (i < te) {
          p = a.order ? a.order[i] : i;
          if (a.valid[p]) {
            T x[D];
#pragma unroll
            for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
            bool in = true;
            int lc[D];
#pragma unroll
            for (int c = 0; c < D; c++) {
              const int lb = stencil_base<T, D>(g, x, c) + 3 - org[c];
              in = in && lb >= 0 && lb < E;
              lc[c] = lb & (E - 1);
            }
            const uint32_t cl = cell_label<D>(lc);
            if (in) {
              cell = cl;
            } else {
              const uint32_t q = atomicAdd(&a.counts[4], 1u);  // aggregated per wavefront by the compiler
              a.mover_list[q] = p;
              a.mover_tile[q] = t;
            }
          }
        }
        const bool act = cell < 64;
        uint64_t m = __ballot(act);
#pragma unroll
        for (int bb = 0; bb < 6; bb++) {
          const bool bit = (cell >> bb) & 1u;
          const uint64_t bal = __ballot(bit);
          m &= bit ? bal : ~bal;
        }
        rank_j[j] = (uint32_t)__popcll(m & lt_mask);
        if (act && lane == __ffsll((unsigned long long)m) - 1) s_u.wc[j][wv][cell] = (uint32_t)__popcll(m);
        cell_j[j] = cell;
        slot_j[j] = p;
      }
      __syncthreads();
      if (threadIdx.x < 64) {  // wave 0: per-cell offsets over (pass, wave), then cell starts
        uint32_t run = 0;
#pragma unroll
        for (int j = 0; j < JT; j++)
#pragma unroll
          for (int w = 0; w < NW; w++) {
            const uint32_t v = s_u.wc[j][w][threadIdx.x];
            s_u.wc[j][w][threadIdx.x] = run;
            run += v;
          }
        uint32_t x = run, mx = run;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(x, o, 64);
          if (lane >= o) x += y;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, 64));
        s_cb[threadIdx.x] = x - run;
        s_cn[threadIdx.x] = run;
        if (threadIdx.x == 0) s_max = mx;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < JT; j++)
        if (cell_j[j] < 64) s_idx[s_cb[cell_j[j]] + s_u.wc[j][wv][cell_j[j]] + rank_j[j]] = slot_j[j];
      __syncthreads();
      const uint32_t a_cb = s_cb[a_cell], a_cnt = s_cn[a_cell];
      const uint32_t b_cnt = s_cn[lane];
      const uint32_t rounds = (s_max + NW - 1) / NW;
      for (uint32_t k = 0; k < rounds; k++) {
        // ---- phase A
        const uint32_t ja = k * NW + a_j;
        if (ja < a_cnt) {
          const uint32_t p = s_idx[a_cb + ja];
          T x[D], v[D], Bm[D * D], Fm[D * D];
#pragma unroll
          for (int c = 0; c < D; c++) { x[c] = a.X[c * a.n + p]; v[c] = a.V[c * a.n + p]; }
#pragma unroll
          for (int q = 0; q < D * D; q++) { Bm[q] = a.B[q * a.n + p]; Fm[q] = a.F[q * a.n + p]; }
          const T m = a.mass[p];
#pragma unroll
          for (int c = 0; c < D; c++) {
            const int bb = stencil_base<T, D>(g, x, c);
            const T cen = g.lo[c] + ((T)bb + (T)0.5) * g.dx[c];
            s_c[C::XE + c][a_slot] = x[c] - cen;
          }
          T tau[D * D];
          // the unfused kernel runs at 128 registers: the Jacobi branch, which fits
          if (!kirchhoff_fast<T, D, false>(a.model, a.mu[p], a.la[p], a.thr, a.vol[p], Fm, tau)) {
#pragma unroll
            for (int q = 0; q < D * D; q++) tau[q] = 0;
            a.slow_list[atomicAdd(&a.counts[3], 1u)] = p;  // the compiler aggregates this per wavefront
          }
          s_c[C::M][a_slot] = m;
#pragma unroll
          for (int c = 0; c < D; c++) s_c[C::MV + c][a_slot] = m * v[c];
#pragma unroll
          for (int j = 0; j < D; j++)
#pragma unroll
            for (int c = 0; c < D; c++) s_c[C::MC + c + D * j][a_slot] = m * Bm[c + D * j] * dinv[j];
          int q = 0;
#pragma unroll
          for (int c = 0; c < D; c++)
#pragma unroll
            for (int j = c; j < D; j++) s_c[C::Q + q++][a_slot] = a.dt * tau[c + D * j];
        }
        __syncthreads();
        // ---- phase B
        if (k * NW + wv < b_cnt) {
          int lcell[D];
          cell_offsets<D>((uint32_t)lane, lcell);
          const int lane_li = D == 3 ? (lcell[0] * R + lcell[1]) * R + lcell[D - 1] : lcell[0] * R + lcell[D - 1];
          p2g_scatter_slot<T, D>(g, s_c, threadIdx.x, lane_li, s_reg[wv]);
        }
        __syncthreads();
      }
    }
    for (int k = threadIdx.x; k < RN; k += BLOCK) {
      T4 sum = s_reg[0][k];
#pragma unroll
      for (int w = 1; w < NW; w++) sum += s_reg[w][k];
      if (sum.x != (T)0) {
        int off[D], nc[D];
        region_offsets<D>(k, off);
#pragma unroll
        for (int c = 0; c < D; c++) nc[c] = org[c] + off[c];
        Node4<T>* nd = a.grid + node_slot<T, D>(nc, g);
        atomic_add(&nd->m, sum.x);
        atomic_add(&nd->x, sum.y);
        atomic_add(&nd->y, sum.z);
        if constexpr (D == 3) atomic_add(&nd->z, sum.w);
      }
    }
    __syncthreads();
  }
}

// Weight, gradient and x_i - x_p of stencil node k (row-major offsets) from
// the particle's offset xe to its base node (PARTICLE_GRID_WEIGHTS_SPLINE.cpp:
// 33-56 for degree 2), for kernels that give each node its own lane.
template<typename T, int D>
__device__ __forceinline__ void node_weight(const Geom<T, D>& g, const T* xe, int k, int* o, T& wt, T* gr, T* dz) {
  int r = k;
#pragma unroll
  for (int c = D - 1; c >= 0; c--) {
    o[c] = r % 3;
    r /= 3;
  }
  T w1[D], d1[D];
#pragma unroll
  for (int c = 0; c < D; c++) {
    const T inv = g.inv_dx[c];
    const T x = xe[c] * inv;
    const T q = x - (o[c] == 0 ? (T)1.5 : (o[c] == 1 ? (T)1 : (T)0.5));  // q0, q1, q2 as in weights()
    w1[c] = o[c] == 1 ? (T)0.75 - q * q : (T)0.5 * q * q;
    d1[c] = (o[c] == 1 ? (T)-2 : (T)1) * q * inv;
    dz[c] = (T)o[c] * g.dx[c] - xe[c];
  }
  wt = 1;
#pragma unroll
  for (int c = 0; c < D; c++) wt *= w1[c];
#pragma unroll
  for (int c = 0; c < D; c++) {
    T gg = d1[c];
#pragma unroll
    for (int j = 0; j < D; j++)
      if (j != c) gg *= w1[j];
    gr[c] = gg;
  }
}

// P2G of what the tile kernel leaves, straight to HBM with atomics:
//   movers (list [0, counts[4])): the whole APIC transfer and stress impulse;
//   queued particles (slow list [0, counts[3])): the stress impulse, FP64 SVD path.
// Pass 1 gives each particle a thread and stores its transfer coefficients in
// LDS; pass 2 gives each (particle, node) pair a lane, so the 27 atomics of a
// particle go out in one instruction to 27 different nodes (one lane per
// particle and node would make neighbouring particles collide in L2).
template<typename T, int D> struct ScatterCoef {
  static constexpr int M = 0, MV = 1, MC = 1 + D, TAU = 1 + D + D * D, XE = 1 + D + 2 * D * D, N = XE + D;
};

// Pass 1 of the side-path P2G for one particle p: its transfer coefficients
// (mover: m, m v, m B D^-1, dt tau, x - x_base; slow-list particle: only
// dt tau) into column `slot` of s_c, its stencil base into s_b (s_b[0] = -1:
// skipped), its scene into s_sc.
template<typename T, int D, bool FUSED>
__device__ __forceinline__ void side_coefficients(const Args<T, D>& a, uint32_t p, bool mover,
                                                  T (*s_c)[BLOCK], int (*s_b)[BLOCK], uint16_t* s_sc, int slot) {
  using C = ScatterCoef<T, D>;
  const Geom<T, D> g = particle_geom(a, p);
  T x[D], Fm[D * D], tau[D * D];
#pragma unroll
  for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
#pragma unroll
  for (int k = 0; k < D * D; k++) Fm[k] = a.F[k * a.n + p];
  bool in = true;
  int base[D];
#pragma unroll
  for (int c = 0; c < D; c++) {
    base[c] = stencil_base<T, D>(g, x, c) + 3;
    in = in && base[c] >= 0 && base[c] + 2 < g.G[c];
  }
  if (FUSED && !a.valid[p]) {
    base[0] = -1;  // left the domain in this step's G2P
  } else if (!in) {
    atomicOr(a.err, 2u);
    base[0] = -1;  // skipped in pass 2
  } else if (mover) {
    kirchhoff_any<T, D>(a.model, a.mu[p], a.la[p], a.thr, a.root, a.vol[p], Fm, tau);
  } else {
    kirchhoff_slow<T, D>(a.model, a.mu[p], a.la[p], a.thr, a.root, a.vol[p], Fm, tau);
  }
  const T m = mover ? a.mass[p] : (T)0;
  s_c[C::M][slot] = m;
#pragma unroll
  for (int c = 0; c < D; c++) {
    s_c[C::MV + c][slot] = mover ? m * a.V[c * a.n + p] : (T)0;
    s_c[C::XE + c][slot] = x[c] - (g.lo[c] + ((T)(base[c] - 3) + (T)0.5) * g.dx[c]);
    s_b[c][slot] = base[c];
  }
#pragma unroll
  for (int j = 0; j < D; j++)
#pragma unroll
    for (int c = 0; c < D; c++) {
      s_c[C::MC + c + D * j][slot] = mover ? m * a.B[(c + D * j) * a.n + p] * ((T)4 * g.inv_dx[j] * g.inv_dx[j]) : (T)0;
      s_c[C::TAU + c + D * j][slot] = a.dt * tau[c + D * j];
    }
  s_sc[slot] = a.pscene ? a.pscene[p] : 0;
}

// Pass 2 of the side-path P2G: one thread per (particle, stencil node) of the
// cnt particles of s_c, atomics into target.
template<typename T, int D>
__device__ __forceinline__ void side_scatter(const Args<T, D>& a, Node4<T>* target, const T (*s_c)[BLOCK],
                                             const int (*s_b)[BLOCK], const uint16_t* s_sc, uint32_t cnt) {
  using C = ScatterCoef<T, D>;
  for (uint32_t u = threadIdx.x; u < cnt * Tile<D>::K; u += BLOCK) {
    const uint32_t sl = u / Tile<D>::K;
    const int k = (int)(u - sl * Tile<D>::K);
    if (s_b[0][sl] < 0) continue;
    const Geom<T, D> g = a.scenes[s_sc[sl]];
    T xe[D], wt, gr[D], dz[D];
    int o[D], nc[D];
#pragma unroll
    for (int c = 0; c < D; c++) xe[c] = s_c[C::XE + c][sl];
    node_weight<T, D>(g, xe, k, o, wt, gr, dz);
#pragma unroll
    for (int c = 0; c < D; c++) nc[c] = s_b[c][sl] + o[c];
    Node4<T>* nd = target + node_slot<T, D>(nc, g);
    T f[D];
#pragma unroll
    for (int c = 0; c < D; c++) {
      T st = 0, ap = s_c[C::MV + c][sl];
#pragma unroll
      for (int j = 0; j < D; j++) {
        st += s_c[C::TAU + c + D * j][sl] * gr[j];
        ap += s_c[C::MC + c + D * j][sl] * dz[j];
      }
      f[c] = wt * ap - st;
    }
    const T m = s_c[C::M][sl];
    if (m != (T)0) atomic_add(&nd->m, wt * m);
    atomic_add(&nd->x, f[0]);
    atomic_add(&nd->y, f[1]);
    if constexpr (D == 3) atomic_add(&nd->z, f[D - 1]);
  }
}

// The side path of the unfused P2G: the movers p2g_kernel listed (counts[4])
// and the particles whose stress needs the FP64 SVD (counts[3]), with global
// atomics into a.grid.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void p2g_extra_kernel(Args<T, D> a) {
  __shared__ T s_c[ScatterCoef<T, D>::N][BLOCK];
  __shared__ int s_b[D][BLOCK];
  __shared__ uint16_t s_sc[BLOCK];
  const uint32_t nm = a.counts[4], ns = a.counts[3], ntot = nm + ns;
  // particles per block and pass: short lists are spread over every block
  const uint32_t per = min((uint32_t)BLOCK, max(1u, (ntot + gridDim.x - 1) / gridDim.x));
  for (uint32_t c0 = blockIdx.x * per; c0 < ntot; c0 += gridDim.x * per) {
    // ---- pass 1: one thread per particle
    const uint32_t q = c0 + threadIdx.x;
    if (threadIdx.x < per && q < ntot) {
      const bool mover = q < nm;
      uint32_t p;
      if (mover) {
        p = a.mover_list[spread_index(q, nm)];
      } else {
        p = a.slow_list[spread_index(q - nm, ns)];
      }
      side_coefficients<T, D, false>(a, p, mover, s_c, s_b, s_sc, (int)threadIdx.x);
    }
    __syncthreads();
    // ---- pass 2: one thread per (particle, stencil node)
    side_scatter<T, D>(a, a.grid, s_c, s_b, s_sc, min(per, ntot - c0));
    __syncthreads();
  }
}

// Collision objects in order on a node velocity (MPM_OBJECTIVE.cpp:150-182):
// applies them to v and returns the number of hits; with want >= 0 it stops at
// hit number `want` (before applying it) and returns its type, normal and
// object.  Replaying instead of recording keeps per-hit arrays (which would
// live in scratch memory) out of the kernels.
template<typename T, int D>
__device__ __forceinline__ int walls_apply(const Args<T, D>& a, const T* X0, T* v, int want, T* nrm_out, int* obj_out) {
  int hits = 0;
  const T inv_dt = (T)1 / a.dt;
  for (int o = 0; o < a.nwalls; o++) {
    const WallDev<T>& wl = a.walls[o];
    T X[D];
#pragma unroll
    for (int c = 0; c < D; c++) X[c] = X0[c] + a.dt * v[c];
    const T phi0 = box_phi<T, D>(wl, X0);
    const T phi = box_phi<T, D>(wl, X) - (phi0 < 0 ? phi0 : (T)0);
    if (wl.type == 0 && phi0 > 0) continue;
    if (wl.type == 1 && phi0 > 0) continue;
    if (wl.type == 2 && phi > (T)1e-15) continue;
    if (wl.type != 0 && phi > (T)3.402823466e38) continue;
    T nrm[D];
    box_normal<T, D>(wl, X, nrm);
    if (hits == want) {
#pragma unroll
      for (int c = 0; c < D; c++) nrm_out[c] = nrm[c];
      *obj_out = o;
      return wl.type;
    }
#pragma unroll
    for (int c = 0; c < D; c++) {
      X[c] -= phi * nrm[c];
      v[c] = (X[c] - X0[c]) * inv_dt;
    }
    if (wl.type == 0) {
#pragma unroll
      for (int c = 0; c < D; c++) v[c] = 0;
    }
    hits++;
  }
  return hits;
}

// ----------------------------------------------------------- grid update ---
// v = p/m + dt g, then each collision object in order (MPM_OBJECTIVE.cpp:150-182),
// then Coulomb friction at colliding nodes (MPM_DRIVER.cpp:1064-1094).
template<typename T, int D, bool FRICTION>
__global__ __launch_bounds__(BLOCK) void grid_kernel(Args<T, D> a) {
  if (a.clear_grid) {
    const uint64_t tc = (uint64_t)(*a.clear_count) * 64u;
    Node4<T> z;
    z.m = z.x = z.y = z.z = 0;
    for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < tc; idx += (uint64_t)gridDim.x * BLOCK)
      a.clear_grid[(uint64_t)a.clear_tiles[idx >> 6] * 64u + (idx & 63u)] = z;
  }
  const uint64_t total = (uint64_t)a.counts[1] * 64u;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t t = a.gtiles[idx >> 6];
    const uint32_t l = (uint32_t)(idx & 63u);
    const uint32_t slot = t * 64u + l;
    Node4<T> nd = a.grid[slot];
    if (nd.m == (T)0) continue;
    const T im = (T)1 / nd.m;
    T v[D];
    v[0] = nd.x * im + a.dt * a.grav[0];
    v[1] = nd.y * im + a.dt * a.grav[1];
    if constexpr (D == 3) v[2] = nd.z * im + a.dt * a.grav[2];
    if (a.gridu) {
      Node4<T> u;
      u.m = nd.m;
      u.x = v[0];
      u.y = v[1];
      u.z = D == 3 ? v[D - 1] : (T)0;
      a.gridu[slot] = u;
    }
    T vf[D];
    if (a.nwalls > 0) {
      const Geom<T, D> g = tile_geom(a, t);
      int org[D], nc[D];
      tile_origin<T, D>(t, g, org);
      uint32_t r = l;
#pragma unroll
      for (int c = D - 1; c >= 0; c--) {
        nc[c] = org[c] + (int)(r % Tile<D>::E);
        r /= Tile<D>::E;
      }
      T X0[D];
#pragma unroll
      for (int c = 0; c < D; c++) X0[c] = g.lo[c] + ((T)(nc[c] - 3) + (T)0.5) * g.dx[c];
      T vpre[D];
#pragma unroll
      for (int c = 0; c < D; c++) vpre[c] = v[c];
      T nd_[D];
      int ob_;
      const int ncol = walls_apply<T, D>(a, X0, v, -1, nd_, &ob_);
#pragma unroll
      for (int c = 0; c < D; c++) vf[c] = v[c];
      if constexpr (FRICTION) {
        // Apply_Friction (MPM_DRIVER.cpp:1064-1094) over this node's collisions in order
        if (ncol > 0) {
          bool stuck = false;
          for (int q = 0; q < ncol; q++) {
            T vq[D], nq[D];
            int oq;
#pragma unroll
            for (int c = 0; c < D; c++) vq[c] = vpre[c];
            if (walls_apply<T, D>(a, X0, vq, q, nq, &oq) == 0) stuck = true;
          }
          T v1[D];
#pragma unroll
          for (int c = 0; c < D; c++) v1[c] = stuck ? (T)0 : v[c];
          for (int q = 0; q < ncol; q++) {
            T vq[D], cn[D];
            int oq;
#pragma unroll
            for (int c = 0; c < D; c++) vq[c] = vpre[c];
            walls_apply<T, D>(a, X0, vq, q, cn, &oq);
            T nf = 0, vn = 0;
#pragma unroll
            for (int c = 0; c < D; c++) {
              nf += cn[c] * (v[c] - vpre[c]);
              vn += v1[c] * cn[c];
            }
            T tv[D], tm = 0;
#pragma unroll
            for (int c = 0; c < D; c++) {
              tv[c] = v1[c] - vn * cn[c];
              tm += tv[c] * tv[c];
            }
            tm = tsqrt<T>(tm);
            const T k = a.walls[oq].friction * nf;
            if (tm <= k) {
#pragma unroll
              for (int c = 0; c < D; c++) v1[c] = vn * cn[c];
            } else {
              const T it = tm > 0 ? (T)1 / tm : (T)0;
#pragma unroll
              for (int c = 0; c < D; c++) v1[c] -= k * (tm > 0 ? tv[c] * it : (c == 0 ? (T)1 : (T)0));
            }
          }
#pragma unroll
          for (int c = 0; c < D; c++) vf[c] = stuck ? (T)0 : v1[c];
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < D; c++) vf[c] = v[c];
    }
    Node4<T> out;
    out.m = nd.m;
    out.x = v[0];
    out.y = v[1];
    out.z = D == 3 ? v[D - 1] : (T)0;
    a.grid[slot] = out;
    if constexpr (FRICTION) {
      Node4<T> of;
      of.m = nd.m;
      of.x = vf[0];
      of.y = vf[1];
      of.z = D == 3 ? vf[D - 1] : (T)0;
      a.gridf[slot] = of;
    }
  }
}

// ------------------------------------------------------------------ G2P ----
// Gather of one particle from node velocities nv(ox, oy, oz, c) and friction
// velocities nf(...) (the same values without friction): v (vb), the
// friction-save velocity (vfb), grad v (G) and B (Bs), column-major.  B uses
// the algebraically equal, cancellation-free form of MPM_DRIVER.cpp:485,
// :504-505:  B = sum w vf (x_i - x_p)^T
//              + dt/2 [sum w (vf v^T - v vf^T) - (vf_p v_p^T - v_p vf_p^T)].
template<typename T, int D, bool FRICTION, class NV, class NF>
__device__ __forceinline__ void g2p_gather(const Geom<T, D>& gm, T dt, const Wts<T, D>& w, NV nv, NF nf, T* vb,
                                           T* vfb, T* G, T* Bs) {
  T dz[3][D];
#pragma unroll
  for (int o = 0; o < 3; o++)
#pragma unroll
    for (int c = 0; c < D; c++) dz[o][c] = (T)o * gm.dx[c] - w.xe[c];
  T An[D * D];
#pragma unroll
  for (int c = 0; c < D; c++) { vb[c] = 0; vfb[c] = 0; }
#pragma unroll
  for (int k = 0; k < D * D; k++) { G[k] = 0; Bs[k] = 0; An[k] = 0; }
  if constexpr (D == 3 && !FRICTION) {
    // Without friction every sum factors by axis:
    //   v_p = sum w_x w_y w_z v_i,   (grad v)_c0 = sum dw_x w_y w_z v_c, ...,
    //   B_cj = sum w v_c dz_j,  dz_j = o_j dx - xe_j,
    // so the z sums are formed per (x, y) column, the y sums per x plane.
    T wzd[3], wyd[3], wxd[3];
#pragma unroll
    for (int o = 0; o < 3; o++) {
      wzd[o] = w.w[o][2] * dz[o][2];
      wyd[o] = w.w[o][1] * dz[o][1];
      wxd[o] = w.w[o][0] * dz[o][0];
    }
#pragma unroll
    for (int ox = 0; ox < 3; ox++) {
      // y-level sums of this x plane, per component:
      // Y00 = sum w_y Z0, Y10 = sum dw_y Z0, Y20 = sum w_y dz_y Z0, Y01 = sum w_y Z1, Y02 = sum w_y Z2
      T Y00[3], Y10[3], Y20[3], Y01[3], Y02[3];
#pragma unroll
      for (int c = 0; c < 3; c++) Y00[c] = Y10[c] = Y20[c] = Y01[c] = Y02[c] = 0;
#pragma unroll
      for (int oy = 0; oy < 3; oy++) {
        // z-level sums of this column: Z0 = sum w_z v, Z1 = sum dw_z v, Z2 = sum w_z dz_z v
        T Z0[3], Z1[3], Z2[3];
#pragma unroll
        for (int c = 0; c < 3; c++) Z0[c] = Z1[c] = Z2[c] = 0;
#pragma unroll
        for (int oz = 0; oz < 3; oz++)
#pragma unroll
          for (int c = 0; c < 3; c++) {
            const T vn = nv(ox, oy, oz, c);
            Z0[c] += w.w[oz][2] * vn;
            Z1[c] += w.dw[oz][2] * vn;
            Z2[c] += wzd[oz] * vn;
          }
#pragma unroll
        for (int c = 0; c < 3; c++) {
          Y00[c] += w.w[oy][1] * Z0[c];
          Y10[c] += w.dw[oy][1] * Z0[c];
          Y20[c] += wyd[oy] * Z0[c];
          Y01[c] += w.w[oy][1] * Z1[c];
          Y02[c] += w.w[oy][1] * Z2[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        vb[c] += w.w[ox][0] * Y00[c];
        G[c + 0] += w.dw[ox][0] * Y00[c];
        Bs[c + 0] += wxd[ox] * Y00[c];
        G[c + 3] += w.w[ox][0] * Y10[c];
        Bs[c + 3] += w.w[ox][0] * Y20[c];
        G[c + 6] += w.w[ox][0] * Y01[c];
        Bs[c + 6] += w.w[ox][0] * Y02[c];
      }
    }
  } else if constexpr (D == 3) {
#pragma unroll
    for (int ox = 0; ox < 3; ox++)
#pragma unroll
      for (int oy = 0; oy < 3; oy++) {
        const T wxy = w.w[ox][0] * w.w[oy][1];
#pragma unroll
        for (int oz = 0; oz < 3; oz++) {
          const T wt = wxy * w.w[oz][2];
          T g[3];
          g[0] = w.dw[ox][0] * w.w[oy][1] * w.w[oz][2];
          g[1] = w.w[ox][0] * w.dw[oy][1] * w.w[oz][2];
          g[2] = wxy * w.dw[oz][2];
          T vn[3], vf[3];
#pragma unroll
          for (int c = 0; c < 3; c++) {
            vn[c] = nv(ox, oy, oz, c);
            vf[c] = nf(ox, oy, oz, c);
          }
          const T zz[3] = {dz[ox][0], dz[oy][1], dz[oz][2]};
#pragma unroll
          for (int c = 0; c < 3; c++) {
            vb[c] += wt * vn[c];
            vfb[c] += wt * vf[c];
            const T wvf = wt * vf[c];
#pragma unroll
            for (int j = 0; j < 3; j++) {
              G[c + 3 * j] += vn[c] * g[j];
              Bs[c + 3 * j] += wvf * zz[j];
              An[c + 3 * j] += wt * (vf[c] * vn[j] - vn[c] * vf[j]);
            }
          }
        }
      }
  } else {
#pragma unroll
    for (int ox = 0; ox < 3; ox++)
#pragma unroll
      for (int oy = 0; oy < 3; oy++) {
        const T wt = w.w[ox][0] * w.w[oy][1];
        T g[2];
        g[0] = w.dw[ox][0] * w.w[oy][1];
        g[1] = w.w[ox][0] * w.dw[oy][1];
        T vn[2], vf[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
          vn[c] = nv(ox, oy, 0, c);
          vf[c] = FRICTION ? nf(ox, oy, 0, c) : vn[c];
        }
        const T zz[2] = {dz[ox][0], dz[oy][1]};
#pragma unroll
        for (int c = 0; c < 2; c++) {
          vb[c] += wt * vn[c];
          if (FRICTION) vfb[c] += wt * vf[c];
          const T wvf = wt * vf[c];
#pragma unroll
          for (int j = 0; j < 2; j++) {
            G[c + 2 * j] += vn[c] * g[j];
            Bs[c + 2 * j] += wvf * zz[j];
            if (FRICTION) An[c + 2 * j] += wt * (vf[c] * vn[j] - vn[c] * vf[j]);
          }
        }
      }
  }
  if (!FRICTION) {
#pragma unroll
    for (int c = 0; c < D; c++) vfb[c] = vb[c];
  } else {
    const T h = (T)0.5 * dt;
#pragma unroll
    for (int c = 0; c < D; c++)
#pragma unroll
      for (int j = 0; j < D; j++) Bs[c + D * j] += h * (An[c + D * j] - (vfb[c] * vb[j] - vb[c] * vfb[j]));
  }
}

// F <- (I + dt grad v) F, x <- x + dt v, v <- vf, B, and invalidation outside
// the domain (MPM_DRIVER.cpp:450-515, :1219-1239).  A particle whose new
// stencil base leaves its binned tile marks the tile it moves to, so that the
// next step clears and updates the grid tiles around it.
// REFLECT: sides with a reflection condition mirror particles back (only the
// unfused kernels, which reflection boundary conditions take).
template<typename T, int D, bool REFLECT>
__device__ __forceinline__ bool g2p_store(const Args<T, D>& a, const Geom<T, D>& g, uint32_t p, const T* x,
                                          const T* vb, const T* vfb, const T* G, const T* Bs, uint32_t bin_tile) {
  T Fm[D * D], Fn[D * D];
#pragma unroll
  for (int k = 0; k < D * D; k++) Fm[k] = a.F[k * a.n + p];
#pragma unroll
  for (int j = 0; j < D; j++)
#pragma unroll
    for (int c = 0; c < D; c++) {
      T s = Fm[c + D * j];
#pragma unroll
      for (int q = 0; q < D; q++) s += a.dt * G[c + D * q] * Fm[q + D * j];
      Fn[c + D * j] = s;
    }
  T xn[D];
  bool ok = true;
#pragma unroll
  for (int c = 0; c < D; c++) {
    xn[c] = x[c] + a.dt * vb[c];
    if (xn[c] < g.vlo[c] || xn[c] > g.vhi[c]) ok = false;
  }
  if constexpr (REFLECT) {
    // Reflect_Or_Invalidate_Particle (MPM_DRIVER.cpp:1219-1239): a side with
    // a reflection condition mirrors the particle back, others invalidate it
    T vn[D];
    ok = true;
#pragma unroll
    for (int c = 0; c < D; c++) {
      vn[c] = vfb[c];
      if (xn[c] < g.vlo[c]) {
        if ((a.reflect_sides >> (2 * c)) & 1u) {
          xn[c] = 2 * g.vlo[c] - xn[c];
          vn[c] = -vn[c];
        } else {
          ok = false;
        }
      } else if (xn[c] > g.vhi[c]) {
        if ((a.reflect_sides >> (2 * c + 1)) & 1u) {
          xn[c] = 2 * g.vhi[c] - xn[c];
          vn[c] = -vn[c];
        } else {
          ok = false;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < D; c++) { a.X[c * a.n + p] = xn[c]; a.V[c * a.n + p] = vn[c]; }
  } else {
#pragma unroll
    for (int c = 0; c < D; c++) { a.X[c * a.n + p] = xn[c]; a.V[c * a.n + p] = vfb[c]; }
  }
#pragma unroll
  for (int k = 0; k < D * D; k++) { a.F[k * a.n + p] = Fn[k]; a.B[k * a.n + p] = Bs[k]; }
  if (!ok) {
    a.valid[p] = 0;
    return false;
  }
  const uint32_t nt = base_tile<T, D>(g, xn);
  if (nt == bin_tile) return false;
  if (nt < a.ntotal) a.mstamp[nt] = a.next_stamp;
  return true;
}

// The tile's stencil region of the grid nodes of a.grid and, with FRICTION,
// of a.gridf (second half), staged in LDS as 4-vectors (v_x, v_y, v_z, m):
// the velocity is the first 12 bytes of a 16-byte-aligned node, so a gather
// reads it with one 12-byte LDS load (16-byte node stride across the lanes;
// single-component loads at that stride would hit every bank four times).
template<typename T, int D, bool FRICTION, int NT = BLOCK>
__device__ __forceinline__ void g2p_stage(const Args<T, D>& a, const Geom<T, D>& g, const int* org,
                                          typename Vec4Type<T>::type* sv) {
  using T4 = typename Vec4Type<T>::type;
  for (int k = threadIdx.x; k < Tile<D>::RN; k += NT) {
    int off[D], nc[D];
    region_offsets<D>(k, off);
#pragma unroll
    for (int c = 0; c < D; c++) nc[c] = org[c] + off[c];
    const uint32_t slot = node_slot<T, D>(nc, g);
    const Node4<T> nd = a.grid[slot];
    sv[k] = T4{nd.x, nd.y, nd.z, nd.m};
    if constexpr (FRICTION) {
      const Node4<T> nf = a.gridf[slot];
      sv[Tile<D>::RN + k] = T4{nf.x, nf.y, nf.z, nf.m};
    }
  }
}

// G2P of the particles binned to tile t from the staged region; movers
// (base outside the tile) are left to g2p_mover_kernel.
template<typename T, int D, bool FRICTION>
__device__ __forceinline__ void g2p_tile_pass(const Args<T, D>& a, uint32_t t, const Geom<T, D>& g, const int* org,
                                              const typename Vec4Type<T>::type* sv) {
  constexpr int R = Tile<D>::R, E = Tile<D>::E, RN = Tile<D>::RN;
  const uint32_t b = a.tile_begin[t], e = a.tile_end[t];
  for (uint32_t i = b + threadIdx.x; i < e; i += BLOCK) {
    const uint32_t p = a.order ? a.order[i] : i;
    if (!a.valid[p]) continue;
    T x[D];
#pragma unroll
    for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
    Wts<T, D> w;
    weights<T, D>(g, x, w);
    bool in = true;
    int lb[D];
#pragma unroll
    for (int c = 0; c < D; c++) {
      lb[c] = w.base[c] - org[c];
      in = in && lb[c] >= 0 && lb[c] < E;
    }
    if (!in) continue;  // mover
    const int li0 = D == 3 ? (lb[0] * R + lb[1]) * R + lb[D - 1] : lb[0] * R + lb[D - 1];
    auto comp = [](const typename Vec4Type<T>::type& q, int c) -> T { return c == 0 ? q.x : (c == 1 ? q.y : q.z); };
    auto nv = [&](int ox, int oy, int oz, int c) -> T {
      return comp(sv[li0 + (D == 3 ? (ox * R + oy) * R + oz : ox * R + oy)], c);
    };
    auto nf = [&](int ox, int oy, int oz, int c) -> T {
      if constexpr (FRICTION) return comp(sv[RN + li0 + (D == 3 ? (ox * R + oy) * R + oz : ox * R + oy)], c);
      else return nv(ox, oy, oz, c);
    };
    T vb[D], vfb[D], G[D * D], Bs[D * D];
    g2p_gather<T, D, FRICTION>(g, a.dt, w, nv, nf, vb, vfb, G, Bs);
    g2p_store<T, D, true>(a, g, p, x, vb, vfb, G, Bs, t);
  }
}

// One workgroup per binned tile: the tile's stencil region of v (and of the
// friction velocity) is staged in LDS; movers are left to g2p_mover_kernel.
template<typename T, int D, bool FRICTION>
__global__ __launch_bounds__(BLOCK) void g2p_kernel(Args<T, D> a) {
  __shared__ typename Vec4Type<T>::type s_v[FRICTION ? 2 : 1][Tile<D>::RN];
  const uint32_t ntile = a.counts[0];
  for (uint32_t seg = blockIdx.x; seg < ntile; seg += gridDim.x) {
    const uint32_t t = a.ptiles[seg];
    const Geom<T, D> g = tile_geom(a, t);
    int org[D];
    tile_origin<T, D>(t, g, org);
    g2p_stage<T, D, FRICTION>(a, g, org, &s_v[0][0]);
    __syncthreads();
    g2p_tile_pass<T, D, FRICTION>(a, t, g, org, &s_v[0][0]);
    __syncthreads();
  }
}

// ------------------------------------------------------------- fused step --
// One workgroup of FB threads (W = FB / 64 wavefronts) per binned tile
// (grid-stride over the tile list): this step's G2P and the next step's P2G
// of the tile's particles, whose state stays on chip in between.  The
// particles are taken FB at a time, one per thread, in the stratified bin
// order (a chunk holds about W particles of each cell):
//   G2P:      thread t gathers its particle from the tile's staged grid
//             region, stores X, V, F, B and, if the new stencil base stays in
//             the tile, computes the particle's P2G coefficients (stress from
//             the new F, m v, m B D^-1, dt tau, x - x_base) in registers;
//   direct:   a wavefront whose lane l holds a particle of the cell labelled
//             l (or none) adds its particles from registers into its copy of
//             the tile's region;
//   grouping: for the other particles, wave-ballot ranks of the new base
//             cells (stable), per-wave counts and offsets in LDS;
//   rounds:   in round k the threads whose rank r in their cell has r / W = k
//             write their coefficients to slot (r % W) * 64 + cell, then lane
//             l of wave w adds slot w * 64 + l into the wave's region copy
//             (p2g_scatter_slot: the 64 lanes of an instruction hold 64
//             distinct cells, so no two touch the same node);
//   flush:    the copies are summed and added to grid_next, one atomic per
//             touched node and component.
// Particles of the current mover list (base outside the tile) are left to
// fused_side_kernel; particles whose new base leaves the tile are appended to
// the next mover list and particles whose stress needs the FP64 SVD to the
// slow list (with tau = 0 here): fused_side_kernel adds their P2G.
template<typename T, int D, bool FRICTION, int FB = BLOCK>
__global__ __launch_bounds__(FB) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? (FRICTION ? 2 : 3) : 1))) void g2p2g_kernel(
    Args<T, D> a) {
  constexpr int R = Tile<D>::R, RN = Tile<D>::RN, NW = FB / 64, E = Tile<D>::E;
  using C = P2GCoef<T, D>;
  using T4 = typename Vec4Type<T>::type;
  __shared__ T4 s_reg[NW][RN];              // per-wave copies of the region: (m, p_x, p_y, p_z)
  __shared__ T4 s_v[FRICTION ? 2 : 1][RN];  // staged grid of the G2P half (and friction velocities)
  __shared__ T s_c[C::N][FB];            // coefficients of one round: slot (r % 4) * 64 + cell
  __shared__ uint32_t s_wc[NW][64];         // per-wave cell counts, then offsets
  __shared__ uint32_t s_cn[64], s_max;
  const uint32_t ntile = a.counts[0];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (uint32_t seg = blockIdx.x; seg < ntile; seg += gridDim.x) {
    const uint32_t t = a.ptiles[seg];
    const Geom<T, D> g = tile_geom(a, t);
    int org[D];
    tile_origin<T, D>(t, g, org);
    g2p_stage<T, D, FRICTION, FB>(a, g, org, &s_v[0][0]);
    for (int k = threadIdx.x; k < NW * RN; k += FB) {
      T4 z = {0, 0, 0, 0};
      s_reg[k / RN][k % RN] = z;
    }
    __syncthreads();
    const uint32_t tb = a.tile_begin[t], te = a.tile_end[t];
    for (uint32_t c0 = tb; c0 < te; c0 += FB) {
      // ---- G2P of one particle per thread; P2G coefficients kept in registers
      uint32_t cell = 64;
      T cf[C::N];
      const uint32_t i = c0 + threadIdx.x;
      if (i < te) {
        const uint32_t p = a.order ? a.order[i] : i;
        if (a.valid[p]) {
          T x[D];
#pragma unroll
          for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
          Wts<T, D> w;
          weights<T, D>(g, x, w);
          bool in = true;
          int lb[D];
#pragma unroll
          for (int c = 0; c < D; c++) {
            lb[c] = w.base[c] - org[c];
            in = in && lb[c] >= 0 && lb[c] < E;
          }
          if (in) {  // otherwise a mover of the current list
            const int li0 = D == 3 ? (lb[0] * R + lb[1]) * R + lb[D - 1] : lb[0] * R + lb[D - 1];
            const T4* sv = &s_v[0][0];
            auto comp = [](const T4& q, int c) -> T { return c == 0 ? q.x : (c == 1 ? q.y : q.z); };
            auto nv = [&](int ox, int oy, int oz, int c) -> T {
              return comp(sv[li0 + (D == 3 ? (ox * R + oy) * R + oz : ox * R + oy)], c);
            };
            auto nf = [&](int ox, int oy, int oz, int c) -> T {
              if constexpr (FRICTION) return comp(sv[RN + li0 + (D == 3 ? (ox * R + oy) * R + oz : ox * R + oy)], c);
              else return nv(ox, oy, oz, c);
            };
            T vb[D], vfb[D], G[D * D], Bs[D * D];
            g2p_gather<T, D, FRICTION>(g, a.dt, w, nv, nf, vb, vfb, G, Bs);
            // F <- (I + dt grad v) F, x <- x + dt v, v <- vf, B; invalidation (MPM_DRIVER.cpp:450-515, :1219-1239)
            T Fn[D * D];
            {
              T Fm[D * D];
#pragma unroll
              for (int k = 0; k < D * D; k++) Fm[k] = a.F[k * a.n + p];
#pragma unroll
              for (int j = 0; j < D; j++)
#pragma unroll
                for (int c = 0; c < D; c++) {
                  T sacc = Fm[c + D * j];
#pragma unroll
                  for (int q = 0; q < D; q++) sacc += a.dt * G[c + D * q] * Fm[q + D * j];
                  Fn[c + D * j] = sacc;
                }
            }
            T xn[D];
            bool ok = true;
#pragma unroll
            for (int c = 0; c < D; c++) {
              xn[c] = x[c] + a.dt * vb[c];
              if (xn[c] < g.vlo[c] || xn[c] > g.vhi[c]) ok = false;
              a.X[c * a.n + p] = xn[c];
              a.V[c * a.n + p] = vfb[c];
            }
#pragma unroll
            for (int k = 0; k < D * D; k++) {
              a.F[k * a.n + p] = Fn[k];
              a.B[k * a.n + p] = Bs[k];
            }
            if (!ok) {
              a.valid[p] = 0;
            } else {
              // the new base: in this tile (P2G here) or a mover of the next step
              bool stay = true;
              uint32_t nt = 0;
              int lc[D];
#pragma unroll
              for (int c = 0; c < D; c++) {
                const int nb = stencil_base<T, D>(g, xn, c) + 3;
                nt = nt * (uint32_t)g.GT[c] + (uint32_t)(nb / E);
                const int q = nb - org[c];
                stay = stay && q >= 0 && q < E;
                lc[c] = q & (E - 1);
              }
              const uint32_t cl = cell_label<D>(lc);
              if (!stay) {
                nt += g.tile0;
                if (nt < a.ntotal) a.mstamp[nt] = a.next_stamp;
                const uint32_t k = atomicAdd(&a.counts[5], 1u);  // aggregated per wavefront by the compiler
                a.mover_list2[k] = p;
                a.mover_tile2[k] = t;
              } else {
                cell = cl;
                const T m = a.mass[p];
                T tau[D * D];
                if (!kirchhoff_fast<T, D>(a.model, a.mu[p], a.la[p], a.thr, a.vol[p], Fn, tau)) {
#pragma unroll
                  for (int q = 0; q < D * D; q++) tau[q] = 0;
                  a.slow_list[atomicAdd(&a.counts[3], 1u)] = p;  // the compiler aggregates this per wavefront
                }
                cf[C::M] = m;
#pragma unroll
                for (int c = 0; c < D; c++) {
                  cf[C::MV + c] = m * vfb[c];
                  const int bb = stencil_base<T, D>(g, xn, c);
                  cf[C::XE + c] = xn[c] - (g.lo[c] + ((T)bb + (T)0.5) * g.dx[c]);
                }
#pragma unroll
                for (int j = 0; j < D; j++) {
                  const T dinv = (T)4 * g.inv_dx[j] * g.inv_dx[j];  // D_p^-1 (PARTICLE_GRID_WEIGHTS_SPLINE.cpp:152-158)
#pragma unroll
                  for (int c = 0; c < D; c++) cf[C::MC + c + D * j] = m * Bs[c + D * j] * dinv;
                }
                int q = 0;
#pragma unroll
                for (int c = 0; c < D; c++)
#pragma unroll
                  for (int j = c; j < D; j++) cf[C::Q + q++] = a.dt * tau[c + D * j];
              }
            }
          }
        }
      }
      // ---- a wavefront whose particles sit each in the cell of its lane's
      // label (the stratified order of an undisturbed tile: one particle per
      // cell, in label order) adds them from registers into its region copy,
      // as the rounds below would with the same lanes
      if (__all(cell == 64 || cell == (uint32_t)lane)) {
        if (cell < 64) {
          int lcell[D];
          cell_offsets<D>((uint32_t)lane, lcell);
          const int lane_li = D == 3 ? (lcell[0] * R + lcell[1]) * R + lcell[D - 1] : lcell[0] * R + lcell[D - 1];
          p2g_scatter_vals<T, D, C>(g, cf, lane_li, s_reg[wv]);
        }
        cell = 64;
      }
      if (!__syncthreads_or(cell < 64)) continue;  // every wavefront done
      // ---- stable ranks of the new base cells
      s_wc[wv][lane] = 0;  // this wave's row only: the wave's own writes below come after it
      const bool act = cell < 64;
      uint64_t mm = __ballot(act);
#pragma unroll
      for (int bb = 0; bb < 6; bb++) {
        const bool bit = (cell >> bb) & 1u;
        const uint64_t bal = __ballot(bit);
        mm &= bit ? bal : ~bal;
      }
      uint32_t rank = (uint32_t)__popcll(mm & lt_mask);
      if (act && lane == __ffsll((unsigned long long)mm) - 1) s_wc[wv][cell] = (uint32_t)__popcll(mm);
      __syncthreads();
      if (threadIdx.x < 64) {  // wave 0: per-cell offsets over the waves, cell totals and the largest
        uint32_t run = 0;
#pragma unroll
        for (int w2 = 0; w2 < NW; w2++) {
          const uint32_t v = s_wc[w2][threadIdx.x];
          s_wc[w2][threadIdx.x] = run;
          run += v;
        }
        uint32_t mx = run;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, 64));
        s_cn[threadIdx.x] = run;
        if (threadIdx.x == 0) s_max = mx;
      }
      __syncthreads();
      if (act) rank += s_wc[wv][cell];
      const uint32_t b_cnt = s_cn[lane];
      const uint32_t rounds = (s_max + NW - 1) / NW;
      for (uint32_t k = 0; k < rounds; k++) {
        if (act && rank / NW == k) {
          const int sl = (int)(rank % NW) * 64 + (int)cell;
#pragma unroll
          for (int q = 0; q < C::N; q++) s_c[q][sl] = cf[q];
        }
        __syncthreads();
        if (k * NW + wv < b_cnt) {
          int lcell[D];
          cell_offsets<D>((uint32_t)lane, lcell);
          const int lane_li = D == 3 ? (lcell[0] * R + lcell[1]) * R + lcell[D - 1] : lcell[0] * R + lcell[D - 1];
          p2g_scatter_slot<T, D, P2GCoef<T, D>, FB>(g, s_c, threadIdx.x, lane_li, s_reg[wv]);
        }
        __syncthreads();
      }
    }
    for (int k = threadIdx.x; k < RN; k += FB) {
      T4 sum = s_reg[0][k];
#pragma unroll
      for (int w2 = 1; w2 < NW; w2++) sum += s_reg[w2][k];
      if (sum.x != (T)0) {
        int off[D], nc[D];
        region_offsets<D>(k, off);
#pragma unroll
        for (int c = 0; c < D; c++) nc[c] = org[c] + off[c];
        Node4<T>* nd = a.grid_next + node_slot<T, D>(nc, g);
        atomic_add(&nd->m, sum.x);
        atomic_add(&nd->x, sum.y);
        atomic_add(&nd->y, sum.z);
        if constexpr (D == 3) atomic_add(&nd->z, sum.w);
      }
    }
    __syncthreads();
  }
}


// Zeroes the grid tiles of a tile list (count at *count).
template<typename T>
__global__ __launch_bounds__(BLOCK) void clear_tiles_kernel(Node4<T>* grid, const uint32_t* tiles, const uint32_t* count) {
  const uint64_t total = (uint64_t)(*count) * 64u;
  Node4<T> z;
  z.m = z.x = z.y = z.z = 0;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK)
    grid[(uint64_t)tiles[idx >> 6] * 64u + (idx & 63u)] = z;
}

// Counter updates of the fused pipeline.  mode 0 (bins rebuilt): the current
// mover list is empty.  mode 2 (end of the priming P2G): the P2G's
// FP64-stress count goes to [11], the slow list is consumed and the next
// mover list starts empty.  (The end-of-step update of a fused step is done
// by the first thread of gcount_kernel.)
__global__ void fused_counts_kernel(uint32_t* counts, int mode) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (mode == 0) {
    counts[4] = 0;
    counts[6] = 0;
    counts[9] = 0;
  } else {
    counts[11] = counts[3];  // FP64-stress particles of the P2G just done
    counts[3] = 0;
    counts[5] = 0;
    counts[7] = 0;
  }
}

// The side path of a fused step in one launch (in place of g2p_mover_kernel
// and p2g_extra_kernel): the current mover list (stencil
// base outside the binned tile at the start of the step: counts[4] entries
// from the bottom of mover_list, counts[6] from the top) is gathered from the
// HBM grid, stored, and then scattered into grid_next together with the
// movers the tile kernel appended (counts[5], bottom of mover_list2) and the
// stress impulse of the slow list (counts[3]).  A current mover still outside
// its binned tile goes to the top of the next list.  The first thread of the
// next kernel (gcount_kernel, lists 2) does the step's list bookkeeping: the
// kernel boundary orders it after every workgroup's appends.
template<typename T, int D, bool FRICTION>
__global__ __launch_bounds__(BLOCK) void fused_side_kernel(Args<T, D> a) {
  __shared__ T s_c[ScatterCoef<T, D>::N][BLOCK];
  __shared__ int s_b[D][BLOCK];
  __shared__ uint16_t s_sc[BLOCK];
  const uint32_t n4 = a.counts[4], n6 = a.counts[6], n5 = a.counts[5];
  const uint32_t nc = n4 + n6, nm = nc + n5, ns = a.counts[3], ntot = nm + ns;
  // each workgroup takes consecutive list entries (the entries one wavefront
  // of the tile kernel appended come from one tile); this measured a little
  // faster than spreading the entries over the workgroups
  const uint32_t per = min((uint32_t)BLOCK, max(1u, (ntot + gridDim.x - 1) / gridDim.x));
  for (uint32_t c0 = blockIdx.x * per; c0 < ntot; c0 += gridDim.x * per) {
    const uint32_t q = c0 + threadIdx.x;
    if (threadIdx.x < per && q < ntot) {
      const bool mover = q < nm;
      uint32_t p;
      if (mover) {
        const uint32_t r = q;
        if (r < nc) {
          // G2P of a current mover from the HBM grid (MPM_DRIVER.cpp:450-515)
          const uint32_t qq = r < n4 ? r : a.mcap - 1u - (r - n4);
          p = a.mover_list[qq];
          const Geom<T, D> g = particle_geom(a, p);
          T x[D];
#pragma unroll
          for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
          Wts<T, D> w;
          weights<T, D>(g, x, w);
          bool in = true;
#pragma unroll
          for (int c = 0; c < D; c++) in = in && w.base[c] >= 0 && w.base[c] + 2 < g.G[c];
          if (in) {  // otherwise flagged by side_coefficients
            auto slot = [&](int ox, int oy, int oz) -> uint32_t {
              int ncd[D];
              ncd[0] = w.base[0] + ox;
              ncd[1] = w.base[1] + oy;
              if constexpr (D == 3) ncd[D - 1] = w.base[D - 1] + oz;
              return node_slot<T, D>(ncd, g);
            };
            auto comp = [](const Node4<T>& nd, int c) -> T { return c == 0 ? nd.x : (c == 1 ? nd.y : nd.z); };
            auto nv = [&](int ox, int oy, int oz, int c) -> T { return comp(a.grid[slot(ox, oy, oz)], c); };
            auto nf = [&](int ox, int oy, int oz, int c) -> T {
              if constexpr (FRICTION) return comp(a.gridf[slot(ox, oy, oz)], c);
              else return nv(ox, oy, oz, c);
            };
            T vb[D], vfb[D], G[D * D], Bs[D * D];
            g2p_gather<T, D, FRICTION>(g, a.dt, w, nv, nf, vb, vfb, G, Bs);
            const uint32_t bt = a.mover_tile[qq];
            if (g2p_store<T, D, false>(a, g, p, x, vb, vfb, G, Bs, bt)) {
              const uint32_t k = a.mcap - 1u - atomicAdd(&a.counts[7], 1u);
              a.mover_list2[k] = p;
              a.mover_tile2[k] = bt;
            }
          }
        } else {
          p = a.mover_list2[r - nc];
        }
      } else {
        p = a.slow_list[q - nm];
      }
      side_coefficients<T, D, true>(a, p, mover, s_c, s_b, s_sc, (int)threadIdx.x);
    }
    __syncthreads();
    side_scatter<T, D>(a, a.grid_next, s_c, s_b, s_sc, min(per, ntot - c0));
    __syncthreads();
  }
}

// G2P of the movers the unfused P2G listed this step (counts[4]), reading
// the grid from HBM.
template<typename T, int D, bool FRICTION>
__global__ __launch_bounds__(BLOCK) void g2p_mover_kernel(Args<T, D> a) {
  const uint32_t nm = a.counts[4];
  // consecutive list entries go to different blocks, so a short list still
  // occupies every compute unit
  for (uint32_t r = threadIdx.x * gridDim.x + blockIdx.x; r < nm; r += gridDim.x * BLOCK) {
    const uint32_t q = spread_index(r, nm);
    const uint32_t p = a.mover_list[q];
    const Geom<T, D> g = particle_geom(a, p);
    T x[D];
#pragma unroll
    for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
    Wts<T, D> w;
    weights<T, D>(g, x, w);
    bool in = true;
#pragma unroll
    for (int c = 0; c < D; c++) in = in && w.base[c] >= 0 && w.base[c] + 2 < g.G[c];
    if (!in) continue;  // flagged by p2g_extra_kernel
    auto slot = [&](int ox, int oy, int oz) -> uint32_t {
      int nc[D];
      nc[0] = w.base[0] + ox;
      nc[1] = w.base[1] + oy;
      if constexpr (D == 3) nc[D - 1] = w.base[D - 1] + oz;
      return node_slot<T, D>(nc, g);
    };
    auto comp = [](const Node4<T>& nd, int c) -> T { return c == 0 ? nd.x : (c == 1 ? nd.y : nd.z); };
    auto nv = [&](int ox, int oy, int oz, int c) -> T { return comp(a.grid[slot(ox, oy, oz)], c); };
    auto nf = [&](int ox, int oy, int oz, int c) -> T {
      if constexpr (FRICTION) return comp(a.gridf[slot(ox, oy, oz)], c);
      else return nv(ox, oy, oz, c);
    };
    T vb[D], vfb[D], G[D * D], Bs[D * D];
    g2p_gather<T, D, FRICTION>(g, a.dt, w, nv, nf, vb, vfb, G, Bs);
    g2p_store<T, D, true>(a, g, p, x, vb, vfb, G, Bs, a.mover_tile[q]);
  }
}

// Marks, for this step, the tiles holding the stencil base of particles that
// are outside their binned tile (used after a state is restored while the
// bins stay: P2G then treats those particles as movers).
// With LIST, movers are appended to mover_list (counts[4]) instead.
template<typename T, int D, bool LIST>
__global__ __launch_bounds__(BLOCK) void mark_movers_kernel(Args<T, D> a, uint32_t stamp) {
  const uint32_t ntile = a.counts[0];
  for (uint32_t seg = blockIdx.x; seg < ntile; seg += gridDim.x) {
    const uint32_t t = a.ptiles[seg];
    const Geom<T, D> g = tile_geom(a, t);
    const uint32_t b = a.tile_begin[t], e = a.tile_end[t];
    for (uint32_t i0 = b; i0 < e; i0 += BLOCK) {  // uniform trip count: the append votes
      const uint32_t i = i0 + threadIdx.x;
      bool mover = false;
      uint32_t p = 0, nt = t;
      if (i < e) {
        p = a.order ? a.order[i] : i;
        if (a.valid[p]) {
          T x[D];
#pragma unroll
          for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
          nt = base_tile<T, D>(g, x);
          mover = nt != t;
        }
      }
      if (LIST) {
        const uint32_t q = wave_append(&a.counts[4], mover);
        if (mover) a.mover_list[q] = p;
      } else if (mover && nt < a.ntotal) {
        a.mstamp[nt] = stamp;
      }
    }
  }
}

// Per-step grid-tile lists kept for the reverse pass.
__global__ __launch_bounds__(BLOCK) void save_tiles_kernel(const uint32_t* gtiles, const uint32_t* counts,
                                                           uint32_t* dst, uint32_t* dst_count) {
  const uint32_t nt = counts[1];
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < nt; i += gridDim.x * BLOCK) dst[i] = gtiles[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst_count = nt;
}

// Reverse step set-up: counts[1] <- the step's tile count, counts[4] <- 0, and
// the node adjoints of the step's tiles cleared.
template<typename T>
__global__ __launch_bounds__(BLOCK) void restore_tiles_kernel(const uint32_t* tiles, const uint32_t* count,
                                                              uint32_t* counts, Node4<T>* agrid) {
  const uint32_t nt = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counts[1] = nt;
    counts[4] = 0;
  }
  Node4<T> z;
  z.m = z.x = z.y = z.z = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < (uint64_t)nt * 64u;
       i += (uint64_t)gridDim.x * BLOCK)
    agrid[tiles[i >> 6] * 64u + (uint32_t)(i & 63u)] = z;
}

// One launch copying X, V, F, B and valid between two particle buffers.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void copy_state_kernel(const T* sX, const T* sV, const T* sF, const T* sB,
                                                           const uint8_t* sv, T* dX, T* dV, T* dF, T* dB, uint8_t* dv,
                                                           uint32_t n) {
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n; p += gridDim.x * BLOCK) {
#pragma unroll
    for (int c = 0; c < D; c++) {
      dX[c * n + p] = sX[c * n + p];
      dV[c * n + p] = sV[c * n + p];
    }
#pragma unroll
    for (int k = 0; k < D * D; k++) {
      dF[k * n + p] = sF[k * n + p];
      dB[k * n + p] = sB[k * n + p];
    }
    dv[p] = sv[p];
  }
}

// ------------------------------------------------------------- adjoint ---
// Reverse of one step (formulas in mpmx_adjoint.h).  Per-particle kernels with
// HBM atomics for the node scatter; the forward P2G and grid update of the
// step are recomputed first, with (m, u) kept in gridu.
template<typename T> struct AdjArgs {
  T* aX; T* aV; T* aF; T* aB;  // state adjoints, SoA like the state
  double* amu; double* ala;    // per-particle Lame adjoints, accumulated over steps
  Node4<T>* agrid;             // vbar (G2P adjoint), then (mbar, Pbar)
  const Node4<T>* gridu;       // (m, u) of the recomputed step
  uint32_t split;              // g2p_adjoint_kernel: workgroups sharing one tile's particle chunks
};

// G2P adjoint of particle p, gather part (formulas in mpmx_adjoint.h): node
// velocities come from nv(index(o), c).  Writes xbar and Fbar of p and returns
// in coef the node-scatter coefficients: vt = v'bar + dt x'bar, B'bar and
// Gbar (ScatterCoef layout: MV = vt, MC = B'bar, TAU = -Gbar, M unused).
template<typename T, int D> struct AdjCoef {
  static constexpr int VT = 0, BB = D, GB = D + D * D, XE = D + 2 * D * D, N = XE + D;
};

template<typename T, int D, class Index, class NV>
__device__ __forceinline__ void g2p_adjoint_gather(const Args<T, D>& a, const Geom<T, D>& g, const AdjArgs<T>& q,
                                                   uint32_t p, const Wts<T, D>& w, Index index, NV nv, T* coef) {
  using AC = AdjCoef<T, D>;
  T Fm[D * D], xb[D], vt[D], Fb[D * D], Bb[D * D], Gb[D * D], G[D * D];
#pragma unroll
  for (int c = 0; c < D; c++) {
    xb[c] = q.aX[c * a.n + p];
    vt[c] = q.aV[c * a.n + p] + a.dt * xb[c];
  }
#pragma unroll
  for (int k = 0; k < D * D; k++) {
    Fm[k] = a.F[k * a.n + p];
    Fb[k] = q.aF[k * a.n + p];
    Bb[k] = q.aB[k * a.n + p];
    G[k] = 0;
  }
#pragma unroll
  for (int c = 0; c < D; c++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T s = 0;
#pragma unroll
      for (int k = 0; k < D; k++) s += Fb[c + D * k] * Fm[j + D * k];
      Gb[c + D * j] = a.dt * s;
    }
#pragma unroll
  for (int k = 0; k < Tile<D>::K; k++) {
    int o[D];
    T wt, gr[D], H[D][D], dz[D], v[D];
    stencil_entry<T, D>(g, w, k, o, wt, gr, H);
#pragma unroll
    for (int c = 0; c < D; c++) dz[c] = (T)o[c] * g.dx[c] - w.xe[c];
    const uint32_t ix = index(o);
#pragma unroll
    for (int c = 0; c < D; c++) v[c] = nv(ix, c);
    T s1 = 0, hv[D], bv[D];
#pragma unroll
    for (int c = 0; c < D; c++) {
      T bdz = 0;
#pragma unroll
      for (int j = 0; j < D; j++) {
        G[c + D * j] += v[c] * gr[j];
        bdz += Bb[c + D * j] * dz[j];
      }
      s1 += v[c] * (vt[c] + bdz);
    }
#pragma unroll
    for (int j = 0; j < D; j++) {
      T h = 0, bb = 0;
#pragma unroll
      for (int c = 0; c < D; c++) {
        h += Gb[c + D * j] * v[c];
        bb += Bb[c + D * j] * v[c];
      }
      hv[j] = h;
      bv[j] = bb;
    }
#pragma unroll
    for (int c = 0; c < D; c++) {
      T hh = 0;
#pragma unroll
      for (int bb = 0; bb < D; bb++) hh += H[c][bb] * hv[bb];
      xb[c] += s1 * gr[c] + hh - wt * bv[c];
    }
  }
#pragma unroll
  for (int c = 0; c < D; c++) q.aX[c * a.n + p] = xb[c];
#pragma unroll
  for (int c = 0; c < D; c++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T s = Fb[c + D * j];
#pragma unroll
      for (int r = 0; r < D; r++) s += a.dt * G[r + D * c] * Fb[r + D * j];
      q.aF[(c + D * j) * a.n + p] = s;
    }
#pragma unroll
  for (int c = 0; c < D; c++) {
    coef[AC::VT + c] = vt[c];
    coef[AC::XE + c] = w.xe[c];
  }
#pragma unroll
  for (int k = 0; k < D * D; k++) {
    coef[AC::BB + k] = Bb[k];
    coef[AC::GB + k] = Gb[k];
  }
}

// Scatter part of the G2P adjoint for stencil node k of a particle whose
// coefficients are column sl of s_c: vbar_i += w vt + Gbar grad w + w B'bar dz.
template<typename T, int D, int NC>
__device__ __forceinline__ void g2p_adjoint_node(const Geom<T, D>& g, const T (*s_c)[BLOCK], uint32_t sl, int k,
                                                 int* o, T* vb) {
  using AC = AdjCoef<T, D>;
  T xe[D], wt, gr[D], dz[D];
#pragma unroll
  for (int c = 0; c < D; c++) xe[c] = s_c[AC::XE + c][sl];
  node_weight<T, D>(g, xe, k, o, wt, gr, dz);
#pragma unroll
  for (int c = 0; c < D; c++) {
    T s = s_c[AC::VT + c][sl], m = 0;
#pragma unroll
    for (int j = 0; j < D; j++) {
      s += s_c[AC::BB + c + D * j][sl] * dz[j];
      m += s_c[AC::GB + c + D * j][sl] * gr[j];
    }
    vb[c] = wt * s + m;
  }
}

// q.split workgroups per binned tile (chunk k of the tile goes to workgroup
// k mod split, so that scenes with few tiles still fill the device): the
// region's velocities are staged in LDS; the particles are grouped by base
// cell in chunks of JT * BLOCK as in p2g_kernel, phase A gathers each particle
// (x-bar, F-bar, ... written to HBM) and leaves its vbar scatter coefficients
// in LDS, phase B scatters them with the P2G scatter (lanes = cells,
// per-wavefront copies of the region, no LDS atomics); one HBM atomic per
// node and component at each workgroup's flush.  Movers are left to
// g2p_adjoint_mover_kernel.
template<typename T, int D, int JT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1))) void g2p_adjoint_kernel(Args<T, D> a, AdjArgs<T> q) {
  using AC = AdjCoef<T, D>;
  using SC = AdjScatterCoef<T, D>;
  using T4 = typename Vec4Type<T>::type;
  static_assert(SC::N == AC::N, "adjoint coefficient layouts");
  constexpr int R = Tile<D>::R, RN = Tile<D>::RN, NW = BLOCK / 64, E = Tile<D>::E;
  constexpr int CH = JT * BLOCK;
  __shared__ T s_v[D][RN];
  __shared__ T4 s_reg[NW][RN];
  __shared__ union {
    T c[SC::N][BLOCK];
    uint32_t wc[JT][NW][64];
  } s_u;
  __shared__ uint32_t s_idx[CH];
  __shared__ uint32_t s_cb[64], s_cn[64], s_max;
  auto& s_c = s_u.c;
  const uint32_t ntile = a.counts[0];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int a_cell = threadIdx.x >> 2, a_j = threadIdx.x & 3;
  const int a_slot = a_j * 64 + a_cell;
  const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t split = q.split;
  for (uint32_t item = blockIdx.x; item < ntile * split; item += gridDim.x) {
    const uint32_t seg = item / split, sub = item - seg * split;
    const uint32_t t = a.ptiles[seg];
    const uint32_t tb = a.tile_begin[t], te = a.tile_end[t];
    if (tb + sub * CH >= te) continue;  // uniform over the workgroup: no chunk here
    const Geom<T, D> g = tile_geom(a, t);
    int org[D];
    tile_origin<T, D>(t, g, org);
    for (int k = threadIdx.x; k < RN; k += BLOCK) {
      int off[D], nc[D];
      region_offsets<D>(k, off);
#pragma unroll
      for (int c = 0; c < D; c++) nc[c] = org[c] + off[c];
      const Node4<T> nd = a.grid[node_slot<T, D>(nc, g)];
      s_v[0][k] = nd.x;
      s_v[1][k] = nd.y;
      if constexpr (D == 3) s_v[2][k] = nd.z;
    }
    for (int k = threadIdx.x; k < NW * RN; k += BLOCK) {
      T4 z = {0, 0, 0, 0};
      s_reg[k / RN][k % RN] = z;
    }
    for (uint32_t c0 = tb + sub * CH; c0 < te; c0 += split * CH) {
      for (int k = threadIdx.x; k < JT * NW * 64; k += BLOCK) (&s_u.wc[0][0][0])[k] = 0;
      __syncthreads();
      // ---- group the chunk by base cell (movers are skipped)
      uint32_t cell_j[JT], rank_j[JT], slot_j[JT];
#pragma unroll
      for (int j = 0; j < JT; j++) {
        const uint32_t i = c0 + j * BLOCK + threadIdx.x;
        uint32_t cell = 64, p = 0;
        if (i < te) {
          p = a.order ? a.order[i] : i;
          if (a.valid[p]) {
            T x[D];
#pragma unroll
            for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
            bool in = true;
            int lc[D];
#pragma unroll
            for (int c = 0; c < D; c++) {
              const int lb = stencil_base<T, D>(g, x, c) + 3 - org[c];
              in = in && lb >= 0 && lb < E;
              lc[c] = lb & (E - 1);
            }
            const uint32_t cl = cell_label<D>(lc);
            if (in) cell = cl;
          }
        }
        const bool act = cell < 64;
        uint64_t m = __ballot(act);
#pragma unroll
        for (int bb = 0; bb < 6; bb++) {
          const bool bit = (cell >> bb) & 1u;
          const uint64_t bal = __ballot(bit);
          m &= bit ? bal : ~bal;
        }
        rank_j[j] = (uint32_t)__popcll(m & lt_mask);
        if (act && lane == __ffsll((unsigned long long)m) - 1) s_u.wc[j][wv][cell] = (uint32_t)__popcll(m);
        cell_j[j] = cell;
        slot_j[j] = p;
      }
      __syncthreads();
      if (threadIdx.x < 64) {
        uint32_t run = 0;
#pragma unroll
        for (int j = 0; j < JT; j++)
#pragma unroll
          for (int w = 0; w < NW; w++) {
            const uint32_t v = s_u.wc[j][w][threadIdx.x];
            s_u.wc[j][w][threadIdx.x] = run;
            run += v;
          }
        uint32_t x = run, mx = run;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(x, o, 64);
          if (lane >= o) x += y;
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, 64));
        s_cb[threadIdx.x] = x - run;
        s_cn[threadIdx.x] = run;
        if (threadIdx.x == 0) s_max = mx;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < JT; j++)
        if (cell_j[j] < 64) s_idx[s_cb[cell_j[j]] + s_u.wc[j][wv][cell_j[j]] + rank_j[j]] = slot_j[j];
      __syncthreads();
      const uint32_t a_cb = s_cb[a_cell], a_cnt = s_cn[a_cell];
      const uint32_t b_cnt = s_cn[lane];
      const uint32_t rounds = (s_max + NW - 1) / NW;
      for (uint32_t k = 0; k < rounds; k++) {
        // ---- phase A: gather of particle 4k + a_j of cell a_cell
        const uint32_t ja = k * NW + a_j;
        if (ja < a_cnt) {
          const uint32_t p = s_idx[a_cb + ja];
          T x[D];
#pragma unroll
          for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
          Wts<T, D> w;
          weights<T, D>(g, x, w);
          int lb[D];
#pragma unroll
          for (int c = 0; c < D; c++) lb[c] = w.base[c] - org[c];
          const int li0 = D == 3 ? (lb[0] * R + lb[1]) * R + lb[D - 1] : lb[0] * R + lb[D - 1];
          auto index = [&](const int* o) -> uint32_t {
            return (uint32_t)(li0 + (D == 3 ? (o[0] * R + o[1]) * R + o[D - 1] : o[0] * R + o[D - 1]));
          };
          auto nv = [&](uint32_t ix, int c) -> T { return s_v[c][ix]; };
          T coef[AC::N];
          g2p_adjoint_gather<T, D>(a, g, q, p, w, index, nv, coef);
#pragma unroll
          for (int kk = 0; kk < AC::N; kk++)
            s_c[kk][a_slot] = kk >= AC::GB && kk < AC::GB + D * D ? -coef[kk] : coef[kk];
        }
        __syncthreads();
        // ---- phase B
        if (k * NW + wv < b_cnt) {
          int lcell[D];
          cell_offsets<D>((uint32_t)lane, lcell);
          const int lane_li = D == 3 ? (lcell[0] * R + lcell[1]) * R + lcell[D - 1] : lcell[0] * R + lcell[D - 1];
          p2g_scatter_slot<T, D, SC>(g, s_c, threadIdx.x, lane_li, s_reg[wv]);
        }
        __syncthreads();
      }
    }
    for (int k = threadIdx.x; k < RN; k += BLOCK) {
      T4 sum = s_reg[0][k];
#pragma unroll
      for (int w = 1; w < NW; w++) sum += s_reg[w][k];
      int off[D], nc[D];
      region_offsets<D>(k, off);
#pragma unroll
      for (int c = 0; c < D; c++) nc[c] = org[c] + off[c];
      Node4<T>* an = q.agrid + node_slot<T, D>(nc, g);
      if (sum.y != (T)0) atomic_add(&an->x, sum.y);
      if (sum.z != (T)0) atomic_add(&an->y, sum.z);
      if constexpr (D == 3)
        if (sum.w != (T)0) atomic_add(&an->z, sum.w);
    }
    __syncthreads();
  }
}

// G2P adjoint of this step's movers, straight from and to HBM (same two
// passes as above, the scatter as HBM atomics).
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void g2p_adjoint_mover_kernel(Args<T, D> a, AdjArgs<T> q) {
  using AC = AdjCoef<T, D>;
  __shared__ T s_c[AC::N][BLOCK];
  __shared__ int s_base[D][BLOCK];
  __shared__ uint16_t s_sc[BLOCK];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t nm = a.counts[4];
  for (uint32_t c0 = blockIdx.x * BLOCK; c0 < nm; c0 += gridDim.x * BLOCK) {
    const uint32_t r = c0 + threadIdx.x;
    s_base[0][threadIdx.x] = -1;
    if (r < nm) {
      const uint32_t p = a.mover_list[spread_index(r, nm)];
      const Geom<T, D> g = particle_geom(a, p);
      T x[D];
#pragma unroll
      for (int c = 0; c < D; c++) x[c] = a.X[c * a.n + p];
      Wts<T, D> w;
      weights<T, D>(g, x, w);
      bool in = true;
#pragma unroll
      for (int c = 0; c < D; c++) in = in && w.base[c] >= 0 && w.base[c] + 2 < g.G[c];
      if (in) {
        auto index = [&](const int* o) -> uint32_t {
          int nc[D];
#pragma unroll
          for (int c = 0; c < D; c++) nc[c] = w.base[c] + o[c];
          return node_slot<T, D>(nc, g);
        };
        auto nv = [&](uint32_t ix, int c) -> T {
          const Node4<T>& nd = a.grid[ix];
          return c == 0 ? nd.x : (c == 1 ? nd.y : nd.z);
        };
        T coef[AC::N];
        g2p_adjoint_gather<T, D>(a, g, q, p, w, index, nv, coef);
#pragma unroll
        for (int k = 0; k < AC::N; k++) s_c[k][threadIdx.x] = coef[k];
#pragma unroll
        for (int c = 0; c < D; c++) s_base[c][threadIdx.x] = w.base[c];
        s_sc[threadIdx.x] = a.pscene ? a.pscene[p] : 0;
      }
    }
    __syncthreads();
    const uint32_t cnt = min((uint32_t)BLOCK, nm - c0);
    for (uint32_t pp = wv * 2; pp < cnt; pp += 2 * (BLOCK / 64)) {
      const uint32_t sl = pp + (lane >> 5);
      const int k = lane & 31;
      if (sl >= cnt || k >= Tile<D>::K || s_base[0][sl] < 0) continue;
      const Geom<T, D> g = a.scenes[s_sc[sl]];
      int o[D], nc[D];
      T vb[D];
      g2p_adjoint_node<T, D, AC::N>(g, s_c, sl, k, o, vb);
#pragma unroll
      for (int c = 0; c < D; c++) nc[c] = s_base[c][sl] + o[c];
      Node4<T>* an = q.agrid + node_slot<T, D>(nc, g);
      atomic_add(&an->x, vb[0]);
      atomic_add(&an->y, vb[1]);
      if constexpr (D == 3) atomic_add(&an->z, vb[D - 1]);
    }
    __syncthreads();
  }
}

// Node adjoint: vbar back through the collision objects (replayed from u, in
// reverse order), then Pbar = ubar / m and mbar = -ubar . (u - dt g) / m.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void grid_adjoint_kernel(Args<T, D> a, AdjArgs<T> q) {
  const uint64_t total = (uint64_t)a.counts[1] * 64u;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t t = a.gtiles[idx >> 6];
    const uint32_t l = (uint32_t)(idx & 63u);
    const uint32_t slot = t * 64u + l;
    const Node4<T> u = q.gridu[slot];
    Node4<T> out;
    out.m = out.x = out.y = out.z = 0;
    if (u.m == (T)0) {
      q.agrid[slot] = out;
      continue;
    }
    const Node4<T> vbn = q.agrid[slot];
    T v[D], vb[D];
    v[0] = u.x;
    v[1] = u.y;
    if constexpr (D == 3) v[2] = u.z;
    vb[0] = vbn.x;
    vb[1] = vbn.y;
    if constexpr (D == 3) vb[2] = vbn.z;
    if (a.nwalls > 0) {
      const Geom<T, D> g = tile_geom(a, t);
      int org[D], nc[D];
      tile_origin<T, D>(t, g, org);
      uint32_t r = l;
#pragma unroll
      for (int c = D - 1; c >= 0; c--) {
        nc[c] = org[c] + (int)(r % Tile<D>::E);
        r /= Tile<D>::E;
      }
      T X0[D];
#pragma unroll
      for (int c = 0; c < D; c++) X0[c] = g.lo[c] + ((T)(nc[c] - 3) + (T)0.5) * g.dx[c];
      T u0[D], dummy[D];
      int ob;
#pragma unroll
      for (int c = 0; c < D; c++) u0[c] = v[c];
      const int nh = walls_apply<T, D>(a, X0, u0, -1, dummy, &ob);
      for (int h = nh - 1; h >= 0; h--) {
        T vq[D], nrm[D];
#pragma unroll
        for (int c = 0; c < D; c++) vq[c] = v[c];
        if (walls_apply<T, D>(a, X0, vq, h, nrm, &ob) == 0) {
#pragma unroll
          for (int c = 0; c < D; c++) vb[c] = 0;
        } else {
          T dn = 0;
#pragma unroll
          for (int c = 0; c < D; c++) dn += nrm[c] * vb[c];
#pragma unroll
          for (int c = 0; c < D; c++) vb[c] -= dn * nrm[c];
        }
      }
    }
    const T im = (T)1 / u.m;
    T mb = 0;
    T ub[3] = {0, 0, 0};
    const T uu[3] = {u.x, u.y, u.z};
#pragma unroll
    for (int c = 0; c < D; c++) {
      ub[c] = vb[c] * im;
      mb -= ub[c] * (uu[c] - a.dt * a.grav[c]);
    }
    out.m = mb;
    out.x = ub[0];
    out.y = ub[1];
    out.z = ub[2];
    q.agrid[slot] = out;
  }
}

template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void p2g_adjoint_kernel(Args<T, D> a, AdjArgs<T> q) {
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < a.n; p += gridDim.x * BLOCK) {
    if (!a.valid[p]) continue;
    const Geom<T, D> g = particle_geom(a, p);
    T x[D], v[D], Cm[D * D], Fm[D * D], tau[D * D], Dinv[D];
#pragma unroll
    for (int c = 0; c < D; c++) {
      x[c] = a.X[c * a.n + p];
      v[c] = a.V[c * a.n + p];
      Dinv[c] = (T)4 * g.inv_dx[c] * g.inv_dx[c];
    }
#pragma unroll
    for (int c = 0; c < D; c++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        Cm[c + D * j] = a.B[(c + D * j) * a.n + p] * Dinv[j];
        Fm[c + D * j] = a.F[(c + D * j) * a.n + p];
      }
    const T m = a.mass[p], vol = a.vol[p], mu = a.mu[p], la = a.la[p];
    kirchhoff_any<T, D>(a.model, mu, la, a.thr, a.root, vol, Fm, tau);
    Wts<T, D> w;
    weights<T, D>(g, x, w);
    bool in = true;
#pragma unroll
    for (int c = 0; c < D; c++) in = in && w.base[c] >= 0 && w.base[c] + 2 < g.G[c];
    if (!in) continue;
    T vbar[D], Cbar[D * D], tbar[D * D], xbar[D];
#pragma unroll
    for (int c = 0; c < D; c++) vbar[c] = xbar[c] = 0;
#pragma unroll
    for (int k = 0; k < D * D; k++) Cbar[k] = tbar[k] = 0;
#pragma unroll
    for (int k = 0; k < Tile<D>::K; k++) {
      int o[D], nc[D];
      T wt, gr[D], H[D][D], dz[D];
      stencil_entry<T, D>(g, w, k, o, wt, gr, H);
#pragma unroll
      for (int c = 0; c < D; c++) {
        nc[c] = w.base[c] + o[c];
        dz[c] = (T)o[c] * g.dx[c] - w.xe[c];
      }
      const Node4<T> nb = q.agrid[node_slot<T, D>(nc, g)];
      const T Pb[3] = {nb.x, nb.y, nb.z};
      T s = nb.m * m, tp[D], ct[D];
#pragma unroll
      for (int c = 0; c < D; c++) {
        vbar[c] += m * wt * Pb[c];
        T ap = v[c];
#pragma unroll
        for (int j = 0; j < D; j++) {
          Cbar[c + D * j] += m * wt * Pb[c] * dz[j];
          tbar[c + D * j] -= a.dt * Pb[c] * gr[j];
          ap += Cm[c + D * j] * dz[j];
        }
        s += m * Pb[c] * ap;
      }
#pragma unroll
      for (int j = 0; j < D; j++) {
        T t1 = 0, t2 = 0;
#pragma unroll
        for (int c = 0; c < D; c++) {
          t1 += tau[j + D * c] * Pb[c];
          t2 += Cm[c + D * j] * Pb[c];
        }
        tp[j] = t1;
        ct[j] = t2;
      }
#pragma unroll
      for (int c = 0; c < D; c++) {
        T hh = 0;
#pragma unroll
        for (int b = 0; b < D; b++) hh += H[c][b] * tp[b];
        xbar[c] += s * gr[c] - m * wt * ct[c] - a.dt * hh;
      }
    }
#pragma unroll
    for (int c = 0; c < D; c++) {
      q.aV[c * a.n + p] = vbar[c];
      q.aX[c * a.n + p] += xbar[c];
    }
#pragma unroll
    for (int c = 0; c < D; c++)
#pragma unroll
      for (int j = 0; j < D; j++) q.aB[(c + D * j) * a.n + p] = Cbar[c + D * j] * Dinv[j];
    // the stress adjoint in the context precision; FP64 (the SVD form) only
    // for an inverted or nearly singular F
    double mub = 0, lab = 0;
    T Fbt[D * D];
    int st = 1;
    if constexpr (sizeof(T) == 4) st = kirchhoff_adjoint_fast<T, D>(a.model, mu, la, a.thr, vol, Fm, tbar, Fbt, mub, lab);
    if (st == 2) {
      atomicOr(a.err + 2, 1u);
      continue;
    }
    if (st == 1) {
      double Fd[D * D], tb[D * D], Fbar[D * D];
#pragma unroll
      for (int k = 0; k < D * D; k++) {
        Fd[k] = (double)Fm[k];
        tb[k] = (double)tbar[k];
      }
      if (!kirchhoff_adjoint<D>(a.model, (double)mu, (double)la, (double)a.thr, (double)vol, Fd, tb, Fbar, mub, lab)) {
        atomicOr(a.err + 2, 1u);
        continue;
      }
#pragma unroll
      for (int k = 0; k < D * D; k++) Fbt[k] = (T)Fbar[k];
    }
#pragma unroll
    for (int k = 0; k < D * D; k++) q.aF[k * a.n + p] += Fbt[k];
    q.amu[p] += mub;
    q.ala[p] += lab;
  }
}

template<typename T>
__global__ __launch_bounds__(BLOCK) void soa_from_aos_pid_kernel(const T* aos, T* soa, const uint32_t* pid, uint32_t n,
                                                                 int comps) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t q = pid[i];
  for (int c = 0; c < comps; c++) soa[(size_t)c * n + i] = aos ? aos[(size_t)q * comps + c] : (T)0;
}

// -------------------------------------------------------------- reorder ----
template<typename T, int D> struct ParticleBufs {
  T* X; T* V; T* F; T* B; T* mass; T* vol; T* mu; T* la;
  uint8_t* valid;
  uint32_t* pid;
  uint16_t* pscene;  // scene of each particle (batches only)
  T* Fp; T* mu0; T* la0;  // plastic deformation, initial Lame parameters (plasticity only)
  T* pdef; T* rhoF; T* coh; T* vc;  // Drucker-Prager: plastic_def, rho_F, cohesion, volume correction
};

template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void reorder_kernel(ParticleBufs<T, D> src, ParticleBufs<T, D> dst,
                                                        const uint32_t* order, uint32_t n) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = order[i];
#pragma unroll
  for (int c = 0; c < D; c++) {
    dst.X[c * n + i] = src.X[c * n + p];
    dst.V[c * n + i] = src.V[c * n + p];
  }
#pragma unroll
  for (int k = 0; k < D * D; k++) {
    dst.F[k * n + i] = src.F[k * n + p];
    dst.B[k * n + i] = src.B[k * n + p];
  }
  dst.mass[i] = src.mass[p];
  dst.vol[i] = src.vol[p];
  dst.mu[i] = src.mu[p];
  dst.la[i] = src.la[p];
  dst.valid[i] = src.valid[p];
  dst.pid[i] = src.pid[p];
  if (src.pscene) dst.pscene[i] = src.pscene[p];
  if (src.Fp) {
#pragma unroll
    for (int k = 0; k < D * D; k++) dst.Fp[k * n + i] = src.Fp[k * n + p];
    dst.mu0[i] = src.mu0[p];
    dst.la0[i] = src.la0[p];
    dst.pdef[i] = src.pdef[p];
    dst.rhoF[i] = src.rhoF[p];
    dst.coh[i] = src.coh[p];
    dst.vc[i] = src.vc[p];
  }
}

// ---------------------------------------------------------- plasticity ----
// Plasticity after G2P (MPM_PLASTICITY_MODEL::Update_Particles,
// MPM_PLASTICITY_MODEL.cpp:34-48) for the particles simulated in the step
// (valid at its start: `simulated`), computed in R: F = U S V^T, the model maps S
// to Sc, and if it changed, F = U Sc V^T and Fp = V Sc^-1 U^T F Fp.
//   clamp (MPM_PLASTICITY_CLAMP.cpp:31-45): Sc = clamp(S, 1 - theta_c,
//     1 + theta_s); h = max(1e-8, exp(min(max_hardening, xi (1 - prod(S) /
//     prod(Sc) det Fp)))), mu = mu0 h, lambda = lambda0 h;
//   Drucker-Prager (MPM_DRUCKER_PRAGER.cpp:64-94): return mapping of ln S
//     onto the cone, with hardening of the friction angle.
// R is the particle precision for the clamp, which is continuous in F; the
// Drucker-Prager mapping runs in FP64: with a volume correction vc != 0 it
// jumps at the cone's edge (dg = 0: unprojected S against S exp(vc / d)), so
// FP32 rounding of dg would move particles across the jump.
template<typename T> struct PlasticDev {
  int kind;                           // 1 clamp, 2 Drucker-Prager
  double lo, hi, max_hardening, xi;   // clamp
  double a0, a1, a3, a4, cohesion;    // Drucker-Prager (friction angle in degrees)
  int use_vc;
};

// MPM_DRUCKER_PRAGER::Update_Hardening (MPM_DRUCKER_PRAGER.cpp:123-131): the
// friction angle phi = a0 + (a1 q - a4) exp(-a3 q) degrees at plastic
// deformation q, rho_F = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)).
template<typename T, typename R = double> __device__ __forceinline__ R dp_rho_f(const PlasticDev<T>& pl, R q) {
  const R phi = (R)(3.14159265358979323846 / 180) * ((R)pl.a0 + ((R)pl.a1 * q - (R)pl.a4) * exp(-(R)pl.a3 * q));
  const R sp = sin(phi);
  return sqrt((R)(2.0 / 3.0)) * 2 * sp / (3 - sp);
}

template<typename T, int D, typename R>
__global__ __launch_bounds__(BLOCK) void plasticity_kernel(ParticleBufs<T, D> b, uint32_t n, const uint8_t* simulated,
                                                           PlasticDev<T> pl) {
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n; p += gridDim.x * BLOCK) {
    if (!simulated[p]) continue;
    R Fe[D * D], U[D * D], S[D], V[D * D];
#pragma unroll
    for (int k = 0; k < D * D; k++) Fe[k] = (R)b.F[(size_t)k * n + p];
    svd<R, D>(Fe, U, S, V);
    R Sc[D];
    R Fp[D * D];
#pragma unroll
    for (int k = 0; k < D * D; k++) Fp[k] = (R)b.Fp[(size_t)k * n + p];
    if (pl.kind == 1) {
      bool same = true;
#pragma unroll
      for (int a = 0; a < D; a++) {
        Sc[a] = S[a] <= (R)pl.lo ? (R)pl.lo : (S[a] >= (R)pl.hi ? (R)pl.hi : S[a]);
        same = same && Sc[a] == S[a];
      }
      if (same) continue;
      R ps = 1, pc = 1;
#pragma unroll
      for (int a = 0; a < D; a++) {
        ps *= S[a];
        pc *= Sc[a];
      }
      R detp;
      if constexpr (D == 3)
        detp = Fp[0] * (Fp[4] * Fp[8] - Fp[7] * Fp[5]) + Fp[3] * (Fp[7] * Fp[2] - Fp[1] * Fp[8]) +
               Fp[6] * (Fp[1] * Fp[5] - Fp[4] * Fp[2]);
      else
        detp = Fp[0] * Fp[3] - Fp[1] * Fp[2];
      const R det = ps / pc * detp;
      R h = exp(fmin((R)pl.max_hardening, (R)pl.xi * (1 - det)));
      if (h < (R)1e-8) h = (R)1e-8;
      b.mu[p] = (T)((R)b.mu0[p] * h);
      b.la[p] = (T)((R)b.la0[p] * h);
    } else {
      // MPM_DRUCKER_PRAGER::Compute (MPM_DRUCKER_PRAGER.cpp:64-94): return
      // mapping of the log strain onto the Drucker-Prager cone
      const R mu = (R)b.mu[p], la = (R)b.la[p];
      const R g = 1 / (D * la + 2 * mu), bb = 1 / (2 * mu);
      const R rho = (R)b.rhoF[p], coh = (R)b.coh[p];
      R vcp = (R)b.vc[p], pdef = (R)b.pdef[p];
      const R beta = -coh * g / (rho * D);
      R lf[D], sh[D], k = 0;
#pragma unroll
      for (int a = 0; a < D; a++) {
        lf[a] = log(fabs(S[a]));
        sh[a] = lf[a] + beta + vcp / D;
        k += sh[a];
      }
      R q2 = 0;
#pragma unroll
      for (int a = 0; a < D; a++) {
        sh[a] -= k / D;
        q2 += sh[a] * sh[a];
      }
      const R q = sqrt(q2), r = rho * k / g, dg = q + bb * r;
      if (dg <= 0) continue;
      R inc;
      if (q == 0 || k > 0) {  // tension: to the tip of the cone
#pragma unroll
        for (int a = 0; a < D; a++) Sc[a] = coh != 0 ? exp(-beta) : (R)1;
        R ls = 0, l2 = 0;
#pragma unroll
        for (int a = 0; a < D; a++) {
          ls += lf[a];
          l2 += lf[a] * lf[a];
        }
        if (pl.use_vc) vcp += ls + beta * D;
        inc = sqrt(l2);
      } else {
        const R pp = r * bb / q;
#pragma unroll
        for (int a = 0; a < D; a++) Sc[a] = exp(k / D - pp * sh[a] - beta);
        vcp = 0;
        inc = dg;
      }
      pdef += inc;
      b.vc[p] = (T)vcp;
      b.pdef[p] = (T)pdef;
      b.rhoF[p] = (T)dp_rho_f<T, R>(pl, pdef);
    }
    // F = U Sc V^T;  M = V Sc^-1 U^T F_e;  Fp = M Fp  (column-major)
    R M[D * D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        R f = 0, m = 0;
#pragma unroll
        for (int k = 0; k < D; k++) {
          f += U[i + D * k] * Sc[k] * V[j + D * k];
          R ut = 0;  // (U^T Fe)_kj
#pragma unroll
          for (int l = 0; l < D; l++) ut += U[l + D * k] * Fe[l + D * j];
          m += V[i + D * k] / Sc[k] * ut;
        }
        b.F[(size_t)(i + D * j) * n + p] = (T)f;
        M[i + D * j] = m;
      }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        R f = 0;
#pragma unroll
        for (int k = 0; k < D; k++) f += M[i + D * k] * Fp[k + D * j];
        b.Fp[(size_t)(i + D * j) * n + p] = (T)f;
      }
  }
}

// Fp = I, mu0 = mu, lambda0 = lambda; Drucker-Prager state as
// MPM_DRUCKER_PRAGER::Initialize_Particles (MPM_DRUCKER_PRAGER.cpp:40-58):
// plastic_def = 0, vc = 0, cohesion = sigma_Y0, rho_F at zero deformation.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void plastic_init_kernel(ParticleBufs<T, D> b, uint32_t n, PlasticDev<T> pl) {
  const T rho0 = (T)dp_rho_f<T>(pl, 0.0);
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n; p += gridDim.x * BLOCK) {
#pragma unroll
    for (int k = 0; k < D * D; k++) b.Fp[(size_t)k * n + p] = k % (D + 1) == 0 ? (T)1 : (T)0;
    b.mu0[p] = b.mu[p];
    b.la0[p] = b.la[p];
    b.pdef[p] = 0;
    b.vc[p] = 0;
    b.coh[p] = (T)pl.cohesion;
    b.rhoF[p] = rho0;
  }
}

// Host layout (per particle contiguous, user order) <-> device SoA (slot order).
template<typename T>
__global__ __launch_bounds__(BLOCK) void soa_from_aos_kernel(const T* aos, T* soa, uint32_t n, int comps) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < comps; c++) soa[(size_t)c * n + i] = aos[(size_t)i * comps + c];
}

template<typename T>
__global__ __launch_bounds__(BLOCK) void aos_from_soa_kernel(const T* soa, T* aos, const uint32_t* pid, uint32_t n,
                                                             int comps) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t q = pid[i];
  for (int c = 0; c < comps; c++) aos[(size_t)q * comps + c] = soa[(size_t)c * n + i];
}

__global__ __launch_bounds__(BLOCK) void bytes_from_slots_kernel(const uint8_t* src, uint8_t* dst, const uint32_t* pid,
                                                                 uint32_t n) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) dst[pid[i]] = src[i];
}

__global__ __launch_bounds__(BLOCK) void iota_kernel(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) p[i] = i;
}

// Dense export of the active tiles, MPM_EXAMPLE::mass / ::velocity layout.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void export_grid_kernel(Args<T, D> a, T* dmass, T* dvel, int momentum) {
  const uint64_t total = (uint64_t)a.counts[1] * 64u;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t t = a.gtiles[idx >> 6];
    const uint32_t l = (uint32_t)(idx & 63u);
    int org[D], nc[D];
    tile_origin<T, D>(t, a.g, org);
    uint32_t r = l;
    bool in = true;
#pragma unroll
    for (int c = D - 1; c >= 0; c--) {
      nc[c] = org[c] + (int)(r % Tile<D>::E);
      r /= Tile<D>::E;
      in = in && nc[c] < a.g.G[c];
    }
    if (!in) continue;
    const Node4<T> nd = a.grid[t * 64u + l];
    if (nd.m == (T)0) continue;
    uint64_t f = 0;
#pragma unroll
    for (int c = 0; c < D; c++) f = f * (uint64_t)a.g.G[c] + (uint64_t)nc[c];
    dmass[f] = nd.m;
    // between P2G and the grid update the nodes hold momentum: v = p / m
    dvel[f * D + 0] = momentum ? nd.x / nd.m : nd.x;
    dvel[f * D + 1] = momentum ? nd.y / nd.m : nd.y;
    if constexpr (D == 3) dvel[f * D + 2] = momentum ? nd.z / nd.m : nd.z;
  }
}

// ---------------------------------------------------- conservation totals --
// FP64 totals as -print_stats reports them (MPM_DRIVER.cpp:1159-1213), in the
// layout of mpmx_totals: mass [0], linear momentum [1..3], angular momentum
// [4..6] (2D: [4]), kinetic energy [7], potential energy [8] = elastic [9] +
// gravity [10], count [11].  Deterministic: a fixed number of blocks strides
// over the items, each block reduces in a fixed tree, and one wavefront sums
// the block partials in block order.
constexpr int NTOT = 12;
constexpr int TOTALS_BLOCKS = 1024;  // block partials of one totals pass

template<int D> __device__ __forceinline__ void cross_d(const double* x, const double* v, double* c) {
  if constexpr (D == 3) {
    c[0] = x[1] * v[2] - x[2] * v[1];
    c[1] = x[2] * v[0] - x[0] * v[2];
    c[2] = x[0] * v[1] - x[1] * v[0];
  } else {
    c[0] = x[0] * v[1] - x[1] * v[0];
    c[1] = c[2] = 0;
  }
}

// Block partial of NTOT accumulators into partial[blockIdx.x * NTOT + c].
__device__ __forceinline__ void totals_block_partial(const double* acc, double* partial) {
  __shared__ double s[BLOCK / 64][NTOT];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NTOT; c++) {
    double x = acc[c];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) s[wv][c] = x;
  }
  __syncthreads();
  if (threadIdx.x < NTOT) {
    double t = 0;
    for (int q = 0; q < BLOCK / 64; q++) t += s[q][threadIdx.x];
    partial[blockIdx.x * NTOT + threadIdx.x] = t;
  }
}

// Particle totals over the valid particles: Total_Particle_Linear_Momentum,
// Total_Particle_Angular_Momentum with the -m B:eps term, Total_Particle_
// Kinetic_Energy with the m/2 (B D^-1):B term (D^-1 = 4/dx^2 per axis)
// (MPM_EXAMPLE.cpp:209-225, :248-263, :302-315) and Potential_Energy of
// MPM_FINITE_ELEMENTS (vol Psi over the singular values of F, FP64 SVD) and
// MPM_GRAVITY (-m x.g) (MPM_EXAMPLE.cpp:123-133, MPM_FINITE_ELEMENTS.cpp:72-80,
// MPM_GRAVITY.cpp:47-55).  X, V, F, B in the particles' scene coordinates.
template<typename T, int D> struct TotalsIn {
  const T* X; const T* V; const T* F; const T* B; const T* mass; const T* vol; const T* mu; const T* la;
  const uint8_t* valid;
  const uint16_t* pscene;
  const Geom<T, D>* scenes;
  uint32_t n;
  int model;
  double grav[3];
};

template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void particle_totals_kernel(TotalsIn<T, D> in, double* partial) {
  double acc[NTOT];
#pragma unroll
  for (int c = 0; c < NTOT; c++) acc[c] = 0;
  const uint32_t n = in.n;
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n; p += gridDim.x * BLOCK) {
    if (!in.valid[p]) continue;
    const Geom<T, D> g = in.scenes[in.pscene ? in.pscene[p] : 0];
    const double m = (double)in.mass[p];
    double x[D], v[D], Fd[D * D], Bd[D * D];
#pragma unroll
    for (int c = 0; c < D; c++) {
      x[c] = (double)in.X[(size_t)c * n + p];
      v[c] = (double)in.V[(size_t)c * n + p];
    }
#pragma unroll
    for (int k = 0; k < D * D; k++) {
      Fd[k] = (double)in.F[(size_t)k * n + p];
      Bd[k] = (double)in.B[(size_t)k * n + p];
    }
    acc[0] += m;
    double vv = 0;
#pragma unroll
    for (int c = 0; c < D; c++) {
      acc[1 + c] += m * v[c];
      vv += v[c] * v[c];
    }
    double cr[3], e[3];
    cross_d<D>(x, v, cr);
    // MATRIX::Contract_Permutation_Tensor (MATRIX_3X3.h:392-393, MATRIX_2X2.h:339-340), column-major
    if constexpr (D == 3) {
      e[0] = Bd[7] - Bd[5];
      e[1] = Bd[2] - Bd[6];
      e[2] = Bd[3] - Bd[1];
    } else {
      e[0] = Bd[2] - Bd[1];
    }
    constexpr int NS = D == 3 ? 3 : 1;
#pragma unroll
    for (int c = 0; c < NS; c++) acc[4 + c] += m * cr[c] - m * e[c];
    double bb = 0;
#pragma unroll
    for (int j = 0; j < D; j++) {
      const double dinv = 4.0 * (double)g.inv_dx[j] * (double)g.inv_dx[j];
#pragma unroll
      for (int i = 0; i < D; i++) bb += Bd[i + D * j] * Bd[i + D * j] * dinv;
    }
    acc[7] += 0.5 * m * vv + 0.5 * m * bb;
    double U[D * D], sg[D], W[D * D];
    svd<double, D>(Fd, U, sg, W);
    acc[9] += (double)in.vol[p] * energy_density<double, D>(in.model, (double)in.mu[p], (double)in.la[p], sg);
    double xg = 0;
#pragma unroll
    for (int c = 0; c < D; c++) xg += x[c] * in.grav[c];
    acc[10] -= m * xg;
    acc[11] += 1;
  }
  totals_block_partial(acc, partial);
}

// Grid totals over the nodes of the grid-tile list with nonzero mass, node
// locations the MAC cell centres lo + (i + 1/2) dx (MPM_DRIVER.cpp:115-126):
// Total_Grid_Linear_Momentum, Total_Grid_Angular_Momentum (no midpoint) and
// Total_Grid_Kinetic_Energy (MPM_EXAMPLE.cpp:226-241, :268-284, :289-297).
// MOMENTUM: the nodes hold P2G's (m, m v) (the velocity after Particle_To_Grid
// is v = p / m); otherwise (m, v) after the grid update.
template<typename T, int D, bool MOMENTUM>
__global__ __launch_bounds__(BLOCK) void grid_totals_kernel(Args<T, D> a, const Node4<T>* grid, double* partial) {
  double acc[NTOT];
#pragma unroll
  for (int c = 0; c < NTOT; c++) acc[c] = 0;
  const uint64_t total = (uint64_t)a.counts[1] * 64u;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t t = a.gtiles[idx >> 6];
    const uint32_t l = (uint32_t)(idx & 63u);
    const Node4<T> nd = grid[(size_t)t * 64u + l];
    if (nd.m == (T)0) continue;
    const Geom<T, D> g = tile_geom(a, t);
    int org[D];
    tile_origin<T, D>(t, g, org);
    uint32_t r = l;
    double x[D];
#pragma unroll
    for (int c = D - 1; c >= 0; c--) {
      const int nc = org[c] + (int)(r % Tile<D>::E);
      r /= Tile<D>::E;
      x[c] = (double)(g.lo[c] + ((T)(nc - 3) + (T)0.5) * g.dx[c]);  // GRID::Center in T
    }
    const double m = (double)nd.m;
    const T comp[3] = {nd.x, nd.y, nd.z};
    double p[D], vv = 0;
#pragma unroll
    for (int c = 0; c < D; c++) {
      p[c] = MOMENTUM ? (double)comp[c] : m * (double)comp[c];
      vv += MOMENTUM ? p[c] * p[c] / m : m * (double)comp[c] * (double)comp[c];
    }
    double cr[3];
    cross_d<D>(x, p, cr);
    acc[0] += m;
#pragma unroll
    for (int c = 0; c < D; c++) acc[1 + c] += p[c];
#pragma unroll
    for (int c = 0; c < (D == 3 ? 3 : 1); c++) acc[4 + c] += cr[c];
    acc[7] += 0.5 * vv;
    acc[11] += 1;
  }
  totals_block_partial(acc, partial);
}

// Sums the block partials in block order (one thread per component) into
// out[0..NTOT), potential = elastic + gravity.
__global__ void totals_finish_kernel(const double* partial, int nblocks, double* out) {
  __shared__ double s[NTOT];
  const int c = threadIdx.x;
  if (c < NTOT) {
    double t = 0;
    for (int b = 0; b < nblocks; b++) t += partial[b * NTOT + c];
    s[c] = t;
  }
  __syncthreads();
  if (c < NTOT) out[c] = c == 8 ? s[9] + s[10] : s[c];
}

// ------------------------------------------------------- adaptive dt ----
// Bit pattern of a non-negative float or double: its unsigned ordering is the
// ordering of the values, so atomicMax on the pattern is a max of the values.
template<typename T> struct OrderedBits;
template<> struct OrderedBits<float> {
  using U = uint32_t;
  __device__ static U of(float x) { return __float_as_uint(x); }
};
template<> struct OrderedBits<double> {
  using U = unsigned long long;
  __device__ static U of(double x) { return (U)__double_as_longlong(x); }
};

// xi = 6 sqrt(d) min_a (1/dx_a) in T, as MPM_DRIVER.cpp:1130 forms it.
template<typename T, int D> __device__ __forceinline__ T scene_xi(const Geom<T, D>& g) {
#pragma clang fp contract(off)
  T odx = g.inv_dx[0];
#pragma unroll
  for (int a = 1; a < D; a++) odx = g.inv_dx[a] < odx ? g.inv_dx[a] : odx;
  return (T)6 * sqrt((T)D) * odx;
}

// |v| + |B|_F xi of particle p, in the order and rounding of
// MPM_DRIVER::Grid_V_Upper_Bound (MPM_DRIVER.cpp:1131-1136): squares summed
// left to right, no fused multiply-add, correctly rounded square roots.
template<typename T, int D>
__device__ __noinline__ T particle_speed_bound(const T* V, const T* B, uint32_t n, uint32_t p, T xi) {
#pragma clang fp contract(off)
  T v2 = V[p] * V[p];
#pragma unroll
  for (int c = 1; c < D; c++) v2 = v2 + V[(size_t)c * n + p] * V[(size_t)c * n + p];
  T b2 = B[p] * B[p];
#pragma unroll
  for (int k = 1; k < D * D; k++) b2 = b2 + B[(size_t)k * n + p] * B[(size_t)k * n + p];
  return sqrt(v2) + sqrt(b2) * xi;
}

// Per-scene max over valid particles of |v| + 6 sqrt(d) / max dx |B|_F
// (Grid_V_Upper_Bound with use_affine_cfl and stored B).  out[s] holds the
// bit pattern of scene s's bound and starts at zero, the reference's initial
// result; NaN particles leave it unchanged as std::max does there.
template<typename T, int D, bool BATCH>
__global__ __launch_bounds__(BLOCK) void speed_bound_kernel(const T* V, const T* B, const uint8_t* valid,
                                                            const uint16_t* pscene, const Geom<T, D>* scenes,
                                                            uint32_t n, typename OrderedBits<T>::U* out) {
  const int lane = threadIdx.x & 63;
  T run = 0;
  T xi0 = 0;
  if (!BATCH) xi0 = scene_xi(scenes[0]);
  for (uint32_t p0 = blockIdx.x * BLOCK; p0 < n; p0 += gridDim.x * BLOCK) {
    const uint32_t p = p0 + threadIdx.x;
    T r = -1;
    int s = -1;
    if (p < n && valid[p]) {
      s = BATCH ? (int)pscene[p] : 0;
      const T xi = BATCH ? scene_xi(scenes[s]) : xi0;
      r = particle_speed_bound<T, D>(V, B, n, p, xi);
      if (!(r == r)) r = -1;
    }
    if (!BATCH) {
      run = r > run ? r : run;
      continue;
    }
    // batches: one atomic per wave when the wave's valid particles share a scene
    int s0 = s;
    for (int o = 32; o > 0; o >>= 1) s0 = max(s0, __shfl_xor(s0, o, 64));
    const bool uniform = __all(s < 0 || s == s0);
    if (uniform) {
      T m = r;
      for (int o = 32; o > 0; o >>= 1) {
        const T q = __shfl_xor(m, o, 64);
        m = q > m ? q : m;
      }
      if (lane == 0 && s0 >= 0 && m >= 0) atomicMax(out + s0, OrderedBits<T>::of(m));
    } else if (s >= 0 && r >= 0) {
      atomicMax(out + s, OrderedBits<T>::of(r));
    }
  }
  if (!BATCH) {
    for (int o = 32; o > 0; o >>= 1) {
      const T q = __shfl_xor(run, o, 64);
      run = q > run ? q : run;
    }
    __shared__ T s_max[BLOCK / 64];
    if (lane == 0) s_max[threadIdx.x >> 6] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
      T m = s_max[0];
      for (int q = 1; q < BLOCK / 64; q++) m = s_max[q] > m ? s_max[q] : m;
      if (m > 0) atomicMax(out, OrderedBits<T>::of(m));
    }
  }
}

// Stand-alone kernels for component parity tests.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void stress_kernel(uint32_t n, const T* F, const T* vol, const T* mu, const T* la,
                                                       T* tau, int model, T thr, T root) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  T Fm[D * D], t[D * D];
#pragma unroll
  for (int k = 0; k < D * D; k++) Fm[k] = F[(size_t)i * D * D + k];
  kirchhoff_any<T, D>(model, mu[i], la[i], thr, root, vol[i], Fm, t);
#pragma unroll
  for (int k = 0; k < D * D; k++) tau[(size_t)i * D * D + k] = t[k];
}

template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void svd_kernel(uint32_t n, const T* F, T* U, T* S, T* V) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  T Fm[D * D], u[D * D], s[D], v[D * D];
#pragma unroll
  for (int k = 0; k < D * D; k++) Fm[k] = F[(size_t)i * D * D + k];
  svd<T, D>(Fm, u, s, v);
#pragma unroll
  for (int k = 0; k < D * D; k++) { U[(size_t)i * D * D + k] = u[k]; V[(size_t)i * D * D + k] = v[k]; }
#pragma unroll
  for (int k = 0; k < D; k++) S[(size_t)i * D + k] = s[k];
}

// ------------------------------------------------------------------ host ---
inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------- Reduce_Dt ----
// MPM_DRIVER::Reduce_Dt (MPM_DRIVER.cpp:519-545) between the grid update and
// G2P: the strong CFL (Grid_To_Particle_Limit_Dt, :572-620, without the
// midpoint rule) and the sound-speed CFL (Limit_Dt_Sound_Speed, :624-637,
// Compute_Max_Sound_Speed, :893-918).  velocity_save, the P2G velocity before
// forces, comes from a second P2G of the step with dt = 0 (no stress impulse)
// into its own grid, normalised by the node mass.

// Enforce_Limit_Max (MPM_DRIVER.cpp:549-559): lower s so that
// |t (a + t b)| <= bound for every 0 <= t <= s.
template<typename T> __device__ __forceinline__ void enforce_limit_max(T& s, T bound, T a, T b) {
  if (a < 0) {
    a = -a;
    b = -b;
  }
  if (a * a + 4 * b * bound >= 0) {
    if (s * (a + s * b) > bound) s = (2 * bound) / (a + sqrt(a * a + 4 * b * bound));
  } else {
    if (s * (a + s * b) < -bound) s = -(a + sqrt(a * a - 4 * b * bound)) / (2 * b);
  }
}

// v <- m v / m on the nodes of the listed tiles that hold mass.
template<typename T>
__global__ __launch_bounds__(BLOCK) void normalize_nodes_kernel(Node4<T>* g, const uint32_t* tiles,
                                                                const uint32_t* count) {
  const uint64_t total = (uint64_t)count[0] * 64u;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t slot = tiles[idx >> 6] * 64u + (uint32_t)(idx & 63u);
    Node4<T> nd = g[slot];
    if (nd.m == (T)0) continue;
    const T im = (T)1 / nd.m;
    nd.x *= im;
    nd.y *= im;
    nd.z *= im;
    g[slot] = nd;
  }
}

// Reduce_Dt's rescale on the valid nodes: v = (v - vs) s + vs, and the same
// for the friction velocity.
template<typename T>
__global__ __launch_bounds__(BLOCK) void rescale_nodes_kernel(Node4<T>* grid, Node4<T>* gridf, const Node4<T>* vsave,
                                                              const uint32_t* tiles, const uint32_t* count, T sc) {
  const uint64_t total = (uint64_t)count[0] * 64u;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t slot = tiles[idx >> 6] * 64u + (uint32_t)(idx & 63u);
    const Node4<T> vs = vsave[slot];
    if (vs.m == (T)0) continue;
    Node4<T> v = grid[slot];
    v.x = (v.x - vs.x) * sc + vs.x;
    v.y = (v.y - vs.y) * sc + vs.y;
    v.z = (v.z - vs.z) * sc + vs.z;
    grid[slot] = v;
    if (gridf) {
      Node4<T> f = gridf[slot];
      f.x = (f.x - vs.x) * sc + vs.x;
      f.y = (f.y - vs.y) * sc + vs.y;
      f.z = (f.z - vs.z) * sc + vs.z;
      gridf[slot] = f;
    }
  }
}

template<typename T> __global__ void dt_limit_init_kernel(typename OrderedBits<T>::U* out) {
  out[0] = OrderedBits<T>::of((T)0);
  out[1] = OrderedBits<T>::of((T)100);
}

// Over the valid particles: out[0] = max sound speed (from the singular values
// of F at the start of the step, density m / V0), out[1] = min over particles
// of the strong-CFL factor s (starting at 100).  Both hold bit patterns of
// non-negative values, so integer max/min order them.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void dt_limit_kernel(Args<T, D> a, const Node4<T>* vsave, int strong, int sound,
                                                         T cfl, T cfl_F, typename OrderedBits<T>::U* out) {
  T cmax = 0, smin = 100;
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < a.n; p += gridDim.x * BLOCK) {
    if (!a.valid[p]) continue;
    if (sound) {
      T Fm[D * D], U[D * D], S[D], V[D * D];
#pragma unroll
      for (int k = 0; k < D * D; k++) Fm[k] = a.F[(size_t)k * a.n + p];
      svd<T, D>(Fm, U, S, V);
      const T c = sound_speed<T, D>(a.model, a.mu[p], a.la[p], a.thr, a.root, S, a.mass[p] / a.vol[p]);
      if (c > cmax) cmax = c;
    }
    if (strong) {
      const Geom<T, D> g = particle_geom(a, p);
      T x[D];
#pragma unroll
      for (int c = 0; c < D; c++) x[c] = a.X[(size_t)c * a.n + p];
      Wts<T, D> w;
      weights<T, D>(g, x, w);
      bool in = true;
#pragma unroll
      for (int c = 0; c < D; c++) in = in && w.base[c] >= 0 && w.base[c] + 2 < g.G[c];
      if (!in) continue;  // flagged by the P2G side path
      T vp[D], vd[D], G[D * D], Gd[D * D];
#pragma unroll
      for (int c = 0; c < D; c++) vp[c] = vd[c] = 0;
#pragma unroll
      for (int k = 0; k < D * D; k++) G[k] = Gd[k] = 0;
      constexpr int K = D == 3 ? 27 : 9;
#pragma unroll 1
      for (int k = 0; k < K; k++) {
        int o[D], r = k;
#pragma unroll
        for (int c = D - 1; c >= 0; c--) {
          o[c] = r % 3;
          r /= 3;
        }
        T wt = w.w[o[0]][0], gr[D];
#pragma unroll
        for (int c = 1; c < D; c++) wt *= w.w[o[c]][c];
#pragma unroll
        for (int i = 0; i < D; i++) {
          T q = 1;
#pragma unroll
          for (int j = 0; j < D; j++) q *= i == j ? w.dw[o[j]][j] : w.w[o[j]][j];
          gr[i] = q;
        }
        int nc[D];
#pragma unroll
        for (int c = 0; c < D; c++) nc[c] = w.base[c] + o[c];
        const uint32_t slot = node_slot<T, D>(nc, g);
        const Node4<T> nn = a.grid[slot], no = vsave[slot];
        const T vn[3] = {nn.x, nn.y, nn.z}, vo[3] = {no.x, no.y, no.z};
#pragma unroll
        for (int i = 0; i < D; i++) {
          const T dv = vn[i] - vo[i];
          vp[i] += wt * vo[i];
          vd[i] += wt * dv;
#pragma unroll
          for (int j = 0; j < D; j++) {
            G[i + D * j] += vo[i] * gr[j];
            Gd[i + D * j] += dv * gr[j];
          }
        }
      }
      T sp = 100;
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) enforce_limit_max<T>(sp, cfl_F, a.dt * G[i + D * j], a.dt * Gd[i + D * j]);
#pragma unroll
      for (int i = 0; i < D; i++) enforce_limit_max<T>(sp, cfl, a.dt * vp[i], a.dt * vd[i]);
      smin = sp < smin ? sp : smin;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const T c = __shfl_xor(cmax, o, 64), q = __shfl_xor(smin, o, 64);
    cmax = c > cmax ? c : cmax;
    smin = q < smin ? q : smin;
  }
  __shared__ T s_c[BLOCK / 64], s_s[BLOCK / 64];
  if ((threadIdx.x & 63) == 0) {
    s_c[threadIdx.x >> 6] = cmax;
    s_s[threadIdx.x >> 6] = smin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < BLOCK / 64; q++) {
      cmax = s_c[q] > cmax ? s_c[q] : cmax;
      smin = s_s[q] < smin ? s_s[q] : smin;
    }
    if (cmax > 0) atomicMax(out, OrderedBits<T>::of(cmax));
    if (smin < 100) atomicMin(out + 1, OrderedBits<T>::of(smin));
  }
}

// ------------------------------------------------------- reflection BC ----
// Reflection boundary conditions (MPM_DRIVER.cpp:310-446, example.reflection_bc).
// The reference visits the ghost nodes side by side (-x, +x, -y, +y, -z, +z),
// each side over its whole slab with the corners (RANGE_ITERATOR with ghost |
// duplicate_corners), pairing ghost node o with its mirror image i across
// the side (index 2 r - o - 1 along the axis, r = 0 or N).  The pairs of one
// axis are disjoint (N >= 6), so one launch handles both sides of an axis and
// the axes run in order.  Side types: 0 free, 1 slip, 2 noslip, 3 separate.
enum { RBC_FREE = 0, RBC_SLIP = 1, RBC_NOSLIP = 2, RBC_SEP = 3 };
struct ReflectAxis {
  int axis;
  int type[2];  // low, high side
};

// Pair q of axis r.axis: ghost node o and image i slots, the side (0 low, 1 high).
template<typename T, int D>
__device__ __forceinline__ bool reflect_pair(const Geom<T, D>& g, const ReflectAxis& r, uint32_t q, uint32_t* io,
                                             uint32_t* oo, int* hi) {
  const int k = r.axis;
  // runtime axis: selected per component (a dynamic index would put the
  // arrays in scratch memory)
  int N = 0, other = 1;
#pragma unroll
  for (int c = 0; c < D; c++) {
    if (c == k) N = g.G[c] - 6;
    else other *= g.G[c];
  }
  if (q >= (uint32_t)(6 * other)) return false;
  uint32_t rest = q % (uint32_t)other;
  const int lay = (int)(q / (uint32_t)other);  // 0..5: side * 3 + layer
  *hi = lay / 3;
  const int o = *hi ? N + (lay % 3) : (lay % 3) - 3;  // node index (ghost excluded)
  const int i = 2 * (*hi ? N : 0) - o - 1;
  int no[D], ni[D];
#pragma unroll
  for (int c = D - 1; c >= 0; c--) {
    if (c == k) {
      no[c] = o + 3;
      ni[c] = i + 3;
    } else {
      no[c] = ni[c] = (int)(rest % (uint32_t)g.G[c]);
      rest /= (uint32_t)g.G[c];
    }
  }
  *oo = node_slot<T, D>(no, g);
  *io = node_slot<T, D>(ni, g);
  return true;
}
template<typename T> __device__ __forceinline__ void reflect_vectors(int type, int axis, T* vi, T* vo) {
  T id[3], od[3], in[3], on[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    id[c] = vi[c] - vo[c];
    od[c] = -id[c];
    in[c] = vi[c] + vo[c];
    on[c] = in[c];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {  // selects, not a dynamic index (which would live in scratch)
    const bool normal = type == RBC_SLIP && c == axis;
    vi[c] = type == RBC_NOSLIP || normal ? id[c] : in[c];
    vo[c] = type == RBC_NOSLIP || normal ? od[c] : on[c];
  }
}
// Reflect_Boundary_Mass_Momentum (:310-342) on the P2G sums (m, m v + dt f):
// the force reflection of Reflect_Boundary_Force (:422-446) is the same map,
// so it is applied once to the sum (gravity included beforehand).
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void reflect_mass_momentum_kernel(Node4<T>* grid, const Geom<T, D>* scene,
                                                                      ReflectAxis r) {
  const Geom<T, D> g = scene[0];
  uint32_t i, o;
  int hi;
  if (!reflect_pair<T, D>(g, r, blockIdx.x * BLOCK + threadIdx.x, &i, &o, &hi)) return;
  Node4<T> a = grid[i], b = grid[o];
  T vi[3] = {a.x, a.y, a.z}, vo[3] = {b.x, b.y, b.z};
  reflect_vectors<T>(r.type[hi], r.axis, vi, vo);
  const T mt = a.m + b.m;
  grid[i] = Node4<T>{mt, vi[0], vi[1], vi[2]};
  grid[o] = Node4<T>{mt, vo[0], vo[1], vo[2]};
}
// Reflect_Boundary_Velocity (:395-420): a ghost node with mass takes its
// image's velocity, normal part negated for slip and all of it for noslip.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void reflect_velocity_kernel(Node4<T>* grid, Node4<T>* gridf,
                                                                 const Geom<T, D>* scene, ReflectAxis r) {
  const Geom<T, D> g = scene[0];
  uint32_t i, o;
  int hi;
  if (!reflect_pair<T, D>(g, r, blockIdx.x * BLOCK + threadIdx.x, &i, &o, &hi)) return;
  const int type = r.type[hi];
  for (int f = 0; f < (gridf ? 2 : 1); f++) {
    Node4<T>* G = f ? gridf : grid;
    const Node4<T> a = G[i];
    Node4<T> b = G[o];
    if (grid[o].m == (T)0) return;
    T vi[3] = {a.x, a.y, a.z}, v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] = type == RBC_NOSLIP || (type == RBC_SLIP && c == r.axis) ? -vi[c] : vi[c];
    b.x = v[0];
    b.y = v[1];
    b.z = D == 3 ? v[2] : b.z;
    G[o] = b;
  }
}
// Reflect_Helper (:343-369) without a boundary velocity: mass-weighted mixing
// of the pair's velocities; separate sides keep a separating pair and apply
// Coulomb friction mu to an approaching one.
template<typename T, int D>
__device__ __forceinline__ void reflect_helper(T m_i, T m_o, T* v_i, T* v_o, int type, T mu, int a, int s) {
  const T m_t = m_i + m_o, a_i = m_i / m_t, a_o = 1 - a_i;
  T v_in[D], v_on[D], v_id[D], v_od[D];
#pragma unroll
  for (int c = 0; c < D; c++) {
    v_in[c] = a_i * v_i[c] + a_o * v_o[c];
    v_on[c] = v_in[c];
    v_id[c] = a_i * v_i[c] + a_o * ((T)0 - v_o[c]);
    v_od[c] = a_o * v_o[c] + a_i * ((T)0 - v_i[c]);
  }
  if (type == RBC_FREE) {
#pragma unroll
    for (int c = 0; c < D; c++) { v_i[c] = v_in[c]; v_o[c] = v_on[c]; }
    return;
  }
  if (type == RBC_NOSLIP) {
#pragma unroll
    for (int c = 0; c < D; c++) { v_i[c] = v_id[c]; v_o[c] = v_od[c]; }
    return;
  }
  if (type == RBC_SLIP) {
#pragma unroll
    for (int c = 0; c < D; c++) {
      v_i[c] = c == a ? v_id[c] : v_in[c];
      v_o[c] = c == a ? v_od[c] : v_on[c];
    }
    return;
  }
  T via = 0, voa = 0, vida = 0, voda = 0;  // the components along axis a
#pragma unroll
  for (int c = 0; c < D; c++)
    if (c == a) {
      via = v_i[c];
      voa = v_o[c];
      vida = v_id[c];
      voda = v_od[c];
    }
  const T d_v = 2 * a_o * a_i * ((T)0 - via - voa);
  const T d_v_n = s ? -d_v : d_v;
  if (d_v_n < 0) return;
#pragma unroll
  for (int c = 0; c < D; c++) v_in[c] = c == a ? (T)0 : v_in[c];
  T m2 = v_in[0] * v_in[0];
#pragma unroll
  for (int c = 1; c < D; c++) m2 += v_in[c] * v_in[c];
  const T mag = sqrt(m2);  // VECTOR::Normalize: x axis when zero
  T t[D];
#pragma unroll
  for (int c = 0; c < D; c++) t[c] = mag ? v_in[c] * ((T)1 / mag) : (c == 0 ? (T)1 : (T)0);
  const T rr = mag - mu * d_v_n;
#pragma unroll
  for (int c = 0; c < D; c++) {
    const T ti = c == a ? vida : (T)0, to = c == a ? voda : (T)0;
    v_i[c] = rr < 0 ? ti : ti + rr * t[c];
    v_o[c] = rr < 0 ? to : to + rr * t[c];
  }
}
// Reflect_Boundary_Friction (:371-393) on the pairs where both nodes hold mass.
template<typename T, int D>
__global__ __launch_bounds__(BLOCK) void reflect_friction_kernel(Node4<T>* grid, const Geom<T, D>* scene,
                                                                 ReflectAxis r, T mu) {
  const Geom<T, D> g = scene[0];
  uint32_t i, o;
  int hi;
  if (!reflect_pair<T, D>(g, r, blockIdx.x * BLOCK + threadIdx.x, &i, &o, &hi)) return;
  Node4<T> a = grid[i], b = grid[o];
  if (a.m == (T)0 || b.m == (T)0) return;
  T vi[D], vo[D];
  vi[0] = a.x; vi[1] = a.y;
  vo[0] = b.x; vo[1] = b.y;
  if constexpr (D == 3) { vi[2] = a.z; vo[2] = b.z; }
  reflect_helper<T, D>(a.m, b.m, vi, vo, r.type[hi], mu, r.axis, hi);
  a.x = vi[0]; a.y = vi[1];
  b.x = vo[0]; b.y = vo[1];
  if constexpr (D == 3) { a.z = vi[2]; b.z = vo[2]; }
  grid[i] = a;
  grid[o] = b;
}
// m v += dt m g on the listed tiles (the gravity impulse of MPM_GRAVITY::
// Add_Forces taken into the P2G sums before they are reflected).
template<typename T>
__global__ __launch_bounds__(BLOCK) void add_gravity_kernel(Node4<T>* g, const uint32_t* tiles, const uint32_t* count,
                                                            T gx, T gy, T gz) {
  const uint64_t total = (uint64_t)count[0] * 64u;
  for (uint64_t idx = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; idx < total; idx += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t slot = tiles[idx >> 6] * 64u + (uint32_t)(idx & 63u);
    Node4<T> nd = g[slot];
    if (nd.m == (T)0) continue;
    nd.x += gx * nd.m;
    nd.y += gy * nd.m;
    nd.z += gz * nd.m;
    g[slot] = nd;
  }
}

// Core/Math_Tools/Robust_Arithmetic.h:22-26 (FLT_MAX in T for both precisions).
template<typename T> T robust_divide(T a, T b) {
  const T fm = (T)std::numeric_limits<float>::max();
  const T aa = std::fabs(a), ab = std::fabs(b);
  if (ab == fm) return T();
  if (ab > 1 || aa < fm * ab) return a / b;
  return (a > 0) == (b >= 0) ? fm : -fm;
}

template<typename X> struct DevBuf {
  X* p = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), count(o.count) { o.p = nullptr; o.count = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      count = o.count;
      o.p = nullptr;
      o.count = 0;
    }
    return *this;
  }
  void alloc(size_t n) {
    if (n <= count && p) return;
    release();
    if (n == 0) return;
    HIPCHECK(hipMalloc((void**)&p, n * sizeof(X)));
    count = n;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    count = 0;
  }
  ~DevBuf() { release(); }
};

struct PhaseTimer {
  struct Pending {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Pending> pending;
  static constexpr int NPH = 11;
  double total[NPH] = {};
  int64_t count[NPH] = {};
  bool on = false;         // every phase
  bool kernel_on = false;  // the fused tile kernel (begin_kernel), also when `on` is off
  hipEvent_t get() {
    if (pool.empty()) {
      hipEvent_t e;
      HIPCHECK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t begin(hipStream_t s) {
    if (!on) return nullptr;
    hipEvent_t e = get();
    HIPCHECK(hipEventRecord(e, s));
    return e;
  }
  hipEvent_t begin_kernel(hipStream_t s) {
    if (!kernel_on) return nullptr;
    hipEvent_t e = get();
    HIPCHECK(hipEventRecord(e, s));
    return e;
  }
  void end(int phase, hipEvent_t a, hipStream_t s) {
    if (!a) return;
    hipEvent_t b = get();
    HIPCHECK(hipEventRecord(b, s));
    pending.push_back({phase, a, b});
  }
  void resolve() {
    for (auto& q : pending) {
      HIPCHECK(hipEventSynchronize(q.b));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, q.a, q.b));
      total[q.phase] += ms;
      count[q.phase] += 1;
      pool.push_back(q.a);
      pool.push_back(q.b);
    }
    pending.clear();
  }
  ~PhaseTimer() {
    for (auto& q : pending) {
      (void)hipEventDestroy(q.a);
      (void)hipEventDestroy(q.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
enum { PH_BIN = 0, PH_P2G, PH_GRID, PH_G2P, PH_REORDER, PH_STEP, PH_REBIN, PH_ADJOINT, PH_DT, PH_G2P2G, PH_FUSED };
static const char* kPhaseNames[PhaseTimer::NPH] = {"bin",  "p2g",   "grid",    "g2p", "reorder",     "step",
                                                   "rebin", "adjoint", "dt",  "g2p2g", "g2p2g_kernel"};

template<typename T, int D>
class Engine final : public EngineBase {
 public:
  // One scene, or a batch of independent scenes (each with its own grid) that
  // share every launch: their tiles are numbered consecutively.
  Engine(const mpmx_grid_desc* descs, int nscenes) {
    if (nscenes < 1 || nscenes > 65535) throw Error{MPMX_EINVAL, "scene count must be in [1, 65535]"};
    descs_.assign(descs, descs + nscenes);
    uint64_t total = 0;
    for (int s = 0; s < nscenes; s++) {
      const mpmx_grid_desc& d = descs[s];
      if (d.dim != D || d.precision != (int)sizeof(T))
        throw Error{MPMX_EINVAL, "all scenes of a batch need the same dim and precision"};
      Geom<T, D> g{};
      for (int a = 0; a < D; a++) {
        if (d.cells[a] < 1 || d.cells[a] > (1 << 14)) throw Error{MPMX_EINVAL, "cells out of range"};
        if (!(d.hi[a] > d.lo[a])) throw Error{MPMX_EINVAL, "empty domain"};
        g.lo[a] = (T)d.lo[a];
        g.hi[a] = (T)d.hi[a];
        g.dx[a] = (g.hi[a] - g.lo[a]) / (T)d.cells[a];  // GRID.cpp:23-24
        g.inv_dx[a] = (T)1 / g.dx[a];
        g.G[a] = d.cells[a] + 6;
        g.GT[a] = (g.G[a] + Tile<D>::E - 1) / Tile<D>::E + 1;
        // Reflect_Or_Invalidate_Particle bounds: eps = 4 machine-eps * (B - A)
        const T eps = std::numeric_limits<T>::epsilon() * (T)4 * (g.hi[a] - g.lo[a]);
        g.vlo[a] = g.lo[a] + eps;
        g.vhi[a] = g.hi[a] - eps;
      }
      uint64_t nt = 1;
      for (int a = 0; a < D; a++) nt *= (uint64_t)g.GT[a];
      g.ntiles = (int)std::min<uint64_t>(nt, 0x7fffffff);
      g.tile0 = (uint32_t)total;
      total += nt;
      if (total >= 0xffffffffull / 64) throw Error{MPMX_EINVAL, "grid too large"};
      scenes_.push_back(g);
    }
    g_ = scenes_[0];
    ntotal_ = (uint32_t)total;
    const uint64_t nt = total;
    key_bits_ = 1;
    while ((1ull << key_bits_) < nt + 1) key_bits_++;
    for (int a = 0; a < D; a++) grav_[a] = 0;
    HIPCHECK(hipGetDevice(&device_));
    HIPCHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    tile_begin_.alloc(nt);
    tile_end_.alloc(nt);
    tile_stamp_.alloc(nt);
    mstamp_.alloc(nt);
    gblock_.alloc(cdiv((uint32_t)nt, BLOCK));
    ptiles_.alloc(nt);
    gtiles_.alloc(nt);
    gtiles2_.alloc(nt);
    counts_.alloc(16);
    lastcnt_.alloc(8);
    err_.alloc(4);
    grid_.alloc(nt * 64);
    HIPCHECK(hipMemsetAsync(tile_stamp_.p, 0, nt * sizeof(uint32_t), stream_));
    HIPCHECK(hipMemsetAsync(mstamp_.p, 0, nt * sizeof(uint32_t), stream_));
    HIPCHECK(hipHostMalloc((void**)&rb_host_, 4 * sizeof(uint32_t), hipHostMallocDefault));
    for (int i = 0; i < 4; i++) {
      HIPCHECK(hipEventCreateWithFlags(&rb_ev_[i], hipEventDisableTiming));
      rb_meta_[i] = Readback{-1, 0};
      rb_host_[i] = 0;
    }
    HIPCHECK(hipMemsetAsync(grid_.p, 0, nt * 64 * sizeof(Node4<T>), stream_));
    grid2_.alloc(nt * 64);
    HIPCHECK(hipMemsetAsync(grid2_.p, 0, nt * 64 * sizeof(Node4<T>), stream_));
    HIPCHECK(hipMemsetAsync(lastcnt_.p, 0, 8 * sizeof(uint32_t), stream_));
    gcur_ = grid_.p;
    gother_ = grid2_.p;
    gt_cur_ = gtiles_.p;
    gt_other_ = gtiles2_.p;
    HIPCHECK(hipMemsetAsync(counts_.p, 0, 16 * sizeof(uint32_t), stream_));
    HIPCHECK(hipMemsetAsync(err_.p, 0, 4 * sizeof(uint32_t), stream_));
    stats_blocks_ = (int)std::min<uint64_t>(std::max<uint64_t>(cdiv(nt * 64, BLOCK), 1), 2048);
    stats_partial_.alloc((size_t)TOTALS_BLOCKS * NTOT);
    stats_out_.alloc(2 * NTOT);
    HIPCHECK(hipMemsetAsync(stats_out_.p, 0, 2 * NTOT * sizeof(double), stream_));
    scenes_dev_.alloc(scenes_.size());
    HIPCHECK(hipMemcpyAsync(scenes_dev_.p, scenes_.data(), scenes_.size() * sizeof(Geom<T, D>), hipMemcpyHostToDevice,
                            stream_));
    if (nscenes > 1) {
      std::vector<uint16_t> ts(nt);
      for (size_t s = 0; s < scenes_.size(); s++)
        std::fill(ts.begin() + scenes_[s].tile0, ts.begin() + scenes_[s].tile0 + scenes_[s].ntiles, (uint16_t)s);
      tscene_.alloc(nt);
      HIPCHECK(hipMemcpyAsync(tscene_.p, ts.data(), nt * sizeof(uint16_t), hipMemcpyHostToDevice, stream_));
    }
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  ~Engine() override {
    if (stream_) (void)hipStreamSynchronize(stream_);
    release_graph();
    if (stream_) (void)hipStreamDestroy(stream_);
    for (int i = 0; i < 4; i++)
      if (rb_ev_[i]) (void)hipEventDestroy(rb_ev_[i]);
    for (hipEvent_t e : {seg_done_[0], seg_done_[1], seg_free_[0], seg_free_[1], rep_go_})
      if (e) (void)hipEventDestroy(e);
    if (rb_host_) (void)hipHostFree(rb_host_);
    if (vb_host_) (void)hipHostFree(vb_host_);
    if (lim_host_) (void)hipHostFree(lim_host_);
  }

  int dim() const override { return D; }
  int precision() const override { return (int)sizeof(T); }
  int64_t particle_count() const override { return n_; }
  double time() const override { return time_; }

  void set_particles(int64_t n, const void* X, const void* V, const void* F, const void* B, const void* mass,
                     const void* volume, const void* mu, const void* lambda, const uint8_t* valid) override {
    if (scenes_.size() > 1) throw Error{MPMX_ESTATE, "a batch context takes mpmx_set_batch_particles"};
    upload_particles(n, X, V, F, B, mass, volume, mu, lambda, valid, nullptr);
  }
  void set_batch_particles(int nscenes, const int64_t* counts, const void* X, const void* V, const void* F,
                           const void* B, const void* mass, const void* volume, const void* mu, const void* lambda,
                           const uint8_t* valid) override {
    if (nscenes != (int)scenes_.size()) throw Error{MPMX_EINVAL, "scene count differs from the context's"};
    if (!counts) throw Error{MPMX_EINVAL, "null counts"};
    std::vector<uint16_t> ids;
    for (int s = 0; s < nscenes; s++) {
      if (counts[s] < 0) throw Error{MPMX_EINVAL, "negative particle count"};
      if ((int64_t)ids.size() + counts[s] > (int64_t)0x7fffffff) throw Error{MPMX_EINVAL, "too many particles"};
      ids.insert(ids.end(), (size_t)counts[s], (uint16_t)s);
    }
    upload_particles((int64_t)ids.size(), X, V, F, B, mass, volume, mu, lambda, valid,
                     scenes_.size() > 1 ? ids.data() : nullptr);
  }
  int scene_count() const override { return (int)scenes_.size(); }

  // The recorded steps are replayed from the scene as it is when the reverse
  // pass runs, so nothing it depends on may change while a tape is open.
  void check_no_tape(const char* what) const {
    if (taping_ && !replaying_) throw Error{MPMX_ESTATE, std::string(what) + " cannot change while a tape is recorded"};
  }

  void upload_particles(int64_t n, const void* X, const void* V, const void* F, const void* B, const void* mass,
                        const void* volume, const void* mu, const void* lambda, const uint8_t* valid,
                        const uint16_t* scene_ids) {
    check_no_tape("particles");
    if (n < 0 || n > (int64_t)0x7fffffff) throw Error{MPMX_EINVAL, "particle count out of range"};
    if (n > 0 && (!X || !V || !F || !B || !mass || !volume || !mu || !lambda || !valid))
      throw Error{MPMX_EINVAL, "null particle array"};
    n_ = (uint32_t)n;
    allocate_particles();
    if (n_ > 0) {
      upload_soa((const T*)X, cur_.X, D);
      upload_soa((const T*)V, cur_.V, D);
      upload_soa((const T*)F, cur_.F, D * D);
      upload_soa((const T*)B, cur_.B, D * D);
      upload_soa((const T*)mass, cur_.mass, 1);
      upload_soa((const T*)volume, cur_.vol, 1);
      upload_soa((const T*)mu, cur_.mu, 1);
      upload_soa((const T*)lambda, cur_.la, 1);
      HIPCHECK(hipMemcpyAsync(cur_.valid, valid, n_, hipMemcpyHostToDevice, stream_));
      if (scene_ids) HIPCHECK(hipMemcpyAsync(cur_.pscene, scene_ids, n_ * sizeof(uint16_t), hipMemcpyHostToDevice, stream_));
      iota_kernel<<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(cur_.pid, n_);
      user_order_ = true;
      if (plastic_)
        plastic_init_kernel<T, D><<<std::min<uint32_t>(cdiv(n_, BLOCK), 4096), BLOCK, 0, stream_>>>(cur_, n_, pl_);
      HIPCHECK(hipMemsetAsync(err_.p, 0, 4 * sizeof(uint32_t), stream_));
      Args<T, D> a = make_args(0);
      key_kernel<T, D><<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(a, err_.p + 1);
      uint32_t errs[4];
      HIPCHECK(hipMemcpyAsync(errs, err_.p, sizeof(errs), hipMemcpyDeviceToHost, stream_));
      HIPCHECK(hipStreamSynchronize(stream_));
      if (errs[1]) {
        n_ = 0;
        throw Error{MPMX_ERANGE, std::to_string(errs[1]) + " valid particle(s) outside the ghosted grid"};
      }
    }
    force_rebin_ = true;
    force_reorder_ = true;
    has_grid_ = false;
    fused_ = false;
    fused_last_ = false;
  }

  void get_particles(void* X, void* V, void* F, void* B, uint8_t* valid) override {
    if (n_ == 0) return;
    download_aos(cur_.X, (T*)X, D);
    download_aos(cur_.V, (T*)V, D);
    download_aos(cur_.F, (T*)F, D * D);
    download_aos(cur_.B, (T*)B, D * D);
    if (valid) {
      stage_bytes_.alloc(n_);
      bytes_from_slots_kernel<<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(cur_.valid, stage_bytes_.p, cur_.pid, n_);
      HIPCHECK(hipMemcpyAsync(valid, stage_bytes_.p, n_, hipMemcpyDeviceToHost, stream_));
    }
    HIPCHECK(hipStreamSynchronize(stream_));
  }

  void get_material(void* mass, void* volume, void* mu, void* lambda) override {
    if (n_ == 0) return;
    download_aos(cur_.mass, (T*)mass, 1);
    download_aos(cur_.vol, (T*)volume, 1);
    download_aos(cur_.mu, (T*)mu, 1);
    download_aos(cur_.la, (T*)lambda, 1);
  }
  void set_time(double t) override { time_ = t; }
  int plastic_kind() const override { return plastic_; }
  void get_lame0(void* mu0, void* lambda0) override {
    if (!plastic_) throw Error{MPMX_ESTATE, "plasticity is not enabled"};
    if (n_ == 0) return;
    download_aos(cur_.mu0, (T*)mu0, 1);
    download_aos(cur_.la0, (T*)lambda0, 1);
  }
  // Plastic state of the particles just given to set_particles (user order),
  // as a restart reads it; null arrays keep the values set_particles gave.
  void restore_plastic(const void* Fp, const void* mu0, const void* lambda0, const void* plastic_def,
                       const void* rho_F, const void* cohesion, const void* vc) override {
    if (!plastic_) throw Error{MPMX_ESTATE, "plasticity is not enabled"};
    if (!user_order_) throw Error{MPMX_ESTATE, "plastic state is restored right after mpmx_set_particles"};
    if (n_ == 0) return;
    if (Fp) upload_soa((const T*)Fp, cur_.Fp, D * D);
    if (mu0) upload_soa((const T*)mu0, cur_.mu0, 1);
    if (lambda0) upload_soa((const T*)lambda0, cur_.la0, 1);
    if (plastic_ == 2) {
      if (plastic_def) upload_soa((const T*)plastic_def, cur_.pdef, 1);
      if (rho_F) upload_soa((const T*)rho_F, cur_.rhoF, 1);
      if (cohesion) upload_soa((const T*)cohesion, cur_.coh, 1);
      if (vc) upload_soa((const T*)vc, cur_.vc, 1);
    }
  }

  void set_constitutive(int model, double thr) override {
    check_no_tape("the constitutive model");
    if (model != 0 && model != 1 && model != 2) throw Error{MPMX_EINVAL, "unknown constitutive model"};
    if (model == 1 && !(thr > 0)) throw Error{MPMX_EINVAL, "neo-Hookean threshold must be positive"};
    model_ = model;
    thr_ = (T)thr;
    fused_ = false;  // the pending P2G used the previous model
    root_ = D == 2 ? (T)std::sqrt(thr) : (T)std::cbrt(thr);  // pow<1,d>(failure_threshold)
  }
  void set_gravity(const double* g) override {
    check_no_tape("gravity");
    if (!g) throw Error{MPMX_EINVAL, "null gravity"};
    for (int a = 0; a < D; a++) grav_[a] = (T)g[a];
  }
  void add_collision_box(int type, double friction, const double* lo, const double* hi) override {
    check_no_tape("collision objects");
    if (type < 0 || type > 2) throw Error{MPMX_EINVAL, "unknown collision type"};
    if (!lo || !hi) throw Error{MPMX_EINVAL, "null box corner"};
    if (nwalls_ >= MAX_WALLS) throw Error{MPMX_EINVAL, "too many collision objects (max 16)"};
    WallDev<T> w{};
    w.type = type;
    w.friction = (T)friction;
    for (int a = 0; a < 3; a++) {
      w.lo[a] = a < D ? (T)lo[a] : 0;
      w.hi[a] = a < D ? (T)hi[a] : 0;
    }
    for (int a = 0; a < D; a++)
      if (!(w.hi[a] > w.lo[a])) throw Error{MPMX_EINVAL, "degenerate collision box"};
    walls_[nwalls_++] = w;
  }
  void clear_collision_objects() override {
    check_no_tape("collision objects");
    nwalls_ = 0;
  }

  void step(double dt, int nsteps, bool wait) override {
    if (!(dt > 0)) throw Error{MPMX_EINVAL, "dt must be positive"};
    if (stage_next_ != 0) throw Error{MPMX_ESTATE, "a step run stage by stage is open: finish its stages first"};
    if (nsteps < 0) throw Error{MPMX_EINVAL, "negative step count"};
    // recorded steps too (the reverse pass recomputes its own); plasticity
    // projects F between G2P and the next P2G, which the fused kernel does not
    // grid statistics read each step's P2G grid, which the fused kernel never
    // holds on its own
    const bool fuse = n_ > 0 && !plastic_ && !limits_on() && !refl_on_ && !stats_on_;
    if (fuse && use_graph_ && !timer_.on && !timer_.kernel_on && !taping_) {
      step_graph((T)dt, nsteps);
    } else {
      for (int s = 0; s < nsteps; s++) {
        if (fuse)
          step_fused((T)dt);
        else
          step_once((T)dt);
      }
    }
    if (wait) synchronize();
  }

  void synchronize() override {
    HIPCHECK(hipStreamSynchronize(stream_));
    timer_.resolve();
    uint32_t errs[4];
    HIPCHECK(hipMemcpy(errs, err_.p, sizeof(errs), hipMemcpyDeviceToHost));
    if (errs[0]) throw Error{MPMX_EINTERNAL, "particle stencil outside the grid"};
  }

  void enable_grid_stats(bool on) override { stats_on_ = on; }
  void grid_stats(double* mass, double* momentum) override {
    double out[NTOT];
    grid_totals(1, out);
    if (mass) *mass = out[0];
    if (momentum)
      for (int a = 0; a < D; a++) momentum[a] = out[1 + a];
  }
  // Grid totals of the last step (after P2G: which = 0; after the grid
  // update: which = 1), recorded when grid stats are enabled.
  void grid_totals(int which, double* out) override {
    if (!stats_on_ || !has_grid_) throw Error{MPMX_ESTATE, "grid totals need enable_grid_stats and a step"};
    if (which != 0 && which != 1) throw Error{MPMX_EINVAL, "which must be 0 (after P2G) or 1 (after the grid update)"};
    HIPCHECK(hipStreamSynchronize(stream_));
    HIPCHECK(hipMemcpy(out, stats_out_.p + which * NTOT, NTOT * sizeof(double), hipMemcpyDeviceToHost));
  }
  // Grid totals pass over the current tile list into stats_out_[which].
  void record_grid_totals(const Args<T, D>& a, const Node4<T>* grid, int which) {
    if (which == 0)
      grid_totals_kernel<T, D, true><<<TOTALS_BLOCKS, BLOCK, 0, stream_>>>(a, grid, stats_partial_.p);
    else
      grid_totals_kernel<T, D, false><<<TOTALS_BLOCKS, BLOCK, 0, stream_>>>(a, grid, stats_partial_.p);
    totals_finish_kernel<<<1, 64, 0, stream_>>>(stats_partial_.p, TOTALS_BLOCKS, stats_out_.p + which * NTOT);
  }
  TotalsIn<T, D> totals_in() const {
    TotalsIn<T, D> in;
    in.X = cur_.X; in.V = cur_.V; in.F = cur_.F; in.B = cur_.B;
    in.mass = cur_.mass; in.vol = cur_.vol; in.mu = cur_.mu; in.la = cur_.la;
    in.valid = cur_.valid;
    in.pscene = scenes_.size() > 1 ? cur_.pscene : nullptr;
    in.scenes = scenes_dev_.p;
    in.n = n_;
    in.model = model_;
    for (int c = 0; c < 3; c++) in.grav[c] = c < D ? (double)grav_[c] : 0.0;
    return in;
  }
  // Particle totals of the current state into the device array out[NTOT]
  // (enqueued on the context stream), block partials in `partial`.
  void enqueue_particle_totals(double* partial, double* out) {
    if (n_ == 0) {
      fill_u32_kernel<<<1, BLOCK, 0, stream_>>>((uint32_t*)out, 2 * NTOT, 0u);
    } else {
      particle_totals_kernel<T, D><<<TOTALS_BLOCKS, BLOCK, 0, stream_>>>(totals_in(), partial);
      totals_finish_kernel<<<1, 64, 0, stream_>>>(partial, TOTALS_BLOCKS, out);
    }
    HIPCHECK(hipGetLastError());
  }
  void particle_totals(double* out) override {
    ptot_partial_.alloc((size_t)TOTALS_BLOCKS * NTOT + NTOT);
    enqueue_particle_totals(ptot_partial_.p, ptot_partial_.p + (size_t)TOTALS_BLOCKS * NTOT);
    HIPCHECK(hipMemcpyAsync(out, ptot_partial_.p + (size_t)TOTALS_BLOCKS * NTOT, NTOT * sizeof(double),
                            hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  void particle_stats(double* mass, double* momentum, int64_t* valid_count) override {
    double out[NTOT];
    particle_totals(out);
    if (mass) *mass = out[0];
    if (momentum)
      for (int a = 0; a < D; a++) momentum[a] = out[1 + a];
    if (valid_count) *valid_count = (int64_t)out[11];
  }

  // ---- adaptive dt (MPM_DRIVER.cpp:1098-1138, :173-197) ----
  // Per-scene Grid_V_Upper_Bound of the current particles; waits for the stream.
  void speed_bounds(T* out) {
    const int ns = (int)scenes_.size();
    if (!vb_host_) {
      vbound_.alloc(ns);
      HIPCHECK(hipHostMalloc((void**)&vb_host_, ns * sizeof(Ubits), hipHostMallocDefault));
    }
    hipEvent_t e0 = timer_.begin(stream_);
    const uint32_t words = (uint32_t)(ns * sizeof(Ubits) / sizeof(uint32_t));
    fill_u32_kernel<<<cdiv(words, BLOCK), BLOCK, 0, stream_>>>((uint32_t*)vbound_.p, words, 0u);
    if (n_ > 0) {
      const uint32_t nb = std::min<uint32_t>(cdiv(n_, BLOCK), 2048);
      if (ns > 1)
        speed_bound_kernel<T, D, true><<<nb, BLOCK, 0, stream_>>>(cur_.V, cur_.B, cur_.valid, cur_.pscene,
                                                                  scenes_dev_.p, n_, vbound_.p);
      else
        speed_bound_kernel<T, D, false><<<nb, BLOCK, 0, stream_>>>(cur_.V, cur_.B, cur_.valid, nullptr,
                                                                   scenes_dev_.p, n_, vbound_.p);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(vb_host_, vbound_.p, ns * sizeof(Ubits), hipMemcpyDeviceToHost, stream_));
    timer_.end(PH_DT, e0, stream_);
    HIPCHECK(hipStreamSynchronize(stream_));
    for (int s = 0; s < ns; s++) std::memcpy(out + s, vb_host_ + s, sizeof(T));
  }
  // MPM_DRIVER::Compute_Dt for scene s and speed bound v, in T.
  T scene_dt(int s, T v, T cfl, T max_dt) const {
    const Geom<T, D>& g = scenes_[s];
    T dxmin = g.dx[0];
    for (int a = 1; a < D; a++) dxmin = g.dx[a] < dxmin ? g.dx[a] : dxmin;
    const T critical_speed = cfl * dxmin / max_dt;
    return v > critical_speed ? cfl * dxmin / v : max_dt;
  }
  void grid_v_upper_bound(double* out) override {
    std::vector<T> v(scenes_.size());
    speed_bounds(v.data());
    for (size_t s = 0; s < v.size(); s++) out[s] = (double)v[s];
  }
  void compute_dt(double cfl, double max_dt, double* dt) override {
    if (!(cfl > 0) || !(max_dt > 0)) throw Error{MPMX_EINVAL, "cfl and max_dt must be positive"};
    std::vector<T> v(scenes_.size());
    speed_bounds(v.data());
    for (size_t s = 0; s < v.size(); s++) dt[s] = (double)scene_dt((int)s, v[s], (T)cfl, (T)max_dt);
  }
  // The substep loop of MPM_DRIVER::Simulate_To_Frame in T: dt from
  // Compute_Dt, clamped to [min_dt, max_dt] (Core/Math_Tools/clamp.h:13-16),
  // the last substep of a frame lands on the frame time and the one before it
  // splits the remainder in half.  With the strong or sound-speed CFL on,
  // Reduce_Dt uses this cfl and min_dt.
  void advance_frames(int nframes, double frame_dt, double cfl, double min_dt, double max_dt, double* dts,
                      int64_t capacity, int64_t* nsub) override {
    if (scenes_.size() > 1) throw Error{MPMX_ESTATE, "adaptive frames take a single-scene context"};
    if (nframes < 0) throw Error{MPMX_EINVAL, "negative frame count"};
    if (!(frame_dt > 0) || !(cfl > 0) || !(max_dt > 0) || !(min_dt >= 0) || min_dt > max_dt)
      throw Error{MPMX_EINVAL, "need frame_dt, cfl, max_dt > 0 and 0 <= min_dt <= max_dt"};
    const T fdt = (T)frame_dt, tcfl = (T)cfl, tmin = (T)min_dt, tmax = (T)max_dt;
    int64_t count = 0;
    if (nsub) *nsub = 0;
    T time = (T)time_;
    for (int f = 0; f < nframes; f++) {
      const T time_at_frame = time + fdt;
      for (bool done = false; !done;) {
        T v;
        speed_bounds(&v);
        T dt = scene_dt(0, v, tcfl, tmax);
        dt = dt <= tmin ? tmin : (dt >= tmax ? tmax : dt);
        T next_time = time + dt;
        if (next_time + (T).0001 * dt > time_at_frame) {
          next_time = time_at_frame;
          done = true;
        } else if (next_time + dt > time_at_frame) {
          next_time = (time + time_at_frame) / 2;
        }
        dt = next_time - time;
        if (!(dt > 0)) throw Error{MPMX_ESTATE, "substep dt is not positive (raise min_dt)"};
        const T taken = step_once(dt, tcfl, tmin);
        if (taken < dt) {  // Reduce_Dt lowered it: the frame goes on (MPM_DRIVER.cpp:190-194)
          next_time = time + taken;
          done = false;
        }
        time = next_time;
        time_ = (double)time;
        if (dts && count < capacity) dts[count] = (double)taken;
        count++;
        if (nsub) *nsub = count;
      }
    }
    synchronize();
  }

  int64_t grid_node_count() const override {
    int64_t c = 1;
    for (int a = 0; a < D; a++) c *= g_.G[a];
    return c;
  }
  void get_grid(void* mass, void* velocity) override {
    if (scenes_.size() > 1) throw Error{MPMX_ESTATE, "mpmx_get_grid takes a single-scene context"};
    const bool momentum = grid_holds_momentum();
    if (!has_grid_ && !momentum && stage_next_ <= ST_P2G) throw Error{MPMX_ESTATE, "no grid yet: call mpmx_step first"};
    const int64_t nn = grid_node_count();
    DevBuf<T> dm, dv;
    dm.alloc(nn);
    dv.alloc(nn * D);
    HIPCHECK(hipMemsetAsync(dm.p, 0, nn * sizeof(T), stream_));
    HIPCHECK(hipMemsetAsync(dv.p, 0, nn * D * sizeof(T), stream_));
    Args<T, D> a = make_args(0);
    if (fused_last_) {  // the last step's grid is the second buffer of the fused pipeline
      a.grid = gother_;
      a.gtiles = gt_other_;
      a.counts = lastcnt_.p;
    }
    if (momentum) a.grid = vsave_.p;  // between "p2g" and "forces": Particle_To_Grid's own (m, m v)
    export_grid_kernel<T, D><<<grid_blocks(), BLOCK, 0, stream_>>>(a, dm.p, dv.p, momentum ? 1 : 0);
    if (mass) HIPCHECK(hipMemcpyAsync(mass, dm.p, nn * sizeof(T), hipMemcpyDeviceToHost, stream_));
    if (velocity) HIPCHECK(hipMemcpyAsync(velocity, dv.p, nn * D * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }

  void set_launch_config(const mpmx_launch_config& cfg) override {
    if (cfg.graph != 0 && cfg.graph != 1) throw Error{MPMX_EINVAL, "graph must be 0 or 1"};
    if (cfg.tile_workgroups < 0 || cfg.tile_workgroups > (1 << 20))
      throw Error{MPMX_EINVAL, "tile_workgroups must be in [0, 2^20]"};
    if (cfg.tile_threads != 0 && cfg.tile_threads != 128 && cfg.tile_threads != 256)
      throw Error{MPMX_EINVAL, "tile_threads must be 0, 128 or 256"};
    use_graph_ = cfg.graph != 0;
    tile_wg_ = (uint32_t)cfg.tile_workgroups;
    fused_threads_ = cfg.tile_threads ? cfg.tile_threads : 128;
    release_graph();
  }
  mpmx_launch_config launch_config() const override {
    mpmx_launch_config c;
    c.graph = use_graph_ ? 1 : 0;
    c.tile_workgroups = (int)tile_wg_;
    c.tile_threads = fused_threads_;
    return c;
  }
  // the fused tile kernel at the configured workgroup size
  void launch_fused_kernel(const Args<T, D>& a, bool fric) {
    if (fused_threads_ == 256) {
      if (fric)
        g2p2g_kernel<T, D, true, 256><<<tile_blocks(), 256, 0, stream_>>>(a);
      else
        g2p2g_kernel<T, D, false, 256><<<tile_blocks(), 256, 0, stream_>>>(a);
    } else {
      if (fric)
        g2p2g_kernel<T, D, true, 128><<<tile_blocks(), 128, 0, stream_>>>(a);
      else
        g2p2g_kernel<T, D, false, 128><<<tile_blocks(), 128, 0, stream_>>>(a);
    }
  }
  // Times each candidate configuration on the context's own particles (saved
  // and restored), applies the fastest.
  void autotune(double dt, int trial, mpmx_launch_config* chosen, double* times, int max_times,
                int* ntimes) override {
    if (!(dt > 0)) throw Error{MPMX_EINVAL, "dt must be positive"};
    if (trial < 8) throw Error{MPMX_EINVAL, "trial_steps must be at least 8"};
    if (n_ == 0) throw Error{MPMX_ESTATE, "autotuning needs particles"};
    if (taping_ || stage_next_ != 0) throw Error{MPMX_ESTATE, "autotuning needs no tape and no open staged step"};
    if (plastic_ || limits_on() || refl_on_ || stats_on_)
      throw Error{MPMX_ESTATE, "autotuning times the fused step: no plasticity, dt limits, reflection or grid stats"};
    int cus = 0, per128 = 0, per256 = 0;
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_));
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per128, g2p2g_kernel<T, D, false, 128>, 128, 0));
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per256, g2p2g_kernel<T, D, false, 256>, 256, 0));
    std::vector<mpmx_launch_config> cand;
    for (int gr = 0; gr < 2; gr++)
      for (int th : {128, 256}) {
        const uint32_t resident = (uint32_t)std::max(1, th == 128 ? per128 : per256) * (uint32_t)std::max(1, cus);
        for (uint32_t w : {0u, resident, 2 * resident, 4 * resident})
          cand.push_back(mpmx_launch_config{gr, (int)w, th});
      }
    const mpmx_launch_config before = launch_config();
    const bool timing = timer_.on, ktiming = timer_.kernel_on;
    const double t_before = time_, last_dt_before = last_dt_;
    // storage in bin order first (as the steps keep it), then saved: the trials
    // rebuild the bins but do not reorder, so the saved state stays in slot
    // order and every trial reads nearly coalesced particles
    force_reorder_ = true;
    rebin();
    fused_ = false;
    fused_last_ = false;
    StateBuf saved;
    save_state(saved);
    std::vector<double> ms(cand.size(), 0.0);
    tuning_ = true;
    try {
      timer_.on = false;
      timer_.kernel_on = false;
      for (size_t c = 0; c < cand.size(); c++) {
        set_launch_config(cand[c]);
        load_state(saved);
        step(dt, 4, true);
        const auto t0 = std::chrono::steady_clock::now();
        step(dt, trial, true);
        const auto t1 = std::chrono::steady_clock::now();
        ms[c] = std::chrono::duration<double, std::milli>(t1 - t0).count() / trial;
      }
    } catch (...) {
      set_launch_config(before);
      timer_.on = timing;
      timer_.kernel_on = ktiming;
      tuning_ = false;
      throw;
    }
    size_t best = 0;
    for (size_t c = 1; c < cand.size(); c++)
      if (ms[c] < ms[best]) best = c;
    set_launch_config(cand[best]);
    load_state(saved);
    HIPCHECK(hipStreamSynchronize(stream_));
    time_ = t_before;
    last_dt_ = last_dt_before;
    timer_.on = timing;
    timer_.kernel_on = ktiming;
    tuning_ = false;
    if (chosen) *chosen = cand[best];
    for (int i = 0; i < max_times && i < (int)cand.size(); i++) times[i] = ms[i];
    if (ntimes) *ntimes = (int)cand.size();
  }

  void set_reorder_interval(int k) override {
    if (k < 1) throw Error{MPMX_EINVAL, "reorder interval must be >= 1"};
    reorder_interval_ = k;
  }
  void set_rebin_policy(double mover_fraction, int max_interval) override {
    if (!(mover_fraction >= 0)) throw Error{MPMX_EINVAL, "mover fraction must be >= 0"};
    if (max_interval < 0) throw Error{MPMX_EINVAL, "max interval must be >= 0"};
    rebin_fraction_ = mover_fraction;
    rebin_max_interval_ = max_interval;
  }
  // Plastic arrays of one particle buffer set: Fp, mu0, lambda0 and the
  // Drucker-Prager state.
  static void add_plastic_arrays(ParticleBufs<T, D>& pb, std::vector<DevBuf<T>>& fb, size_t n) {
    fb.resize(15);
    fb[8].alloc(n * D * D);
    for (int k = 9; k < 15; k++) fb[k].alloc(n);
    pb.Fp = fb[8].p;
    pb.mu0 = fb[9].p;
    pb.la0 = fb[10].p;
    pb.pdef = fb[11].p;
    pb.rhoF = fb[12].p;
    pb.coh = fb[13].p;
    pb.vc = fb[14].p;
  }
  // Switches plasticity model `kind` (0 off, 1 clamp, 2 Drucker-Prager) on,
  // starting from Fp = I, mu0 = mu, lambda0 = lambda and the initial
  // Drucker-Prager state.
  void enable_plastic(int kind, const PlasticDev<T>& pl) {
    if (taping_) throw Error{MPMX_ESTATE, "plasticity cannot change while a tape is recorded"};
    fused_ = false;
    if (kind == 0) {
      plastic_ = 0;
      cur_.Fp = cur_.mu0 = cur_.la0 = cur_.pdef = cur_.rhoF = cur_.coh = cur_.vc = nullptr;
      alt_.Fp = alt_.mu0 = alt_.la0 = alt_.pdef = alt_.rhoF = alt_.coh = alt_.vc = nullptr;
      return;
    }
    pl_ = pl;
    pl_.kind = kind;
    plastic_ = kind;
    if (n_ > 0) {
      // cur_ always uses fa_ and alt_ fb_ (a reorder swaps both pairs)
      add_plastic_arrays(cur_, fa_, n_);
      add_plastic_arrays(alt_, fb_, n_);
      plastic_init_kernel<T, D><<<std::min<uint32_t>(cdiv(n_, BLOCK), 4096), BLOCK, 0, stream_>>>(cur_, n_, pl_);
      HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  void set_plasticity(bool enable, double theta_c, double theta_s, double max_hardening, double xi) override {
    if (!enable) {
      if (plastic_ == 1) enable_plastic(0, pl_);
      return;
    }
    if (!(theta_c >= 0 && theta_c < 1) || !(theta_s >= 0) || !std::isfinite(max_hardening) || !std::isfinite(xi))
      throw Error{MPMX_EINVAL, "need 0 <= theta_c < 1, theta_s >= 0 and finite hardening parameters"};
    PlasticDev<T> pl{};
    // the clamp bounds as the reference forms them in T
    pl.lo = (double)((T)1 - (T)theta_c);
    pl.hi = (double)((T)1 + (T)theta_s);
    pl.max_hardening = (double)(T)max_hardening;
    pl.xi = (double)(T)xi;
    enable_plastic(1, pl);
  }
  void set_drucker_prager(bool enable, double a0, double a1, double a3, double a4, double cohesion,
                          bool volume_correction) override {
    if (!enable) {
      if (plastic_ == 2) enable_plastic(0, pl_);
      return;
    }
    if (!std::isfinite(a0) || !std::isfinite(a1) || !std::isfinite(a3) || !std::isfinite(a4) || !(cohesion >= 0))
      throw Error{MPMX_EINVAL, "need finite friction-angle parameters and cohesion >= 0"};
    PlasticDev<T> pl{};
    pl.a0 = (double)(T)a0;
    pl.a1 = (double)(T)a1;
    pl.a3 = (double)(T)a3;
    pl.a4 = (double)(T)a4;
    pl.cohesion = (double)(T)cohesion;
    pl.use_vc = volume_correction ? 1 : 0;
    if (!(dp_rho_host(pl, 0.0) > 0)) throw Error{MPMX_EINVAL, "the initial friction angle must be positive"};
    enable_plastic(2, pl);
  }
  static double dp_rho_host(const PlasticDev<T>& pl, double q) {
    const double phi = 3.14159265358979323846 / 180 * (pl.a0 + (pl.a1 * q - pl.a4) * std::exp(-pl.a3 * q));
    const double sp = std::sin(phi);
    return std::sqrt(2.0 / 3.0) * 2 * sp / (3 - sp);
  }
  void get_drucker_prager_state(void* plastic_def, void* rho_F, void* cohesion, void* vc) override {
    if (plastic_ != 2) throw Error{MPMX_ESTATE, "Drucker-Prager plasticity is not enabled"};
    if (n_ == 0) return;
    download_aos(cur_.pdef, (T*)plastic_def, 1);
    download_aos(cur_.rhoF, (T*)rho_F, 1);
    download_aos(cur_.coh, (T*)cohesion, 1);
    download_aos(cur_.vc, (T*)vc, 1);
  }
  void get_plastic_state(void* Fp, void* mu, void* lambda) override {
    if (!plastic_) throw Error{MPMX_ESTATE, "plasticity is not enabled"};
    if (n_ == 0) return;
    download_aos(cur_.Fp, (T*)Fp, D * D);
    download_aos(cur_.mu, (T*)mu, 1);
    download_aos(cur_.la, (T*)lambda, 1);
  }
  void bin_stats(int64_t* rebins, int64_t* movers, int64_t* slow) override {
    synchronize();
    uint32_t c[16];
    HIPCHECK(hipMemcpy(c, counts_.p, sizeof(c), hipMemcpyDeviceToHost));
    if (rebins) *rebins = rebins_;
    if (movers) *movers = has_grid_ ? (int64_t)(fused_last_ ? c[10] : c[4]) : 0;
    if (slow) *slow = has_grid_ ? (int64_t)(fused_last_ ? c[11] : c[3]) : 0;
  }
  void enable_timing(int mode) override {
    synchronize();
    timer_.on = mode == 1;
    timer_.kernel_on = mode != 0;
  }
  void reset_timing() override {
    synchronize();
    for (int i = 0; i < PhaseTimer::NPH; i++) {
      timer_.total[i] = 0;
      timer_.count[i] = 0;
    }
  }
  bool phase_time(const char* phase, double* total_ms, int64_t* count) override {
    synchronize();
    if (phase && std::strcmp(phase, "graph") == 0) {  // graph replays (groups of four steps) so far; no time
      if (total_ms) *total_ms = 0;
      if (count) *count = graph_groups_;
      return true;
    }
    for (int i = 0; i < PhaseTimer::NPH; i++)
      if (phase && std::strcmp(phase, kPhaseNames[i]) == 0) {
        if (total_ms) *total_ms = timer_.total[i];
        if (count) *count = timer_.count[i];
        return true;
      }
    return false;
  }

  // ---- differentiation: checkpointed tape and reverse pass
  // One saved particle state; one recomputed step's grids and touched tiles.
  struct StateBuf {
    DevBuf<T> s;
    DevBuf<uint8_t> valid;
  };
  struct StepGrid {
    DevBuf<Node4<T>> grid, gridu;
    DevBuf<uint32_t> tiles;
  };

  void tape_begin(int k) override {
    if (k < 1) throw Error{MPMX_EINVAL, "checkpoint interval must be >= 1"};
    if (n_ == 0) throw Error{MPMX_ESTATE, "no particles"};
    if (any_friction()) throw Error{MPMX_EINVAL, "the adjoint step has no friction: remove wall friction"};
    if (plastic_) throw Error{MPMX_EINVAL, "the adjoint step has no plasticity"};
    if (model_ == 2) throw Error{MPMX_EINVAL, "the adjoint step has no Hencky-strain model"};
    if (limits_on()) throw Error{MPMX_EINVAL, "the adjoint step has a fixed dt: turn the strong and sound CFL off"};
    if (refl_on_) throw Error{MPMX_EINVAL, "the adjoint step has no reflection boundary conditions"};
    if (force_reorder_ || force_rebin_) rebin();  // settle the storage order before recording
    particle_stats(nullptr, nullptr, &tape_valid_);
    fused_ = false;
    taping_ = true;
    tape_k_ = k;
    tape_dt_.clear();
    nckpt_ = 0;
    save_state(next_checkpoint());
  }

  void backward(const void* dX, const void* dV, const void* dF, const void* dB, void* gX, void* gV, void* gF, void* gB,
                double* gmu, double* gla) override {
    if (!taping_) throw Error{MPMX_ESTATE, "no tape: call mpmx_tape_begin, then mpmx_step"};
    const size_t n = n_;
    aX_.alloc(n * D); aV_.alloc(n * D); aF_.alloc(n * D * D); aB_.alloc(n * D * D);
    amu_.alloc(n); ala_.alloc(n);
    agrid_.alloc((size_t)ntotal_ * 64);
    upload_adjoint((const T*)dX, aX_.p, D);
    upload_adjoint((const T*)dV, aV_.p, D);
    upload_adjoint((const T*)dF, aF_.p, D * D);
    upload_adjoint((const T*)dB, aB_.p, D * D);
    HIPCHECK(hipMemsetAsync(amu_.p, 0, n * sizeof(double), stream_));
    HIPCHECK(hipMemsetAsync(ala_.p, 0, n * sizeof(double), stream_));
    HIPCHECK(hipMemsetAsync(err_.p, 0, 4 * sizeof(uint32_t), stream_));
    StateBuf fin;
    save_state(fin);
    // On any exception the live state pointers come back, the flags clear and
    // the final state is copied back (best effort); the tape is dropped.
    struct Restore {
      Engine& e;
      const StateBuf& fin;
      ParticleBufs<T, D> live;
      bool armed = true;
      ~Restore() {
        if (!armed) return;
        e.cur_.X = live.X;
        e.cur_.V = live.V;
        e.cur_.F = live.F;
        e.cur_.B = live.B;
        e.cur_.valid = live.valid;
        e.grid_override_ = nullptr;
        e.gridu_override_ = nullptr;
        e.replaying_ = false;
        e.taping_ = false;
        e.nckpt_ = 0;
        e.fused_ = false;
        e.fused_last_ = false;
        e.force_rebin_ = true;
        try {
          e.load_state(fin);
          (void)hipStreamSynchronize(e.stream_);
        } catch (...) {
        }
      }
    } restore{*this, fin, cur_};
    const double time0 = time_;
    int64_t valid_now = 0;
    particle_stats(nullptr, nullptr, &valid_now);
    const bool invalidated = valid_now != tape_valid_;
    const int ns = (int)tape_dt_.size(), k = tape_k_;
    const size_t kk = (size_t)std::min(k, std::max(ns, 1));
    if (seg_.size() < kk) seg_.resize(kk);
    if (sgrid_.size() < kk) sgrid_.resize(kk);
    std::vector<StateBuf>& seg = seg_;
    std::vector<StepGrid>& sgrid = sgrid_;
    seg_counts_.alloc(kk);
    const size_t nodes = (size_t)ntotal_ * 64;
    replaying_ = true;
    hipEvent_t ea = timer_.begin(stream_);
    const int nseg = ns > 0 ? (ns - 1) / k + 1 : 0;
    if (!invalidated && nseg >= 2) {
      reverse_overlapped(nseg, kk);
      timer_.end(PH_ADJOINT, ea, stream_);
      replaying_ = false;
      restore.armed = false;
      finish_backward(fin, time0, gX, gV, gF, gB, gmu, gla);
      return;
    }
    for (int sg = nseg - 1; sg >= 0; sg--) {
      const int s0 = sg * k, s1 = std::min(ns, s0 + k);
      // forward from the checkpoint, keeping each step's state, grid and tiles
      load_state(ckpts_[(size_t)sg]);  // sg < nckpt_
      for (int j = s0; j < s1; j++) {
        const size_t i = (size_t)(j - s0);
        save_state(seg[i]);
        sgrid[i].grid.alloc(nodes);
        sgrid[i].gridu.alloc(nodes);
        sgrid[i].tiles.alloc((size_t)ntotal_);
        grid_override_ = sgrid[i].grid.p;
        gridu_override_ = sgrid[i].gridu.p;
        step_once((T)tape_dt_[(size_t)j]);
        save_tiles_kernel<<<copy_blocks(), BLOCK, 0, stream_>>>(gt_cur_, counts_.p, sgrid[i].tiles.p,
                                                                seg_counts_.p + i);
        grid_override_ = nullptr;
        gridu_override_ = nullptr;
      }
      const ParticleBufs<T, D> live = cur_;
      for (int j = s1 - 1; j >= s0; j--) {
        // The bins of a later state cover an earlier one (movers are listed)
        // unless a particle was invalidated in between: then re-bin.
        const size_t i = (size_t)(j - s0);
        if (invalidated) {
          load_state(seg[i], true);
        } else {
          // the reverse kernels only read X, V, F, B and valid: they read the
          // saved state where it lies (no copy); a re-bin while taping sorts
          // keys but moves no particle data
          T* q = seg[i].s.p;
          cur_.X = q;
          cur_.V = q + n * D;
          cur_.F = q + 2 * n * D;
          cur_.B = q + n * (2 * D + D * D);
          cur_.valid = seg[i].valid.p;
          fused_ = false;
          fused_last_ = false;
        }
        backward_step((T)tape_dt_[(size_t)j], sgrid[i], seg_counts_.p + i);
      }
      cur_.X = live.X;
      cur_.V = live.V;
      cur_.F = live.F;
      cur_.B = live.B;
      cur_.valid = live.valid;
    }
    timer_.end(PH_ADJOINT, ea, stream_);
    replaying_ = false;
    restore.armed = false;
    finish_backward(fin, time0, gX, gV, gF, gB, gmu, gla);
  }

  // The reverse pass's results to the host and the live state back.
  void finish_backward(const StateBuf& fin, double time0, void* gX, void* gV, void* gF, void* gB, double* gmu,
                       double* gla) {
    download_aos(aX_.p, (T*)gX, D);
    download_aos(aV_.p, (T*)gV, D);
    download_aos(aF_.p, (T*)gF, D * D);
    download_aos(aB_.p, (T*)gB, D * D);
    download_param(amu_.p, gmu);
    download_param(ala_.p, gla);
    uint32_t errs[4];
    HIPCHECK(hipMemcpy(errs, err_.p, sizeof(errs), hipMemcpyDeviceToHost));
    load_state(fin);
    time_ = time0;
    taping_ = false;
    nckpt_ = 0;
    HIPCHECK(hipStreamSynchronize(stream_));
    if (errs[2]) throw Error{MPMX_EINVAL, "adjoint: a neo-Hookean particle is in the clamp branch (J < threshold)"};
    if (errs[0]) throw Error{MPMX_EINTERNAL, "particle stencil outside the grid"};
  }

  // Reverse pass over nseg >= 2 segments with no particle invalidated on the
  // tape.  A second engine on its own stream recomputes segment s-1 from its
  // checkpoint (states, grids, touched tiles into storage set (s-1) mod 2)
  // while this stream runs the reverse kernels of segment s on set s mod 2;
  // events order each set's writer after its last reader.  The reverse
  // kernels read the saved states in place and bin them on this engine.
  void reverse_overlapped(int nseg, size_t kk) {
    const int k = tape_k_, ns = (int)tape_dt_.size();
    const size_t n = n_;
    for (hipEvent_t* e : {&seg_done_[0], &seg_done_[1], &seg_free_[0], &seg_free_[1], &rep_go_})
      if (!*e) HIPCHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    std::vector<StateBuf>* segs[2] = {&seg_, &segb_};
    std::vector<StepGrid>* grids[2] = {&sgrid_, &sgridb_};
    DevBuf<uint32_t>* cnts[2] = {&seg_counts_, &seg_countsb_};
    for (int x = 0; x < 2; x++) {
      if (segs[x]->size() < kk) segs[x]->resize(kk);
      if (grids[x]->size() < kk) grids[x]->resize(kk);
      cnts[x]->alloc(kk);
    }
    Engine& r = replayer();
    auto replay = [&](int sg) {
      const int x = sg & 1, s0 = sg * k, s1 = std::min(ns, s0 + k);
      r.replay_segment(ckpts_[(size_t)sg], tape_dt_.data() + s0, s1 - s0, segs[x]->data(), grids[x]->data(),
                       cnts[x]->p, seg_free_[x], seg_done_[x]);
    };
    replay(nseg - 1);
    const ParticleBufs<T, D> live = cur_;
    for (int sg = nseg - 1; sg >= 0; sg--) {
      const int x = sg & 1, s0 = sg * k, s1 = std::min(ns, s0 + k);
      HIPCHECK(hipStreamWaitEvent(stream_, seg_done_[x], 0));
      force_rebin_ = true;  // bins of this engine, from the segment's last saved state
      for (int j = s1 - 1; j >= s0; j--) {
        const size_t i = (size_t)(j - s0);
        T* q = (*segs[x])[i].s.p;
        cur_.X = q;
        cur_.V = q + n * D;
        cur_.F = q + 2 * n * D;
        cur_.B = q + n * (2 * D + D * D);
        cur_.valid = (*segs[x])[i].valid.p;
        fused_ = false;
        fused_last_ = false;
        backward_step((T)tape_dt_[(size_t)j], (*grids[x])[i], cnts[x]->p + i);
      }
      HIPCHECK(hipEventRecord(seg_free_[x], stream_));
      if (sg > 0) replay(sg - 1);  // runs beside the reverse kernels just queued
    }
    cur_.X = live.X;
    cur_.V = live.V;
    cur_.F = live.F;
    cur_.B = live.B;
    cur_.valid = live.valid;
  }

  // The replay engine, set up like this one (gravity, model, walls, rebin
  // policy, materials in this engine's slot order) in taping mode, so that a
  // re-bin sorts keys but moves no particle data.
  Engine& replayer() {
    if (!rep_) rep_.reset(new Engine(descs_.data(), (int)descs_.size()));
    Engine& r = *rep_;
    for (int c = 0; c < D; c++) r.grav_[c] = grav_[c];
    r.model_ = model_;
    r.thr_ = thr_;
    r.root_ = root_;
    r.nwalls_ = nwalls_;
    for (int w = 0; w < MAX_WALLS; w++) r.walls_[w] = walls_[w];
    r.rebin_fraction_ = rebin_fraction_;
    r.rebin_max_interval_ = rebin_max_interval_;
    if (r.n_ != n_ || !r.cur_.X) {
      r.n_ = n_;
      r.allocate_particles();
    }
    r.taping_ = true;
    r.replaying_ = true;
    r.fused_ = false;
    r.fused_last_ = false;
    HIPCHECK(hipEventRecord(rep_go_, stream_));  // checkpoints and materials are written on this stream
    HIPCHECK(hipStreamWaitEvent(r.stream_, rep_go_, 0));
    const size_t n = n_;
    HIPCHECK(hipMemcpyAsync(r.cur_.mass, cur_.mass, n * sizeof(T), hipMemcpyDeviceToDevice, r.stream_));
    HIPCHECK(hipMemcpyAsync(r.cur_.vol, cur_.vol, n * sizeof(T), hipMemcpyDeviceToDevice, r.stream_));
    HIPCHECK(hipMemcpyAsync(r.cur_.mu, cur_.mu, n * sizeof(T), hipMemcpyDeviceToDevice, r.stream_));
    HIPCHECK(hipMemcpyAsync(r.cur_.la, cur_.la, n * sizeof(T), hipMemcpyDeviceToDevice, r.stream_));
    if (cur_.pscene && r.cur_.pscene)
      HIPCHECK(hipMemcpyAsync(r.cur_.pscene, cur_.pscene, n * sizeof(uint16_t), hipMemcpyDeviceToDevice, r.stream_));
    return r;
  }

  // On the replay engine: after free_ev, the cnt steps from checkpoint ck,
  // each step's starting state, grids and touched tiles into seg, sg and
  // counts_out; done_ev marks the end.
  void replay_segment(const StateBuf& ck, const double* dts, int cnt, StateBuf* seg, StepGrid* sg, uint32_t* counts_out,
                      hipEvent_t free_ev, hipEvent_t done_ev) {
    HIPCHECK(hipStreamWaitEvent(stream_, free_ev, 0));
    load_state(ck);
    const size_t nodes = (size_t)ntotal_ * 64;
    for (int i = 0; i < cnt; i++) {
      save_state(seg[i]);
      sg[i].grid.alloc(nodes);
      sg[i].gridu.alloc(nodes);
      sg[i].tiles.alloc((size_t)ntotal_);
      grid_override_ = sg[i].grid.p;
      gridu_override_ = sg[i].gridu.p;
      step_once((T)dts[i]);
      save_tiles_kernel<<<copy_blocks(), BLOCK, 0, stream_>>>(gt_cur_, counts_.p, sg[i].tiles.p, counts_out + i);
      grid_override_ = nullptr;
      gridu_override_ = nullptr;
    }
    HIPCHECK(hipEventRecord(done_ev, stream_));
  }

  void stress(int64_t n, const void* F, const void* volume, const void* mu, const void* lambda, void* tau) override {
    if (n <= 0) return;
    if (n > 0x7fffffff) throw Error{MPMX_EINVAL, "too many matrices"};
    DevBuf<T> dF, dv, dm, dl, dt;
    dF.alloc(n * D * D); dv.alloc(n); dm.alloc(n); dl.alloc(n); dt.alloc(n * D * D);
    HIPCHECK(hipMemcpyAsync(dF.p, F, n * D * D * sizeof(T), hipMemcpyHostToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(dv.p, volume, n * sizeof(T), hipMemcpyHostToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(dm.p, mu, n * sizeof(T), hipMemcpyHostToDevice, stream_));
    HIPCHECK(hipMemcpyAsync(dl.p, lambda, n * sizeof(T), hipMemcpyHostToDevice, stream_));
    stress_kernel<T, D><<<cdiv(n, BLOCK), BLOCK, 0, stream_>>>((uint32_t)n, dF.p, dv.p, dm.p, dl.p, dt.p, model_,
                                                                thr_, root_);
    HIPCHECK(hipMemcpyAsync(tau, dt.p, n * D * D * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  void svd(int64_t n, const void* F, void* U, void* s, void* V) override {
    if (n <= 0) return;
    if (n > 0x7fffffff) throw Error{MPMX_EINVAL, "too many matrices"};
    DevBuf<T> dF, dU, dS, dV;
    dF.alloc(n * D * D); dU.alloc(n * D * D); dS.alloc(n * D); dV.alloc(n * D * D);
    HIPCHECK(hipMemcpyAsync(dF.p, F, n * D * D * sizeof(T), hipMemcpyHostToDevice, stream_));
    svd_kernel<T, D><<<cdiv(n, BLOCK), BLOCK, 0, stream_>>>((uint32_t)n, dF.p, dU.p, dS.p, dV.p);
    HIPCHECK(hipMemcpyAsync(U, dU.p, n * D * D * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipMemcpyAsync(s, dS.p, n * D * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipMemcpyAsync(V, dV.p, n * D * D * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  void sort_keys(int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out, uint32_t* order_out) override {
    if (n <= 0) return;
    if (n > 0x7fffffff) throw Error{MPMX_EINVAL, "too many keys"};
    if (key_bits < 1 || key_bits > 32) throw Error{MPMX_EINVAL, "key_bits must be in [1, 32]"};
    DevBuf<uint32_t> k, ko, vo, kt, vt;
    DevBuf<char> scratch;
    k.alloc(n); ko.alloc(n); vo.alloc(n); kt.alloc(n); vt.alloc(n);
    scratch.alloc(sort_scratch_bytes((uint32_t)n, key_bits));
    HIPCHECK(hipMemcpyAsync(k.p, keys, n * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
    radix_sort_pairs(k.p, ko.p, vo.p, kt.p, vt.p, (uint32_t)n, key_bits, scratch.p, stream_);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(keys_out, ko.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipMemcpyAsync(order_out, vo.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  static_assert(SORT_NOALLOC_MAX_KEYS == MPMX_SORT_NOALLOC_MAX_KEYS, "the header's limit is the sort's");
  // The same sort on device arrays, queued on the context stream, with its
  // scratch in the sort's static per-GPU array: no allocation, no wait.
  void sort_keys_noalloc(int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out,
                         uint32_t* order_out) override {
    if (n < 0) throw Error{MPMX_EINVAL, "negative key count"};
    if (n > (int64_t)SORT_NOALLOC_MAX_KEYS)
      throw Error{MPMX_EINVAL, "mpmx_sort_keys_noalloc takes at most " + std::to_string(SORT_NOALLOC_MAX_KEYS) +
                                   " keys (got " + std::to_string(n) + ")"};
    if (key_bits < 1 || key_bits > 32) throw Error{MPMX_EINVAL, "key_bits must be in [1, 32]"};
    radix_sort_pairs_noalloc(keys, keys_out, order_out, (uint32_t)n, key_bits, device_, stream_);
    HIPCHECK(hipGetLastError());
  }

 private:
  void allocate_particles() {
    const bool batch = scenes_.size() > 1;
    auto alloc_set = [&](ParticleBufs<T, D>& pb, std::vector<DevBuf<T>>& fb, DevBuf<uint8_t>& vb,
                         DevBuf<uint32_t>& ib, DevBuf<uint16_t>& sb) {
      fb.resize(8);
      const size_t n = std::max<uint32_t>(n_, 1);
      fb[0].alloc(n * D); fb[1].alloc(n * D); fb[2].alloc(n * D * D); fb[3].alloc(n * D * D);
      fb[4].alloc(n); fb[5].alloc(n); fb[6].alloc(n); fb[7].alloc(n);
      vb.alloc(n);
      ib.alloc(n);
      if (batch) sb.alloc(n);
      pb = ParticleBufs<T, D>{fb[0].p, fb[1].p, fb[2].p, fb[3].p, fb[4].p, fb[5].p, fb[6].p, fb[7].p,
                              vb.p,    ib.p,    batch ? sb.p : nullptr};
      if (plastic_) add_plastic_arrays(pb, fb, n);
    };
    // Buffers are sized exactly to n so that the SoA stride equals n.
    for (auto* fbv : {&fa_, &fb_})
      for (auto& b : *fbv) b.release();
    va_.release(); vb_.release(); ia_.release(); ib_.release(); sa_.release(); sb_.release();
    alloc_set(cur_, fa_, va_, ia_, sa_);
    alloc_set(alt_, fb_, vb_, ib_, sb_);
    const size_t n = std::max<uint32_t>(n_, 1);
    keys_.release(); skeys_.release(); order_.release(); ktmp_.release(); vtmp_.release();
    keys_.alloc(n); skeys_.alloc(n); order_.alloc(n); ktmp_.alloc(n); vtmp_.alloc(n);
    sort_scratch_.release();
    sort_scratch_.alloc(sort_scratch_bytes((uint32_t)n, key_bits_ + 6));
    stage_.release();
    stage_.alloc(n * D * D);
    slow_list_.release();
    slow_list_.alloc(n);
    mover_list_.release();
    mover_list_.alloc(n);
    mover_list2_.release();
    mover_list2_.alloc(n);
    mover_tile_.release();
    mover_tile_.alloc(n);
    mover_tile2_.release();
    mover_tile2_.alloc(n);
    ml_cur_ = mover_list_.p;
    mt_cur_ = mover_tile_.p;
    ml_next_ = mover_list2_.p;
    mt_next_ = mover_tile2_.p;
  }

  // X, V, F, B and valid of the current state, in slot order.
  void save_state(StateBuf& b) {
    const size_t n = n_;
    b.s.alloc(n * (2 * D + 2 * D * D));
    b.valid.alloc(n);
    T* q = b.s.p;
    copy_state_kernel<T, D><<<copy_blocks(), BLOCK, 0, stream_>>>(cur_.X, cur_.V, cur_.F, cur_.B, cur_.valid, q,
                                                                 q + n * D, q + 2 * n * D, q + n * (2 * D + D * D),
                                                                 b.valid.p, n_);
  }
  // Restores a saved state.  With rebin the positions are re-binned before the
  // next step; without, the bins stay and backward_step marks the movers.
  void load_state(const StateBuf& b, bool rebin_next = true) {
    const size_t n = n_;
    const T* q = b.s.p;
    copy_state_kernel<T, D><<<copy_blocks(), BLOCK, 0, stream_>>>(q, q + n * D, q + 2 * n * D, q + n * (2 * D + D * D),
                                                                 b.valid.p, cur_.X, cur_.V, cur_.F, cur_.B,
                                                                 cur_.valid, n_);
    if (rebin_next) force_rebin_ = true;
    fused_ = false;
    fused_last_ = false;
  }
  uint32_t copy_blocks() const { return std::max<uint32_t>(1, std::min<uint32_t>(cdiv(n_, BLOCK), 2048)); }
  // Host adjoint (user order, NULL = zero) -> device SoA in slot order.
  void upload_adjoint(const T* host, T* dev, int comps) {
    if (host)
      HIPCHECK(hipMemcpyAsync(stage_.p, host, (size_t)n_ * comps * sizeof(T), hipMemcpyHostToDevice, stream_));
    soa_from_aos_pid_kernel<T><<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(host ? stage_.p : nullptr, dev, cur_.pid, n_,
                                                                        comps);
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  void download_param(const double* dev, double* host) {
    if (!host) return;
    paramstage_.alloc(n_);
    aos_from_soa_kernel<double><<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(dev, paramstage_.p, cur_.pid, n_, 1);
    HIPCHECK(hipMemcpyAsync(host, paramstage_.p, (size_t)n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }

  // Grid tiles touched this step (binned tiles and this step's mover tiles,
  // each with its -octant neighbours) -> gtiles, counts[1]; their nodes in
  // z0 (and z1, z2 if set) are cleared.  Also clears counts[3], counts[4].
  void build_grid_tiles(uint32_t ms, Node4<T>* z0, Node4<T>* z1, Node4<T>* z2) {
    const uint32_t nb = cdiv(ntotal_, BLOCK);
    const Args<T, D> a = make_args(0);
    gcount_kernel<T, D, true><<<nb, BLOCK, 0, stream_>>>(a, tile_stamp_.p, bin_stamp_, ms, gblock_.p, 1, nullptr);
    glist_kernel<T, D, true><<<nb, BLOCK, 0, stream_>>>(a, tile_stamp_.p, bin_stamp_, ms, gblock_.p, gt_cur_,
                                                        counts_.p + 1, z0, z1, z2);
  }

  // Reverse of one step from the state it started from (in cur_) and the grid
  // the forward step built (velocities, (m, u) and touched tiles).
  void backward_step(T dt, const StepGrid& sgr, const uint32_t* seg_count) {
    if (rebin_due()) rebin();
    Args<T, D> a = make_args(dt);
    a.order = identity_order_ ? nullptr : order_.p;
    a.grid = sgr.grid.p;
    a.gtiles = sgr.tiles.p;
    restore_tiles_kernel<T><<<grid_blocks(), BLOCK, 0, stream_>>>(sgr.tiles.p, seg_count, counts_.p, agrid_.p);
    mark_movers_kernel<T, D, true><<<tile_blocks(), BLOCK, 0, stream_>>>(a, 0u);
    // Below ~512 tiles of 512 particles two workgroups of 256-particle chunks
    // share each tile: twice the workgroups in flight at the same total work.
    const uint32_t split = n_ < (size_t)512 * 512 ? 2u : 1u;
    AdjArgs<T> q{aX_.p, aV_.p, aF_.p, aB_.p, amu_.p, ala_.p, agrid_.p, sgr.gridu.p, split};
    const uint32_t pb = std::min<uint32_t>(cdiv(n_, BLOCK), 4096);
    const uint32_t tb = std::min<uint32_t>(tile_blocks() * split, 8192);
    if (split == 2)
      g2p_adjoint_kernel<T, D, 1><<<tb, BLOCK, 0, stream_>>>(a, q);
    else
      g2p_adjoint_kernel<T, D, 2><<<tb, BLOCK, 0, stream_>>>(a, q);
    g2p_adjoint_mover_kernel<T, D><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(a, q);
    grid_adjoint_kernel<T, D><<<grid_blocks(), BLOCK, 0, stream_>>>(a, q);
    p2g_adjoint_kernel<T, D><<<pb, BLOCK, 0, stream_>>>(a, q);
    HIPCHECK(hipGetLastError());
    step_count_++;
    steps_since_rebin_++;
  }

  void upload_soa(const T* host, T* dev, int comps) {
    HIPCHECK(hipMemcpyAsync(stage_.p, host, (size_t)n_ * comps * sizeof(T), hipMemcpyHostToDevice, stream_));
    soa_from_aos_kernel<T><<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(stage_.p, dev, n_, comps);
    // the staging buffer is reused: wait before the next upload overwrites it
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  void download_aos(const T* dev, T* host, int comps) {
    if (!host) return;
    aos_from_soa_kernel<T><<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(dev, stage_.p, cur_.pid, n_, comps);
    HIPCHECK(hipMemcpyAsync(host, stage_.p, (size_t)n_ * comps * sizeof(T), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }

  Args<T, D> make_args(T dt) {
    Args<T, D> a;
    std::memset(&a, 0, sizeof(a));  // padding too: graph_signature compares the bytes
    a.g = g_;
    const bool batch = scenes_.size() > 1;
    a.scenes = scenes_dev_.p;
    a.tscene = batch ? tscene_.p : nullptr;
    a.pscene = batch ? cur_.pscene : nullptr;
    a.ntotal = ntotal_;
    for (int c = 0; c < D; c++) a.grav[c] = grav_[c];
    a.dt = dt;
    a.model = model_;
    a.thr = thr_;
    a.root = root_;
    a.nwalls = nwalls_;
    for (int i = 0; i < nwalls_; i++) a.walls[i] = walls_[i];
    a.X = cur_.X; a.V = cur_.V; a.F = cur_.F; a.B = cur_.B;
    a.mass = cur_.mass; a.vol = cur_.vol; a.mu = cur_.mu; a.la = cur_.la;
    a.valid = cur_.valid;
    a.keys = keys_.p;
    a.n = n_;
    a.order = order_.p;
    a.tile_begin = tile_begin_.p;
    a.tile_end = tile_end_.p;
    a.ptiles = ptiles_.p;
    a.gtiles = gt_cur_;
    a.counts = counts_.p;
    a.slow_list = slow_list_.p;
    a.mover_list = ml_cur_;
    a.mover_tile = mt_cur_;
    a.mover_list2 = ml_next_;
    a.mover_tile2 = mt_next_;
    a.mcap = n_;
    a.grid_next = gother_;
    a.mstamp = mstamp_.p;
    a.grid = grid_override_ ? grid_override_ : gcur_;
    a.gridf = gridf_.p;
    a.gridu = gridu_override_;
    a.err = err_.p;
    a.band = refl_on_;
    a.reflect_sides = 0;
    if (refl_on_)
      for (int i = 0; i < 2 * D; i++)
        if (refl_side_[i] != RBC_FREE) a.reflect_sides |= 1u << i;
    return a;
  }

  bool any_friction() const {
    for (int i = 0; i < nwalls_; i++)
      if (walls_[i].friction != (T)0 && walls_[i].type != 0) return true;
    return false;
  }
  // workgroups of the tile kernels (grid-stride over the tile list): the
  // launch configuration's count, else one per tile up to 8192
  uint32_t tile_blocks() const {
    return tile_wg_ ? tile_wg_ : std::min<uint32_t>(std::max<uint32_t>(ntotal_, 1), 8192);
  }
  uint32_t grid_blocks() const { return (uint32_t)stats_blocks_; }

  // Rebuild the bins when forced, when max_interval steps have passed, or when
  // the mover count two steps back (read without stalling the stream) exceeds
  // mover_fraction * n.  Using a fixed lag keeps the decision independent of
  // host/GPU timing.
  bool rebin_due() {
    if (force_rebin_) return true;
    if (rebin_max_interval_ > 0 && steps_since_rebin_ >= rebin_max_interval_) return true;
    const int64_t want = step_count_ - 2;
    if (want < 0) return false;
    const Readback& r = rb_meta_[want & 3];
    if (r.step != want || r.gen != bin_stamp_) return false;
    HIPCHECK(hipEventSynchronize(rb_ev_[want & 3]));
    return (double)rb_host_[want & 3] > rebin_fraction_ * (double)n_;
  }

  void rebin() {
    hipEvent_t e0 = timer_.begin(stream_);
    Args<T, D> a = make_args(0);
    key_kernel<T, D><<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(a, err_.p + 1);
    radix_sort_pairs(keys_.p, skeys_.p, vtmp_.p, ktmp_.p, order_.p, n_, key_bits_ + 6, sort_scratch_.p, stream_);
    bin_stamp_++;
    bounds_kernel<<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(skeys_.p, n_, tile_begin_.p, tile_end_.p, tile_stamp_.p,
                                                          bin_stamp_);
    stratify_kernel<<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(skeys_.p, vtmp_.p, order_.p, tile_begin_.p, tile_end_.p,
                                                            n_);
    const uint32_t nb = cdiv(ntotal_, BLOCK);
    gcount_kernel<T, D, false><<<nb, BLOCK, 0, stream_>>>(a, tile_stamp_.p, bin_stamp_, 0u, gblock_.p, 0, nullptr);
    glist_kernel<T, D, false><<<nb, BLOCK, 0, stream_>>>(a, tile_stamp_.p, bin_stamp_, 0u, gblock_.p, ptiles_.p,
                                                         counts_.p + 0, nullptr, nullptr, nullptr);
    timer_.end(PH_REBIN, e0, stream_);
    identity_order_ = false;
    // storage stays in place while a tape or the autotuner holds saved states
    if (!taping_ && !tuning_ && (force_reorder_ || since_reorder_ >= reorder_interval_)) {
      hipEvent_t er = timer_.begin(stream_);
      reorder_kernel<T, D><<<cdiv(n_, BLOCK), BLOCK, 0, stream_>>>(cur_, alt_, order_.p, n_);
      user_order_ = false;
      std::swap(cur_, alt_);
      std::swap(fa_, fb_);
      std::swap(va_, vb_);
      std::swap(ia_, ib_);
      std::swap(sa_, sb_);
      identity_order_ = true;
      since_reorder_ = 0;
      force_reorder_ = false;
      timer_.end(PH_REORDER, er, stream_);
    }
    force_rebin_ = false;
    steps_since_rebin_ = 0;
    rebins_++;
  }

  // The unfused step as the named steps of MPM_DRIVER::Advance_One_Time_Step
  // (MPM_DRIVER.cpp:141-157), in order; step_once() runs them all, and
  // step_stage() one at a time so that a caller can fire its callbacks
  // around each (MPM_DRIVER::Step, :1380-1393).
  enum {
    ST_SIMULATED = 0,  // "simulated-particles": invalid particles carry INVALID_KEY
    ST_WEIGHTS,        // "update-weights": the kernels recompute the weights from X
    ST_SCATTER,        // "prepare-scatter": bins (rebuilt when due) and the grid-tile list
    ST_P2G,            // "p2g"
    ST_FORCES,         // "forces": grid update with gravity, collision objects, friction
    ST_REDUCE_DT,      // "reduce-dt": the strong and sound-speed CFL, if on
    ST_G2P,            // "g2p"
    ST_PLASTICITY,     // "plasticity-hardening", then the end of the step
    NSTAGES
  };
  struct OpenStep {
    Args<T, D> a;
    bool fric = false, empty = false;
    bool p2g_copy = false;  // vsave_ holds Particle_To_Grid's (m, m v) of this step
    T dt = 0, cfl = 0, min_dt = 0;
    hipEvent_t e_step = nullptr;
    uint32_t ms = 0;
  };
  OpenStep os_;
  int stage_next_ = 0;  // next stage of an open step run stage by stage (0: none open)

  void run_stage(int k) {
    OpenStep& s = os_;
    switch (k) {
      case ST_SIMULATED: {
        fused_ = false;
        fused_last_ = false;
        s.fric = any_friction();
        if (refl_on_ && refl_full_ && s.fric)
          throw Error{MPMX_ESTATE, "frictional reflection does not combine with collision objects that have friction"};
        if (s.fric) gridf_.alloc((size_t)ntotal_ * 64);
        s.e_step = timer_.begin(stream_);
        s.empty = n_ == 0;
        s.a.dt = s.dt;
        break;
      }
      case ST_WEIGHTS:
        break;
      case ST_SCATTER: {
        if (s.empty) break;
        if (rebin_due()) rebin();
        hipEvent_t e0 = timer_.begin(stream_);
        s.ms = (uint32_t)(step_count_ + 1);  // mover stamp of this step
        build_grid_tiles(s.ms, grid_override_ ? grid_override_ : gcur_, s.fric ? gridf_.p : nullptr, nullptr);
        timer_.end(PH_BIN, e0, stream_);
        s.a = make_args(s.dt);
        s.a.order = identity_order_ ? nullptr : order_.p;
        s.a.next_stamp = s.ms + 1;
        break;
      }
      case ST_P2G: {
        if (s.empty) break;
        const Args<T, D>& a = s.a;
        hipEvent_t e1 = timer_.begin(stream_);
        p2g_kernel<T, D><<<tile_blocks(), BLOCK, 0, stream_>>>(a);
        p2g_extra_kernel<T, D><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(a);
        // The grid above holds P2G's momentum plus the elastic impulse of the
        // forces step.  Grid statistics and a step run stage by stage need
        // Particle_To_Grid's own result (MPM_DRIVER.cpp:245-306, reflection
        // of m and m v included): the same P2G with dt = 0 into vsave_.
        s.p2g_copy = stats_on_ || stage_next_ != 0;
        if (s.p2g_copy) {
          p2g_momentum_copy(a);
          if (stats_on_) record_grid_totals(a, vsave_.p, 0);
        }
        timer_.end(PH_P2G, e1, stream_);
        if (!replaying_) {  // the mover count for the rebin policy; a reverse pass's replay keeps the host free
          const int slot = (int)(step_count_ & 3);
          HIPCHECK(hipMemcpyAsync(rb_host_ + slot, counts_.p + 4, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
          HIPCHECK(hipEventRecord(rb_ev_[slot], stream_));
          rb_meta_[slot] = Readback{step_count_, bin_stamp_};
        }
        break;
      }
      case ST_FORCES: {
        if (s.empty) break;
        const Args<T, D>& a = s.a;
        const T dt = s.dt;
        hipEvent_t e2 = timer_.begin(stream_);
        Args<T, D> ag = a;  // the grid update's arguments
        if (refl_on_ && !refl_full_) {
          // gravity into the P2G sums, then Reflect_Boundary_Mass_Momentum with
          // the force reflection folded in (the same linear map on m v and dt f)
          add_gravity_kernel<T><<<grid_blocks(), BLOCK, 0, stream_>>>(a.grid, a.gtiles, counts_.p + 1, dt * a.grav[0],
                                                                      dt * a.grav[1], D == 3 ? dt * a.grav[D - 1] : (T)0);
          for (int k2 = 0; k2 < D; k2++)
            reflect_mass_momentum_kernel<T, D><<<reflect_blocks(k2), BLOCK, 0, stream_>>>(a.grid, scenes_dev_.p,
                                                                                            reflect_axis(k2));
          for (int c = 0; c < D; c++) ag.grav[c] = 0;
        }
        if (s.fric)
          grid_kernel<T, D, true><<<grid_blocks(), BLOCK, 0, stream_>>>(ag);
        else
          grid_kernel<T, D, false><<<grid_blocks(), BLOCK, 0, stream_>>>(ag);
        if (refl_on_ && refl_full_) {  // Reflect_Boundary_Friction after the collision objects (MPM_DRIVER.cpp:999-1019)
          for (int k2 = 0; k2 < D; k2++)
            reflect_friction_kernel<T, D><<<reflect_blocks(k2), BLOCK, 0, stream_>>>(a.grid, scenes_dev_.p,
                                                                                       reflect_axis(k2), refl_fric_);
        } else if (refl_on_) {  // Reflect_Boundary_Velocity of velocity and velocity_friction_save (:1057-1059)
          for (int k2 = 0; k2 < D; k2++)
            reflect_velocity_kernel<T, D><<<reflect_blocks(k2), BLOCK, 0, stream_>>>(
                a.grid, s.fric ? a.gridf : nullptr, scenes_dev_.p, reflect_axis(k2));
        }
        if (stats_on_) record_grid_totals(a, a.grid, 1);
        timer_.end(PH_GRID, e2, stream_);
        break;
      }
      case ST_REDUCE_DT: {
        if (s.empty || !limits_on()) break;
        hipEvent_t e4 = timer_.begin(stream_);
        s.a.dt = reduce_dt(s.a, s.fric, s.cfl, s.min_dt);
        timer_.end(PH_DT, e4, stream_);
        break;
      }
      case ST_G2P: {
        if (s.empty) break;
        const Args<T, D>& a = s.a;
        hipEvent_t e3 = timer_.begin(stream_);
        if (plastic_) {  // the particles simulated in this step, before G2P invalidates any
          simmask_.alloc(n_);
          HIPCHECK(hipMemcpyAsync(simmask_.p, cur_.valid, n_, hipMemcpyDeviceToDevice, stream_));
        }
        if (s.fric) {
          g2p_kernel<T, D, true><<<tile_blocks(), BLOCK, 0, stream_>>>(a);
          g2p_mover_kernel<T, D, true><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(a);
        } else {
          g2p_kernel<T, D, false><<<tile_blocks(), BLOCK, 0, stream_>>>(a);
          g2p_mover_kernel<T, D, false><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(a);
        }
        timer_.end(PH_G2P, e3, stream_);
        break;
      }
      case ST_PLASTICITY: {
        if (s.empty) {
          time_ += (double)s.dt;
          last_dt_ = (double)s.dt;
          timer_.end(PH_STEP, s.e_step, stream_);
          break;
        }
        if (plastic_) {
          hipEvent_t e5 = timer_.begin(stream_);
          const uint32_t pb = std::min<uint32_t>(cdiv(n_, BLOCK), 4096);
          if (pl_.kind == 1)
            plasticity_kernel<T, D, T><<<pb, BLOCK, 0, stream_>>>(cur_, n_, simmask_.p, pl_);
          else
            plasticity_kernel<T, D, double><<<pb, BLOCK, 0, stream_>>>(cur_, n_, simmask_.p, pl_);
          timer_.end(PH_G2P, e5, stream_);
        }
        step_count_++;
        steps_since_rebin_++;
        since_reorder_++;
        if (taping_ && !replaying_) {
          tape_dt_.push_back((double)s.a.dt);
          if (tape_dt_.size() % (size_t)tape_k_ == 0) save_state(next_checkpoint());
        }
        HIPCHECK(hipGetLastError());
        has_grid_ = true;
        time_ += (double)s.a.dt;
        last_dt_ = (double)s.a.dt;
        timer_.end(PH_STEP, s.e_step, stream_);
        break;
      }
    }
  }

  // One unfused step; returns the dt it took (dt unless Reduce_Dt lowered it,
  // with cfl and min_dt as example.cfl and example.min_dt).
  T step_once(T dt) { return step_once(dt, lim_cfl_, lim_min_dt_); }
  T step_once(T dt, T cfl, T min_dt) {
    if (stage_next_ != 0) throw Error{MPMX_ESTATE, "a step run stage by stage is open: finish its stages first"};
    os_.dt = dt;
    os_.cfl = cfl;
    os_.min_dt = min_dt;
    for (int k = 0; k < NSTAGES; k++) run_stage(k);
    return os_.empty ? dt : os_.a.dt;
  }
  // One named step of an unfused step: stage 0 opens a step with dt (cfl and
  // min_dt of the dt limits), the stages follow in order, and the last one
  // closes it.  Returns after the stage's GPU work.
  void step_stage(double dt, int k) override {
    if (k < 0 || k >= NSTAGES) throw Error{MPMX_EINVAL, "stage must be in [0, 8)"};
    if (k != stage_next_)
      throw Error{MPMX_ESTATE, "stage " + std::to_string(k) + " out of order: the next stage is " +
                                   std::to_string(stage_next_)};
    if (k == 0) {
      if (!(dt > 0)) throw Error{MPMX_EINVAL, "dt must be positive"};
      os_.dt = (T)dt;
      os_.cfl = lim_cfl_;
      os_.min_dt = lim_min_dt_;
    }
    stage_next_ = (k + 1) % NSTAGES;
    try {
      run_stage(k);
      synchronize();
    } catch (...) {
      stage_next_ = 0;  // the step is abandoned
      throw;
    }
  }
  // The grid holds P2G's momentum between the "p2g" and "forces" stages.
  bool grid_holds_momentum() const { return stage_next_ == ST_FORCES; }

  bool limits_on() const { return lim_strong_ || lim_sound_; }
  ReflectAxis reflect_axis(int k) const {
    ReflectAxis r;
    r.axis = k;
    r.type[0] = refl_side_[2 * k];
    r.type[1] = refl_side_[2 * k + 1];
    return r;
  }
  // one thread per ghost node of axis k: 2 sides x 3 layers x the other axes' nodes
  uint32_t reflect_blocks(int k) const {
    uint64_t other = 1;
    for (int c = 0; c < D; c++)
      if (c != k) other *= (uint64_t)g_.G[c];
    return (uint32_t)cdiv(6 * other, (uint64_t)BLOCK);
  }
  // Reflection boundary conditions on the domain sides (-use_reflect,
  // -use_reflect_friction: MPM_EXAMPLE.h:106-109, :134; STANDARD_TESTS_BASE.cpp:487-505).
  void set_reflection(bool enable, bool full, double friction, const int* sides) override {
    if (taping_) throw Error{MPMX_ESTATE, "boundary conditions cannot change while a tape is recorded"};
    if (!enable) {
      refl_on_ = 0;
      fused_ = false;
      return;
    }
    if (scenes_.size() > 1) throw Error{MPMX_ESTATE, "reflection boundary conditions take a single-scene context"};
    if (!sides) throw Error{MPMX_EINVAL, "null side types"};
    if (!(friction >= 0)) throw Error{MPMX_EINVAL, "friction must be >= 0"};
    if (limits_on()) throw Error{MPMX_ESTATE, "reflection boundary conditions do not combine with the dt limits"};
    for (int c = 0; c < D; c++)
      if (g_.G[c] - 6 < 6) throw Error{MPMX_EINVAL, "reflection boundary conditions need at least 6 cells per axis"};
    for (int i = 0; i < 2 * D; i++)
      if (sides[i] < RBC_FREE || sides[i] > RBC_SEP) throw Error{MPMX_EINVAL, "side type must be 0..3"};
    for (int i = 0; i < 2 * D; i++) refl_side_[i] = sides[i];
    refl_full_ = full;
    refl_fric_ = (T)friction;
    refl_on_ = 1;
    fused_ = false;
  }
  // Max sound speed (out[0]) and strong-CFL factor (out[1]) of the current
  // particles against the grid of a.grid and vsave_; waits for the stream.
  void dt_limit_values(const Args<T, D>& a, int strong, int sound, T cfl, T cfl_F, T* out) {
    lim_.alloc(2);
    if (!lim_host_) HIPCHECK(hipHostMalloc((void**)&lim_host_, 2 * sizeof(Ubits), hipHostMallocDefault));
    dt_limit_init_kernel<T><<<1, 1, 0, stream_>>>(lim_.p);
    const uint32_t nb = std::min<uint32_t>(cdiv(n_, BLOCK), 4096);
    dt_limit_kernel<T, D><<<nb, BLOCK, 0, stream_>>>(a, vsave_.p, strong, sound, cfl, cfl_F, lim_.p);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(lim_host_, lim_.p, 2 * sizeof(Ubits), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    std::memcpy(out, lim_host_, sizeof(T));
    std::memcpy(out + 1, lim_host_ + 1, sizeof(T));
  }
  // Particle_To_Grid's (m, m v) of the step whose arguments are `a` into
  // vsave_ (the P2G again with dt = 0: no elastic impulse), with
  // Reflect_Boundary_Mass_Momentum applied as the reference applies it inside
  // Particle_To_Grid (MPM_DRIVER.cpp:276-280).  The P2G rebuilds the slow and
  // mover lists, so they start empty.
  void p2g_momentum_copy(const Args<T, D>& a) {
    vsave_.alloc((size_t)ntotal_ * 64);
    clear_tiles_kernel<T><<<grid_blocks(), BLOCK, 0, stream_>>>(vsave_.p, a.gtiles, counts_.p + 1);
    fill_u32_kernel<<<1, BLOCK, 0, stream_>>>(counts_.p + 3, 2u, 0u);
    Args<T, D> b = a;
    b.dt = 0;
    b.grid = vsave_.p;
    p2g_kernel<T, D><<<tile_blocks(), BLOCK, 0, stream_>>>(b);
    p2g_extra_kernel<T, D><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(b);
    if (refl_on_ && !refl_full_)
      for (int k2 = 0; k2 < D; k2++)
        reflect_mass_momentum_kernel<T, D><<<reflect_blocks(k2), BLOCK, 0, stream_>>>(vsave_.p, scenes_dev_.p,
                                                                                        reflect_axis(k2));
    HIPCHECK(hipGetLastError());
  }
  // MPM_DRIVER::Reduce_Dt (MPM_DRIVER.cpp:519-545) in T after the grid
  // update of a step with a.dt; returns the dt G2P takes and, when it is
  // lowered below 0.99 dt, rescales the valid nodes' velocities.
  T reduce_dt(const Args<T, D>& a, bool fric, T cfl, T min_dt) {
    vsave_.alloc((size_t)ntotal_ * 64);
    // velocity_save: the step's P2G again, with dt = 0 (no stress impulse); the
    // slow and mover lists are rebuilt by it, so they start empty
    clear_tiles_kernel<T><<<grid_blocks(), BLOCK, 0, stream_>>>(vsave_.p, a.gtiles, counts_.p + 1);
    fill_u32_kernel<<<1, BLOCK, 0, stream_>>>(counts_.p + 3, 2u, 0u);
    Args<T, D> b = a;
    b.dt = 0;
    b.grid = vsave_.p;
    p2g_kernel<T, D><<<tile_blocks(), BLOCK, 0, stream_>>>(b);
    p2g_extra_kernel<T, D><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(b);
    normalize_nodes_kernel<T><<<grid_blocks(), BLOCK, 0, stream_>>>(vsave_.p, a.gtiles, counts_.p + 1);
    T v[2];
    dt_limit_values(a, lim_strong_, lim_sound_, cfl, lim_cfl_F_, v);
    const T dt0 = a.dt;
    T dt = dt0;
    if (lim_strong_) {
      const T new_dt = dt0 * v[1];
      if (new_dt < dt) dt = new_dt;
    }
    T dts = dt0;
    if (lim_sound_) {
      T dxmin = a.g.dx[0];
      for (int c = 1; c < D; c++) dxmin = std::min(dxmin, a.g.dx[c]);
      dts = std::min(dts, robust_divide(dxmin, v[0]) * lim_cfl_sound_);
    }
    dt = std::min(dt, dts);
    dt = std::max(dt, min_dt);
    if (dt >= dt0 * (T)0.99) return dt0;
    const T sc = dt / dt0;
    rescale_nodes_kernel<T><<<grid_blocks(), BLOCK, 0, stream_>>>(a.grid, fric ? a.gridf : nullptr, vsave_.p, a.gtiles,
                                                                 counts_.p + 1, sc);
    HIPCHECK(hipGetLastError());
    return dt;
  }
  void set_dt_limits(bool strong, bool sound, double cfl, double cfl_F, double cfl_sound, double min_dt) override {
    if (!(cfl > 0) || !(cfl_F > 0) || !(cfl_sound > 0) || !(min_dt >= 0))
      throw Error{MPMX_EINVAL, "need cfl, cfl_F, cfl_sound > 0 and min_dt >= 0"};
    if ((strong || sound) && scenes_.size() > 1)
      throw Error{MPMX_ESTATE, "the strong and sound-speed CFL take a single-scene context"};
    if ((strong || sound) && refl_on_)
      throw Error{MPMX_ESTATE, "the dt limits do not combine with reflection boundary conditions"};
    if (taping_) throw Error{MPMX_ESTATE, "dt limits cannot change while a tape is recorded"};
    lim_strong_ = strong;
    lim_sound_ = sound;
    lim_cfl_ = (T)cfl;
    lim_cfl_F_ = (T)cfl_F;
    lim_cfl_sound_ = (T)cfl_sound;
    lim_min_dt_ = (T)min_dt;
    fused_ = false;
  }
  // MPM_DRIVER::Compute_Max_Sound_Speed (MPM_DRIVER.cpp:893-918) of the current particles.
  void max_sound_speed(double* out) override {
    if (scenes_.size() > 1) throw Error{MPMX_ESTATE, "the sound speed takes a single-scene context"};
    T v[2] = {0, 100};
    if (n_ > 0) {
      Args<T, D> a = make_args(0);
      dt_limit_values(a, 0, 1, 1, 1, v);
    }
    *out = (double)v[0];
  }
  double last_dt() const override { return last_dt_; }

  // P2G of the current state into gcur_, starting the fused pipeline: both
  // grids are cleared in full (the fused step clears only the tiles it used).
  void prime_fused(T dt, bool fric) {
    const size_t bytes = (size_t)ntotal_ * 64 * sizeof(Node4<T>);
    HIPCHECK(hipMemsetAsync(gcur_, 0, bytes, stream_));
    HIPCHECK(hipMemsetAsync(gother_, 0, bytes, stream_));
    HIPCHECK(hipMemsetAsync(lastcnt_.p, 0, 8 * sizeof(uint32_t), stream_));
    if (rebin_due()) rebin();
    hipEvent_t e0 = timer_.begin(stream_);
    const uint32_t ms = (uint32_t)(step_count_ + 1);
    build_grid_tiles(ms, gcur_, fric ? gridf_.p : nullptr, nullptr);
    timer_.end(PH_BIN, e0, stream_);
    Args<T, D> a = make_args(dt);
    a.order = identity_order_ ? nullptr : order_.p;
    a.next_stamp = ms + 1;
    hipEvent_t e1 = timer_.begin(stream_);
    p2g_kernel<T, D><<<tile_blocks(), BLOCK, 0, stream_>>>(a);
    p2g_extra_kernel<T, D><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(a);
    fused_counts_kernel<<<1, 64, 0, stream_>>>(counts_.p, 2);
    timer_.end(PH_P2G, e1, stream_);
    HIPCHECK(hipGetLastError());
    fused_ = true;
    fused_dt_ = dt;
  }

  // One step with the G2P of this step and the P2G of the next one fused
  // (g2p2g_kernel).  The P2G uses this step's dt; a later step with another
  // dt, or any step through step_once, starts the pipeline again.
  void step_fused(T dt) {
    const bool fric = any_friction();
    if (fric) gridf_.alloc((size_t)ntotal_ * 64);
    hipEvent_t e_step = timer_.begin(stream_);
    if (!fused_ || fused_dt_ != dt) prime_fused(dt, fric);
    if (rebin_due()) {
      rebin();
      fused_counts_kernel<<<1, 64, 0, stream_>>>(counts_.p, 0);
    }
    fused_body(dt, fric, false);
    timer_.end(PH_STEP, e_step, stream_);
  }

  // The GPU work of one fused step on the stream, then its host-side advance.
  // While the stream is captured into a graph (capturing) no event is
  // recorded; the replay records the readback events itself.
  void fused_body(T dt, bool fric, bool capturing) {
    Args<T, D> a = make_args(dt);
    a.order = identity_order_ ? nullptr : order_.p;
    const uint32_t ms_next = (uint32_t)(step_count_ + 2);  // mover stamp of the next step's grid
    a.next_stamp = ms_next;

    hipEvent_t e2 = timer_.begin(stream_);
    {
      Args<T, D> ag = a;  // the update also zeroes the other grid's tiles for the fused kernel's P2G
      ag.clear_grid = gother_;
      ag.clear_tiles = gt_other_;
      ag.clear_count = lastcnt_.p + 1;
      if (fric)
        grid_kernel<T, D, true><<<grid_blocks(), BLOCK, 0, stream_>>>(ag);
      else
        grid_kernel<T, D, false><<<grid_blocks(), BLOCK, 0, stream_>>>(ag);
    }
    timer_.end(PH_GRID, e2, stream_);

    hipEvent_t e3 = timer_.begin(stream_);
    hipEvent_t ek = timer_.begin_kernel(stream_);  // the fused kernel alone
    launch_fused_kernel(a, fric);
    timer_.end(PH_FUSED, ek, stream_);
    if (fric)
      fused_side_kernel<T, D, true><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(a);
    else
      fused_side_kernel<T, D, false><<<EXTRA_BLOCKS, BLOCK, 0, stream_>>>(a);
    timer_.end(PH_G2P2G, e3, stream_);

    // tile list of the next step's grid (now in gother_): binned tiles and
    // the tiles marked by this step's movers; the friction grid is cleared
    // over it.  gcount_kernel also does the step's list bookkeeping.
    hipEvent_t e0 = timer_.begin(stream_);
    const uint32_t nb = cdiv(ntotal_, BLOCK);
    gcount_kernel<T, D, true><<<nb, BLOCK, 0, stream_>>>(a, tile_stamp_.p, bin_stamp_, ms_next, gblock_.p, 2,
                                                         lastcnt_.p);
    glist_kernel<T, D, true><<<nb, BLOCK, 0, stream_>>>(a, tile_stamp_.p, bin_stamp_, ms_next, gblock_.p, gt_other_,
                                                        counts_.p + 1, fric ? gridf_.p : nullptr, nullptr, nullptr);
    timer_.end(PH_BIN, e0, stream_);
    // the mover count of this step (counts[9], set by the bookkeeping) for
    // the host's rebin policy, read without stalling the stream
    const int slot = (int)(step_count_ & 3);
    HIPCHECK(hipMemcpyAsync(rb_host_ + slot, counts_.p + 9, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    if (!capturing) {
      HIPCHECK(hipEventRecord(rb_ev_[slot], stream_));
      rb_meta_[slot] = Readback{step_count_, bin_stamp_};
    }
    HIPCHECK(hipGetLastError());
    fused_advance(dt);
  }
  // Host-side state of one fused step: the two grids, their tile lists and
  // the mover lists swap roles; counters and time advance.
  void fused_advance(T dt) {
    std::swap(gcur_, gother_);
    std::swap(gt_cur_, gt_other_);
    std::swap(ml_cur_, ml_next_);
    std::swap(mt_cur_, mt_next_);
    step_count_++;
    steps_since_rebin_++;
    since_reorder_++;
    if (taping_) {
      tape_dt_.push_back((double)dt);
      if (tape_dt_.size() % (size_t)tape_k_ == 0) {
        save_state(next_checkpoint());
      }
    }
    has_grid_ = true;
    fused_last_ = true;
    time_ += (double)dt;
    last_dt_ = (double)dt;
  }

  // ---------------------------------------------------------------- graphs --
  // Fused steps replayed four at a time from one captured hipGraph (the
  // launch configuration's `graph`): the four steps bring the swapped grids and
  // lists back to their places and fill the four readback slots, so one graph
  // serves every group that starts at a step count divisible by four with the
  // same arguments.  The graph's mover stamps repeat every four steps: a tile
  // marked four steps earlier and not since may join a grid-tile list again,
  // which adds a zero tile (cleared, skipped by the update) and never drops
  // one.  Bins are rebuilt between groups when due.
  std::vector<unsigned char> graph_signature(T dt, bool fric) {
    Args<T, D> a = make_args(dt);
    a.order = identity_order_ ? nullptr : order_.p;
    a.next_stamp = 0;
    std::vector<unsigned char> sig(sizeof(a) + 64, 0);
    std::memcpy(sig.data(), &a, sizeof(a));
    unsigned char* q = sig.data() + sizeof(a);
    const uint64_t extra[7] = {bin_stamp_, (uint64_t)fric, tile_blocks(), (uint64_t)(uintptr_t)gother_,
                               (uint64_t)(uintptr_t)gt_other_, (uint64_t)(uintptr_t)gridf_.p, (uint64_t)fused_threads_};
    std::memcpy(q, extra, sizeof(extra));
    return sig;
  }
  void release_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_sig_.clear();
  }
  // nsteps fused steps, groups of four from the graph, the rest one by one
  void step_graph(T dt, int nsteps) {
    while (nsteps > 0) {
      if (!fused_ || fused_dt_ != dt || (step_count_ & 3) != 0 || nsteps < 4) {
        step_fused(dt);
        nsteps--;
        continue;
      }
      const bool fric = any_friction();
      if (fric) gridf_.alloc((size_t)ntotal_ * 64);
      if (rebin_due()) {
        rebin();
        fused_counts_kernel<<<1, 64, 0, stream_>>>(counts_.p, 0);
      }
      std::vector<unsigned char> sig = graph_signature(dt, fric);
      const int64_t s0 = step_count_;
      if (!graph_exec_ || sig != graph_sig_) {
        release_graph();
        HIPCHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
        try {
          for (int k = 0; k < 4; k++) fused_body(dt, fric, true);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(stream_, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHECK(hipStreamEndCapture(stream_, &graph_));
        HIPCHECK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        graph_sig_ = sig;
      } else {
        for (int k = 0; k < 4; k++) fused_advance(dt);
      }
      HIPCHECK(hipGraphLaunch(graph_exec_, stream_));
      for (int k = 0; k < 4; k++) {
        const int64_t st = s0 + k;
        HIPCHECK(hipEventRecord(rb_ev_[st & 3], stream_));
        rb_meta_[st & 3] = Readback{st, bin_stamp_};
      }
      graph_groups_++;
      nsteps -= 4;
    }
  }

  Geom<T, D> g_{};
  std::vector<Geom<T, D>> scenes_;
  DevBuf<Geom<T, D>> scenes_dev_;
  DevBuf<uint16_t> tscene_;
  uint32_t ntotal_ = 0;
  T grav_[D];
  int model_ = 0;
  T thr_ = (T)0.25, root_ = (T)(D == 2 ? 0.5 : 0.6299605249474366);
  int nwalls_ = 0;
  WallDev<T> walls_[MAX_WALLS];
  int key_bits_ = 1;
  int device_ = 0;  // the GPU current at creation: the stream's
  hipStream_t stream_ = nullptr;
  uint32_t n_ = 0;
  uint32_t stamp_ = 0;
  int since_reorder_ = 0;
  int reorder_interval_ = 16;
  bool force_reorder_ = true;
  bool force_rebin_ = true;
  bool identity_order_ = false;
  bool taping_ = false, replaying_ = false;
  int tape_k_ = 1;
  int64_t tape_valid_ = 0;
  std::vector<double> tape_dt_;
  // Tape storage, kept across tapes so that a training loop allocates once:
  // checkpoints (the first nckpt_ are live) and, for the reverse pass, the
  // states and grids of one segment.
  std::vector<StateBuf> ckpts_;
  size_t nckpt_ = 0;
  std::vector<StateBuf> seg_;
  std::vector<StepGrid> sgrid_;
  StateBuf& next_checkpoint() {
    if (nckpt_ == ckpts_.size()) ckpts_.emplace_back();
    return ckpts_[nckpt_++];
  }
  DevBuf<T> aX_, aV_, aF_, aB_;
  DevBuf<double> amu_, ala_, paramstage_;
  DevBuf<Node4<T>> agrid_;
  DevBuf<uint32_t> seg_counts_;
  // Second storage set and the replay engine of the overlapped reverse pass.
  std::vector<StateBuf> segb_;
  std::vector<StepGrid> sgridb_;
  DevBuf<uint32_t> seg_countsb_;
  std::vector<mpmx_grid_desc> descs_;
  std::unique_ptr<Engine> rep_;
  hipEvent_t seg_done_[2] = {}, seg_free_[2] = {}, rep_go_ = nullptr;
  Node4<T>* grid_override_ = nullptr;  // the grid arrays of a recomputed step
  Node4<T>* gridu_override_ = nullptr;
  uint32_t bin_stamp_ = 0;
  int64_t step_count_ = 0, rebins_ = 0;
  int steps_since_rebin_ = 0;
  double rebin_fraction_ = 1.0 / 128;
  int rebin_max_interval_ = 0;
  struct Readback {
    int64_t step;
    uint32_t gen;
  };
  Readback rb_meta_[4];
  hipEvent_t rb_ev_[4] = {};
  uint32_t* rb_host_ = nullptr;
  using Ubits = typename OrderedBits<T>::U;
  DevBuf<Ubits> vbound_;      // per-scene speed bound (ordered bit patterns)
  Ubits* vb_host_ = nullptr;  // pinned readback of vbound_
  // Reduce_Dt settings (example.use_strong_cfl, use_sound_speed_cfl, cfl,
  // cfl_F, cfl_sound, min_dt; MPM_EXAMPLE.h:95-104) and its buffers
  int lim_strong_ = 0, lim_sound_ = 0;
  T lim_cfl_ = 1, lim_cfl_F_ = (T)0.5, lim_cfl_sound_ = (T)0.9, lim_min_dt_ = 0;
  DevBuf<Node4<T>> vsave_;
  DevBuf<Ubits> lim_;
  Ubits* lim_host_ = nullptr;
  double last_dt_ = 0;
  bool has_grid_ = false;
  bool stats_on_ = false;
  // launch configuration: fused steps replayed from a captured graph, and the
  // tile kernels' workgroup count (0: one per tile up to 8192)
  bool use_graph_ = false;
  int fused_threads_ = 128;  // workgroup size of the fused tile kernel (128 or 256)
  bool tuning_ = false;  // mpmx_autotune is timing candidates
  uint32_t tile_wg_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<unsigned char> graph_sig_;
  int64_t graph_groups_ = 0;
  double time_ = 0;
  int stats_blocks_ = 1;
  ParticleBufs<T, D> cur_{}, alt_{};
  std::vector<DevBuf<T>> fa_, fb_;
  DevBuf<uint8_t> va_, vb_, stage_bytes_;
  DevBuf<uint32_t> ia_, ib_;
  DevBuf<uint16_t> sa_, sb_;
  DevBuf<uint32_t> keys_, skeys_, order_, ktmp_, vtmp_;
  DevBuf<char> sort_scratch_;
  DevBuf<uint32_t> tile_begin_, tile_end_, tile_stamp_, mstamp_, gblock_, ptiles_, gtiles_, counts_, err_;
  DevBuf<Node4<T>> grid_, grid2_, gridf_;
  DevBuf<uint32_t> gtiles2_, lastcnt_, mover_list2_, mover_tile2_;
  // Fused steps alternate two grids: gcur_ holds the P2G of the current state
  // (tile list gt_cur_, count counts[1]) and gother_ the last step's grid
  // (gt_other_, count lastcnt_[1]), cleared before it takes the next P2G.
  Node4<T>* gcur_ = nullptr;
  Node4<T>* gother_ = nullptr;
  uint32_t* gt_cur_ = nullptr;
  uint32_t* gt_other_ = nullptr;
  uint32_t *ml_cur_ = nullptr, *mt_cur_ = nullptr, *ml_next_ = nullptr, *mt_next_ = nullptr;
  int plastic_ = 0;  // plasticity after every G2P (unfused steps): 1 clamp, 2 Drucker-Prager
  bool user_order_ = true;  // particle slots are in user order (no reorder since set_particles)
  // reflection boundary conditions (example.reflection_bc, use_full_reflection,
  // reflection_bc_friction, side_bc_type)
  int refl_on_ = 0, refl_full_ = 0;
  T refl_fric_ = 0;
  int refl_side_[6] = {0, 0, 0, 0, 0, 0};
  PlasticDev<T> pl_{};
  DevBuf<uint8_t> simmask_;
  bool fused_ = false;       // gcur_ holds the P2G of the current state with dt fused_dt_
  bool fused_last_ = false;  // gother_ holds the last step's grid (for mpmx_get_grid)
  T fused_dt_ = 0;
  DevBuf<double> stats_partial_, stats_out_;  // grid totals: block partials, [after P2G, after update]
  DevBuf<double> ptot_partial_;                // particle totals: block partials, then the result
  DevBuf<T> stage_;
  DevBuf<uint32_t> slow_list_, mover_list_, mover_tile_;
  PhaseTimer timer_;
};

EngineBase* make_engine(const mpmx_grid_desc* d, int nscenes) {
  if (!d || nscenes < 1) throw Error{MPMX_EINVAL, "no grid description"};
  if (d[0].dim == 3 && d[0].precision == MPMX_FP32) return new Engine<float, 3>(d, nscenes);
  if (d[0].dim == 3 && d[0].precision == MPMX_FP64) return new Engine<double, 3>(d, nscenes);
  if (d[0].dim == 2 && d[0].precision == MPMX_FP32) return new Engine<float, 2>(d, nscenes);
  if (d[0].dim == 2 && d[0].precision == MPMX_FP64) return new Engine<double, 2>(d, nscenes);
  throw Error{MPMX_EINVAL, "dim must be 2 or 3 and precision MPMX_FP32 or MPMX_FP64"};
}

}  // namespace mpmx
