// DFSPH substep on gfx950 (src/fluid/dfsph.rs:79-103): density and factors,
// the divergence and density Jacobi solves as one persistent launch each, viscosity with gravity, and the advance.
//
// Neighbours: the density pass scans the 27 cells once per substep and
// records, for every particle, which particles of each of the 27 cells are
// its neighbours, as one bit mask per cell (16 or 32 bits; bit b is the
// cell's b-th particle in id order). Positions do not move until the
// advance, so every later pass of the substep walks the masks in the
// reference's order (cells dz, dy, dx; ids ascending) and rebuilds each
// pair's kernel gradient as (x_i - x_j) * f with f = slope(d)/d recomputed
// by the same expression: the same IEEE values the reference caches per pair
// (dfsph.rs:24-41, 108-136), for 54 mask bytes per particle instead of 8 list
// bytes per neighbour. Walls keep explicit lists of (wall, f). A cell with
// more particles than its mask has bits moves the world to 32-bit masks,
// and past that to explicit lists of every neighbour (the list format).
//
// Stop test: the reference averages the per-particle errors with an f32 left
// fold in particle order (dfsph.rs:263-267), which no parallel sum can
// reproduce. The solve instead sums them in FP64 with a fixed shape that the
// CPU oracle replays exactly, over the solve tiles of the fluid cell box
// (4 x 2 cell columns by tz cells in z, tz = ceil(sz / ceil(sz / 16)), numbered
// x-major, then y, then z). A tile's particles in particle order are its
// processed order q = 0, 1, ...: thread t of 1024 adds the errors of
// q = t, t + 1024, ... in order from 0.0; each wave's 64 threads fold by
// halving (s_l += s_{l+h}, h = 32..1), then the 16 wave sums (h = 8..1): the
// tile sum. Lane l of 64 adds tile sums l, l + 64, ... in order; the lanes
// fold by halving (h = 32..1). The tiled rate pass forms each tile's sum as
// it goes, so after the pass's barrier every wave folds the tile sums itself
// and all blocks take the same decision with no further barrier.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "device.hpp"
#include "launch.hpp"

namespace sfl {

constexpr uint32_t kChunk = 512;          // particles of one chunk of an untiled pass
constexpr uint32_t kSolveThreads = 1024;  // threads of a solve block: two chunks at a time
constexpr uint32_t kSolveWaves = kSolveThreads / 64;

// List slots are read with raw buffer loads: a buffer descriptor based at
// row k of the column-major array sits in scalar registers, rows k + 1 ..
// k + 3 are scalar offsets from it, and the lane's offset is 32-bit, so no
// load needs 64-bit vector address arithmetic. The arrays carry
// kListSlack rows past their capacity, so loads ahead of a list's end stay
// unconditional (a load under a branch makes the compiler wait for every
// load in flight at the join) and in bounds; their values are discarded.
// The entries stream through once per pass and load non-temporally,
// keeping them from evicting the data the gathers reuse. Word 3 of the
// descriptor: a raw buffer of 32-bit format, as the gfx9 family reads it.
constexpr int kRawBufferWord3 = 0x00020000;
constexpr int kStreamLoad = 2;  // cache policy bits: non-temporal

__device__ __forceinline__ __amdgpu_buffer_rsrc_t list_rows(const void* base, uint32_t elem, const StepParams& prm,
                                                            uint32_t k) {
  const size_t row = (size_t)k * prm.pair_stride * elem;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(static_cast<const char*>(base)) + row, 0, -1,
                                           kRawBufferWord3);
}

// list slots k .. k + 3 of particle i: (index, bits of the factor)
__device__ __forceinline__ void list_entries(const StepParams& prm, uint32_t k, uint32_t i, uint2 (&e)[4]) {
  const __amdgpu_buffer_rsrc_t rows = list_rows(prm.pairs, 8u, prm, k);
  const uint32_t step = prm.pair_stride * 8u;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    auto v = __builtin_amdgcn_raw_buffer_load_b64(rows, i * 8u, u * step, kStreamLoad);
    e[u] = make_uint2(v[0], v[1]);
  }
}

// Data that other blocks of the persistent solve write during the launch
// (velocities, stiffnesses, tile sums) is read by the tiled kernel with
// device-scope (sc1) buffer loads. Each XCD keeps its own L2: a plain load
// can return a line this XCD's L2 still holds from before another XCD wrote
// it back, which the barrier's agent-scope acquire would otherwise have to
// drop by invalidating the whole L2 under every block of the XCD. A
// device-scope load sees every store written back before the barrier and
// leaves the L2 alone.
constexpr int kCoherentLoad = 16;  // cache policy bits: device scope
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t whole(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, kRawBufferWord3);
}

__device__ __forceinline__ float4 coherent4(const float4* base, uint32_t i) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(whole(base), i * 16u, 0, kCoherentLoad);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

__device__ __forceinline__ double coherent_f64(const double* base, uint32_t i) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(whole(base), i * 8u, 0, kCoherentLoad);
  return __hiloint2double((int)v[1], (int)v[0]);
}

__device__ __forceinline__ uint32_t coherent_u32(const uint32_t* base, uint32_t i) {
  return __builtin_amdgcn_raw_buffer_load_b32(whole(base), i * 4u, 0, kCoherentLoad);
}

// A velocity or stiffness of particle i: device-scope in the tiled kernel
// (COHERENT), a plain load in the untiled ones, whose barrier acquires.
template <bool COHERENT>
__device__ __forceinline__ float4 shared4(const float4* base, uint32_t i) {
  if constexpr (COHERENT)
    return coherent4(base, i);
  else
    return base[i];
}

__device__ __forceinline__ bool list_aborted(const Control* ctl) {
  return ctl->abort != 0 || (ctl->flags & (kFlagPairs | kFlagWide)) != 0;
}

template <int FORMAT>
struct MaskOf {
  using T = uint16_t;
  static constexpr uint32_t bits = 16;
};
template <>
struct MaskOf<kMasks32> {
  using T = uint32_t;
  static constexpr uint32_t bits = 32;
};

// f = slope(d)/d of a pair with |o|^2 = sq, 0 at d = 0: the expression the
// density pass evaluates, so the same bits.
__device__ __forceinline__ float pair_factor(float sq, const StepParams& prm) {
  float distance = sqrtf(sq);
  return distance <= 0.0f ? 0.0f : kernel_slope(distance, prm.radius, prm.rinv, prm.scale) / distance;
}

// Walks particle i's fluid neighbours from its masks in the reference's
// order, four at a time: body(j[4], ok[4]) gets the next up to four
// neighbours (slots past the end point at i itself and are not ok). The
// masks and cell starts of one dz plane (9 cells) are loaded together.
template <int FORMAT, class Body>
__device__ __forceinline__ void mask_walk(const StepParams& prm, uint32_t i, uint32_t own, int64_t sysz, int64_t sz,
                                          Body&& body) {
  using Mask = typename MaskOf<FORMAT>::T;
  const Mask* __restrict__ masks = static_cast<const Mask*>(prm.masks) + i;
  const size_t stride = prm.pair_stride;
  const uint32_t* __restrict__ start = prm.cell_start;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
    Mask mk[9];
    uint32_t cs[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) mk[q] = masks[(size_t)((dz + 1) * 9 + q) * stride];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      int64_t dy = q / 3 - 1, dx = q % 3 - 1;
      cs[q] = mk[q] ? start[(int64_t)own + dx * sysz + dy * sz + dz] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      uint32_t m = mk[q];
      while (m) {
        uint32_t j[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          ok[u] = m != 0u;
          j[u] = ok[u] ? cs[q] + (uint32_t)__builtin_ctz(m) : i;
          m = ok[u] ? m & (m - 1u) : m;
        }
        body(j, ok);
      }
    }
  }
}

// Visits particle i's list: fluid(j, f) for its fluid neighbours, then
// wall(w, f) for its walls, in the reference's order.
template <class Fluid, class Wall>
__device__ __forceinline__ void visit_list(const StepParams& prm, uint32_t i, Fluid&& fluid, Wall&& wall) {
  uint32_t c = prm.pair_counts[i];
  uint32_t nf = c & 0xffffu, nw = c >> 16;
  const uint2* p = prm.pairs + i;
  const uint32_t stride = prm.pair_stride;
  uint32_t k = 0;
  for (; k < nf; ++k) {
    uint2 e = p[(size_t)k * stride];
    fluid(list_index(e.x), __uint_as_float(e.y));
  }
  for (; k < nf + nw; ++k) {
    uint2 e = p[(size_t)k * stride];
    wall(e.x, __uint_as_float(e.y));
  }
}

// Compact entries k = 4g .. 4g + 3 of particle i form group g: near[(g * N +
// i) * 4 + k % 4] and factor[(g * N + i) * 4 + k % 4], so a lane moves a
// group as one 8-byte and one 16-byte access.
__device__ __forceinline__ void store_group(const StepParams& prm, uint32_t i, uint32_t g, unsigned long long near,
                                            float4 factor) {
  const size_t at = (size_t)g * prm.pair_stride + i;
  reinterpret_cast<unsigned long long*>(prm.near)[at] = near;
  reinterpret_cast<float4*>(prm.factor)[at] = factor;
}

// compute_densities (src/fluid/mod.rs:621-675) fused with gradients + factors
// (dfsph.rs:108-163), the neighbour count the divergence solve uses
// (dfsph.rs:249-255) and the substep's neighbours (masks or lists). rho goes
// to V.w; alpha to `alpha` with its sign bit set for particles with fewer
// than 20 fluid + wall neighbours. A particle with more listed neighbours
// than the lists hold, or a neighbour past its cell mask's width, flags the
// world; later launches of the substep then do nothing and the host regrows
// or widens. KEYED: the walls' table is searched.
template <uint32_t THREADS>
__device__ void flow_tiles_block(const StepParams& prm, const FlowSetup& f);

template <int FORMAT, bool KEYED>
__global__ __launch_bounds__(256, 8) void k_dfsph_density(const float4* __restrict__ P, float4* __restrict__ V,
                                                       float* __restrict__ alpha, StepParams prm, FlowSetup setup) {
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  // the flow solve's tile set-up: the launch's logical last block, one past
  // the particles' blocks
  if (setup.tile_done && xcd_block() == gridDim.x - 1) {
    flow_tiles_block<256>(prm, setup);
    return;
  }
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  bool live = i < prm.n;
  float comp = 0.0f;
  uint32_t need = 0;
  if (live) {
    int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
    uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
    float4 at = P[i];
    float sum = 0.0f, wsum = 0.0f, squares = 0.0f;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    uint32_t nf = 0, nw = 0;
    const uint32_t cap = prm.pair_cap;
    uint2* out = prm.pairs + i;
    const size_t stride = prm.pair_stride;
    unsigned long long gnear = 0;  // compact: the open group of four entries
    float4 gfac = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    auto fluid = [&](uint32_t j, float ox, float oy, float oz, float sq) {
      float distance = sqrtf(sq);
      sum += kernel_value(distance, prm.radius, prm.rinv, prm.scale);
      float f = distance <= 0.0f ? 0.0f : kernel_slope(distance, prm.radius, prm.rinv, prm.scale) / distance;
      float gx = ox * f, gy = oy * f, gz = oz * f;
      float qx = gx * prm.mass, qy = gy * prm.mass, qz = gz * prm.mass;
      sx = sx + qx;
      sy = sy + qy;
      sz = sz + qz;
      squares += dot3(qx, qy, qz, qx, qy, qz);
      if constexpr (FORMAT == kLists) {
        if (nf < cap) out[nf * stride] = make_uint2(j, __float_as_uint(f));
      } else if constexpr (FORMAT == kCompact) {
        // entries go out four at a time (8 + 16 bytes a lane), far fewer
        // partial-line writes than one 2-byte and one 4-byte store each
        gnear = gnear >> 16 | (unsigned long long)j << 48;
        gfac = make_float4(gfac.y, gfac.z, gfac.w, f);
        if ((nf & 3u) == 3u && nf < cap) store_group(prm, i, nf >> 2, gnear, gfac);
      }
      ++nf;
    };
    bool wide = false;
    if constexpr (FORMAT == kLists || FORMAT == kCompact) {
      // visit_near's walk, tagging each entry with its neighbour column
      int32_t cx = cell_coord(at.x, prm.radius), cy = cell_coord(at.y, prm.radius), cz = cell_coord(at.z, prm.radius);
      // compact: offsets count from z0 - 1 of the particle's solve tile
      int64_t za = 0;
      if constexpr (FORMAT == kCompact) {
        const uint32_t depth = tile_depth(span[2]);
        const uint32_t z0 = (uint32_t)((int64_t)cz - lo[2]) / depth * depth;
        za = z0 > 0 ? z0 - 1 : 0;
      }
      // a DFSPH world's water table is always dense
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int64_t c = cell_slot((int64_t)cx + dx, (int64_t)cy + dy, (int64_t)cz + dz, lo, span);
            if (c < 0) continue;
            const uint32_t code = FORMAT == kCompact ? (uint32_t)((dx + 1) * 4 + (dy + 1))
                                                     : (uint32_t)((dx + 1) * 3 + (dy + 1));
            uint32_t first = 0;
            if constexpr (FORMAT == kCompact) first = prm.cell_start[c - ((int64_t)cz + dz - lo[2]) + za];
            uint32_t s0 = prm.cell_start[c], e0 = prm.cell_start[c + 1];
            // four candidates' loads in flight at once, tested in order
            for (uint32_t j0 = s0; j0 < e0; j0 += 4) {
              float4 q[4];
#pragma unroll
              for (uint32_t u = 0; u < 4; ++u) q[u] = P[min(j0 + u, e0 - 1u)];
#pragma unroll
              for (uint32_t u = 0; u < 4; ++u) {
                if (j0 + u >= e0) break;
                const uint32_t j = j0 + u;
                float ox = at.x - q[u].x, oy = at.y - q[u].y, oz = at.z - q[u].z;
                float sq = dot3(ox, oy, oz, ox, oy, oz);
                if (sq < prm.reach) {
                  if constexpr (FORMAT == kLists) {
                    fluid(j | code << kListIndexBits, ox, oy, oz, sq);
                  } else {
                    const uint32_t off = j - first;
                    wide = wide || off > kNearMask;
                    fluid(code << kNearBits | (off & kNearMask), ox, oy, oz, sq);
                  }
                }
              }
            }
          }
      if (wide) atomicOr(&ctl->flags, kFlagWide);
      if constexpr (FORMAT == kCompact) {
        // the last, partial group, its entries shifted down to slot 0
        if ((nf & 3u) != 0u && nf <= cap) {
          for (uint32_t r = nf & 3u; r < 4u; ++r) {
            gnear >>= 16;
            gfac = make_float4(gfac.y, gfac.z, gfac.w, 0.0f);
          }
          store_group(prm, i, nf >> 2, gnear, gfac);
        }
      }
    } else {
      // visit_near's walk, recording each cell's neighbours as a mask
      using Mask = typename MaskOf<FORMAT>::T;
      Mask* masks = static_cast<Mask*>(prm.masks) + i;
      int32_t cx = cell_coord(at.x, prm.radius), cy = cell_coord(at.y, prm.radius), cz = cell_coord(at.z, prm.radius);
      prm.own_slot[i] = (uint32_t)cell_slot(cx, cy, cz, lo, span);
      int k = 0;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx, ++k) {
            int64_t c = cell_slot((int64_t)cx + dx, (int64_t)cy + dy, (int64_t)cz + dz, lo, span);
            uint32_t m = 0;
            if (c >= 0) {
              uint32_t s0 = prm.cell_start[c], e0 = prm.cell_start[c + 1];
              for (uint32_t j = s0; j < e0; ++j) {
                float4 pj = P[j];
                float ox = at.x - pj.x, oy = at.y - pj.y, oz = at.z - pj.z;
                float sq = dot3(ox, oy, oz, ox, oy, oz);
                if (sq < prm.reach) {
                  fluid(j, ox, oy, oz, sq);
                  uint32_t b = j - s0;
                  if (b < MaskOf<FORMAT>::bits)
                    m |= 1u << b;
                  else
                    wide = true;
                }
              }
            }
            masks[(size_t)k * stride] = (Mask)m;
          }
      if (wide) atomicOr(&ctl->flags, kFlagWide);
    }
    const uint32_t wall0 = FORMAT == kLists ? nf : 0u;
    if (prm.wall_cells)
      visit_walls<KEYED>(prm, at,
                 [&](uint32_t w, float4, float ox, float oy, float oz, float sq) {
                   float distance = sqrtf(sq);
                   wsum += kernel_value(distance, prm.radius, prm.rinv, prm.scale);
                   float f = distance <= 0.0f ? 0.0f : kernel_slope(distance, prm.radius, prm.rinv, prm.scale) / distance;
                   float gx = ox * f, gy = oy * f, gz = oz * f;
                   sx = sx + gx * prm.mass;
                   sy = sy + gy * prm.mass;
                   sz = sz + gz * prm.mass;
                   uint32_t k = wall0 + nw;
                   if (k < cap) out[k * stride] = make_uint2(w, __float_as_uint(f));
                   ++nw;
                 });
    uint32_t listed = FORMAT == kCompact ? max(nf, nw) : wall0 + nw;
    if (listed > cap) atomicOr(&ctl->flags, kFlagPairs);
    need = listed;
    prm.pair_counts[i] = nf | (nw << 16);
    float density = prm.mass * (sum + wsum);
    float total = dot3(sx, sy, sz, sx, sy, sz) + squares;
    float a = total > 0.0f ? 1.0f / total : 0.0f;
    alpha[i] = nf + nw < 20u ? -a : a;
    V[i].w = density;
    comp = density / prm.rest - 1.0f;
  }
  wave_max_to(&stripe_of(ctl).comp_bits, comp);
  for (int off = 32; off >= 1; off >>= 1) need = max(need, (uint32_t)__shfl_xor((int)need, off, 64));
  if ((threadIdx.x & 63) == 0 && need) atomicMax(&stripe_of(ctl).pairs_max, need);
}

// ---- persistent Jacobi solve ------------------------------------------------

// Wall-clock limit of one barrier wait: wall_clock64() counts at the
// device's constant 100 MHz reference clock.
constexpr uint64_t kBarrierTicks = 200000000ull;  // 2 s

// Grid-wide barrier of the persistent solve (every block resident, the grid
// a multiple of 8). Every wave waits until its own stores have reached the
// L2 (a block barrier alone does not); then thread 0 of the block writes
// its XCD's L2 back (one agent-scope release fence) and arrives at its XCD
// group's count (blockIdx % 8), and the last block of a group arrives at
// the group count; the last group bumps the generation. The arrivals need
// no release of their own: each block's stores were in memory before it
// arrived. The other blocks poll the generation with relaxed loads and,
// with `acquire`, take one agent-scope acquire when it moves (an acquire per
// poll would invalidate the L2 the blocks still working gather from). The
// tiled kernel passes acquire = false: it reads what other blocks wrote
// with device-scope loads (coherent4), so its L2 is never invalidated.
// 32-deep instead of 256-deep atomic chains, each word on its own cache
// line: 5.4 (7.1 with the acquire) instead of 29 microseconds per barrier
// over 256 blocks. Returns false if it waited past the limit, having
// flagged the world; the limit counts from the last change of `heartbeat`
// when one is given (a block that has no work in the flow solve waits for
// the whole solve, which moves the word at every stop decision).
__device__ bool grid_barrier(Control* ctl, bool acquire, const uint32_t* heartbeat = nullptr) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  __shared__ int ok_s;
  if (threadIdx.x == 0) {
    int ok = 1;
    const uint32_t group = blockIdx.x % kXcds, per_group = gridDim.x / kXcds;
    uint32_t* mine = &ctl->bar_xcd[group][0];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    uint32_t gen = __hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t arrived = __hip_atomic_fetch_add(mine, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool last = false;
    if (arrived == per_group - 1) {
      __hip_atomic_store(mine, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t groups = __hip_atomic_fetch_add(&ctl->bar_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = groups == kXcds - 1;
    }
    if (last) {
      __hip_atomic_store(&ctl->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&ctl->bar_gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      uint64_t start = wall_clock64();
      uint32_t beat = heartbeat ? __hip_atomic_load(heartbeat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
      while (__hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if (heartbeat) {
          const uint32_t now = __hip_atomic_load(heartbeat, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (now != beat) {
            beat = now;
            start = wall_clock64();
          }
        }
        if (wall_clock64() - start > kBarrierTicks) {
          ok = 0;
          atomicOr(&ctl->flags, kFlagBarrier);
          break;
        }
      }
    }
    if (acquire) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok_s = ok;
  }
  __syncthreads();
  return ok_s != 0;
}

struct SolveArgs {
  StepParams prm;
  const float4* P;      // positions, fixed during the solve
  float4* V;            // velocities (updated in place); w = density
  const float* alpha;   // factors, sign bit = surface particle
  float4* kappa;        // (x, stiffness) of the last rate pass: the push gathers both at once
  float* errors;        // error of the last rate pass (untiled passes)
  double* tile_sum;     // per solve tile error sums
  uint32_t* tile_count; // per solve tile count of errors > 0
  uint32_t tile_cap;    // room of both
  uint32_t chunks;      // 512-particle chunks the untiled passes walk
  int32_t which;        // 0 divergence, 1 density
  float tolerance;
  uint32_t min_iterations;
  uint32_t max_iterations;
  uint32_t lds_cap;     // particles one block stages in LDS; 0: no tiles
  const float4* Vin;    // tiled density solve: velocities before the viscosity pass, which it runs first
  float* forces;        // that pass's accelerations, or null
  // the flow solve (tiled, no grid barriers between passes; null V2: the
  // passes meet at grid barriers): the second velocity buffer, per tile its
  // passes finished (a half of tile_cap for each solve of the substep) and
  // its weight; tile_sum and tile_count then hold two sets of tile_cap, one
  // per iteration parity
  float4* V2;
  uint32_t* tile_done;
  uint32_t* tile_work;
  uint32_t* flow_list;  // the non-empty tiles in order
  uint32_t through;     // stores other blocks read go through the L2 (the flow solve)
};

// A store another block reads in the same launch: through the L2 to memory
// (device scope) in the flow solve, where no L2 write-back follows each
// task, a plain store otherwise (the grid barrier writes the L2 back).
__device__ __forceinline__ void put4(float4* base, uint32_t i, float4 v, bool through) {
  if (through) {
    const u32x4 bits = {__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)};
    __builtin_amdgcn_raw_buffer_store_b128(bits, whole(base), i * 16u, 0, kCoherentLoad);
  } else {
    base[i] = v;
  }
}

// ---- one particle of a pass ------------------------------------------------
//
// density_rate (dfsph.rs:166-181) and the stiffness and error of solve
// (dfsph.rs:222-262) for particle i, then push_apart (dfsph.rs:184-216).
// Neighbour data comes from global memory (`Glob`) or from the block's LDS
// tile (`Lds`); both give the same values, so the same bits.

struct Glob {
  const float4* __restrict__ P;
  const float4* __restrict__ V;
  const float4* __restrict__ K;
};

// A tile's staged neighbourhood in LDS: x and vx in A, vy and vz in B for
// the rate pass; (x, stiffness) in Q for the push pass. Staged slot of list
// entry e of a particle in processed column p: index(e) + delta[p][column(e)].
// A tile's state: the staged columns' first indexes and lengths and the
// column stride of its LDS (its longest staged column; column c starts at
// slot c * stride), the processed ranges' starts and prefix, and for the
// list format the slot offsets per (processed column, entry column).
constexpr uint32_t kNoTile = 0xffffffffu;
constexpr uint32_t kStagedColumns = 24;    // 6 (x) by 4 (y)
constexpr uint32_t kProcessedColumns = 8;  // 4 (x) by 2 (y)
struct TileState {
  uint32_t ss[kStagedColumns];
  uint32_t len[kStagedColumns];
  uint32_t stride;
  uint32_t ps[kProcessedColumns];
  uint32_t ppre[kProcessedColumns + 1];
  int32_t delta[kProcessedColumns][9];  // list format: c * stride - ss[c], c the entry's staged column
  uint32_t tile;                        // its number, kNoTile for none
};
struct Lds {
  const float4* A;
  const float2* B;
  const float4* Q;
  const TileState* ts;  // the tile's ranges and slot tables
  uint32_t p;           // the particle's processed column
  uint32_t col0;        // staged column of its (-1, -1) neighbour column
  uint32_t stride;      // staged column stride
};

// A fluid entry of either list format: neighbour reference and factor.
struct Ent {
  uint32_t r;
  float f;
};

// fluid entries k .. k + 3 of particle i, either list format
template <int FORMAT>
__device__ __forceinline__ void fluid_entries(const StepParams& prm, uint32_t i, uint32_t k, Ent (&e)[4]) {
  if constexpr (FORMAT == kCompact) {
    // k is a multiple of 4: group k / 4, one 8-byte and one 16-byte load
    const uint32_t g = k >> 2;
    auto n = __builtin_amdgcn_raw_buffer_load_b64(list_rows(prm.near, 8u, prm, g), i * 8u, 0, kStreamLoad);
    auto f = __builtin_amdgcn_raw_buffer_load_b128(list_rows(prm.factor, 16u, prm, g), i * 16u, 0, kStreamLoad);
    e[0] = Ent{n[0] & 0xffffu, __uint_as_float(f[0])};
    e[1] = Ent{n[0] >> 16, __uint_as_float(f[1])};
    e[2] = Ent{n[1] & 0xffffu, __uint_as_float(f[2])};
    e[3] = Ent{n[1] >> 16, __uint_as_float(f[3])};
  } else {
    uint2 v[4];
    list_entries(prm, k, i, v);
#pragma unroll
    for (int u = 0; u < 4; ++u) e[u] = Ent{v[u].x, __uint_as_float(v[u].y)};
  }
}

// LDS slot of an entry of the tile's staged neighbourhood
// The compact code (dx + 1) * 4 + (dy + 1) is the staged column less col0.
template <int FORMAT>
__device__ __forceinline__ int32_t entry_slot(const Lds& l, uint32_t r) {
  if constexpr (FORMAT == kCompact)
    return (int32_t)(__umul24(l.col0 + (r >> kNearBits), l.stride) + (r & kNearMask));
  else
    return (int32_t)list_index(r) + l.ts->delta[l.p][list_column(r)];
}

// particle index of an entry
template <int FORMAT>
__device__ __forceinline__ uint32_t entry_index(const Lds* l, uint32_t r) {
  if constexpr (FORMAT == kCompact)
    return l->ts->ss[l->col0 + (r >> kNearBits)] + (r & kNearMask);
  else
    return list_index(r);
}

// DEEP: the three-stage pipeline below; otherwise one batch of entries
// ahead, for the tiled kernel's fallback, whose registers go to its tiles.
template <int FORMAT, bool DEEP = true>
__device__ __forceinline__ float fluid_rate(const SolveArgs& a, uint32_t i, float4 at, float4 vi, uint32_t nf,
                                            const Glob& g, const Lds* l, int64_t sysz, int64_t szz) {
  const StepParams& prm = a.prm;
  float rate = 0.0f;
  if constexpr (FORMAT == kMasks16 || FORMAT == kMasks32) {
    mask_walk<FORMAT>(prm, i, prm.own_slot[i], sysz, szz, [&](const uint32_t(&j)[4], const bool(&ok)[4]) {
      float4 pj[4], vj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        pj[u] = g.P[j[u]];
        vj[u] = g.V[j[u]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (ok[u]) {
          float ox = at.x - pj[u].x, oy = at.y - pj[u].y, oz = at.z - pj[u].z;
          float f = pair_factor(dot3(ox, oy, oz, ox, oy, oz), prm);
          float gx = ox * f, gy = oy * f, gz = oz * f;
          rate += dot3(vi.x - vj[u].x, vi.y - vj[u].y, vi.z - vj[u].z, gx, gy, gz);
        }
      }
    });
  } else if constexpr (!DEEP) {
    // one batch of entries ahead; slots past the list gather the particle
    // itself and are skipped
    Ent nxt[4];
    fluid_entries<FORMAT>(prm, i, 0, nxt);
    for (uint32_t k = 0; k < nf; k += 4) {
      Ent cur[4];
      float4 pj[4], vj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) cur[u] = nxt[u];
      fluid_entries<FORMAT>(prm, i, k + 4, nxt);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t j = k + u < nf ? entry_index<FORMAT>(l, cur[u].r) : i;
        pj[u] = g.P[j];
        vj[u] = coherent4(g.V, j);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k + u < nf) {
          float f = cur[u].f;
          float gx = (at.x - pj[u].x) * f, gy = (at.y - pj[u].y) * f, gz = (at.z - pj[u].z) * f;
          rate += dot3(vi.x - vj[u].x, vi.y - vj[u].y, vi.z - vj[u].z, gx, gy, gz);
        }
      }
    }
  } else {
    static_assert(FORMAT == kLists, "the deep pipeline reads the list format");
    // Entries go four at a time, in a three-stage pipeline: while batch k
    // is summed in list order, the neighbours of batch k + 1 are being
    // gathered and the list entries of batch k + 2 loaded. Slots past the
    // list point at the particle itself, which keeps every load in bounds;
    // their terms are skipped.
    uint2 cur[4], nxt[4];
    float4 pj[4], vj[4];
    list_entries(prm, 0, i, cur);
    list_entries(prm, 4, i, nxt);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if ((uint32_t)u >= nf) cur[u].x = i;
      if (4u + u >= nf) nxt[u].x = i;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      pj[u] = g.P[list_index(cur[u].x)];
      vj[u] = g.V[list_index(cur[u].x)];
    }
    for (uint32_t k = 0; k < nf; k += 4) {
      float4 pn[4], vn[4];
      uint2 after[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        pn[u] = g.P[list_index(nxt[u].x)];
        vn[u] = g.V[list_index(nxt[u].x)];
      }
      list_entries(prm, k + 8, i, after);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k + 8 + u >= nf) after[u].x = i;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k + u < nf) {
          float f = __uint_as_float(cur[u].y);
          float gx = (at.x - pj[u].x) * f, gy = (at.y - pj[u].y) * f, gz = (at.z - pj[u].z) * f;
          rate += dot3(vi.x - vj[u].x, vi.y - vj[u].y, vi.z - vj[u].z, gx, gy, gz);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cur[u] = nxt[u];
        nxt[u] = after[u];
        pj[u] = pn[u];
        vj[u] = vn[u];
      }
    }
  }
  return rate;
}

// Four entries of particle i from list slot k on; the loops below skip
// those past the list.
template <int FORMAT>
__device__ __forceinline__ void load_batch(const StepParams& prm, uint32_t i, uint32_t k, uint32_t nf, Ent (&e)[4]) {
  fluid_entries<FORMAT>(prm, i, k, e);
}

// Staged slot of a batch entry; entries past the list read slot 0.
template <int FORMAT>
__device__ __forceinline__ int32_t batch_slot(const Lds& l, const Ent& e, bool live) {
  return entry_slot<FORMAT>(l, e.r) & -(int32_t)live;
}

// The rate pass's fluid sum over the tile's LDS: batches of four entries,
// three in flight from global memory, each batch's LDS reads issued
// together before its terms are added in list order.
// `first`, `second`: the particle's first two batches, already loaded.
template <int FORMAT>
__device__ __forceinline__ float fluid_rate_lds(const SolveArgs& a, uint32_t i, float4 at, float4 vi, uint32_t nf,
                                                const Lds& l, const Ent (&first)[4], const Ent (&second)[4]) {
  const StepParams& prm = a.prm;
  float rate = 0.0f;
  Ent b0[4], b1[4], b2[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    b0[u] = first[u];
    b1[u] = second[u];
  }
  load_batch<FORMAT>(prm, i, 8, nf, b2);
  auto sum = [&](const Ent(&e)[4], uint32_t k) {
    float4 pv[4];
    float2 w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t s = batch_slot<FORMAT>(l, e[u], k + u < nf);
      pv[u] = l.A[s];
      w[u] = l.B[s];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float f = e[u].f;
      const float gx = (at.x - pv[u].x) * f, gy = (at.y - pv[u].y) * f, gz = (at.z - pv[u].z) * f;
      const float term = dot3(vi.x - pv[u].w, vi.y - w[u].x, vi.z - w[u].y, gx, gy, gz);
      // past the list: add -0, the identity of IEEE addition, branch-free
      rate = rate + (k + u < nf ? term : -0.0f);
    }
  };
  for (uint32_t k = 0; k < nf; k += 12) {
    sum(b0, k);
    load_batch<FORMAT>(prm, i, k + 12, nf, b0);
    if (k + 4 >= nf) break;
    sum(b1, k + 4);
    load_batch<FORMAT>(prm, i, k + 16, nf, b1);
    if (k + 8 >= nf) break;
    sum(b2, k + 8);
    load_batch<FORMAT>(prm, i, k + 20, nf, b2);
  }
  return rate;
}

// Wall entries w0 .. nt - 1 of particle i, four at a time in list order:
// entries load two batches ahead and wall positions are gathered one batch
// ahead, so no batch waits on a load issued just before it. body(f, pw,
// live) sees each entry; `live` is false past the list.
template <class Body>
__device__ __forceinline__ void walk_walls(const StepParams& prm, uint32_t i, uint32_t w0, uint32_t nt, Body&& body) {
  if (w0 >= nt) return;
  const float4* __restrict__ W = prm.walls;
  uint2 e0[4], e1[4];
  float4 p0[4];
  list_entries(prm, w0, i, e0);
  list_entries(prm, w0 + 4, i, e1);
#pragma unroll
  for (int u = 0; u < 4; ++u) p0[u] = W[w0 + u < nt ? e0[u].x : 0u];
  for (uint32_t k = w0; k < nt; k += 4) {
    uint2 e2[4];
    float4 p1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) p1[u] = W[k + 4 + u < nt ? e1[u].x : 0u];
    list_entries(prm, k + 8, i, e2);
#pragma unroll
    for (int u = 0; u < 4; ++u) body(__uint_as_float(e0[u].y), p0[u], k + u < nt);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      e0[u] = e1[u];
      e1[u] = e2[u];
      p0[u] = p1[u];
    }
  }
}

// Walls of the rate pass, then stiffness and error; writes kappa, returns the
// error.
template <int FORMAT>
__device__ __forceinline__ float finish_rate(const SolveArgs& a, uint32_t i, float4 at, float4 vi, float rate,
                                            uint32_t nf, uint32_t nw) {
  const StepParams& prm = a.prm;
  // walls in list order (a particle at a corner lists ~40 of them)
  const uint32_t w0 = FORMAT == kLists ? nf : 0u;
  walk_walls(prm, i, w0, w0 + nw, [&](float f, float4 pw, bool live) {
    float gx = (at.x - pw.x) * f, gy = (at.y - pw.y) * f, gz = (at.z - pw.z) * f;
    float term = dot3(vi.x, vi.y, vi.z, gx, gy, gz);
    rate = rate + (live ? term : -0.0f);
  });
  const float dt = prm.dt;
  rate = rate * prm.mass;
  float fa = a.alpha[i];
  float target, error;
  if (a.which == 1) {
    float over = rmax(vi.w + dt * rate - prm.rest, 0.0f);
    target = over / (dt * dt);
    error = over / prm.rest;
  } else {
    float r = signbit(fa) ? 0.0f : rmax(rate, 0.0f);
    target = r / dt;
    error = dt * r / prm.rest;
  }
  put4(a.kappa, i, make_float4(at.x, at.y, at.z, target * fabsf(fa)), a.through != 0);
  return error;
}

template <int FORMAT, bool DEEP = true>
__device__ __forceinline__ float rate_one(const SolveArgs& a, uint32_t i, const Glob& g, const Lds* l, bool staged,
                                         int64_t sysz, int64_t szz) {
  float4 at = a.P[i];
  float4 vi = shared4<!DEEP>(g.V, i);
  uint32_t counts = a.prm.pair_counts[i];
  uint32_t nf = counts & 0xffffu, nw = counts >> 16;
  float rate;
  if constexpr (FORMAT == kLists || FORMAT == kCompact) {
    if (staged) {
      Ent b0[4], b1[4];
      load_batch<FORMAT>(a.prm, i, 0, nf, b0);
      load_batch<FORMAT>(a.prm, i, 4, nf, b1);
      rate = fluid_rate_lds<FORMAT>(a, i, at, vi, nf, *l, b0, b1);
    } else {
      rate = fluid_rate<FORMAT, DEEP>(a, i, at, vi, nf, g, l, sysz, szz);
    }
  } else {
    rate = fluid_rate<FORMAT>(a, i, at, vi, nf, g, l, sysz, szz);
  }
  return finish_rate<FORMAT>(a, i, at, vi, rate, nf, nw);
}

// The push pass's fluid sum over the tile's LDS (at = (x, own stiffness)),
// as the rate pass's; `first`, `second`: the first two batches, loaded.
template <int FORMAT>
__device__ __forceinline__ void push_fluid_lds(const SolveArgs& a, uint32_t i, float4 at, uint32_t nf, const Lds& l,
                                               const Ent (&first)[4], const Ent (&second)[4], float& cx, float& cy,
                                               float& cz) {
  const StepParams& prm = a.prm;
  const float mass = prm.mass, own = at.w;
  Ent b0[4], b1[4], b2[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    b0[u] = first[u];
    b1[u] = second[u];
  }
  load_batch<FORMAT>(prm, i, 8, nf, b2);
  auto sum = [&](const Ent(&e)[4], uint32_t k) {
    float4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = l.Q[batch_slot<FORMAT>(l, e[u], k + u < nf)];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float f = e[u].f;
      const float gx = (at.x - q[u].x) * f, gy = (at.y - q[u].y) * f, gz = (at.z - q[u].z) * f;
      const float k2 = mass * (own + q[u].w);
      // past the list: subtract +0, the identity of IEEE subtraction
      const bool live = k + u < nf;
      cx = cx - (live ? gx * k2 : 0.0f);
      cy = cy - (live ? gy * k2 : 0.0f);
      cz = cz - (live ? gz * k2 : 0.0f);
    }
  };
  for (uint32_t k = 0; k < nf; k += 12) {
    sum(b0, k);
    load_batch<FORMAT>(prm, i, k + 12, nf, b0);
    if (k + 4 >= nf) break;
    sum(b1, k + 4);
    load_batch<FORMAT>(prm, i, k + 16, nf, b1);
    if (k + 8 >= nf) break;
    sum(b2, k + 8);
    load_batch<FORMAT>(prm, i, k + 20, nf, b2);
  }
}

// The push's walls and velocity update of particle i after its fluid sum.
template <int FORMAT>
__device__ __forceinline__ void push_finish(const SolveArgs& a, uint32_t i, float4 at, uint32_t nf, uint32_t nw,
                                            float4 vi, float cx, float cy, float cz, float4* vout) {
  const StepParams& prm = a.prm;
  const uint32_t w0 = FORMAT == kLists ? nf : 0u;
  const float kw = prm.mass * at.w;
  walk_walls(prm, i, w0, w0 + nw, [&](float f, float4 pw, bool live) {
    float gx = (at.x - pw.x) * f, gy = (at.y - pw.y) * f, gz = (at.z - pw.z) * f;
    cx = cx - (live ? gx * kw : 0.0f);
    cy = cy - (live ? gy * kw : 0.0f);
    cz = cz - (live ? gz * kw : 0.0f);
  });
  const float dt = prm.dt;
  vi.x = vi.x + cx * dt;
  vi.y = vi.y + cy * dt;
  vi.z = vi.z + cz * dt;
  put4(vout, i, vi, a.through != 0);
}

// Writes the pushed velocity to `vout` (g.V itself but in the flow solve).
template <int FORMAT, bool DEEP = true>
__device__ __forceinline__ void push_one(const SolveArgs& a, uint32_t i, const Glob& g, const Lds* l, bool staged,
                                         int64_t sysz, int64_t szz, float4* vout) {
  const StepParams& prm = a.prm;
  const float mass = prm.mass;
  float4 at = shared4<!DEEP>(a.kappa, i);  // x as P holds it, own stiffness in w
  float own = at.w;
  float cx = 0.0f, cy = 0.0f, cz = 0.0f;
  uint32_t counts = prm.pair_counts[i];
  uint32_t nf = counts & 0xffffu, nw = counts >> 16;
  if constexpr (FORMAT == kMasks16 || FORMAT == kMasks32) {
    mask_walk<FORMAT>(prm, i, prm.own_slot[i], sysz, szz, [&](const uint32_t(&j)[4], const bool(&ok)[4]) {
      float4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) pj[u] = g.K[j[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (ok[u]) {
          float ox = at.x - pj[u].x, oy = at.y - pj[u].y, oz = at.z - pj[u].z;
          float f = pair_factor(dot3(ox, oy, oz, ox, oy, oz), prm);
          float gx = ox * f, gy = oy * f, gz = oz * f;
          float k2 = mass * (own + pj[u].w);
          cx = cx - gx * k2;
          cy = cy - gy * k2;
          cz = cz - gz * k2;
        }
      }
    });
  } else if (staged) {
    Ent b0[4], b1[4];
    load_batch<FORMAT>(prm, i, 0, nf, b0);
    load_batch<FORMAT>(prm, i, 4, nf, b1);
    push_fluid_lds<FORMAT>(a, i, at, nf, *l, b0, b1, cx, cy, cz);
  } else if constexpr (!DEEP) {
    Ent nxt[4];
    fluid_entries<FORMAT>(prm, i, 0, nxt);
    for (uint32_t k = 0; k < nf; k += 4) {
      Ent cur[4];
      float4 pj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) cur[u] = nxt[u];
      fluid_entries<FORMAT>(prm, i, k + 4, nxt);
#pragma unroll
      for (int u = 0; u < 4; ++u) pj[u] = coherent4(g.K, k + u < nf ? entry_index<FORMAT>(l, cur[u].r) : i);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k + u < nf) {
          float f = cur[u].f;
          float gx = (at.x - pj[u].x) * f, gy = (at.y - pj[u].y) * f, gz = (at.z - pj[u].z) * f;
          float k2 = mass * (own + pj[u].w);
          cx = cx - gx * k2;
          cy = cy - gy * k2;
          cz = cz - gz * k2;
        }
      }
    }
  } else {
    static_assert(FORMAT == kLists, "the deep pipeline reads the list format");
    uint2 cur[4], nxt[4];
    float4 pj[4];
    list_entries(prm, 0, i, cur);
    list_entries(prm, 4, i, nxt);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if ((uint32_t)u >= nf) cur[u].x = i;
      if (4u + u >= nf) nxt[u].x = i;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) pj[u] = g.K[list_index(cur[u].x)];
    for (uint32_t k = 0; k < nf; k += 4) {
      float4 pn[4];
      uint2 after[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) pn[u] = g.K[list_index(nxt[u].x)];
      list_entries(prm, k + 8, i, after);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k + 8 + u >= nf) after[u].x = i;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k + u < nf) {
          float f = __uint_as_float(cur[u].y);
          float gx = (at.x - pj[u].x) * f, gy = (at.y - pj[u].y) * f, gz = (at.z - pj[u].z) * f;
          float k2 = mass * (own + pj[u].w);
          cx = cx - gx * k2;
          cy = cy - gy * k2;
          cz = cz - gz * k2;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cur[u] = nxt[u];
        nxt[u] = after[u];
        pj[u] = pn[u];
      }
    }
  }
  push_finish<FORMAT>(a, i, at, nf, nw, shared4<!DEEP>(g.V, i), cx, cy, cz, vout);
}

// ---- LDS tiles --------------------------------------------------------------
//
// The cell box is cut into tiles of 4 (x) by 2 (y) cell columns by
// tile_depth cells in z, numbered x-major like the cells. A tile's
// particles are eight contiguous ranges (one per column); the cells that can
// hold their neighbours are the 6 x 4 columns around them over the tile's z
// range widened by one cell: 24 contiguous ranges, which the block stages
// into its LDS with coalesced loads. Staged column c = 4 * (x offset) + (y
// offset), so a compact entry's column code (dx + 1) * 4 + (dy + 1) plus
// col0 = 4 * px + py of the particle's processed column (px, py) names its
// staged column.

struct TileShape {
  uint32_t sx, sy, sz, ntx, nty, ntz, tz, count;
};

__device__ __forceinline__ TileShape tile_shape(const Control* ctl) {
  TileShape t;
  t.sx = ctl->span[0];
  t.sy = ctl->span[1];
  t.sz = ctl->span[2];
  t.ntx = (t.sx + 3) / 4;
  t.nty = (t.sy + 1) / 2;
  t.ntz = (t.sz + 15) / 16;
  t.tz = (t.sz + t.ntz - 1) / t.ntz;
  t.count = t.ntx * t.nty * t.ntz;
  return t;
}

// A tile's column ranges, in two steps so the loads of the next tile's can
// be in flight while this tile's particles are summed. Wave 0 (all 64
// lanes, `have` uniform): lanes 0..23 load the staged columns' ranges,
// 24..31 the processed columns'.
struct TileLoad {
  uint32_t s0, n, tile;
};

__device__ __forceinline__ TileLoad tile_load(const TileShape& sh, uint32_t tile, bool have,
                                              const uint32_t* __restrict__ start, uint32_t lane) {
  TileLoad r{0u, 0u, have ? tile : kNoTile};
  if (have) {
    uint32_t tx = tile / (sh.nty * sh.ntz), ty = (tile / sh.ntz) % sh.nty, tzi = tile % sh.ntz;
    uint32_t z0 = tzi * sh.tz, z1 = min(z0 + sh.tz, sh.sz);
    if (lane < kStagedColumns) {
      int64_t x = (int64_t)(4 * tx) - 1 + (lane >> 2), y = (int64_t)(2 * ty) - 1 + (lane & 3);
      if (x >= 0 && y >= 0 && x < sh.sx && y < sh.sy) {
        uint64_t base = ((uint64_t)x * sh.sy + (uint64_t)y) * sh.sz;
        uint32_t za = z0 > 0 ? z0 - 1 : 0, zb = min(z1 + 1, sh.sz);
        r.s0 = start[base + za];
        r.n = start[base + zb] - r.s0;
      }
    } else if (lane < kStagedColumns + kProcessedColumns) {
      uint32_t p = lane - kStagedColumns;
      uint64_t x = 4 * tx + (p >> 1), y = 2 * ty + (p & 1);
      if (x < sh.sx && y < sh.sy) {
        uint64_t base = (x * sh.sy + y) * sh.sz;
        r.s0 = start[base + z0];
        r.n = start[base + z1] - r.s0;
      }
    }
  }
  return r;
}

// Then the same wave takes the longest staged column, scans the processed
// prefix, and writes the list format's slot offsets into `ts`.
__device__ __forceinline__ void tile_scan(TileLoad r, TileState& ts, uint32_t lane) {
  uint32_t longest = lane < kStagedColumns ? r.n : 0u;
  for (int d = 16; d >= 1; d >>= 1) longest = max(longest, (uint32_t)__shfl_xor((int)longest, d, 64));
  const uint32_t stride = (uint32_t)__shfl((int)longest, 0, 64);
  uint32_t incl = r.n;
  for (uint32_t d = 1; d < kProcessedColumns; d <<= 1) {
    uint32_t up = (uint32_t)__shfl_up((int)incl, d, 64);
    if (lane < kStagedColumns + kProcessedColumns && lane >= kStagedColumns + d) incl += up;
  }
  if (lane < kStagedColumns) {
    ts.ss[lane] = r.s0;
    ts.len[lane] = r.n;
    if (lane == 0) ts.stride = stride;
  } else if (lane < kStagedColumns + kProcessedColumns) {
    ts.ps[lane - kStagedColumns] = r.s0;
    ts.ppre[lane - kStagedColumns] = incl - r.n;
    if (lane == kStagedColumns + kProcessedColumns - 1) ts.ppre[kProcessedColumns] = incl;
  }
  if (lane == 0) ts.tile = r.tile;
  // list format: staged slot of entry e of processed column p is
  // index(e) + delta[p][column(e)]; 72 entries, lanes take k and k + 64
  for (uint32_t k = lane; k < kProcessedColumns * 9; k += 64) {
    const uint32_t p = k / 9, c = k % 9;
    const int dx = (int)(c / 3) - 1, dy = (int)(c % 3) - 1;
    const uint32_t col = (uint32_t)((1 + (int)(p >> 1) + dx) * 4 + (1 + (int)(p & 1) + dy));
    ts.delta[p][c] = (int32_t)(col * stride) - (int32_t)ts.ss[col];
  }
}

__device__ __forceinline__ uint32_t processed_index(const TileState& ts, uint32_t q, uint32_t& p) {
  p = 0;
  while (p < kProcessedColumns - 1 && ts.ppre[p + 1] <= q) ++p;
  return ts.ps[p] + (q - ts.ppre[p]);
}

// Tiles [begin, end) of XCD group v: the v-th eighth, numbered x-major, so
// one group's tiles are one slab of the box.
__device__ __forceinline__ uint32_t group_begin(const TileShape& sh, uint32_t v) {
  return (uint32_t)((uint64_t)sh.count * v / kXcds);
}

// Each pass starts with a fixed tile per block, the same in every pass (its
// ranges are set up once per launch): tile b + rank of the block's group, b
// the group's first, rank = blockIdx / 8. The rest of the group's tiles are
// its queue.
__device__ __forceinline__ uint32_t fixed_tiles(const TileShape& sh, uint32_t v) {
  return min(gridDim.x / kXcds, group_begin(sh, v + 1) - group_begin(sh, v));
}

// Wave 0 takes the block's next tile for this pass: from its own group's
// queue while it lasts, then from the others in turn. `done` holds the
// groups this block found empty; each failed take adds one, so the loop
// ends. Returns kNoTile when every queue is empty.
__device__ __forceinline__ uint32_t take_tile(Control* ctl, const TileShape& sh, uint32_t parity, uint32_t& done) {
  const uint32_t lane = threadIdx.x & 63u, group = blockIdx.x % kXcds;
  for (;;) {
    uint32_t v = (group + lane) % kXcds;
    bool left = false;
    if (lane < kXcds && !((done >> v) & 1u)) {
      uint32_t next = __hip_atomic_load(&ctl->tile_next[parity][v][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      left = fixed_tiles(sh, v) + next < group_begin(sh, v + 1) - group_begin(sh, v);
    }
    uint64_t m = __ballot(left);
    if (m == 0) return kNoTile;
    v = (group + (uint32_t)__builtin_ctzll(m)) % kXcds;
    const uint32_t b = group_begin(sh, v) + fixed_tiles(sh, v), e = group_begin(sh, v + 1);
    uint32_t k = 0;
    if (lane == 0) k = __hip_atomic_fetch_add(&ctl->tile_next[parity][v][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    k = (uint32_t)__shfl((int)k, 0, 64);
    if (b + k < e) return b + k;
    done |= 1u << v;
  }
}

// ---- the viscosity pass (dfsph.rs:90-99) -------------------------------------
//
// viscous (src/fluid/step.rs:229-251) for one fluid pair: only when the
// particles approach, in the reference's operation order.
__device__ __forceinline__ void viscous_pair(const StepParams& prm, float4 at, float4 vi, float3 pj, float4 vj,
                                             float f, float& px, float& py, float& pz) {
  float ox = at.x - pj.x, oy = at.y - pj.y, oz = at.z - pj.z;
  float toward = dot3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, ox, oy, oz);
  if (toward < 0.0f) {
    float squeeze = toward / (dot3(ox, oy, oz, ox, oy, oz) + prm.soften);
    float damp = -prm.nu * squeeze * 2.0f / (vi.w + vj.w);
    float k = prm.mass * damp;
    px = px - (ox * f) * k;
    py = py - (oy * f) * k;
    pz = pz - (oz * f) * k;
  }
}

// The walls' part and the result: the walls' pull in list order, then
// wall_viscous (step.rs:262-301: normalize_or(away, ZERO), motion into the
// walls only), then v + (fluid + walls + gravity) * dt into VA (density
// carried along) and `forces`.
template <int FORMAT>
__device__ __forceinline__ void viscous_finish(const StepParams& prm, uint32_t i, float4 at, float4 vi, float px,
                                               float py, float pz, uint32_t nf, uint32_t nw, float4* VA,
                                               float* forces, bool through = false) {
  const float density = vi.w;
  const uint32_t w0 = FORMAT == kLists ? nf : 0u;
  float wx = 0.0f, wy = 0.0f, wz = 0.0f;
  if (nw) {
    // both walks in list order, their loads run ahead (walk_walls)
    float awx = 0.0f, awy = 0.0f, awz = 0.0f;
    walk_walls(prm, i, w0, w0 + nw, [&](float f, float4 pw, bool live) {
      if (live) {
        awx = awx + -((at.x - pw.x) * f);
        awy = awy + -((at.y - pw.y) * f);
        awz = awz + -((at.z - pw.z) * f);
      }
    });
    float len = sqrtf(dot3(awx, awy, awz, awx, awy, awz));
    float rcp = 1.0f / len;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    if (isfinite(rcp) && rcp > 0.0f) {
      nx = awx * rcp;
      ny = awy * rcp;
      nz = awz * rcp;
    }
    float tn = rmin(dot3(vi.x, vi.y, vi.z, nx, ny, nz), 0.0f);
    float ix = nx * tn, iy = ny * tn, iz = nz * tn;
    walk_walls(prm, i, w0, w0 + nw, [&](float f, float4 pw, bool live) {
      float ox = at.x - pw.x, oy = at.y - pw.y, oz = at.z - pw.z;
      float toward = dot3(ix, iy, iz, ox, oy, oz);
      if (live && toward < 0.0f) {
        float squeeze = toward / (dot3(ox, oy, oz, ox, oy, oz) + prm.soften);
        float damp = -prm.wall_nu * squeeze / density;
        float k2 = prm.mass * damp;
        wx = wx - (ox * f) * k2;
        wy = wy - (oy * f) * k2;
        wz = wz - (oz * f) * k2;
      }
    });
  }
  float ax = px + wx + prm.gx;
  float ay = py + wy + prm.gy;
  float az = pz + wz + prm.gz;
  if (forces) {
    forces[3 * i] = ax;
    forces[3 * i + 1] = ay;
    forces[3 * i + 2] = az;
  }
  put4(VA, i, make_float4(vi.x + ax * prm.dt, vi.y + ay * prm.dt, vi.z + az * prm.dt, density), through);
}

// The viscosity pass for particle i of a tile, inside the density solve's
// launch: neighbours from the tile's LDS (A = (x, y, z, vx), C = (vy, vz,
// density)) or from global memory; the same values as k_dfsph_viscous.
// STAGED: the neighbours come from the tile's LDS (a loop of its own, so
// that its loads are LDS reads rather than generic ones chosen at run time).
template <int FORMAT, bool STAGED>
__device__ __forceinline__ void viscous_walk(const SolveArgs& a, uint32_t i, const Lds* l, const float4* C) {
  const StepParams& prm = a.prm;
  const float4 at = a.P[i];
  const float4 vi = a.Vin[i];
  const uint32_t counts = prm.pair_counts[i];
  const uint32_t nf = counts & 0xffffu, nw = counts >> 16;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  // batches of four entries, three in flight (as the rate pass)
  Ent b0[4], b1[4], b2[4];
  load_batch<FORMAT>(prm, i, 0, nf, b0);
  load_batch<FORMAT>(prm, i, 4, nf, b1);
  load_batch<FORMAT>(prm, i, 8, nf, b2);
  auto sum = [&](const Ent(&e)[4], uint32_t k) {
    float3 pj[4];
    float4 vj[4];
    uint32_t j[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool live = k + u < nf;
      j[u] = live ? entry_index<FORMAT>(l, e[u].r) : i;
      if constexpr (STAGED) {
        const int32_t s = batch_slot<FORMAT>(*l, e[u], live);
        const float4 q = l->A[s], c = C[s];
        pj[u] = make_float3(q.x, q.y, q.z);
        vj[u] = make_float4(q.w, c.x, c.y, c.z);
      } else {
        const float4 q = a.P[j[u]];
        pj[u] = make_float3(q.x, q.y, q.z);
        vj[u] = a.Vin[j[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k + u < nf && j[u] != i) viscous_pair(prm, at, vi, pj[u], vj[u], e[u].f, px, py, pz);
  };
  for (uint32_t k = 0; k < nf; k += 12) {
    sum(b0, k);
    load_batch<FORMAT>(prm, i, k + 12, nf, b0);
    if (k + 4 >= nf) break;
    sum(b1, k + 4);
    load_batch<FORMAT>(prm, i, k + 16, nf, b1);
    if (k + 8 >= nf) break;
    sum(b2, k + 8);
    load_batch<FORMAT>(prm, i, k + 20, nf, b2);
  }
  viscous_finish<FORMAT>(prm, i, at, vi, px, py, pz, nf, nw, a.V, a.forces, a.through != 0);
}

template <int FORMAT>
__device__ __forceinline__ void viscous_one(const SolveArgs& a, uint32_t i, const Lds* l, const float4* C,
                                            bool staged) {
  if (staged)
    viscous_walk<FORMAT, true>(a, i, l, C);
  else
    viscous_walk<FORMAT, false>(a, i, l, C);
}

// ---- the stop sum over solve tiles ------------------------------------------

// Thread sums (value, count) of a wave folded by halving, h = 32..1; lane 0
// holds the wave's.
__device__ __forceinline__ void wave_fold(double& v, uint32_t& c) {
  for (int h = 32; h >= 1; h >>= 1) {
    v = v + __shfl_down(v, h, 64);
    c += (uint32_t)__shfl_down((int)c, h, 64);
  }
}

// One wave: the block's 16 wave sums folded by halving (h = 8..1) into the
// tile's sum, stored at `tile`.
__device__ __forceinline__ void store_tile_sum(const SolveArgs& a, uint32_t tile, const double* wsum,
                                               const uint32_t* wcnt, uint32_t lane) {
  double v = lane < kSolveWaves ? wsum[lane] : 0.0;
  uint32_t c = lane < kSolveWaves ? wcnt[lane] : 0u;
  for (int h = kSolveWaves / 2; h >= 1; h >>= 1) {
    v = v + __shfl_down(v, h, 64);
    c += (uint32_t)__shfl_down((int)c, h, 64);
  }
  if (lane == 0 && tile < a.tile_cap) {
    a.tile_sum[tile] = v;
    a.tile_count[tile] = c;
  }
}

// Every wave: lane l adds tile sums l, l + 64, ... in order (eight loads in
// flight), the lanes fold by halving; returns the total and its count.
__device__ __forceinline__ double fold_tiles(const SolveArgs& a, uint32_t tiles, uint32_t& loaded) {
  const uint32_t lane = threadIdx.x & 63u;
  double s = 0.0;
  uint32_t cn = 0;
  uint32_t k = lane;
  for (; k + 7u * 64u < tiles; k += 8u * 64u) {
    double v[8];
    uint32_t w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      v[u] = a.tile_sum[k + u * 64u];
      w[u] = a.tile_count[k + u * 64u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      s += v[u];
      cn += w[u];
    }
  }
  for (; k < tiles; k += 64u) {
    s += a.tile_sum[k];
    cn += a.tile_count[k];
  }
  wave_fold(s, cn);
  loaded = (uint32_t)__shfl((int)cn, 0, 64);
  return __shfl(s, 0, 64);
}

// The same fold, for the tiled solve: the block's threads copy the tile
// sums into its staging LDS (free between passes) `room` at a time, room a
// multiple of 64, so lane l's tiles keep their order; every wave then adds
// them from LDS.
__device__ __forceinline__ double fold_tiles_lds(const SolveArgs& a, uint32_t tiles, float4* lds, uint32_t room,
                                                 uint32_t& loaded) {
  double* L = reinterpret_cast<double*>(lds);
  uint32_t* C = reinterpret_cast<uint32_t*>(L + room);
  const uint32_t t = threadIdx.x, lane = t & 63u;
  double s = 0.0;
  uint32_t cn = 0;
  for (uint32_t base = 0; base < tiles; base += room) {
    const uint32_t m = min(room, tiles - base);
    for (uint32_t k = t; k < m; k += kSolveThreads) {
      L[k] = coherent_f64(a.tile_sum, base + k);
      C[k] = coherent_u32(a.tile_count, base + k);
    }
    __syncthreads();
    for (uint32_t k = lane; k < m; k += 64u) {
      s += L[k];
      cn += C[k];
    }
    // the LDS is overwritten next (next room, or the push pass's staging)
    __syncthreads();
  }
  wave_fold(s, cn);
  loaded = (uint32_t)__shfl((int)cn, 0, 64);
  return __shfl(s, 0, 64);
}

// The untiled passes' tile sums, from the errors they stored: block b takes
// tiles b, b + grid, ...; wave 0 loads a tile's ranges, then every thread
// adds its errors in processed order.
__device__ void tile_sums_from_errors(const SolveArgs& a, const TileShape& sh, TileState& ts, double* wsum,
                                      uint32_t* wcnt) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  for (uint32_t tile = blockIdx.x; tile < sh.count; tile += gridDim.x) {
    if (wave == 0) tile_scan(tile_load(sh, tile, true, a.prm.cell_start, lane), ts, lane);
    __syncthreads();
    const uint32_t count = ts.ppre[kProcessedColumns];
    double acc = 0.0;
    uint32_t cnt = 0;
    for (uint32_t q = t; q < count; q += kSolveThreads) {
      uint32_t p;
      const float e = a.errors[processed_index(ts, q, p)];
      acc += (double)e;
      cnt += e > 0.0f ? 1u : 0u;
    }
    wave_fold(acc, cnt);
    if (lane == 0) {
      wsum[wave] = acc;
      wcnt[wave] = cnt;
    }
    __syncthreads();
    if (wave == 0) store_tile_sum(a, tile, wsum, wcnt, lane);
  }
}

constexpr int kRatePass = 0, kPushPass = 1, kViscousPass = 2;

// Stages tile ts's neighbourhood for a pass into the block's LDS: wave w
// stages columns w and w + 16 (when below 24), lanes on consecutive
// particles, two rows of 64 of each column per trip: eight loads in flight
// (past a column's end they read particle 0 and are not stored). The rate
// pass stages (x, y, z, vx | vy, vz) from P and g.V, the push pass (x,
// stiffness) from g.K, the viscosity pass (x, y, z, vx | vy, vz, density)
// from P and Vin.
__device__ __forceinline__ void stage_tile(const SolveArgs& a, const Glob& g, int mode, const TileState& ts,
                                           uint32_t stride, float4* A, float2* B, float4* Q, float4* C) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t c0 = wave, c1 = wave + 16u;
  const bool second = c1 < kStagedColumns;
  const uint32_t f0 = ts.ss[c0], n0 = ts.len[c0];
  const uint32_t f1 = second ? ts.ss[c1] : 0u, n1 = second ? ts.len[c1] : 0u;
  const uint32_t to0 = c0 * stride, to1 = c1 * stride;
  const uint32_t rows = max(n0, n1);
  for (uint32_t s = lane; s < rows; s += 128) {
    const uint32_t sb = s + 64;
    const bool h0 = s < n0, h1 = s < n1, h2 = sb < n0, h3 = sb < n1;
    const uint32_t j0 = h0 ? f0 + s : 0u, j1 = h1 ? f1 + s : 0u;
    const uint32_t j2 = h2 ? f0 + sb : 0u, j3 = h3 ? f1 + sb : 0u;
    if (mode == kViscousPass) {
      const float4 p0 = a.P[j0], p1 = a.P[j1], p2 = a.P[j2], p3 = a.P[j3];
      const float4 v0 = a.Vin[j0], v1 = a.Vin[j1], v2 = a.Vin[j2], v3 = a.Vin[j3];
      if (h0) {
        A[to0 + s] = make_float4(p0.x, p0.y, p0.z, v0.x);
        C[to0 + s] = make_float4(v0.y, v0.z, v0.w, 0.0f);
      }
      if (h1) {
        A[to1 + s] = make_float4(p1.x, p1.y, p1.z, v1.x);
        C[to1 + s] = make_float4(v1.y, v1.z, v1.w, 0.0f);
      }
      if (h2) {
        A[to0 + sb] = make_float4(p2.x, p2.y, p2.z, v2.x);
        C[to0 + sb] = make_float4(v2.y, v2.z, v2.w, 0.0f);
      }
      if (h3) {
        A[to1 + sb] = make_float4(p3.x, p3.y, p3.z, v3.x);
        C[to1 + sb] = make_float4(v3.y, v3.z, v3.w, 0.0f);
      }
    } else if (mode == kRatePass) {
      const float4 p0 = a.P[j0], p1 = a.P[j1], p2 = a.P[j2], p3 = a.P[j3];
      const float4 v0 = coherent4(g.V, j0), v1 = coherent4(g.V, j1), v2 = coherent4(g.V, j2),
                   v3 = coherent4(g.V, j3);
      if (h0) {
        A[to0 + s] = make_float4(p0.x, p0.y, p0.z, v0.x);
        B[to0 + s] = make_float2(v0.y, v0.z);
      }
      if (h1) {
        A[to1 + s] = make_float4(p1.x, p1.y, p1.z, v1.x);
        B[to1 + s] = make_float2(v1.y, v1.z);
      }
      if (h2) {
        A[to0 + sb] = make_float4(p2.x, p2.y, p2.z, v2.x);
        B[to0 + sb] = make_float2(v2.y, v2.z);
      }
      if (h3) {
        A[to1 + sb] = make_float4(p3.x, p3.y, p3.z, v3.x);
        B[to1 + sb] = make_float2(v3.y, v3.z);
      }
    } else {
      const float4 q0 = coherent4(g.K, j0), q1 = coherent4(g.K, j1), q2 = coherent4(g.K, j2),
                   q3 = coherent4(g.K, j3);
      if (h0) Q[to0 + s] = q0;
      if (h1) Q[to1 + s] = q1;
      if (h2) Q[to0 + sb] = q2;
      if (h3) Q[to1 + sb] = q3;
    }
  }
}

// One pass over the tiles (rate, push or viscosity pass), queued
// per XCD group with stealing (tile_next[parity]). Each round the block
// takes a tile: stages its neighbourhood in LDS, then every particle of it
// sums from LDS (a tile past the LDS gathers from global memory instead).
// The next round's tile is taken and its column ranges loaded by the last
// wave while the others sum (both double buffered), so neither latency
// sits between rounds.
//
// The rate pass also forms each tile's error sum: every thread adds the
// errors of its particles, the waves fold theirs into `wsum`, and wave 0
// folds those into the tile's sum at the start of the next round.
template <int FORMAT>
__device__ void tiled_pass(const SolveArgs& a, const TileShape& sh, int mode, uint32_t parity, float4* lds,
                           TileState (&tstate)[3], uint32_t (&grab_s)[2], const Glob& g, double* wsum,
                           uint32_t* wcnt) {
  const StepParams& prm = a.prm;
  Control* ctl = prm.ctl;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const bool rate = mode == kRatePass, viscous = mode == kViscousPass;
  // the LDS: A (float4) then B (float2) for the rate pass, Q (float4) for
  // the push pass, A then C (float4: vy, vz, density) for the viscosity
  // pass, which stages 32 bytes a particle instead of 24
  const uint32_t cap = viscous ? a.lds_cap * kTileBytes / 32u : a.lds_cap;
  float4* A = lds;
  float2* B = reinterpret_cast<float2*>(lds + cap);
  float4* Q = lds;
  float4* C = lds + cap;
  uint32_t done = 0;
  // round 0: the fixed tile (state 2); round r > 0: the tile taken during
  // round r - 1 (state (r - 1) & 1)
  const TileState* tsp = &tstate[2];
  bool any = blockIdx.x / kXcds < fixed_tiles(sh, blockIdx.x % kXcds);
  uint32_t summed = kNoTile;  // the last round's tile (rate pass), its wave sums in wsum
  for (uint32_t r = 0;; ++r) {
    if (rate && wave == 0 && summed != kNoTile) store_tile_sum(a, summed, wsum, wcnt, lane);
    const uint32_t nxt = r & 1u;
    if (!any && r > 0) break;
    const TileState& ts = *tsp;
    summed = __builtin_amdgcn_readfirstlane(ts.tile);
    const uint32_t stride = __builtin_amdgcn_readfirstlane(ts.stride);
    const uint32_t count = ts.ppre[kProcessedColumns];
    const bool fits = kStagedColumns * stride <= cap;
    if (fits && count) stage_tile(a, g, mode, ts, stride, A, B, Q, C);
    __syncthreads();
    // the last wave (whose particles, if any, come last) takes the next
    // tile and loads its ranges while the others sum
    TileLoad next{0u, 0u};
    if (wave == kSolveWaves - 1) {
      const uint32_t tile = take_tile(ctl, sh, parity, done);
      if (lane == 0) grab_s[nxt] = tile;
      next = tile_load(sh, tile, tile != kNoTile, prm.cell_start, lane);
    }
    double acc = 0.0;
    uint32_t cnt = 0;
    for (uint32_t q = t; q < count; q += kSolveThreads) {
      uint32_t p;
      uint32_t i = processed_index(ts, q, p);
      const Lds l{A, B, Q, &ts, p, 4u * (p >> 1) + (p & 1u), stride};
      if (rate) {
        const float e = rate_one<FORMAT, false>(a, i, g, &l, fits, 0, 0);
        acc += (double)e;
        cnt += e > 0.0f ? 1u : 0u;
      } else if (viscous) {
        viscous_one<FORMAT>(a, i, &l, C, fits);
      } else {
        push_one<FORMAT, false>(a, i, g, &l, fits, 0, 0, a.V);
      }
    }
    if (rate) {
      wave_fold(acc, cnt);
      if (lane == 0) {
        wsum[wave] = acc;
        wcnt[wave] = cnt;
      }
    }
    if (wave == kSolveWaves - 1) tile_scan(next, tstate[nxt], lane);
    __syncthreads();
    tsp = &tstate[nxt];
    any = grab_s[nxt] != kNoTile;
  }
}

// ---- the dataflow solve -------------------------------------------------------
//
// The tiled solve without grid barriers between its passes. Pass p of tile
// T is a task. The density solve's launch starts with the viscosity pass
// (p = 0) and its Jacobi passes follow at p = 1 + q; the divergence solve's
// Jacobi passes are p = q. Jacobi pass q = 2k is iteration k's rate pass,
// q = 2k + 1 its push. A task reads only its tile's neighbourhood, so it
// waits only for the tiles around it: (p, T) starts once every tile within
// one tile of T in x, y and z has finished pass p - 1 (tile_done >= p). The
// velocities are double buffered: iteration k's rate pass reads V_k (V when
// k is even, V2 when odd) and its push writes V_k+1 into the other buffer,
// so the push may run before iteration k's stop decision: if iteration k
// stops, V_k is intact. The push of iteration k + 1 overwrites V_k, so it
// waits for the decision of iteration k. Each decision is taken by the block
// that publishes the last rate task of its iteration: it folds the tile
// sums in the same fixed shape and commits the decision in iteration order
// (FlowWords::resolved).
//
// Each XCD group has one queue over its slab of tiles: index j is pass
// j / len, tile begin + j % len; the slabs are cut at equal weights, and a
// group runs its slab's tasks alone. No block waits on a task that no
// running block will take: a task waits only on tasks of lower passes, every
// queue hands its tasks out in pass order, and each queue's own blocks take
// them; past the stop decision, tasks of later passes are dropped, not
// waited for.
//
// Stores that other blocks read (velocities, stiffnesses, tile sums) go
// through the L2 to memory (device scope, put4), and a task is published
// only after every wave of its block has waited for its stores, so no L2
// write-back is needed between tasks; other blocks' data is read with
// device-scope loads. Empty tiles are marked finished for every pass and
// skipped.
//
// The last wave of a block (the 64 particles of every 1024 that come last,
// usually none) runs the control: it publishes the previous task, takes the
// next, loads its ranges, waits for its neighbours, and folds when its
// publication completes an iteration.

constexpr uint32_t kTaskDone = 0xffffffffu;  // tile_done of an empty tile
constexpr uint32_t kTaskWeight = 768;       // a task's cost beside its particles, in particles (measured)
constexpr uint32_t kResolvedStop = 0x80000000u;

__device__ __forceinline__ FlowWords& flow_words(const SolveArgs& a) { return a.prm.ctl->flow[a.which]; }

// Each solve's half of tile_done.
__device__ __forceinline__ uint32_t* done_of(const SolveArgs& a) {
  return a.tile_done + (size_t)a.which * a.tile_cap;
}

// Passes before the Jacobi passes: the density solve's viscosity pass.
__device__ __forceinline__ uint32_t flow_base(const SolveArgs& a) { return a.Vin ? 1u : 0u; }

// Particles of a tile (its eight processed column ranges).
__device__ __forceinline__ uint32_t tile_particles(const TileShape& sh, uint32_t tile, const uint32_t* start) {
  const uint32_t tx = tile / (sh.nty * sh.ntz), ty = (tile / sh.ntz) % sh.nty, tzi = tile % sh.ntz;
  const uint32_t z0 = tzi * sh.tz, z1 = min(z0 + sh.tz, sh.sz);
  uint32_t n = 0;
  for (uint32_t p = 0; p < kProcessedColumns; ++p) {
    const uint64_t x = 4 * tx + (p >> 1), y = 2 * ty + (p & 1);
    if (x < sh.sx && y < sh.sy) {
      const uint64_t base = (x * sh.sy + y) * sh.sz;
      n += start[base + z1] - start[base + z0];
    }
  }
  return n;
}

// The flow solve's tiles for a substep, by one block of THREADS threads (the
// density launch's last block, beside the density pass: the cell table holds
// until the advance, and nothing here reads what the density pass writes):
// the particles of every tile; the non-empty tiles listed in order
// (flow_list); the slabs of the XCD groups cut at equal weights, as ranges of
// that list (flow_begin: slab v is list entries begin[v] .. begin[v + 1]);
// empty tiles marked finished for every pass (tile_done = kTaskDone, in the
// divergence and the density solve's halves; their tile sums zero in both
// sets), the others not started; both solves' control words cleared. Thread
// t works the contiguous tiles of segment t. A substep whose density pass
// then flags its lists reruns, this with it.
template <uint32_t THREADS>
__device__ void flow_tiles_block(const StepParams& prm, const FlowSetup& f) {
  static_assert(THREADS % 64 == 0 && THREADS >= 64 * kXcds / 2, "whole waves, at most two slabs a wave");
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  const TileShape sh = tile_shape(ctl);
  if (sh.count > f.tile_cap) return;  // the solve flags it
  __shared__ uint32_t work_pre[THREADS + 1], live_pre[THREADS + 1];
  const uint32_t t = threadIdx.x, count = sh.count;
  const uint32_t per = (count + THREADS - 1) / THREADS;
  const uint32_t lo = min(t * per, count), hi = min(lo + per, count);
  uint32_t work = 0, live = 0;
  for (uint32_t k = lo; k < hi; ++k) {
    const uint32_t n = tile_particles(sh, k, prm.cell_start);
    f.tile_done[k] = n ? 0u : kTaskDone;
    f.tile_done[f.tile_cap + k] = n ? 0u : kTaskDone;
    if (!n) {
      f.tile_sum[k] = 0.0;
      f.tile_count[k] = 0u;
      f.tile_sum[f.tile_cap + k] = 0.0;
      f.tile_count[f.tile_cap + k] = 0u;
    }
    // a task costs about as much as kTaskWeight particles besides its particles:
    // slabs of sparse tiles take more tasks for the same particles
    const uint32_t w = n ? n + kTaskWeight : 0u;
    f.tile_work[k] = w;
    work += w;
    live += n ? 1u : 0u;
  }
  work_pre[t] = work;
  live_pre[t] = live;
  __syncthreads();
  // exclusive scans of both by one wave, THREADS / 64 segments a lane
  constexpr uint32_t kSeg = THREADS / 64;
  if (t < 64) {
    uint32_t rw = 0, rl = 0;
    for (uint32_t k = t * kSeg; k < t * kSeg + kSeg; ++k) {
      rw += work_pre[k];
      rl += live_pre[k];
    }
    uint32_t iw = rw, il = rl;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t uw = (uint32_t)__shfl_up((int)iw, d, 64), ul = (uint32_t)__shfl_up((int)il, d, 64);
      if ((int)t >= d) {
        iw += uw;
        il += ul;
      }
    }
    uint32_t aw = iw - rw, al = il - rl;
    for (uint32_t k = t * kSeg; k < t * kSeg + kSeg; ++k) {
      const uint32_t vw = work_pre[k], vl = live_pre[k];
      work_pre[k] = aw;
      live_pre[k] = al;
      aw += vw;
      al += vl;
    }
    if (t == 63) {
      work_pre[THREADS] = iw;
      live_pre[THREADS] = il;
    }
  }
  __syncthreads();
  const uint64_t total = work_pre[THREADS];
  const uint32_t pw = work_pre[t], pl = live_pre[t];
  // slab v starts at the first tile whose preceding tiles weigh at least
  // ceil(total * v / 8) (bt); its list entries start after the non-empty
  // tiles before it
  __shared__ uint32_t bt[kXcds + 1], bl[kXcds + 1];
  if (t == 0) {
    bt[0] = 0u;
    bl[0] = 0u;
    bt[kXcds] = count;
    bl[kXcds] = live_pre[THREADS];
  }
  for (uint32_t v = 1; v < kXcds; ++v) {
    const uint64_t target = (total * v + kXcds - 1) / kXcds;
    if (target == 0) {
      if (t == 0) {
        bt[v] = 0u;
        bl[v] = 0u;
      }
    } else if (pw < target && target <= (uint64_t)pw + work) {
      uint64_t run = pw;
      uint32_t k = lo, listed = pl;
      while (run < target) {
        const uint32_t n = f.tile_work[k++];
        run += n;
        listed += n ? 1u : 0u;
      }
      bt[v] = k;
      bl[v] = listed;
    }
  }
  __syncthreads();
  if (t <= kXcds) ctl->flow_begin[t] = bl[t];
  // Each slab's non-empty tiles listed y-major (tile row ty, then x, then
  // z): a task of pass q + 1 then needs the tasks of pass q a row or so
  // earlier in its slab's order (and the neighbouring slabs' at about the
  // same place in theirs), so a pass can start while the last one is still
  // finishing, where x-major order would make it wait for a whole plane.
  // Wave w lists slabs w, w + THREADS / 64, ..., 64 candidates at a time.
  const uint32_t plane = sh.nty * sh.ntz, lane = t & 63u, wave = t >> 6;
  for (uint32_t v = wave; v < kXcds; v += THREADS / 64) {
    if (bt[v] >= bt[v + 1]) continue;
    const uint32_t x0 = bt[v] / plane, x1 = (bt[v + 1] - 1) / plane, nx = x1 - x0 + 1;
    const uint32_t candidates = sh.nty * nx * sh.ntz;
    uint32_t at = bl[v];
    for (uint32_t c0 = 0; c0 < candidates; c0 += 64u) {
      const uint32_t e = c0 + lane;
      uint32_t idx = 0;
      bool keep = false;
      if (e < candidates) {
        const uint32_t ty = e / (nx * sh.ntz), tx = x0 + (e / sh.ntz) % nx, tz = e % sh.ntz;
        idx = (tx * sh.nty + ty) * sh.ntz + tz;
        keep = idx >= bt[v] && idx < bt[v + 1] && f.tile_work[idx] != 0u;
      }
      const uint64_t m = __ballot(keep);
      if (keep) f.tile_list[at + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = idx;
      at += (uint32_t)__popcll(m);
    }
  }
  // both solves' control words
  uint32_t* words = reinterpret_cast<uint32_t*>(&ctl->flow[0]);
  for (uint32_t k = t; k < 2 * sizeof(FlowWords) / 4; k += THREADS) words[k] = 0u;
}

// A task's neighbour wait, by the control wave: every tile within one of T
// has finished pass p - 1, and the push of iteration k >= 1 finds iteration
// k - 1 committed as going on (and its own iteration decided when the last
// decision marked it as likely the last). Returns 1 when the task may start,
// 0 to wait, -1 when the solve stopped before it (the task and every later
// one are dropped).
__device__ __forceinline__ int flow_ready(const SolveArgs& a, const TileShape& sh, uint32_t p, uint32_t tile,
                                          uint32_t lane) {
  const uint32_t base = flow_base(a);
  if (p >= base) {
    const uint32_t q = p - base;
    uint32_t r = 0;
    if (lane == 0) r = __hip_atomic_load(&flow_words(a).resolved, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r = __builtin_amdgcn_readfirstlane(r);
    if (r & kResolvedStop) {
      if (q >= 2u * (r & ~kResolvedStop) + 1u) return -1;
    } else if ((q & 1u) && q >= 3u && r < (q >> 1)) {
      return 0;
    } else if (q & 1u) {
      uint32_t c = 0;
      if (lane == 0) c = __hip_atomic_load(&flow_words(a).cautious, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      c = __builtin_amdgcn_readfirstlane(c);
      if (c == (q >> 1) + 1u && r < c) return 0;
    }
  }
  if (p == 0) return 1;
  const uint32_t* done = done_of(a);
  bool ok = true;
  if (lane < 27) {
    const uint32_t tx = tile / (sh.nty * sh.ntz), ty = (tile / sh.ntz) % sh.nty, tzi = tile % sh.ntz;
    const int64_t x = (int64_t)tx + (int)(lane / 9) - 1, y = (int64_t)ty + (int)((lane / 3) % 3) - 1,
                  z = (int64_t)tzi + (int)(lane % 3) - 1;
    if (x >= 0 && y >= 0 && z >= 0 && x < sh.ntx && y < sh.nty && z < sh.ntz) {
      const uint32_t u = (uint32_t)((x * sh.nty + y) * sh.ntz + z);
      ok = __hip_atomic_load(&done[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= p;
    }
  }
  return __ballot(!ok) == 0 ? 1 : 0;
}

// The control wave takes the block's next task from its group's queue: its
// pass and tile, kNoTile as the tile when the queue is done.
__device__ __forceinline__ uint2 flow_take(const SolveArgs& a, const uint32_t* begin, uint32_t lane) {
  const uint32_t v = blockIdx.x % kXcds;
  const uint32_t b = begin[v], len = begin[v + 1] - b;
  if (len == 0) return make_uint2(0u, kNoTile);
  uint32_t j = 0;
  if (lane == 0) j = __hip_atomic_fetch_add(&flow_words(a).next[v][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  j = __builtin_amdgcn_readfirstlane(j);
  const uint32_t p = j / len;
  if (p > 2u * a.max_iterations + flow_base(a)) return make_uint2(0u, kNoTile);
  return make_uint2(p, a.flow_list[b + j % len]);
}

// Waits until the task may start: 1, or -1 when the solve stopped before it,
// 0 when the wait gave up (the world flagged). A block waits only once it
// has published every task it finished, so no wait hangs on a task that is
// done but unpublished.
__device__ int flow_wait(const SolveArgs& a, const TileShape& sh, uint2 task, uint32_t lane) {
  const uint64_t t0 = wall_clock64();
  for (;;) {
    const int ready = flow_ready(a, sh, task.x, task.y, lane);
    if (ready != 0) return ready;
    __builtin_amdgcn_s_sleep(2);
    if (wall_clock64() - t0 > kBarrierTicks) {
      if (lane == 0) atomicOr(&a.prm.ctl->flags, kFlagBarrier);
      return 0;
    }
  }
}

// Commits decided iterations in order, by lane 0 of the control wave:
// resolved = m (iterations 0 .. m - 1 went on) advances while the decision
// of iteration m is in (decision[m % 2] = (m + 1) << 1 | stop).
__device__ void flow_commit(FlowWords& fw) {
  uint32_t r = __hip_atomic_load(&fw.resolved, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    if (r & kResolvedStop) return;
    const uint32_t d = __hip_atomic_load(&fw.decision[r & 1u][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((d >> 1) != r + 1) return;
    const uint32_t nv = (d & 1u) ? (kResolvedStop | r) : r + 1;
    if (__hip_atomic_compare_exchange_strong(&fw.resolved, &r, nv, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      r = nv;
  }
}

// Publishes a finished task, by the control wave once every wave of the block
// has waited for its stores. A rate task also counts toward its iteration;
// the publication that completes it folds the tile sums (lane l adds sums l,
// l + 64, ... in order, then the lanes by halving: the fixed shape), decides
// and commits.
__device__ void flow_publish(const SolveArgs& a, const TileShape& sh, uint2 task, const uint32_t* begin,
                             uint32_t groups, uint32_t log_base, uint32_t lane) {
  Control* ctl = a.prm.ctl;
  FlowWords& fw = flow_words(a);
  const uint32_t p = task.x, tile = task.y, base = flow_base(a);
  bool fold = false;
  if (lane == 0) {
    __hip_atomic_store(&done_of(a)[tile], p + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p >= base && !((p - base) & 1u)) {
      const uint32_t par = ((p - base) >> 1) & 1u, v = blockIdx.x % kXcds;
      const uint32_t c = __hip_atomic_fetch_add(&fw.rates[par][v][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (c + 1u == begin[v + 1] - begin[v]) {
        // the group's count starts again for iteration k + 2, which no task
        // reaches before this iteration's decision
        __hip_atomic_store(&fw.rates[par][v][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t g = __hip_atomic_fetch_add(&fw.groups[par][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g + 1u == groups) {
          __hip_atomic_store(&fw.groups[par][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          fold = true;
        }
      }
    }
  }
  fold = __shfl((int)fold, 0, 64) != 0;
  if (!fold) return;
  const uint32_t k = (p - base) >> 1, par = k & 1u;
  const double* sums = a.tile_sum + (size_t)par * a.tile_cap;
  const uint32_t* counts = a.tile_count + (size_t)par * a.tile_cap;
  double s = 0.0;
  uint32_t cn = 0;
  uint32_t m = lane;
  for (; m + 7u * 64u < sh.count; m += 8u * 64u) {
    double v[8];
    uint32_t w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      v[u] = coherent_f64(sums, m + u * 64u);
      w[u] = coherent_u32(counts, m + u * 64u);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      s += v[u];
      cn += w[u];
    }
  }
  for (; m < sh.count; m += 64u) {
    s += coherent_f64(sums, m);
    cn += coherent_u32(counts, m);
  }
  wave_fold(s, cn);
  if (lane == 0) {
    const double average = s / (double)(cn > 1u ? cn : 1u);
    const uint32_t at = log_base + k;
    if (at < (uint32_t)kResidualLogCap) {
      ctl->residual_log[at] = a.which == 0 ? average : -average;
      ctl->loaded_log[at] = cn;
    }
    const bool settled = k >= a.min_iterations && average <= (double)a.tolerance;
    const uint32_t stop = (settled || k >= a.max_iterations) ? 1u : 0u;
    // the tolerance is met and only the least number of iterations held the
    // stop back: the next iteration is very likely the last, so its push
    // waits for its own decision instead of running ahead of it (nearer
    // the tolerance alone, every push running ahead measured faster)
    if (!stop && k + 1u >= a.min_iterations && average <= (double)a.tolerance)
      __hip_atomic_store(&fw.cautious, k + 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    __hip_atomic_store(&fw.decision[par][0], (k + 1u) << 1 | stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    flow_commit(fw);
  }
}

// The flow passes of one block. Each round: stage the task's neighbourhood,
// then the control wave publishes the last round's task (every wave has
// waited for its stores), takes the next task and checks once whether it may
// start, while the other waves sum. A next task not found ready is waited
// for at the start of its round, after the last task is published.
template <int FORMAT>
__device__ void flow_passes(const SolveArgs& a, const TileShape& sh, float4* lds, TileState (&tstate)[3],
                            uint2* task_s, int* ready_s, double* wsum, uint32_t* wcnt, const uint32_t* begin,
                            uint32_t log_base) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  constexpr uint32_t kControl = kSolveWaves - 1;
  const uint32_t base = flow_base(a);
  uint32_t groups = 0;
  for (uint32_t v = 0; v < kXcds; ++v) groups += begin[v + 1] > begin[v] ? 1u : 0u;
  if (wave == kControl) {
    uint2 first = flow_take(a, begin, lane);
    if (first.y != kNoTile && flow_wait(a, sh, first, lane) <= 0) first.y = kNoTile;
    tile_scan(tile_load(sh, first.y, first.y != kNoTile, a.prm.cell_start, lane), tstate[0], lane);
    if (lane == 0) {
      task_s[0] = first;
      ready_s[0] = 1;
    }
  }
  __syncthreads();
  uint2 pending = make_uint2(0u, kNoTile);  // the last round's task, not yet published
  uint32_t r = 0;
  // One round; false when the block has no task left. The viscosity rounds
  // (VISCOUS) and the Jacobi rounds are separate loops, so that neither's
  // registers crowd the other's.
  auto round = [&](auto viscous_tag) -> bool {
    constexpr bool viscous = decltype(viscous_tag)::value;
    const uint32_t cur = r & 1u, nxt = cur ^ 1u;
    uint2 task = task_s[cur];
    if (task.y == kNoTile) return false;
    if (!ready_s[cur]) {
      // publish, then wait
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (wave == kControl) {
        if (pending.y != kNoTile) flow_publish(a, sh, pending, begin, groups, log_base, lane);
        if (flow_wait(a, sh, task, lane) <= 0 && lane == 0) task_s[cur].y = kNoTile;
      }
      pending.y = kNoTile;
      __syncthreads();
      task = task_s[cur];
      if (task.y == kNoTile) return false;
    }
    const uint32_t p = task.x;
    const uint32_t q = viscous ? 0u : p - base, k = q >> 1;
    const bool rate = !viscous && !(q & 1u);
    float4* vcur = (k & 1u) ? a.V2 : a.V;
    float4* vnext = (k & 1u) ? a.V : a.V2;
    const Glob g{a.P, vcur, a.kappa};
    // the LDS: A (float4) then B (float2) for the rate pass, Q (float4) for
    // the push pass, A then C (float4: vy, vz, density) for the viscosity
    // pass, which stages 32 bytes a particle instead of 24
    const uint32_t cap = viscous ? a.lds_cap * kTileBytes / 32u : a.lds_cap;
    float4* A = lds;
    float2* B = reinterpret_cast<float2*>(lds + cap);
    float4* Q = lds;
    float4* C = lds + cap;
    const TileState& ts = tstate[cur];
    const uint32_t stride = __builtin_amdgcn_readfirstlane(ts.stride);
    const uint32_t count = ts.ppre[kProcessedColumns];
    const bool fits = kStagedColumns * stride <= cap;
    if (fits && count) stage_tile(a, g, viscous ? kViscousPass : rate ? kRatePass : kPushPass, ts, stride, A, B, Q, C);
    // this round's staging and the last round's stores (tile sum included)
    // complete in every wave before the last round's task is published
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    uint2 next = make_uint2(0u, kNoTile);
    int next_ready = 1;
    TileLoad nl{0u, 0u, kNoTile};
    if (wave == kControl) {
      if (pending.y != kNoTile) flow_publish(a, sh, pending, begin, groups, log_base, lane);
      next = flow_take(a, begin, lane);
      if (next.y != kNoTile) {
        next_ready = flow_ready(a, sh, next.x, next.y, lane);
        if (next_ready < 0) next.y = kNoTile;
      }
      nl = tile_load(sh, next.y, next.y != kNoTile, a.prm.cell_start, lane);
    }
    double acc = 0.0;
    uint32_t cnt = 0;
    for (uint32_t p0 = t; p0 < count; p0 += kSolveThreads) {
      uint32_t pc;
      const uint32_t i = processed_index(ts, p0, pc);
      const Lds l{A, B, Q, &ts, pc, 4u * (pc >> 1) + (pc & 1u), stride};
      if constexpr (viscous) {
        viscous_one<FORMAT>(a, i, &l, C, fits);
      } else if (rate) {
        const float e = rate_one<FORMAT, false>(a, i, g, &l, fits, 0, 0);
        acc += (double)e;
        cnt += e > 0.0f ? 1u : 0u;
      } else {
        push_one<FORMAT, false>(a, i, g, &l, fits, 0, 0, vnext);
      }
    }
    if (rate) {
      wave_fold(acc, cnt);
      if (lane == 0) {
        wsum[wave] = acc;
        wcnt[wave] = cnt;
      }
    }
    if (wave == kControl) {
      tile_scan(nl, tstate[nxt], lane);
      if (lane == 0) {
        task_s[nxt] = next;
        ready_s[nxt] = next_ready > 0 ? 1 : 0;
      }
    }
    __syncthreads();
    if (rate && wave == 0) {
      // the tile's error sum, through the L2, in its iteration's set
      double v = lane < kSolveWaves ? wsum[lane] : 0.0;
      uint32_t c = lane < kSolveWaves ? wcnt[lane] : 0u;
      for (int h = kSolveWaves / 2; h >= 1; h >>= 1) {
        v = v + __shfl_down(v, h, 64);
        c += (uint32_t)__shfl_down((int)c, h, 64);
      }
      if (lane == 0) {
        const uint32_t at = (k & 1u) * a.tile_cap + task.y;
        const uint64_t bits = (uint64_t)__double_as_longlong(v);
        const u32x2 w = {(uint32_t)bits, (uint32_t)(bits >> 32)};
        __builtin_amdgcn_raw_buffer_store_b64(w, whole(a.tile_sum), at * 8u, 0, kCoherentLoad);
        __builtin_amdgcn_raw_buffer_store_b32(c, whole(a.tile_count), at * 4u, 0, kCoherentLoad);
      }
    }
    pending = task;
    ++r;
    return true;
  };
  // the viscosity rounds while the next task is one, then the Jacobi rounds
  while (task_s[r & 1u].y != kNoTile && task_s[r & 1u].x < base && round(std::true_type{})) {
  }
  while (round(std::false_type{})) {
  }
  // the last task: its stores complete, then it is published
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (wave == kControl && pending.y != kNoTile) flow_publish(a, sh, pending, begin, groups, log_base, lane);
}

// The flow solve (tiled, passes waiting tile by tile instead of at grid
// barriers) as a kernel of its own, so that its control code does not take
// registers from the barrier kernel's passes. One grid barrier at the end;
// the iteration the solve stopped at left its velocities in V2 when odd, and
// they are copied to V.
template <int FORMAT>
__global__ __launch_bounds__(1024) void k_dfsph_flow(SolveArgs a) {
  extern __shared__ float4 lds[];
  const StepParams& prm = a.prm;
  Control* ctl = prm.ctl;
  if (list_aborted(ctl)) return;
  const uint32_t t = threadIdx.x;
  __shared__ double wsum[kSolveWaves];
  __shared__ uint32_t wcnt[kSolveWaves];
  __shared__ TileState tstate[3];
  __shared__ TileShape shape_s;
  __shared__ uint32_t begin_s[kXcds + 1];
  __shared__ uint2 task_s[2];
  __shared__ int ready_s[2];
  __shared__ uint32_t log_base_s;
  if (t == 0) {
    shape_s = tile_shape(ctl);
    log_base_s = ctl->log_len;
  }
  if (t <= kXcds) begin_s[t] = ctl->flow_begin[t];
  __syncthreads();
  const TileShape& sh = shape_s;
  if (sh.count > a.tile_cap) {
    if (blockIdx.x == 0 && t == 0) atomicOr(&ctl->flags, kFlagTiles);
    return;
  }
  const uint32_t log_base = log_base_s;
  flow_passes<FORMAT>(a, sh, lds, tstate, task_s, ready_s, wsum, wcnt, begin_s, log_base);
  if (!grid_barrier(ctl, false, &flow_words(a).resolved)) return;
  const uint32_t r = __hip_atomic_load(&flow_words(a).resolved, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!(r & kResolvedStop)) return;  // a wait gave up; the world is flagged
  const uint32_t stopped = r & ~kResolvedStop;
  if (stopped & 1u)
    for (uint32_t i = blockIdx.x * kSolveThreads + t; i < prm.n; i += gridDim.x * kSolveThreads)
      a.V[i] = coherent4(a.V2, i);
  if (blockIdx.x == 0 && t == 0) {
    ctl->iterations[a.which] = stopped;
    ctl->iterations_step[a.which] += stopped;
    ctl->rates_step += stopped + 1;
    ctl->log_len = min(log_base + stopped + 1, (uint32_t)kResidualLogCap);
  }
}

template <int FORMAT, bool TILED>
__global__ __launch_bounds__(1024) void k_dfsph_solve(SolveArgs a) {
  extern __shared__ float4 lds[];
  const StepParams& prm = a.prm;
  Control* ctl = prm.ctl;
  if (list_aborted(ctl)) return;
  const int64_t szz = ctl->span[2], sysz = (int64_t)ctl->span[1] * szz;
  const uint32_t xcd = blockIdx.x % kXcds, per_xcd = gridDim.x / kXcds, rank = blockIdx.x / kXcds;
  const uint32_t c_begin = (uint32_t)((uint64_t)a.chunks * xcd / kXcds);
  const uint32_t c_end = (uint32_t)((uint64_t)a.chunks * (xcd + 1) / kXcds);
  const uint32_t t = threadIdx.x;
  const uint32_t half = t / kChunk, tc = t % kChunk;
  __shared__ double wsum[kSolveWaves];
  __shared__ uint32_t wcnt[kSolveWaves];
  __shared__ TileState tstate[3];
  __shared__ uint32_t grab_s[2];
  __shared__ TileShape shape_s;  // read once: the passes start without a global load

  // the stop sum's room is sized for the largest cell box the table holds;
  // every block sees the same count and leaves before any barrier
  if (t == 0) shape_s = tile_shape(ctl);
  __syncthreads();
  const TileShape& sh = shape_s;
  if (sh.count > a.tile_cap) {
    if (blockIdx.x == 0 && t == 0) atomicOr(&ctl->flags, kFlagTiles);
    return;
  }
  bool tiled = false;
  if constexpr (TILED) tiled = a.lds_cap > 0;
  const Glob g{a.P, a.V, a.kappa};
  if constexpr (TILED) {
    // the ranges of the block's fixed tile, the same in every pass
    if (tiled && t < 64) {
      const uint32_t group = blockIdx.x % kXcds;
      const bool have = rank < fixed_tiles(sh, group);
      const uint32_t tile = group_begin(sh, group) + rank;
      tile_scan(tile_load(sh, tile, have, prm.cell_start, t), tstate[2], t);
    }
    __syncthreads();
  }
  uint32_t it = 0, pass = 0;
  if constexpr (TILED) {
    if (tiled && a.Vin) {
      // the density solve's launch runs the viscosity pass first
      if (blockIdx.x == 0 && t < kXcds)
        __hip_atomic_store(&ctl->tile_next[(pass + 1) & 1u][t][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      tiled_pass<FORMAT>(a, sh, kViscousPass, pass & 1u, lds, tstate, grab_s, g, wsum, wcnt);
      ++pass;
      if (!grid_barrier(ctl, !tiled)) return;
    }
  }
  for (;;) {
    // rate pass: stiffness and error of every particle, and the tile sums
    if (tiled) {
      // the queues of the next pass start again from zero
      if (blockIdx.x == 0 && t < kXcds)
        __hip_atomic_store(&ctl->tile_next[(pass + 1) & 1u][t][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if constexpr (TILED) tiled_pass<FORMAT>(a, sh, kRatePass, pass & 1u, lds, tstate, grab_s, g, wsum, wcnt);
      ++pass;
      if (!grid_barrier(ctl, !tiled)) return;
    } else {
      for (uint32_t c = c_begin + rank + half * per_xcd; c < c_end; c += 2 * per_xcd) {
        uint32_t i = c * kChunk + tc;
        if (i < prm.n) a.errors[i] = rate_one<FORMAT, !TILED>(a, i, g, nullptr, false, sysz, szz);
      }
      if (!grid_barrier(ctl, !tiled)) return;
      tile_sums_from_errors(a, sh, tstate[0], wsum, wcnt);
      if (!grid_barrier(ctl, !tiled)) return;
    }
    // every wave folds the tile sums the same way
    uint32_t loaded = 0;
    double total;
    if (tiled)
      total = fold_tiles_lds(a, sh.count, lds, a.lds_cap * kTileBytes / 12u / 64u * 64u, loaded);
    else
      total = fold_tiles(a, sh.count, loaded);
    double average = total / (double)(loaded > 1u ? loaded : 1u);
    if (blockIdx.x == 0 && t == 0) {
      uint32_t at = ctl->log_len;
      if (at < (uint32_t)kResidualLogCap) {
        ctl->residual_log[at] = a.which == 0 ? average : -average;
        ctl->loaded_log[at] = loaded;
        ctl->log_len = at + 1;
      }
    }
    bool settled = it >= a.min_iterations && average <= (double)a.tolerance;
    if (settled || it >= a.max_iterations) break;
    // push pass
    if (tiled) {
      if (blockIdx.x == 0 && t < kXcds)
        __hip_atomic_store(&ctl->tile_next[(pass + 1) & 1u][t][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if constexpr (TILED) tiled_pass<FORMAT>(a, sh, kPushPass, pass & 1u, lds, tstate, grab_s, g, wsum, wcnt);
      ++pass;
    } else {
      for (uint32_t c = c_begin + rank + half * per_xcd; c < c_end; c += 2 * per_xcd) {
        uint32_t i = c * kChunk + tc;
        if (i < prm.n) push_one<FORMAT, !TILED>(a, i, g, nullptr, false, sysz, szz, a.V);
      }
    }
    if (!grid_barrier(ctl, !tiled)) return;
    ++it;
  }
  // every block is past the last pass: leave both queues at zero for the
  // next launch
  if (blockIdx.x == 0 && t < 2 * kXcds)
    __hip_atomic_store(&ctl->tile_next[t / kXcds][t % kXcds][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (blockIdx.x == 0 && t == 0) {
    ctl->iterations[a.which] = it;
    ctl->iterations_step[a.which] += it;
    ctl->rates_step += it + 1;
  }
}

// The viscosity pass of dfsph_substep (dfsph.rs:90-99) as a kernel of its
// own, for the solves without tiles: viscous (src/fluid/step.rs:229-251, with
// wall_viscous step.rs:262-301) + gravity, v += push * dt, over the
// substep's lists. Reads VB, writes VA (density carried along). The tiled
// density solve runs the same arithmetic first in its own launch.
template <int FORMAT>
__global__ __launch_bounds__(256) void k_dfsph_viscous(const float4* __restrict__ PB, const float4* __restrict__ VB,
                                                       StepParams prm, float4* __restrict__ VA,
                                                       float* __restrict__ forces) {
  Control* ctl = prm.ctl;
  if (list_aborted(ctl)) return;
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= prm.n) return;
  const float4 at = PB[i];
  const float4 vi = VB[i];
  const uint32_t counts = prm.pair_counts[i];
  const uint32_t nf = counts & 0xffffu, nw = counts >> 16;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  auto fluid = [&](uint32_t j, float f) {
    const float4 pj = PB[j];
    viscous_pair(prm, at, vi, make_float3(pj.x, pj.y, pj.z), VB[j], f, px, py, pz);
  };
  if constexpr (FORMAT == kLists) {
    const uint2* p = prm.pairs + i;
    for (uint32_t k = 0; k < nf; ++k) {
      uint2 e = p[(size_t)k * prm.pair_stride];
      const uint32_t j = list_index(e.x);
      if (j != i) fluid(j, __uint_as_float(e.y));
    }
  } else if constexpr (FORMAT == kCompact) {
    // offsets count from the first particle of each neighbour column's
    // cells from z0 - 1 on, z0 the first cell of the particle's tile: those
    // firsts go to LDS once (a row per lane), then entries go four at a
    // time, each batch's gathers issued together
    __shared__ uint32_t firsts[11][256];
    const int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
    const uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
    const int32_t cx = cell_coord(at.x, prm.radius), cy = cell_coord(at.y, prm.radius);
    const int32_t cz = cell_coord(at.z, prm.radius);
    const uint32_t depth = tile_depth(span[2]);
    const uint32_t z0 = (uint32_t)((int64_t)cz - lo[2]) / depth * depth;
    const int64_t za = (int64_t)lo[2] + (z0 > 0 ? z0 - 1 : 0);
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (uint32_t code = 0; code < 11; ++code) {
      if ((code & 3u) == 3u) continue;  // no column: dy would be 2
      const int64_t first = cell_slot((int64_t)cx + (int64_t)(code >> 2) - 1,
                                      (int64_t)cy + (int64_t)(code & 3u) - 1, za, lo, span);
      firsts[code][t] = first >= 0 ? prm.cell_start[first] : 0u;
    }
    for (uint32_t k = 0; k < nf; k += 4) {
      Ent e[4];
      fluid_entries<kCompact>(prm, i, k, e);
      uint32_t j[4];
      float4 pj[4], vj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) j[u] = k + u < nf ? firsts[e[u].r >> kNearBits][t] + (e[u].r & kNearMask) : i;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        pj[u] = PB[j[u]];
        vj[u] = VB[j[u]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k + u < nf && j[u] != i)
          viscous_pair(prm, at, vi, make_float3(pj[u].x, pj[u].y, pj[u].z), vj[u], e[u].f, px, py, pz);
    }
  } else {
    const int64_t szz = ctl->span[2], sysz = (int64_t)ctl->span[1] * szz;
    mask_walk<FORMAT>(prm, i, prm.own_slot[i], sysz, szz, [&](const uint32_t(&j)[4], const bool(&ok)[4]) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (ok[u] && j[u] != i) {
          float4 pj = PB[j[u]];
          float ox = at.x - pj.x, oy = at.y - pj.y, oz = at.z - pj.z;
          fluid(j[u], pair_factor(dot3(ox, oy, oz, ox, oy, oz), prm));
        }
      }
    });
  }
  viscous_finish<FORMAT>(prm, i, at, vi, px, py, pz, nf, nw, VA, forces);
}

// advance (src/fluid/step.rs:305-339): PB + VA*dt -> PA, VA. The velocity
// before the advance is kept in VKEEP for the repair after a non-finite one.
__global__ __launch_bounds__(256) void k_dfsph_advance(const float4* __restrict__ PB, StepParams prm,
                                                       float4* __restrict__ PA, float4* __restrict__ VA,
                                                       float4* __restrict__ VKEEP) {
  Control* ctl = prm.ctl;
  if (list_aborted(ctl)) return;
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  bool live = i < prm.n;
  float speed = 0.0f;
  uint32_t held = 0;
  if (live) {
    float4 at = PB[i];
    float4 v = VA[i];
    VKEEP[i] = v;
    float vx = v.x, vy = v.y, vz = v.z, nx, ny, nz;
    bool ok = advance_one(prm, at, vx, vy, vz, nx, ny, nz, held);
    if (!ok) {
      atomicOr(&ctl->flags, kFlagNonFinite);
      atomicMin(&ctl->bad_index, i);
    }
    PA[i] = make_float4(nx, ny, nz, at.w);
    VA[i] = make_float4(vx, vy, vz, v.w);
    speed = sqrtf(dot3(vx, vy, vz, vx, vy, vz));
  }
  wave_max_to(&stripe_of(ctl).vmax_bits, speed);
  wave_sum_to(&stripe_of(ctl).held, held);
}

// After a non-finite advance: particles past the first bad one go back to
// their old position and pre-advance velocity (kept in VKEEP); the bad one
// keeps its old position and its zeroed velocity components, as the
// reference leaves them (step.rs:318-335).
__global__ __launch_bounds__(256) void k_dfsph_repair(const float4* __restrict__ PB, const float4* __restrict__ VKEEP,
                                                      StepParams prm, float4* __restrict__ PA,
                                                      float4* __restrict__ VA) {
  Control* ctl = prm.ctl;
  uint32_t bad = ctl->bad_index;
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= prm.n || i < bad) return;
  PA[i] = PB[i];
  if (i > bad) VA[i] = VKEEP[i];
}

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

#define SFL_FORMAT_LAUNCH(KERNEL, format, ...)                                    \
  do {                                                                            \
    if ((format) == kMasks16)                                                     \
      hipLaunchKernelGGL(KERNEL<kMasks16>, __VA_ARGS__);                          \
    else if ((format) == kMasks32)                                                \
      hipLaunchKernelGGL(KERNEL<kMasks32>, __VA_ARGS__);                          \
    else if ((format) == kCompact)                                                \
      hipLaunchKernelGGL(KERNEL<kCompact>, __VA_ARGS__);                          \
    else                                                                          \
      hipLaunchKernelGGL(KERNEL<kLists>, __VA_ARGS__);                            \
  } while (0)

// The same with a second template argument after the format.
#define SFL_FORMAT_LAUNCH_WITH(KERNEL, format, ARG, ...)                          \
  do {                                                                            \
    if ((format) == kMasks16)                                                     \
      hipLaunchKernelGGL((KERNEL<kMasks16, ARG>), __VA_ARGS__);                   \
    else if ((format) == kMasks32)                                                \
      hipLaunchKernelGGL((KERNEL<kMasks32, ARG>), __VA_ARGS__);                   \
    else if ((format) == kCompact)                                                \
      hipLaunchKernelGGL((KERNEL<kCompact, ARG>), __VA_ARGS__);                   \
    else                                                                          \
      hipLaunchKernelGGL((KERNEL<kLists, ARG>), __VA_ARGS__);                     \
  } while (0)

hipError_t launch_dfsph_density(hipStream_t s, const float4* P, float4* V, float* alpha, const StepParams& prm,
                                const FlowSetup& setup) {
  if (prm.n == 0) return hipSuccess;
  const bool ready = setup.tile_done && setup.tile_work && setup.tile_list && setup.tile_sum && setup.tile_count;
  const FlowSetup f = ready ? setup : FlowSetup();
  const uint32_t blocks = blocks_for(prm.n, 256) + (ready ? 1u : 0u);
  if (prm.wall_keys)
    SFL_FORMAT_LAUNCH_WITH(k_dfsph_density, prm.format, true, blocks, 256, 0, s, P, V, alpha, prm, f);
  else
    SFL_FORMAT_LAUNCH_WITH(k_dfsph_density, prm.format, false, blocks, 256, 0, s, P, V, alpha, prm, f);
  return hipGetLastError();
}

// Blocks of the persistent solve. Every block must be resident at once for
// the grid barrier: at most what the occupancy query and the resident
// threads per CU allow, and one 1024-thread block per CU by default: the
// barrier's cost grows with the blocks that meet at it, and one block per
// CU ran the 1M-particle solve fastest. The list format's solve also takes
// the LDS left beside its static arrays for its tiles (`tile_bytes`, 0 when
// `tiles` is false or one block per CU is not the grid).
int solve_grid_limit(int blocks_per_cu, bool tiles, uint32_t* tile_bytes) {
  *tile_bytes = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  int cus = 0, threads = 0, lds = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&threads, hipDeviceAttributeMaxThreadsPerMultiProcessor, dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
  int per = 0;
  // the smallest occupancy of the solve kernels
  const void* kernels[7] = {reinterpret_cast<const void*>(k_dfsph_solve<kMasks16, false>),
                            reinterpret_cast<const void*>(k_dfsph_solve<kMasks32, false>),
                            reinterpret_cast<const void*>(k_dfsph_solve<kLists, false>),
                            reinterpret_cast<const void*>(k_dfsph_solve<kLists, true>),
                            reinterpret_cast<const void*>(k_dfsph_solve<kCompact, true>),
                            reinterpret_cast<const void*>(k_dfsph_flow<kLists>),
                            reinterpret_cast<const void*>(k_dfsph_flow<kCompact>)};
  per = 1 << 30;
  for (const void* k : kernels) {
    int p = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, k, kSolveThreads, 0) != hipSuccess) return 0;
    per = p < per ? p : per;
  }
  per = per < threads / (int)kSolveThreads ? per : threads / (int)kSolveThreads;
  int cap = blocks_per_cu > 0 && blocks_per_cu <= 2 ? blocks_per_cu : 1;
  per = per < cap ? per : cap;
  if (tiles && per == 1) {
    // both tiled kernels (list and compact formats) get the same room
    const void* tiled[4] = {reinterpret_cast<const void*>(k_dfsph_solve<kLists, true>),
                            reinterpret_cast<const void*>(k_dfsph_solve<kCompact, true>),
                            reinterpret_cast<const void*>(k_dfsph_flow<kLists>),
                            reinterpret_cast<const void*>(k_dfsph_flow<kCompact>)};
    int64_t room = (int64_t)lds;
    for (const void* k : tiled) {
      hipFuncAttributes attr;
      if (hipFuncGetAttributes(&attr, k) != hipSuccess) return 0;
      room = std::min<int64_t>(room, (int64_t)lds - (int64_t)attr.sharedSizeBytes);
    }
    // whole staged particles, rounded down to 64
    room = room / kTileBytes / 64 * 64 * kTileBytes;
    if (room >= kTileBytes * 1024) {
      bool fit = true;
      for (const void* k : tiled) {
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)room) != hipSuccess) return 0;
        int p = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, k, kSolveThreads, (size_t)room) != hipSuccess) return 0;
        fit = fit && p >= 1;
      }
      if (fit) *tile_bytes = (uint32_t)room;
    }
  }
  return per * cus;
}

bool dfsph_solve_tiled(const StepParams& prm, uint32_t tile_bytes) {
  return (prm.format == kLists || prm.format == kCompact) && tile_bytes > 0;
}

hipError_t launch_dfsph_solve(hipStream_t s, const StepParams& prm, const float4* P, float4* V, const float* alpha,
                              float4* kappa, float* errors, double* tile_sum, uint32_t* tile_count,
                              uint32_t tile_cap, int which, float tolerance, uint32_t min_it, uint32_t max_it,
                              int grid, uint32_t tile_bytes, const float4* vin, float* forces,
                              const FlowBuffers& flow) {
  if (prm.n == 0) return hipSuccess;
  SolveArgs a;
  a.prm = prm;
  a.P = P;
  a.V = V;
  a.alpha = alpha;
  a.kappa = kappa;
  a.errors = errors;
  a.tile_sum = tile_sum;
  a.tile_count = tile_count;
  a.tile_cap = tile_cap;
  a.chunks = (prm.n + kChunk - 1) / kChunk;
  a.which = which;
  a.tolerance = tolerance;
  a.min_iterations = min_it;
  a.max_iterations = max_it;
  // the compact format is read by the tiled kernel alone
  if (prm.format == kCompact && tile_bytes == 0) return hipErrorInvalidValue;
  const bool tiled = dfsph_solve_tiled(prm, tile_bytes);
  // the viscosity pass runs inside the tiled density solve only
  if (vin && !(tiled && which == 1)) return hipErrorInvalidValue;
  a.Vin = vin;
  a.forces = forces;
  a.lds_cap = tiled ? tile_bytes / kTileBytes : 0u;
  const bool flowing = tiled && flow.v2 && flow.tile_done && flow.tile_work && flow.tile_list;
  a.V2 = flowing ? flow.v2 : nullptr;
  a.tile_done = flow.tile_done;
  a.tile_work = flow.tile_work;
  a.flow_list = flow.tile_list;
  a.through = flowing ? 1u : 0u;
  // a multiple of 8 blocks so every XCD gets the same share; at least one
  // chunk per XCD group or the small grid of a small world
  uint32_t blocks = (uint32_t)grid;
  uint32_t want = (a.chunks + kXcds - 1) / kXcds * kXcds;
  if (want < blocks) blocks = want;
  blocks = blocks / kXcds * kXcds;
  if (blocks == 0) blocks = kXcds;
  // A cooperative launch: the runtime refuses up front a grid whose blocks
  // cannot all be resident at once, which the grid barrier needs (the grid
  // is sized by the occupancy query, so that refusal means the device is
  // not the one the query saw); the barrier's 2 s bound stays as the last
  // guard.
  const void* fn;
  size_t lds = 0;
  if (flowing) {
    fn = prm.format == kCompact ? reinterpret_cast<const void*>(k_dfsph_flow<kCompact>)
                                : reinterpret_cast<const void*>(k_dfsph_flow<kLists>);
    lds = tile_bytes;
  } else if (tiled && prm.format == kCompact) {
    fn = reinterpret_cast<const void*>(k_dfsph_solve<kCompact, true>);
    lds = tile_bytes;
  } else if (tiled) {
    fn = reinterpret_cast<const void*>(k_dfsph_solve<kLists, true>);
    lds = tile_bytes;
  } else if (prm.format == kMasks16) {
    fn = reinterpret_cast<const void*>(k_dfsph_solve<kMasks16, false>);
  } else if (prm.format == kMasks32) {
    fn = reinterpret_cast<const void*>(k_dfsph_solve<kMasks32, false>);
  } else {
    fn = reinterpret_cast<const void*>(k_dfsph_solve<kLists, false>);
  }
  // SOLUNAR_SOLVE_COOPERATIVE=0 launches plainly instead (same grid, same
  // kernel), for tools that do not follow cooperative launches
  static const bool cooperative = [] {
    const char* v = std::getenv("SOLUNAR_SOLVE_COOPERATIVE");
    return !(v && std::atoi(v) == 0);
  }();
  void* args[] = {&a};
  if (cooperative)
    return hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(kSolveThreads), args, (unsigned int)lds, s);
  return hipLaunchKernel(fn, dim3(blocks), dim3(kSolveThreads), args, lds, s);
}

void launch_dfsph_viscous(hipStream_t s, const float4* PB, const float4* VB, const StepParams& prm, float4* VA,
                          float* forces) {
  if (prm.n == 0) return;
  SFL_FORMAT_LAUNCH(k_dfsph_viscous, prm.format, blocks_for(prm.n, 256), 256, 0, s, PB, VB, prm, VA, forces);
}

void launch_dfsph_advance(hipStream_t s, const float4* PB, const StepParams& prm, float4* PA, float4* VA,
                          float4* VKEEP) {
  if (prm.n == 0) return;
  hipLaunchKernelGGL(k_dfsph_advance, blocks_for(prm.n, 256), 256, 0, s, PB, prm, PA, VA, VKEEP);
}

void launch_dfsph_repair(hipStream_t s, const float4* PB, const float4* VKEEP, const StepParams& prm, float4* PA,
                         float4* VA) {
  if (prm.n == 0) return;
  hipLaunchKernelGGL(k_dfsph_repair, blocks_for(prm.n, 256), 256, 0, s, PB, VKEEP, prm, PA, VA);
}

}  // namespace sfl
