// Device-side types and arithmetic shared by the gfx950 kernels of the SPH
// fluid step.
//
// Every floating-point expression here is written in the operation order of
// the reference's Rust code (src/fluid/mod.rs, step.rs, dfsph.rs) and the
// library is compiled with -ffp-contract=off and correctly rounded f32
// division and square root, so each pair term is the same IEEE single value
// the CPU step computes. Sums over a particle's neighbours are left folds in
// the reference's list order (cells dz, dy, dx; ids ascending within a
// cell), one lane per particle, which keeps densities and forces bit-equal
// to the CPU step rather than merely close.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfl {

// Control flags (sticky within a step). Kernels never test them: they test
// `Control::abort`, which only changes between launches (an overflow found by
// the grid setup, a non-finite advance promoted at the next substep's start),
// so every lane of a launch sees one value.
constexpr uint32_t kFlagOverflow = 1u;   // the cell box outgrew the table
constexpr uint32_t kFlagNonFinite = 2u;  // an advance produced a non-finite value
constexpr uint32_t kFlagBarrier = 4u;    // a grid barrier gave up waiting

constexpr uint32_t kFlagPairs = 8u;     // a particle had more neighbours than the list holds
constexpr uint32_t kFlagWide = 32u;     // a neighbour sat past the width of its cell's mask
constexpr uint32_t kFlagTiles = 64u;    // the solve tiles outnumbered the stop sum's room
constexpr uint32_t kFlagMoverBox = 128u;  // the body points' cell box outgrew the sort key
constexpr uint32_t kFlagPairRoom = 256u;  // the body-point pairs outgrew their records

// A drain's or the domain's box, and an emitter's lattice (flows after a
// step, kernels_flow.hip).
struct FlowBox {
  float lo[3], hi[3];
};
struct EmitBox {
  float lo[3];
  float spacing;
  int64_t cells[3];  // lattice slots per axis (x, y, z)
  float clear;       // (0.9 spacing)^2
  float vel[3];
};
// An emitter's lattice slots, at most (the reference sizes its slot flags by
// the box, up to 4096 per axis; past this a box is refused).
constexpr int64_t kEmitterSlotsMax = int64_t(1) << 28;

constexpr int kResidualLogCap = 256;
constexpr int kStripes = 64;

// One of 64 copies of the report reductions; a block folds into stripe
// blockIdx % 64, so no single address takes every wave's atomic.
struct alignas(64) Stripe {
  int32_t bmin[3];
  int32_t bmax[3];
  uint32_t vmax_bits;  // bits of max |v| after the last advance
  uint32_t comp_bits;  // bits of max(rho/rest - 1, 0), last density pass
  uint32_t held;       // particles held back this step
  uint32_t pairs_max;  // most list entries any particle needed, this substep
  uint32_t pad[6];
};

// One flow solve's control words: each XCD group's task queue, its rate
// tasks finished per iteration parity, the groups finished per parity, each
// iteration's decision per parity, the decisions committed in iteration
// order, and the iteration whose push waits for its own decision (m + 1,
// or 0); one cache line each.
struct FlowWords {
  alignas(128) uint32_t next[8][32];
  alignas(128) uint32_t rates[2][8][32];
  alignas(128) uint32_t groups[2][32];
  alignas(128) uint32_t decision[2][32];
  alignas(128) uint32_t resolved;
  alignas(128) uint32_t cautious;
};

// Device-resident control block of one world: cell box, report reductions,
// solve counters and the persistent kernels' barrier.
struct Control {
  int32_t lo[3];
  uint32_t span[3];
  uint32_t volume;
  uint32_t flags;
  uint32_t abort;        // non-zero: every later launch of the step is a no-op
  uint32_t pad0;
  uint64_t need_volume;  // cells the box needed when it overflowed
  uint64_t need_pairs;   // body-point pair records a substep needed past their room
  uint32_t cells;        // searched table: occupied cells of the last sort
  uint32_t pad1;
  uint32_t pairs_used;   // most body-point pair records a substep of this step used
  uint32_t bad_index;    // smallest sorted index whose advance went non-finite
  uint32_t substeps_done;
  uint32_t iterations[2];        // last solve's pushes: divergence, density
  uint32_t iterations_step[2];   // pushes summed over this step
  uint32_t rates_step;           // rate passes summed over this step
  uint32_t log_len;
  // the persistent solve's grid barrier: blocks arrived per XCD group
  // (blockIdx % 8), XCD groups arrived, generation; one cache line each
  alignas(128) uint32_t bar_count;
  alignas(128) uint32_t bar_gen;
  alignas(128) uint32_t bar_xcd[8][32];
  // the list solve's tile queues: next tile of each XCD group's share, per
  // pass parity, one cache line each; zero between launches
  alignas(128) uint32_t tile_next[2][8][32];
  alignas(128) Stripe stripes[kStripes];
  // average error of every rate pass (solve in the sign: + divergence, -
  // density), for parity logs
  double residual_log[kResidualLogCap];
  uint32_t loaded_log[kResidualLogCap];
  // the flow solves of a substep (0 divergence, 1 density), both set up
  // once a substep by the tile set-up
  FlowWords flow[2];
  // the slabs of the groups as ranges of the listed tiles (flow_list), set per substep
  alignas(128) uint32_t flow_begin[9];
};

__device__ __forceinline__ Stripe& stripe_of(Control* ctl) { return ctl->stripes[blockIdx.x & (kStripes - 1)]; }

// MI355X dispatches workgroups round-robin over its 8 XCDs, each with its own
// L2. Consecutive sorted particles are spatial neighbours, so a launch over
// particles maps hardware block b to logical block (b % 8) * (n / 8) + b / 8:
// the blocks one XCD runs then cover one contiguous range (an x-slab of
// cells) instead of every eighth stripe of the domain. A bijection on
// [0, gridDim.x); blocks past the last multiple of 8 keep their index.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
  uint32_t b = blockIdx.x, n = gridDim.x;
  uint32_t q = n / kXcds;
  if (b >= q * kXcds) return b;
  return (b % kXcds) * q + b / kXcds;
}

// A dense cell table over a box of cells: slot of cell (x, y, z) is
// ((x - lo.x) * span.y + (y - lo.y)) * span.z + (z - lo.z); particles of slot
// c are start[c] .. start[c + 1] in the sorted arrays.
struct CellBox {
  int32_t lo[3];
  uint32_t span[3];
};

// A coupled body during a step (src/fluid/couple.rs:15-32): its pose (the
// centre rounded from the rigid world's f64 position, rotation x, y, z, w),
// how far its centre has gone in the current substep, the damping impulse
// and angular impulse summed over the step, and its ranges of points and
// probes in the concatenated body-axis arrays.
struct BodyState {
  float center[3];
  float rotation[4];
  float linear[3];
  float angular[3];
  float moved[3];
  float impulse[3];
  float angular_impulse[3];
  uint32_t first_point, points;
  uint32_t first_probe, probes;
  float probe_spacing;
};

// The cell box of the body points after a carry (Movers::grid,
// src/fluid/couple.rs:293-319). A point's sort key is its cell's slot in the
// box, ((x - lo.x) * span.y + (y - lo.y)) * span.z + (z - lo.z), shifted up
// past the owner bits, or-ed with its owner; `bad` is set when the box's
// slots do not fit beside the owner bits in 64. When the box's volume + 1
// fits the table (`dense`), slot c's points are table[c] .. table[c + 1].
struct MoverGrid {
  int32_t lo[3];
  int32_t hi[3];
  uint64_t span[3];
  uint64_t volume;
  uint32_t bad;
  uint32_t dense;
};

// One body-point pair of the WCSPH force pass, for the damping the body
// feels (mover_push, src/fluid/couple.rs:396-407): slot offsets[i] + k holds
// particle i's k-th pair; key = owner when the pair reaches its body, else
// the body count (never summed).
struct PairRecords {
  uint32_t* counts;         // per particle: body points within reach (density pass)
  const uint32_t* offsets;  // exclusive scan of counts
  uint32_t* key;
  float4* force;   // x, y, z, chunk (bits)
  float4* torque;  // x, y, z, 0
  uint32_t none;   // the key of a pair that does not reach
};

// The body points of a substep, sorted by (cell, owner, previous place), as
// the WCSPH passes see them.
struct MoverView {
  const float4* pos;      // x, y, z, owner bits
  const float4* vel;      // vx, vy, vz, 0
  const uint64_t* keys;   // sort keys, ascending
  const uint32_t* table;  // dense cell table, when the grid says so
  uint32_t count;         // 0: no body points this step
  uint32_t owner_bits;
  const MoverGrid* grid;
  const BodyState* bodies;
  float damping;  // (1 c r).min(0.25 r r / dt), step.rs:163-164
  float settles;  // |g| spacing, couple.rs:387
  PairRecords pairs;
};

// A terrain region (TerrainField, src/fluid/terrain.rs:35-42): signed
// distances at nodes `step` apart from `low`, node (x, y, z) at
// first + (y * nodes.z + z) * nodes.x + x of the distance array.
struct TerrainDesc {
  float low[3];
  uint32_t nodes[3];
  float step;
  uint32_t first;
};

// The terrains of a world and its GroundSites (terrain.rs:173-247): the
// lattice sites within the kernel (offset, density weight) and how far
// each is, and the damping table into flat ground. `slopes` holds each
// particle's ground slope from the density pass, read by the force pass.
struct TerrainView {
  const TerrainDesc* fields;
  const float* distances;
  uint32_t count;  // 0: no terrain
  uint32_t nsites;
  const float4* sites;  // x, y, z, weight
  const float* far;
  const float* damping;  // kMapSamples values
  float low, step;       // of the damping table
  float reach;           // radius + spacing / 2
  float spacing;
  float4* slopes;
};
constexpr int kMapSamples = 256;  // MAP_SAMPLES, terrain.rs:11

// Scalars of one substep, computed on the host in the reference's f32
// arithmetic.
struct StepParams {
  uint32_t n;
  float radius;
  float rinv;       // RN(1 / radius), for over_radius
  float reach;      // radius * radius
  float scale;      // Wendland normalisation
  float mass;
  float rest;
  float stiffness;  // rest * c^2 / 7
  float dt;
  float nu;         // viscosity(dt)
  float wall_nu;    // wall viscosity
  float soften;     // 0.01 r^2
  float gx, gy, gz;
  float half_spacing;
  int32_t has_bounds;
  float blo[3], bhi[3];
  // walls (static, sorted by cell); the table is searched when wall_keys is
  // set (wall_keyed occupied cells)
  const float4* walls;
  const uint32_t* wall_start;
  const uint64_t* wall_keys;
  uint32_t wall_keyed;
  CellBox wall_box;
  uint32_t wall_cells;  // 0: no walls
  // fluid cell table (box read from the control block); searched when
  // cell_keys is set (the occupied cells' slots, count in the control block)
  const uint32_t* cell_start;
  const uint64_t* cell_keys;
  Control* ctl;
  // per-substep neighbour lists (DFSPH): slot k of particle i at
  // pairs[k * pair_stride + i] = (neighbour index, bits of slope(d)/d);
  // counts[i] = fluid | walls << 16. In the mask formats the lists hold the
  // walls only; fluid neighbours are masks[k * pair_stride + i], k the
  // neighbour cell in the reference's order (dz, dy, dx), bit b set when
  // particle start[cell k] + b is a neighbour, and own_slot[i] the
  // particle's own cell. In the list format fluid entries come first. In
  // the compact format the lists hold the walls only and fluid entry k is
  // (near[k * pair_stride + i], factor[k * pair_stride + i]).
  uint2* pairs;
  uint32_t* pair_counts;
  uint32_t pair_stride;
  uint32_t pair_cap;
  void* masks;
  uint32_t* own_slot;
  uint16_t* near;
  float* factor;
  int32_t format;  // kMasks16, kMasks32, kLists or kCompact
  MoverView movers;
  TerrainView terrain;
};

// Rows the neighbour list arrays carry past their capacity: the solve's
// loads run up to 15 slots ahead of a list's end.
constexpr uint32_t kListSlack = 16;

// LDS per particle a tile of the list solve stages: float4 + float2 (rate
// pass), float4 (push pass)
constexpr uint32_t kTileBytes = 24;

// A fluid entry of the list format carries the neighbour's index in its low
// 28 bits and, in the top 4, which of the 3x3 neighbour columns (dx, dy) it
// sits in: (dx + 1) * 3 + (dy + 1). Wall entries carry their index alone.
constexpr uint32_t kListIndexBits = 28;
constexpr uint32_t kListIndexMask = (1u << kListIndexBits) - 1u;
__device__ __forceinline__ uint32_t list_index(uint32_t x) { return x & kListIndexMask; }
__device__ __forceinline__ uint32_t list_column(uint32_t x) { return x >> kListIndexBits; }

// A fluid entry of the compact format names its neighbour in 16 bits: the
// neighbour column (dx, dy) in the top 4, as (dx + 1) * 4 + (dy + 1), and in
// the low 12 its offset from the first particle of that column's cells from
// z0 - 1 on, z0 the first cell in z of the particle's solve tile: the ranges
// the tiled solve stages in LDS. A neighbour past offset 4095 widens to
// lists.
constexpr uint32_t kNearBits = 12;
constexpr uint32_t kNearMask = (1u << kNearBits) - 1u;

constexpr int kMasks16 = 0;
constexpr int kMasks32 = 1;
constexpr int kLists = 2;
constexpr int kCompact = 3;

// The solve's tiles are 4 x 2 cell columns by tile_depth(span z) cells: the
// z span split into pieces of at most 16, as even as they go.
__host__ __device__ __forceinline__ uint32_t tile_depth(uint32_t sz) {
  uint32_t pieces = (sz + 15) / 16;
  return pieces ? (sz + pieces - 1) / pieces : 1u;
}

// ---- Wendland C2 (src/fluid/mod.rs:126-165) ------------------------------

__host__ __device__ __forceinline__ float kernel_value(float distance, float radius, float scale) {
  float q = distance / radius;
  if (q >= 1.0f) return 0.0f;
  float t = 1.0f - q;
  float t2 = t * t;
  float t4 = t2 * t2;
  return scale * t4 * (1.0f + 4.0f * q);
}

__host__ __device__ __forceinline__ float kernel_slope(float distance, float radius, float scale) {
  float q = distance / radius;
  if (q >= 1.0f) return 0.0f;
  float t = 1.0f - q;
  float t3 = t * (t * t);
  return scale * -20.0f * q * t3 / radius;
}

// a / radius, correctly rounded, for a finite a, from rinv = RN(1 / radius):
// q0 = RN(a rinv) is within an ulp of a / radius, the residual a - q0 radius
// is exact in one FMA, and RN(q0 + residual rinv) is RN(a / radius)
// (Markstein's theorem; P. Markstein, IBM J. Res. Dev. 34(1), 1990). Three
// instructions where IEEE division takes thirteen. A zero a keeps its sign
// (the correction would turn -0 into +0).
__device__ __forceinline__ float over_radius(float a, float radius, float rinv) {
  const float q0 = a * rinv;
  const float residual = __builtin_fmaf(-q0, radius, a);
  const float q = __builtin_fmaf(residual, rinv, q0);
  return a == 0.0f ? q0 : q;
}

// kernel_value and kernel_slope with their divisions by the radius through
// over_radius: the same bits.
__device__ __forceinline__ float kernel_value(float distance, float radius, float rinv, float scale) {
  float q = over_radius(distance, radius, rinv);
  if (q >= 1.0f) return 0.0f;
  float t = 1.0f - q;
  float t2 = t * t;
  float t4 = t2 * t2;
  return scale * t4 * (1.0f + 4.0f * q);
}

__device__ __forceinline__ float kernel_slope(float distance, float radius, float rinv, float scale) {
  float q = over_radius(distance, radius, rinv);
  if (q >= 1.0f) return 0.0f;
  float t = 1.0f - q;
  float t3 = t * (t * t);
  return over_radius(scale * -20.0f * q * t3, radius, rinv);
}

// Kernel::gradient (mod.rs:158-164) given the offset and its squared length.
__host__ __device__ __forceinline__ void kernel_gradient(float ox, float oy, float oz, float squared, float radius, float scale,
                                                float& gx, float& gy, float& gz) {
  float distance = sqrtf(squared);
  if (distance <= 0.0f) {
    gx = 0.0f;
    gy = 0.0f;
    gz = 0.0f;
    return;
  }
  float f = kernel_slope(distance, radius, scale) / distance;
  gx = ox * f;
  gy = oy * f;
  gz = oz * f;
}

// kernel_gradient with the slope's divisions by the radius through over_radius
__device__ __forceinline__ void kernel_gradient(float ox, float oy, float oz, float squared, float radius, float rinv,
                                                float scale, float& gx, float& gy, float& gz) {
  float distance = sqrtf(squared);
  if (distance <= 0.0f) {
    gx = 0.0f;
    gy = 0.0f;
    gz = 0.0f;
    return;
  }
  float f = kernel_slope(distance, radius, rinv, scale) / distance;
  gx = ox * f;
  gy = oy * f;
  gz = oz * f;
}

// (x*x' + y*y') + z*z', no contraction
__host__ __device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return ax * bx + ay * by + az * bz;
}

// `x as i32` (saturating, NaN -> 0) of floor(v / radius) (mod.rs:526-533)
__device__ __forceinline__ int32_t cell_coord(float v, float radius) {
  float c = floorf(v / radius);
  if (!(c == c)) return 0;
  if (c >= 2147483648.0f) return 2147483647;
  if (c <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)c;
}

// Slot of cell (x, y, z) in `box`, or -1 outside it.
__device__ __forceinline__ int64_t cell_slot(int64_t x, int64_t y, int64_t z, const int32_t lo[3],
                                             const uint32_t span[3]) {
  int64_t ax = x - lo[0], ay = y - lo[1], az = z - lo[2];
  if (ax < 0 || ay < 0 || az < 0 || ax >= (int64_t)span[0] || ay >= (int64_t)span[1] || az >= (int64_t)span[2])
    return -1;
  return (ax * (int64_t)span[1] + ay) * (int64_t)span[2] + az;
}

__device__ __forceinline__ uint32_t id_of(float4 p) { return __float_as_uint(p.w); }

// f32::max / f32::min with NaN operands yielding the other one
__device__ __forceinline__ float rmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }

__device__ __forceinline__ bool finite3(float x, float y, float z) {
  return isfinite(x) && isfinite(y) && isfinite(z);
}

// Tait pressure (step.rs:151-158).
__device__ __forceinline__ float tait_pressure(float density, float rest, float stiffness) {
  float ratio = rmin(density / rest, 1.1f);
  float r2 = ratio * ratio;
  float r4 = r2 * r2;
  float r7 = ratio * r2 * r4;
  return rmax(stiffness * (r7 - 1.0f), 0.0f);
}

}  // namespace sfl

namespace sfl {

// A rank in cell `slot` for every live lane, distinct within the cell, from
// one atomic per run of consecutive lanes in the same cell: particles come in
// last substep's sorted order, so such runs are long. Every lane of the wave
// calls it (dead lanes end runs and take no rank). The sort ranks cell-mates
// by id afterwards, so which lane gets which rank does not matter.
__device__ __forceinline__ uint32_t wave_rank(uint32_t* counts, uint32_t slot, bool live) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t before = (uint32_t)__shfl_up((int)slot, 1u, 64);
  const bool head = lane == 0 || before != slot || !live;
  const uint64_t heads = __ballot(head);
  const uint64_t upto = (2ull << lane) - 1ull;  // lanes 0..lane
  const uint32_t first = 63u - (uint32_t)__clzll(heads & upto);
  const uint64_t later = heads & ~upto;
  const uint32_t next = later ? (uint32_t)__ffsll((long long)later) - 1u : 64u;
  uint32_t base = 0;
  if (live && head) base = atomicAdd(&counts[slot], next - lane);
  base = (uint32_t)__shfl((int)base, (int)first, 64);
  return base + (lane - first);
}

// A cell table (Grid, src/fluid/mod.rs:168-236) over a box of cells, slot
// ((x - lo.x) * span.y + (y - lo.y)) * span.z + (z - lo.z). Dense (keys
// null): start[slot] .. start[slot + 1] is the cell's range. Searched:
// keys[0, cells) are the occupied cells' slots in ascending order and
// start[k] .. start[k + 1] the k-th one's range, found by binary search as
// the reference's Grid::range does without its dense index. Either gives the
// same ranges, so the same neighbours in the same order.
struct GridView {
  int32_t lo[3];
  uint32_t span[3];
  const uint32_t* start;
  const uint64_t* keys;
  uint32_t cells;
};

// The fluid's table of the last sort (box in the control block).
__device__ __forceinline__ GridView fluid_grid(const StepParams& prm) {
  const Control* ctl = prm.ctl;
  return GridView{{ctl->lo[0], ctl->lo[1], ctl->lo[2]},
                  {ctl->span[0], ctl->span[1], ctl->span[2]},
                  prm.cell_start,
                  prm.cell_keys,
                  prm.cell_keys ? ctl->cells : 0u};
}

__device__ __forceinline__ GridView wall_grid(const StepParams& prm) {
  return GridView{{prm.wall_box.lo[0], prm.wall_box.lo[1], prm.wall_box.lo[2]},
                  {prm.wall_box.span[0], prm.wall_box.span[1], prm.wall_box.span[2]},
                  prm.wall_start,
                  prm.wall_keys,
                  prm.wall_keyed};
}

// The range [s, e) of cell (x, y, z); false for a cell outside the box or
// not in a searched table (an empty range may come back true).
__device__ __forceinline__ bool grid_range(const GridView& g, int64_t x, int64_t y, int64_t z, uint32_t& s,
                                           uint32_t& e) {
  const int64_t c = cell_slot(x, y, z, g.lo, g.span);
  if (c < 0) return false;
  if (!g.keys) {
    s = g.start[c];
    e = g.start[c + 1];
    return true;
  }
  uint32_t a = 0, b = g.cells;
  while (a < b) {
    const uint32_t mid = a + (b - a) / 2;
    if (g.keys[mid] < (uint64_t)c)
      a = mid + 1;
    else
      b = mid;
  }
  if (a >= g.cells || g.keys[a] != (uint64_t)c) return false;
  s = g.start[a];
  e = g.start[a + 1];
  return true;
}

// Calls f(j, pj, ox, oy, oz, squared) for every particle j of `P` within
// reach of `at`, o = at - P[j], in the reference's neighbour order: cells
// dz, dy, dx (Grid::around, src/fluid/mod.rs:250-261), then the cell's range
// in sorted order; membership |o|^2 < reach as in `near` (mod.rs:264-277),
// whose |x_j - x_i|^2 has the same bits.
template <class F>
__device__ __forceinline__ void visit_near(const float4* __restrict__ P, const GridView& g, float4 at, float radius,
                                           float reach, F&& f) {
  int32_t cx = cell_coord(at.x, radius), cy = cell_coord(at.y, radius), cz = cell_coord(at.z, radius);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        uint32_t s, e;
        if (!grid_range(g, (int64_t)cx + dx, (int64_t)cy + dy, (int64_t)cz + dz, s, e)) continue;
        for (uint32_t j = s; j < e; ++j) {
          float4 pj = P[j];
          float ox = at.x - pj.x, oy = at.y - pj.y, oz = at.z - pj.z;
          float sq = dot3(ox, oy, oz, ox, oy, oz);
          if (sq < reach) f(j, pj, ox, oy, oz, sq);
        }
      }
}

// visit_near over a dense table given as its parts: the form the hottest
// kernels use when no table of the step is searched (a GridView built from
// the parameters holds its fields in scalar registers for the whole kernel,
// which spills them there).
template <class F>
__device__ __forceinline__ void visit_dense(const float4* __restrict__ P, const uint32_t* __restrict__ start,
                                            const int32_t lo[3], const uint32_t span[3], float4 at, float radius,
                                            float reach, F&& f) {
  const int32_t cx = cell_coord(at.x, radius), cy = cell_coord(at.y, radius), cz = cell_coord(at.z, radius);
  // cell k of the 27 in the reference's order (dz, dy, dx); an empty range
  // outside the box
  auto range = [&](int k, uint32_t& s, uint32_t& e) {
    const int64_t c = cell_slot((int64_t)cx + (k % 3 - 1), (int64_t)cy + ((k / 3) % 3 - 1),
                                (int64_t)cz + (k / 9 - 1), lo, span);
    s = c < 0 ? 0u : start[c];
    e = c < 0 ? 0u : start[c + 1];
  };
  uint32_t s, e;
  range(0, s, e);
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    // the next cell's range loads while this one is walked
    uint32_t s2 = 0, e2 = 0;
    if (k + 1 < 27) range(k + 1, s2, e2);
    // four candidates' loads in flight at once (the last one repeated past
    // the range), then their tests in order
    for (uint32_t j = s; j < e; j += 4) {
      float4 q[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) q[u] = P[min(j + u, e - 1u)];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        if (j + u >= e) break;
        float ox = at.x - q[u].x, oy = at.y - q[u].y, oz = at.z - q[u].z;
        float sq = dot3(ox, oy, oz, ox, oy, oz);
        if (sq < reach) f(j + u, q[u], ox, oy, oz, sq);
      }
    }
    s = s2;
    e = e2;
  }
}

// A fluid neighbour's data a WCSPH force pass gathers: velocity (density in
// w) and p / rho^2. Nothing: a pass that loads them itself (fewer registers).
struct Neighbour {
  float4 v;
  float s;
};
struct Nothing {};

// visit_dense in two phases, for pair terms much dearer than the distance
// test: the walk queues each lane's accepted candidates in `queue` (LDS,
// slot k of a lane at queue[k * stride]) and the terms run over the queued
// ones whenever a lane's queue could fill within the next batch, and at the
// end. The pair terms then run over accepted pairs only, most lanes busy,
// instead of over every candidate any lane accepts. The queued pairs go four
// at a time: their indices, positions and gather(j) (the neighbour's other
// data) load together, then f(j, pj, o, sq, g) runs for each in order, so f
// sees the same pairs in the same order as from visit_dense.
template <uint32_t Q, class Gather, class F>
__device__ __forceinline__ void visit_dense_queued(const float4* __restrict__ P, const uint32_t* __restrict__ start,
                                                   const int32_t lo[3], const uint32_t span[3], float4 at,
                                                   float radius, float reach, uint32_t* queue, uint32_t stride,
                                                   Gather&& gather, F&& f) {
  static_assert(Q >= 8, "a queue holds at least two batches");
  uint32_t n = 0;
  auto flush = [&]() {
    for (uint32_t k = 0; k < n; k += 4) {
      uint32_t jj[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) jj[u] = queue[min(k + u, n - 1u) * stride];
      float4 pj[4];
      decltype(gather(0u)) g[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        pj[u] = P[jj[u]];
        g[u] = gather(jj[u]);
      }
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        if (k + u >= n) break;
        const float ox = at.x - pj[u].x, oy = at.y - pj[u].y, oz = at.z - pj[u].z;
        f(jj[u], pj[u], ox, oy, oz, dot3(ox, oy, oz, ox, oy, oz), g[u]);
      }
    }
    n = 0;
  };
  const int32_t cx = cell_coord(at.x, radius), cy = cell_coord(at.y, radius), cz = cell_coord(at.z, radius);
  auto range = [&](int k, uint32_t& s, uint32_t& e) {
    const int64_t c = cell_slot((int64_t)cx + (k % 3 - 1), (int64_t)cy + ((k / 3) % 3 - 1),
                                (int64_t)cz + (k / 9 - 1), lo, span);
    s = c < 0 ? 0u : start[c];
    e = c < 0 ? 0u : start[c + 1];
  };
  uint32_t s, e;
  range(0, s, e);
  for (int k = 0; k < 27; ++k) {
    uint32_t s2 = 0, e2 = 0;
    if (k + 1 < 27) range(k + 1, s2, e2);
    for (uint32_t j = s; j < e; j += 4) {
      if (__any(n + 4 > Q)) flush();
      float4 q[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) q[u] = P[min(j + u, e - 1u)];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        if (j + u >= e) break;
        const float ox = at.x - q[u].x, oy = at.y - q[u].y, oz = at.z - q[u].z;
        if (dot3(ox, oy, oz, ox, oy, oz) < reach) queue[n++ * stride] = j + u;
      }
    }
    s = s2;
    e = e2;
  }
  flush();
}

// The water's neighbours of `at` in the last sort's table, and the walls';
// SEARCH = false for a step whose tables are all dense.
template <bool SEARCH, class F>
__device__ __forceinline__ void visit_fluid(const float4* __restrict__ P, const StepParams& prm, float4 at, F&& f) {
  if constexpr (SEARCH) {
    visit_near(P, fluid_grid(prm), at, prm.radius, prm.reach, f);
  } else {
    const Control* ctl = prm.ctl;
    const int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
    const uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
    visit_dense(P, prm.cell_start, lo, span, at, prm.radius, prm.reach, f);
  }
}

template <bool SEARCH, class F>
__device__ __forceinline__ void visit_walls(const StepParams& prm, float4 at, F&& f) {
  if constexpr (SEARCH)
    visit_near(prm.walls, wall_grid(prm), at, prm.radius, prm.reach, f);
  else
    visit_dense(prm.walls, prm.wall_start, prm.wall_box.lo, prm.wall_box.span, at, prm.radius, prm.reach, f);
}

// First index in keys[0, n) whose key is not below `target`.
__device__ __forceinline__ uint32_t lower_bound64(const uint64_t* __restrict__ keys, uint32_t n, uint64_t target) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Calls f(j, pj, ox, oy, oz, squared) for every body point j within reach
// of `at`, o = at - pos[j], in the order `near` lists them over the body
// points' grid (src/fluid/mod.rs:583-589): cells dz, dy, dx, then each
// cell's points in sorted order.
template <class F>
__device__ __forceinline__ void visit_movers(const MoverView& mv, float4 at, float radius, float reach, F&& f) {
  if (mv.count == 0) return;
  const MoverGrid& g = *mv.grid;
  int64_t cx = cell_coord(at.x, radius), cy = cell_coord(at.y, radius), cz = cell_coord(at.z, radius);
  if (cx + 1 < g.lo[0] || cx - 1 > g.hi[0] || cy + 1 < g.lo[1] || cy - 1 > g.hi[1] || cz + 1 < g.lo[2] ||
      cz - 1 > g.hi[2])
    return;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int64_t x = cx + dx, y = cy + dy, z = cz + dz;
        if (x < g.lo[0] || x > g.hi[0] || y < g.lo[1] || y > g.hi[1] || z < g.lo[2] || z > g.hi[2]) continue;
        uint64_t slot = ((uint64_t)(x - g.lo[0]) * g.span[1] + (uint64_t)(y - g.lo[1])) * g.span[2] +
                        (uint64_t)(z - g.lo[2]);
        uint32_t b, e;
        if (g.dense) {
          b = mv.table[slot];
          e = mv.table[slot + 1];
        } else {
          // the box's slots fit beside the owner bits (MoverGrid::bad unset),
          // so slot + 1 shifted does too
          b = lower_bound64(mv.keys, mv.count, slot << mv.owner_bits);
          e = lower_bound64(mv.keys + b, mv.count - b, (slot + 1) << mv.owner_bits) + b;
        }
        for (uint32_t j = b; j < e; ++j) {
          float4 pj = mv.pos[j];
          float ox = at.x - pj.x, oy = at.y - pj.y, oz = at.z - pj.z;
          float sq = dot3(ox, oy, oz, ox, oy, oz);
          if (sq < reach) f(j, pj, ox, oy, oz, sq);
        }
      }
}

// Wave-wide max of a non-negative float, folded into `bits` by lane 0.
__device__ __forceinline__ void wave_max_to(uint32_t* bits, float v) {
  uint32_t b = __float_as_uint(v > 0.0f ? v : 0.0f);
  for (int off = 32; off >= 1; off >>= 1) b = max(b, (uint32_t)__shfl_xor((int)b, off, 64));
  if ((threadIdx.x & 63) == 0 && b) atomicMax(bits, b);
}

__device__ __forceinline__ void wave_sum_to(uint32_t* dst, uint32_t v) {
  for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// ---- terrain ground (src/fluid/terrain.rs) --------------------------------

// TerrainField::read (terrain.rs:101-136): the trilinear distance at
// (px, py, pz) and its change per axis; false outside the region.
__device__ __forceinline__ bool terrain_read(const TerrainDesc& f, const float* __restrict__ dist, float px, float py,
                                             float pz, float& value, float& cx, float& cy, float& cz) {
  float lx = (px - f.low[0]) / f.step, ly = (py - f.low[1]) / f.step, lz = (pz - f.low[2]) / f.step;
  if (!(lx >= 0.0f && lx <= (float)(f.nodes[0] - 1) && ly >= 0.0f && ly <= (float)(f.nodes[1] - 1) && lz >= 0.0f &&
        lz <= (float)(f.nodes[2] - 1)))
    return false;
  uint32_t x = min((uint32_t)floorf(lx), f.nodes[0] - 2);
  uint32_t y = min((uint32_t)floorf(ly), f.nodes[1] - 2);
  uint32_t z = min((uint32_t)floorf(lz), f.nodes[2] - 2);
  float fx = lx - (float)x, fy = ly - (float)y, fz = lz - (float)z;
  const uint32_t nx = f.nodes[0], nz = f.nodes[2];
  auto corner = [&](uint32_t dx, uint32_t dy, uint32_t dz) {
    return dist[f.first + ((size_t)(y + dy) * nz + (z + dz)) * nx + (x + dx)];
  };
  auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
  float a00 = corner(0, 0, 0), b00 = corner(1, 0, 0);
  float a10 = corner(0, 1, 0), b10 = corner(1, 1, 0);
  float a01 = corner(0, 0, 1), b01 = corner(1, 0, 1);
  float a11 = corner(0, 1, 1), b11 = corner(1, 1, 1);
  float v00 = lerp(a00, b00, fx), g00 = b00 - a00;
  float v10 = lerp(a10, b10, fx), g10 = b10 - a10;
  float v01 = lerp(a01, b01, fx), g01 = b01 - a01;
  float v11 = lerp(a11, b11, fx), g11 = b11 - a11;
  float v0 = lerp(v00, v10, fy), v1 = lerp(v01, v11, fy);
  value = lerp(v0, v1, fz);
  cx = lerp(lerp(g00, g10, fy), lerp(g01, g11, fy), fz) / f.step;
  cy = lerp(v10 - v00, v11 - v01, fz) / f.step;
  cz = (v1 - v0) / f.step;
  return true;
}

// TerrainField::sample (terrain.rs:138-142): distance and the way out of the
// ground, normalize_or(change, Y).
__device__ __forceinline__ bool terrain_sample(const TerrainDesc& f, const float* __restrict__ dist, float px, float py,
                                               float pz, float& distance, float& nx, float& ny, float& nz) {
  float cx, cy, cz;
  if (!terrain_read(f, dist, px, py, pz, distance, cx, cy, cz)) return false;
  float rcp = 1.0f / sqrtf(dot3(cx, cy, cz, cx, cy, cz));
  if (isfinite(rcp) && rcp > 0.0f) {
    nx = cx * rcp;
    ny = cy * rcp;
    nz = cz * rcp;
  } else {
    nx = 0.0f;
    ny = 1.0f;
    nz = 0.0f;
  }
  return true;
}

// TerrainField::ground (terrain.rs:144-170) summed over the fields from +0
// as compute_densities folds them (mod.rs:657-662): the density the ground's
// sites add and its slope.
__device__ __forceinline__ void ground_near(const TerrainView& t, float px, float py, float pz, float& density,
                                            float& sx, float& sy, float& sz) {
  density = 0.0f;
  sx = sy = sz = 0.0f;
  for (uint32_t k = 0; k < t.count; ++k) {
    const TerrainDesc& f = t.fields[k];
    float distance, cx, cy, cz;
    if (!terrain_read(f, t.distances, px, py, pz, distance, cx, cy, cz)) continue;
    float d = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
    if (distance < t.reach) {
      float open = distance - t.spacing / 2.0f;
      for (uint32_t q = 0; q < t.nsites; ++q) {
        if (t.far[q] <= open) continue;
        float4 site = t.sites[q];
        float there, tx, ty, tz;
        if (!terrain_read(f, t.distances, px + site.x, py + site.y, pz + site.z, there, tx, ty, tz)) continue;
        float share = 0.5f - there / t.spacing;  // solid, terrain.rs:249-252
        if (share < 0.0f) share = 0.0f;
        if (share > 1.0f) share = 1.0f;
        d += site.w * share;
        if (share > 0.0f && share < 1.0f) {
          float k2 = site.w / t.spacing;
          gx = gx - tx * k2;
          gy = gy - ty * k2;
          gz = gz - tz * k2;
        }
      }
    }
    density = density + d;
    sx = sx + gx;
    sy = sy + gy;
    sz = sz + gz;
  }
}

// GroundSites::damping (terrain.rs:236-245)
__device__ __forceinline__ float ground_damping(const TerrainView& t, float distance) {
  float at = (fmaxf(distance, t.low) - t.low) / t.step;
  float f = floorf(at);
  if (!(f < (float)(kMapSamples - 1))) return 0.0f;  // below + 1 >= MAP_SAMPLES
  uint32_t below = (uint32_t)f;
  float part = at - (float)below;
  return t.damping[below] + (t.damping[below + 1] - t.damping[below]) * part;
}

// terrain_contact (terrain.rs:315-343): water closer than `rest` to a
// surface is put back there and stops moving into it; true when it got
// under the surface.
__device__ __forceinline__ bool terrain_contact(const TerrainView& t, float rest, float& px, float& py, float& pz,
                                                float& vx, float& vy, float& vz) {
  bool under = false;
  for (uint32_t k = 0; k < t.count; ++k) {
    float distance, nx, ny, nz;
    if (!terrain_sample(t.fields[k], t.distances, px, py, pz, distance, nx, ny, nz)) continue;
    if (distance < rest) {
      under = under || distance < 0.0f;
      float lift = rest - distance;
      px = px + nx * lift;
      py = py + ny * lift;
      pz = pz + nz * lift;
      float into = dot3(vx, vy, vz, nx, ny, nz);
      if (into < 0.0f) {
        vx = vx - nx * into;
        vy = vy - ny * into;
        vz = vz - nz * into;
      }
    }
  }
  return under;
}

// advance (src/fluid/step.rs:305-339) of one particle: next position, the
// clamped velocity components zeroed, held-back count, finiteness. `clamped`
// gets bit c set when component c was clamped inside the container.
__device__ __forceinline__ bool advance_one(const StepParams& prm, float4 p, float& vx, float& vy, float& vz,
                                            float& nx, float& ny, float& nz, uint32_t& held, uint32_t& clamped) {
  nx = p.x + vx * prm.dt;
  ny = p.y + vy * prm.dt;
  nz = p.z + vz * prm.dt;
  held = 0;
  clamped = 0;
  if (prm.has_bounds) {
    float h = prm.half_spacing;
    float ix = rmin(rmax(nx, prm.blo[0] + h), prm.bhi[0] - h);
    float iy = rmin(rmax(ny, prm.blo[1] + h), prm.bhi[1] - h);
    float iz = rmin(rmax(nz, prm.blo[2] + h), prm.bhi[2] - h);
    if (ix != nx || iy != ny || iz != nz) {
      float cx = rmin(rmax(nx, prm.blo[0]), prm.bhi[0]);
      float cy = rmin(rmax(ny, prm.blo[1]), prm.bhi[1]);
      float cz = rmin(rmax(nz, prm.blo[2]), prm.bhi[2]);
      if (cx != nx || cy != ny || cz != nz) held = 1;
      if (ix != nx) {
        vx = 0.0f;
        clamped |= 1u;
      }
      if (iy != ny) {
        vy = 0.0f;
        clamped |= 2u;
      }
      if (iz != nz) {
        vz = 0.0f;
        clamped |= 4u;
      }
      nx = ix;
      ny = iy;
      nz = iz;
    }
  }
  if (prm.terrain.count && terrain_contact(prm.terrain, prm.half_spacing, nx, ny, nz, vx, vy, vz)) held += 1;
  return finite3(nx, ny, nz) && finite3(vx, vy, vz);
}

__device__ __forceinline__ bool advance_one(const StepParams& prm, float4 p, float& vx, float& vy, float& vz,
                                            float& nx, float& ny, float& nz, uint32_t& held) {
  uint32_t clamped;
  return advance_one(prm, p, vx, vy, vz, nx, ny, nz, held, clamped);
}

// wall_viscous (src/fluid/step.rs:262-301), no terrains; `away` is the
// finished sum of -gradient over the particle's walls.
template <class Visit>
__device__ __forceinline__ void wall_viscous(const StepParams& prm, float4 at, float vx, float vy, float vz,
                                             float density, float awx, float awy, float awz, Visit&& visit_walls,
                                             float& px, float& py, float& pz) {
  px = 0.0f;
  py = 0.0f;
  pz = 0.0f;
  // normalize_or(away, ZERO)
  float len = sqrtf(dot3(awx, awy, awz, awx, awy, awz));
  float rcp = 1.0f / len;
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  if (isfinite(rcp) && rcp > 0.0f) {
    nx = awx * rcp;
    ny = awy * rcp;
    nz = awz * rcp;
  }
  float t = rmin(dot3(vx, vy, vz, nx, ny, nz), 0.0f);
  float ix = nx * t, iy = ny * t, iz = nz * t;
  float wnu = prm.wall_nu;
  // the ground's part, on motion into it (step.rs:284-290)
  for (uint32_t k = 0; k < prm.terrain.count; ++k) {
    float distance, gx, gy, gz;
    if (!terrain_sample(prm.terrain.fields[k], prm.terrain.distances, at.x, at.y, at.z, distance, gx, gy, gz))
      continue;
    float into = dot3(vx, vy, vz, gx, gy, gz);
    if (into < 0.0f) {
      float weight = ground_damping(prm.terrain, distance);
      float k2 = wnu * into * weight / density;
      px = px - gx * k2;
      py = py - gy * k2;
      pz = pz - gz * k2;
    }
  }
  visit_walls([&](uint32_t, float4, float ox, float oy, float oz, float sq) {
    float toward = dot3(ix, iy, iz, ox, oy, oz);
    if (toward < 0.0f) {
      float squeeze = toward / (sq + prm.soften);
      float damp = -wnu * squeeze / density;
      float gx, gy, gz;
      kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
      float k = prm.mass * damp;
      px = px - gx * k;
      py = py - gy * k;
      pz = pz - gz * k;
    }
  });
}

}  // namespace sfl
