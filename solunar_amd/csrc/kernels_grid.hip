// Neighbour build on gfx950: cell bounds, counting sort by cell, cell-start
// table, in-cell order by id, and the gather into sorted order. Replaces
// update_neighbors' sort, gathers and Grid::build (src/fluid/mod.rs:540-566,
// 182-247); the CSR list emission at the end replaces `near` + Lists::join
// (mod.rs:264-308, 569-607) for the readback and list-density entry points.
//
// The permutation is the reference's: slots of the dense cell table are
// lexicographic in (cx, cy, cz), so the counting sort orders cells exactly as
// sort_unstable on (cell, id, index) does, and within a cell each particle's
// slot is the number of cell-mates with a smaller id (ids are unique), which
// does not depend on the order the counting atomics resolved in.
#include <rocprim/device/device_radix_sort.hpp>

#include "device.hpp"
#include "launch.hpp"

namespace sfl {

// ---- per-step / per-substep resets of the control block ----------------

__global__ void k_step_begin(Control* ctl) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < kStripes; ++k) ctl->stripes[k].held = 0;
  ctl->flags = 0;
  ctl->abort = 0;
  ctl->pairs_used = 0;
  ctl->bad_index = 0xffffffffu;
  ctl->substeps_done = 0;
  ctl->iterations_step[0] = 0;
  ctl->iterations_step[1] = 0;
  ctl->rates_step = 0;
  ctl->log_len = 0;
  // no persistent launch is in flight on this stream
  ctl->bar_count = 0;
  for (int k = 0; k < 8; ++k) {
    ctl->bar_xcd[k][0] = 0;
    ctl->tile_next[0][k][0] = 0;
    ctl->tile_next[1][k][0] = 0;
  }
}

__global__ void k_substep_begin(Control* ctl) {
  if (threadIdx.x != 0) return;
  if (ctl->flags & kFlagNonFinite) ctl->abort = 1;
  if (!ctl->abort) ctl->substeps_done += 1;
  for (int k = 0; k < kStripes; ++k) {
    Stripe& s = ctl->stripes[k];
    s.bmin[0] = s.bmin[1] = s.bmin[2] = 0x7fffffff;
    s.bmax[0] = s.bmax[1] = s.bmax[2] = (int32_t)0x80000000u;
    s.comp_bits = 0;
    s.vmax_bits = 0;
    s.pairs_max = 0;
  }
}

// fastest (src/fluid/step.rs:92-96): max |v| over the particles.
__global__ __launch_bounds__(256) void k_fastest(const float4* __restrict__ V, uint32_t n, Control* ctl) {
  float most = 0.0f;
  for (uint32_t i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 v = V[i];
    most = rmax(most, sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z)));
  }
  wave_max_to(&stripe_of(ctl).vmax_bits, most);
}

// ---- cell bounds ---------------------------------------------------------

__global__ __launch_bounds__(256) void k_bounds(const float4* __restrict__ P, uint32_t n, float radius,
                                                Control* ctl) {
  if (ctl->abort) return;
  // cell_coord is monotonic in the coordinate (a correctly rounded division
  // by radius > 0, floor, saturation), so the cells of the extreme
  // coordinates are the extreme cells; a NaN coordinate's cell is 0
  float mn[3] = {INFINITY, INFINITY, INFINITY};
  float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool nan[3] = {false, false, false};
  for (uint32_t i = xcd_block() * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = P[i];
    const float v[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], v[a]);
      mx[a] = fmaxf(mx[a], v[a]);
      nan[a] = nan[a] || !(v[a] == v[a]);
    }
  }
  // waves, then the block's 4 waves in LDS: one set of atomics a block
  __shared__ int32_t part[4][6];
  const uint32_t w = threadIdx.x >> 6;
  for (int a = 0; a < 3; ++a) {
    for (int off = 32; off >= 1; off >>= 1) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
    }
    // no particle: +inf and -inf give the identities of min and max
    int32_t lo = cell_coord(mn[a], radius), hi = cell_coord(mx[a], radius);
    if (__any(nan[a])) {
      lo = min(lo, 0);
      hi = max(hi, 0);
    }
    if ((threadIdx.x & 63) == 0) {
      part[w][a] = lo;
      part[w][3 + a] = hi;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    int32_t v = part[0][a];
    for (uint32_t k = 1; k < 4; ++k) v = a < 3 ? min(v, part[k][a]) : max(v, part[k][a]);
    Stripe& s = stripe_of(ctl);
    if (a < 3)
      atomicMin(&s.bmin[a], v);
    else
      atomicMax(&s.bmax[a - 3], v);
  }
}

// Box of the cell table from the bounds; flags an overflow when it needs more
// slots than `capacity` (then every later launch of the step does nothing and
// the host grows the table).
__global__ void k_grid_setup(Control* ctl, uint32_t n, uint64_t capacity) {
  if (threadIdx.x != 0) return;
  if (ctl->abort) return;
  if (n == 0) {
    for (int a = 0; a < 3; ++a) {
      ctl->lo[a] = 0;
      ctl->span[a] = 1;
    }
    ctl->volume = 1;
    return;
  }
  int32_t bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int32_t bmax[3] = {(int32_t)0x80000000u, (int32_t)0x80000000u, (int32_t)0x80000000u};
  for (int k = 0; k < kStripes; ++k)
    for (int a = 0; a < 3; ++a) {
      bmin[a] = min(bmin[a], ctl->stripes[k].bmin[a]);
      bmax[a] = max(bmax[a], ctl->stripes[k].bmax[a]);
    }
  uint64_t vol = 1;
  bool big = false;
  uint32_t span[3];
  for (int a = 0; a < 3; ++a) {
    int64_t s = (int64_t)bmax[a] - (int64_t)bmin[a] + 1;
    if (s > 0xffffffffll) {
      big = true;
      s = 0xffffffffll;
    }
    span[a] = (uint32_t)s;
    uint64_t prod = vol * (uint64_t)s;
    if (s != 0 && prod / (uint64_t)s != vol) big = true;
    vol = big ? 0xffffffffffffffffull : prod;
  }
  if (capacity == kSearchedTable) {
    // the searched table: any box whose slots fit a 63-bit key
    if (big || vol >= (1ull << 63)) {
      ctl->flags |= kFlagOverflow;
      ctl->abort = 1;
      ctl->need_volume = 0xffffffffffffffffull;
      return;
    }
    for (int a = 0; a < 3; ++a) {
      ctl->lo[a] = bmin[a];
      ctl->span[a] = span[a];
    }
    ctl->volume = 0;
    ctl->need_volume = vol;  // the box, for the host to go dense again
    return;
  }
  if (big || vol > capacity || vol >= 0xffffffffull) {
    ctl->flags |= kFlagOverflow;
    ctl->abort = 1;
    ctl->need_volume = vol;
    return;
  }
  for (int a = 0; a < 3; ++a) {
    ctl->lo[a] = bmin[a];
    ctl->span[a] = span[a];
  }
  ctl->volume = (uint32_t)vol;
}

// Cell slot of every particle and its arrival rank in the slot.
__global__ __launch_bounds__(256) void k_count(const float4* __restrict__ P, uint32_t n, float radius,
                                               uint32_t* __restrict__ counts, uint2* __restrict__ slot_rank,
                                               Control* ctl) {
  if (ctl->abort) return;
  const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  const bool live = i < n;
  int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
  uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
  uint32_t slot = 0xffffffffu;
  if (live) {
    float4 p = P[i];
    slot = (uint32_t)cell_slot(cell_coord(p.x, radius), cell_coord(p.y, radius), cell_coord(p.z, radius), lo, span);
  }
  const uint32_t rank = wave_rank(counts, slot, live);
  if (live) slot_rank[i] = make_uint2(slot, rank);
}

// ---- exclusive scan: out[k] = sum(in[0..k)), out[len] = total -----------
// Tiles of 256 threads x 8 items. `len` comes from `len_dev` when given (the
// cell table's volume) else `len_host`. With `zero_input` the input is
// cleared as it is read, which keeps the cell counts zeroed between substeps.

constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = 256 * kScanItems;

__device__ __forceinline__ uint32_t scan_len(const uint32_t* len_dev, uint32_t len_host) {
  return len_dev ? *len_dev : len_host;
}

__global__ __launch_bounds__(256) void k_scan_tiles(uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    uint32_t* __restrict__ tile_sums, const uint32_t* len_dev,
                                                    uint32_t len_host, int zero_input, const Control* ctl) {
  if (ctl && ctl->abort) return;
  uint32_t len = scan_len(len_dev, len_host);
  uint32_t base = blockIdx.x * kScanTile;
  if (base >= len) return;
  __shared__ uint32_t warp_sums[4];
  uint32_t v[kScanItems];
  uint32_t local = 0;
  uint32_t first = base + threadIdx.x * kScanItems;
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t idx = first + k;
    uint32_t x = idx < len ? in[idx] : 0u;
    if (zero_input && idx < len) in[idx] = 0u;
    v[k] = local;
    local += x;
  }
  // inclusive scan of `local` across the block
  uint32_t lane = threadIdx.x & 63, warp = threadIdx.x >> 6;
  uint32_t incl = local;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63) warp_sums[warp] = incl;
  __syncthreads();
  uint32_t warp_off = 0;
  for (uint32_t w = 0; w < warp; ++w) warp_off += warp_sums[w];
  uint32_t excl = warp_off + incl - local;
  for (uint32_t k = 0; k < kScanItems; ++k) {
    uint32_t idx = first + k;
    if (idx < len) out[idx] = excl + v[k];
  }
  if (threadIdx.x == 255) tile_sums[blockIdx.x] = warp_off + incl;
}

// One block: exclusive scan of the tile sums in place, total to `total_out`.
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ tile_sums, const uint32_t* len_dev,
                                                    uint32_t len_host, uint32_t* __restrict__ total_out,
                                                    const Control* ctl) {
  if (ctl && ctl->abort) return;
  uint32_t len = scan_len(len_dev, len_host);
  uint32_t tiles = (len + kScanTile - 1) / kScanTile;
  __shared__ uint32_t warp_sums[16];
  __shared__ uint32_t carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  uint32_t lane = threadIdx.x & 63, warp = threadIdx.x >> 6;
  for (uint32_t base = 0; base < tiles; base += 1024) {
    uint32_t idx = base + threadIdx.x;
    uint32_t x = idx < tiles ? tile_sums[idx] : 0u;
    uint32_t incl = x;
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t y = __shfl_up(incl, off, 64);
      if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) warp_sums[warp] = incl;
    __syncthreads();
    uint32_t warp_off = 0;
    for (uint32_t w = 0; w < warp; ++w) warp_off += warp_sums[w];
    uint32_t carry = carry_s;
    if (idx < tiles) tile_sums[idx] = carry + warp_off + incl - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry_s = carry + warp_off + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total_out = carry_s;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ out, const uint32_t* __restrict__ tile_sums,
                                                  const uint32_t* len_dev, uint32_t len_host,
                                                  const uint32_t* __restrict__ total, const Control* ctl) {
  if (ctl && ctl->abort) return;
  uint32_t len = scan_len(len_dev, len_host);
  uint32_t base = blockIdx.x * kScanTile;
  if (base >= len) return;
  uint32_t add = tile_sums[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < kScanTile; k += 256) {
    uint32_t idx = base + k;
    if (idx < len) out[idx] += add;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[len] = *total;
}

// Each particle to its cell's range, in arrival order, with its id.
__global__ __launch_bounds__(256) void k_scatter(const float4* __restrict__ P, uint32_t n,
                                                 const uint2* __restrict__ slot_rank,
                                                 const uint32_t* __restrict__ start, uint2* __restrict__ pairs,
                                                 const Control* ctl) {
  if (ctl->abort) return;
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint2 sr = slot_rank[i];
  uint32_t id = id_of(P[i]);
  pairs[start[sr.x] + sr.y] = make_uint2(i, id);
}

// Sorted slot = cell start + number of cell-mates with a smaller id; moves
// x, id (P) and v, rho (V) there. `src` (optional) keeps each sorted slot's
// index before the sort.
__global__ __launch_bounds__(256) void k_rank_gather(const float4* __restrict__ PA, const float4* __restrict__ VA,
                                                     uint32_t n, float radius, const uint2* __restrict__ pairs,
                                                     const uint32_t* __restrict__ start, float4* __restrict__ PB,
                                                     float4* __restrict__ VB, const Control* ctl,
                                                     uint32_t* __restrict__ src) {
  if (ctl->abort) return;
  uint32_t q = xcd_block() * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
  uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
  uint2 me = pairs[q];
  float4 p = PA[me.x];
  float4 v = VA[me.x];
  int64_t c = cell_slot(cell_coord(p.x, radius), cell_coord(p.y, radius), cell_coord(p.z, radius), lo, span);
  uint32_t s = start[c], e = start[c + 1];
  uint32_t rank = 0;
  for (uint32_t k = s; k < e; ++k) rank += pairs[k].y < me.y ? 1u : 0u;
  PB[s + rank] = p;
  VB[s + rank] = v;
  if (src) src[s + rank] = me.x;
}

// ---- the searched table (src/fluid/mod.rs:214-236 without the dense index) --
//
// A sort on a searched table: every particle's cell slot in the box as a
// 64-bit key (ascending keys are the reference's (cx, cy, cz) order), a
// stable radix sort of (key, index), the run heads of the sorted keys and
// their exclusive scan (the occupied cells, with their first positions),
// then the gather, which ranks each particle among its cell-mates by id as
// the dense sort's gather does.

__global__ __launch_bounds__(256) void k_searched_keys(const float4* __restrict__ P, uint32_t n, float radius,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ values,
                                                       const Control* ctl) {
  if (ctl->abort) return;
  const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
  const uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
  const float4 p = P[i];
  keys[i] = (uint64_t)cell_slot(cell_coord(p.x, radius), cell_coord(p.y, radius), cell_coord(p.z, radius), lo, span);
  values[i] = i;
}

__global__ __launch_bounds__(256) void k_searched_heads(const uint64_t* __restrict__ keys, uint32_t n,
                                                        uint32_t* __restrict__ heads, const Control* ctl) {
  if (ctl->abort) return;
  const uint32_t q = xcd_block() * blockDim.x + threadIdx.x;
  if (q >= n) return;
  heads[q] = q == 0 || keys[q] != keys[q - 1] ? 1u : 0u;
}

// With `scan` the exclusive scan of the heads (scan[n] = occupied cells):
// each cell's key and first position, and start[cells] = n.
__global__ __launch_bounds__(256) void k_searched_cells(const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ heads,
                                                        const uint32_t* __restrict__ scan, uint32_t n,
                                                        uint64_t* __restrict__ cell_keys,
                                                        uint32_t* __restrict__ start, Control* ctl) {
  if (ctl->abort) return;
  const uint32_t q = xcd_block() * blockDim.x + threadIdx.x;
  if (q < n && heads[q]) {
    cell_keys[scan[q]] = keys[q];
    start[scan[q]] = q;
  }
  if (q == 0) {
    start[scan[n]] = n;
    ctl->cells = scan[n];
  }
}

// Sorted position q: its cell's range from the scan, its slot the cell's
// start plus the number of cell-mates with a smaller id. `src` (optional)
// keeps each sorted slot's index before the sort.
__global__ __launch_bounds__(256) void k_searched_gather(const float4* __restrict__ PA, const float4* __restrict__ VA,
                                                         uint32_t n, const uint32_t* __restrict__ values,
                                                         const uint32_t* __restrict__ heads,
                                                         const uint32_t* __restrict__ scan,
                                                         const uint32_t* __restrict__ start, float4* __restrict__ PB,
                                                         float4* __restrict__ VB, const Control* ctl,
                                                         uint32_t* __restrict__ src) {
  if (ctl->abort) return;
  const uint32_t q = xcd_block() * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t cell = scan[q] + heads[q] - 1u;
  const uint32_t s = start[cell], e = start[cell + 1];
  const uint32_t me = values[q];
  const float4 p = PA[me];
  const uint32_t id = id_of(p);
  uint32_t rank = 0;
  for (uint32_t k = s; k < e; ++k) rank += id_of(PA[values[k]]) < id ? 1u : 0u;
  PB[s + rank] = p;
  VB[s + rank] = VA[me];
  if (src) src[s + rank] = me;
}

// ---- CSR neighbour lists (readback / list density only) ----------------

// Visits, in the reference's order, every particle of `P` within reach of
// `at` through the cell table, calling f(j) for each (near, mod.rs:264-277:
// |x_j - x_i|^2 < r^2).
template <class F>
__device__ __forceinline__ void for_each_near(const float4* __restrict__ P, const GridView& g, float4 at, float radius,
                                              float reach, F&& f) {
  int32_t cx = cell_coord(at.x, radius), cy = cell_coord(at.y, radius), cz = cell_coord(at.z, radius);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        uint32_t s, e;
        if (!grid_range(g, (int64_t)cx + dx, (int64_t)cy + dy, (int64_t)cz + dz, s, e)) continue;
        for (uint32_t j = s; j < e; ++j) {
          float4 o = P[j];
          float ox = o.x - at.x, oy = o.y - at.y, oz = o.z - at.z;
          if (dot3(ox, oy, oz, ox, oy, oz) < reach) f(j);
        }
      }
}

// kind 0: fluid neighbours from `lp` (the positions the table was built on);
// kind 1: walls. Counts per particle.
__global__ __launch_bounds__(256) void k_list_count(const float4* __restrict__ lp, StepParams prm,
                                                    uint32_t* __restrict__ counts, int kind) {
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= prm.n) return;
  uint32_t cnt = 0;
  float4 at = lp[i];
  if (kind == 0) {
    for_each_near(lp, fluid_grid(prm), at, prm.radius, prm.reach, [&](uint32_t) { ++cnt; });
  } else if (prm.wall_cells) {
    for_each_near(prm.walls, wall_grid(prm), at, prm.radius, prm.reach, [&](uint32_t) { ++cnt; });
  }
  counts[i] = cnt;
}

__global__ __launch_bounds__(256) void k_list_emit(const float4* __restrict__ lp, StepParams prm,
                                                   const uint32_t* __restrict__ list_start,
                                                   uint32_t* __restrict__ items, int kind) {
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= prm.n) return;
  uint32_t at_k = list_start[i];
  float4 at = lp[i];
  if (kind == 0) {
    for_each_near(lp, fluid_grid(prm), at, prm.radius, prm.reach, [&](uint32_t j) { items[at_k++] = j; });
  } else if (prm.wall_cells) {
    for_each_near(prm.walls, wall_grid(prm), at, prm.radius, prm.reach, [&](uint32_t j) { items[at_k++] = j; });
  }
}

// compute_densities (src/fluid/mod.rs:621-675) over the CSR lists, with the
// current positions, as the reference does after a step.
__global__ __launch_bounds__(256) void k_list_density(const float4* __restrict__ P, float4* __restrict__ V,
                                                      StepParams prm, const uint32_t* __restrict__ fstart,
                                                      const uint32_t* __restrict__ fitems,
                                                      const uint32_t* __restrict__ wstart,
                                                      const uint32_t* __restrict__ witems) {
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= prm.n) return;
  float4 at = P[i];
  float sum = 0.0f;
  for (uint32_t k = fstart[i]; k < fstart[i + 1]; ++k) {
    float4 o = P[fitems[k]];
    float ox = o.x - at.x, oy = o.y - at.y, oz = o.z - at.z;
    sum += kernel_value(sqrtf(dot3(ox, oy, oz, ox, oy, oz)), prm.radius, prm.rinv, prm.scale);
  }
  float wsum = 0.0f;
  for (uint32_t k = wstart[i]; k < wstart[i + 1]; ++k) {
    float4 o = prm.walls[witems[k]];
    float ox = o.x - at.x, oy = o.y - at.y, oz = o.z - at.z;
    wsum += kernel_value(sqrtf(dot3(ox, oy, oz, ox, oy, oz)), prm.radius, prm.rinv, prm.scale);
  }
  float density = prm.mass * (sum + wsum);
  if (prm.terrain.count) {
    // the ground's sites (mod.rs:657-663)
    float ground, sx, sy, sz;
    ground_near(prm.terrain, at.x, at.y, at.z, ground, sx, sy, sz);
    density = density + ground;
  }
  V[i].w = density;
}

// ---- launchers ------------------------------------------------------------

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

void launch_step_begin(hipStream_t s, Control* ctl) { hipLaunchKernelGGL(k_step_begin, 1, 64, 0, s, ctl); }
void launch_substep_begin(hipStream_t s, Control* ctl) { hipLaunchKernelGGL(k_substep_begin, 1, 64, 0, s, ctl); }

void launch_fastest(hipStream_t s, const float4* V, uint32_t n, Control* ctl) {
  uint32_t blocks = blocks_for(n, 256);
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_fastest, blocks, 256, 0, s, V, n, ctl);
}

void launch_bounds(hipStream_t s, const float4* P, uint32_t n, float radius, Control* ctl) {
  uint32_t blocks = blocks_for(n, 256);
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_bounds, blocks, 256, 0, s, P, n, radius, ctl);
}

void launch_grid_setup(hipStream_t s, Control* ctl, uint32_t n, uint64_t capacity) {
  hipLaunchKernelGGL(k_grid_setup, 1, 64, 0, s, ctl, n, capacity);
}

hipError_t searched_sort_bytes(uint32_t n, size_t* bytes) {
  *bytes = 0;
  return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                   (const uint32_t*)nullptr, (uint32_t*)nullptr, n, 0, 64);
}

hipError_t launch_searched_sort(hipStream_t s, const SearchedSort& t, const float4* PA, const float4* VA, uint32_t n,
                                float radius, float4* PB, float4* VB, Control* ctl, uint32_t* src) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = blocks_for(n, 256);
  hipLaunchKernelGGL(k_searched_keys, blocks, 256, 0, s, PA, n, radius, t.keys_in, t.values_in, ctl);
  size_t bytes = t.temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(t.temp, bytes, t.keys_in, t.keys, t.values_in, t.values, n, 0, 64, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_searched_heads, blocks, 256, 0, s, t.keys, n, t.heads, ctl);
  launch_scan(s, t.heads, t.scan, t.tile_sums, t.total, nullptr, n, 0, ctl);
  hipLaunchKernelGGL(k_searched_cells, blocks, 256, 0, s, t.keys, t.heads, t.scan, n, t.cell_keys, t.start, ctl);
  hipLaunchKernelGGL(k_searched_gather, blocks, 256, 0, s, PA, VA, n, t.values, t.heads, t.scan, t.start, PB, VB,
                     ctl, src);
  return hipGetLastError();
}

void launch_count(hipStream_t s, const float4* P, uint32_t n, float radius, uint32_t* counts, uint2* slot_rank,
                  Control* ctl) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_count, blocks_for(n, 256), 256, 0, s, P, n, radius, counts, slot_rank, ctl);
}

void launch_scan(hipStream_t s, uint32_t* in, uint32_t* out, uint32_t* tile_sums, uint32_t* total,
                 const uint32_t* len_dev, uint64_t len_max, int zero_input, const Control* ctl) {
  uint32_t tiles = blocks_for(len_max, kScanTile);
  if (tiles == 0) tiles = 1;
  uint32_t len_host = (uint32_t)len_max;
  hipLaunchKernelGGL(k_scan_tiles, tiles, 256, 0, s, in, out, tile_sums, len_dev, len_host, zero_input, ctl);
  hipLaunchKernelGGL(k_scan_sums, 1, 1024, 0, s, tile_sums, len_dev, len_host, total, ctl);
  hipLaunchKernelGGL(k_scan_add, tiles, 256, 0, s, out, tile_sums, len_dev, len_host, total, ctl);
}

void launch_scatter(hipStream_t s, const float4* P, uint32_t n, const uint2* slot_rank, const uint32_t* start,
                    uint2* pairs, const Control* ctl) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scatter, blocks_for(n, 256), 256, 0, s, P, n, slot_rank, start, pairs, ctl);
}

void launch_rank_gather(hipStream_t s, const float4* PA, const float4* VA, uint32_t n, float radius,
                        const uint2* pairs, const uint32_t* start, float4* PB, float4* VB, const Control* ctl,
                        uint32_t* src) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_rank_gather, blocks_for(n, 256), 256, 0, s, PA, VA, n, radius, pairs, start, PB, VB, ctl,
                     src);
}

void launch_list_count(hipStream_t s, const float4* lp, const StepParams& prm, uint32_t* counts, int kind) {
  if (prm.n == 0) return;
  hipLaunchKernelGGL(k_list_count, blocks_for(prm.n, 256), 256, 0, s, lp, prm, counts, kind);
}

void launch_list_emit(hipStream_t s, const float4* lp, const StepParams& prm, const uint32_t* list_start,
                      uint32_t* items, int kind) {
  if (prm.n == 0) return;
  hipLaunchKernelGGL(k_list_emit, blocks_for(prm.n, 256), 256, 0, s, lp, prm, list_start, items, kind);
}

void launch_list_density(hipStream_t s, const float4* P, float4* V, const StepParams& prm, const uint32_t* fstart,
                         const uint32_t* fitems, const uint32_t* wstart, const uint32_t* witems) {
  if (prm.n == 0) return;
  hipLaunchKernelGGL(k_list_density, blocks_for(prm.n, 256), 256, 0, s, P, V, prm, fstart, fitems, wstart, witems);
}

}  // namespace sfl
