// Batched WCSPH on gfx950: many independent scenes stepped by one sequence
// of launches. Each block of a particle launch belongs to one scene (the
// block table maps logical block b to (scene, block of the scene)), so the
// scene's box, walls, container and substep scalars are wave-uniform.
//
// Per scene the arithmetic is that of a World stepping alone
// (src/fluid/step.rs:102-339, src/fluid/mod.rs:540-675): the same
// expressions in the same order, through the same device functions, over
// the same neighbour order, because the sort keeps scene s in its own index
// range and orders it by (cell, id) exactly as the single world does.
//
// Launch sequence of one step: plan, then `slots` substep slots, then the
// step end. A slot whose index is at or past the step's largest substep
// count does nothing; a scene whose own count is reached skips the slot
// (its cell range holds one slot with all its particles, so the cell table
// offsets of later scenes stay right). Kernels read the substep index from
// the cursor, which only changes between launches.
#include "batch.hpp"

namespace sfl {

namespace {

// Logical block -> (scene, particle index); `live` false past the scene's end.
template <uint32_t BS>
__device__ __forceinline__ uint32_t batch_index(const BatchArgs& b, const uint2* __restrict__ map, uint32_t& scene,
                                                bool& live) {
  uint2 m = map[xcd_block()];
  scene = m.x;
  const SceneDesc& d = b.desc[m.x];
  uint32_t local = m.y * BS + threadIdx.x;
  live = local < d.count;
  return d.first + local;
}

__device__ __forceinline__ bool slot_dead(const BatchArgs& b) {
  return b.ctl->abort || b.cur->k >= b.cur->max_sub;
}

__device__ __forceinline__ bool scene_active(const BatchArgs& b, const SceneState& st) {
  return b.cur->k < st.substeps;
}

__device__ __forceinline__ uint32_t scene_slot(const SceneState& st, float4 p, float radius) {
  int32_t lo[3] = {st.box.lo[0], st.box.lo[1], st.box.lo[2]};
  uint32_t span[3] = {st.box.span[0], st.box.span[1], st.box.span[2]};
  int64_t c = cell_slot(cell_coord(p.x, radius), cell_coord(p.y, radius), cell_coord(p.z, radius), lo, span);
  return st.slot_base + (uint32_t)c;
}

// StepParams of scene s for the current substep (the fields World::params
// fills, src/fluid/step.rs:149-165, 254-259, 282-283).
__device__ __forceinline__ StepParams scene_params(const BatchArgs& b, const SceneState& st, const SceneDesc& d) {
  StepParams p;
  p.movers = MoverView{};
  p.terrain = TerrainView{};
  p.n = b.n;
  p.radius = b.radius;
  p.rinv = b.rinv;
  p.reach = b.reach;
  p.scale = b.scale;
  p.mass = b.mass;
  p.rest = b.rest;
  p.stiffness = b.stiffness;
  p.dt = st.dt;
  p.nu = st.nu;
  p.wall_nu = st.wall_nu;
  p.soften = b.soften;
  p.gx = b.gx;
  p.gy = b.gy;
  p.gz = b.gz;
  p.half_spacing = b.half_spacing;
  p.has_bounds = d.has_bounds;
  for (int a = 0; a < 3; ++a) {
    p.blo[a] = d.blo[a];
    p.bhi[a] = d.bhi[a];
  }
  p.walls = b.walls;
  p.wall_start = b.wall_start + d.wall_base;
  p.wall_keys = nullptr;  // a batch keeps dense wall tables
  p.wall_keyed = 0;
  p.wall_box = d.wall_box;
  p.wall_cells = d.wall_cells;
  p.cell_start = b.cell_start + st.slot_base;
  p.cell_keys = nullptr;
  p.ctl = b.ctl;
  p.pairs = nullptr;
  p.pair_counts = nullptr;
  p.pair_stride = 0;
  p.pair_cap = 0;
  return p;
}

}  // namespace

// fastest (src/fluid/step.rs:92-96) of every scene: max |v|.
template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_b_fastest(BatchArgs b, const float4* __restrict__ V,
                                                  const uint2* __restrict__ map) {
  uint32_t s;
  bool live;
  uint32_t i = batch_index<BS>(b, map, s, live);
  float most = 0.0f;
  if (live) {
    float4 v = V[i];
    most = rmax(most, sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z)));
  }
  wave_max_to(&b.state[s].vmax_bits, most);
}

// Before a run of launches (outside any graph): clears the stop word and
// the flags. A fresh call starts at substep 0 with a new target; a rerun
// after a regrow keeps the cursor where the aborted launch left it.
__global__ void k_b_reset(BatchArgs b, int fresh, uint32_t steps, float dt, uint32_t slots) {
  if (threadIdx.x != 0) return;
  BatchCursor* c = b.cur;
  if (fresh) {
    c->k = 0;
    c->target = c->step + steps;
    c->dt = dt;
  }
  c->slots = slots;
  b.ctl->abort = c->step >= c->target ? 1u : 0u;
  b.ctl->flags = 0;
}

// wcsph / wcsph_steps (src/fluid/step.rs:102-142) for every scene at the
// start of a step: substep count and length from the scene's fastest
// particle in f32, then the viscosities of that length. A step resumed
// after a cell-table regrow (k > 0) keeps its plan.
__global__ __launch_bounds__(1024) void k_b_plan(BatchArgs b) {
  if (b.ctl->abort || b.cur->k != 0) return;
  __shared__ uint32_t most_s;
  if (threadIdx.x == 0) most_s = 0;
  __syncthreads();
  const float c = b.sound_speed, r = b.radius, dt = b.cur->dt;
  uint32_t most = 0;
  for (uint32_t s = threadIdx.x; s < b.scenes; s += blockDim.x) {
    SceneState& st = b.state[s];
    float vmax = __uint_as_float(st.vmax_bits);
    float largest = b.courant * r / (c + vmax);
    float wanted = rmax(ceilf(dt / largest), 1.0f);
    float cap = (float)b.max_substeps;
    float take = rmin(wanted, cap);
    uint32_t n = take >= 4294967295.0f ? 0xffffffffu : (uint32_t)take;
    float small = dt / (float)n;
    st.substeps = b.desc[s].count ? n : 0u;
    st.capped = wanted > cap ? 1u : 0u;
    st.dt = small;
    st.nu = rmin(b.viscosity * c * r, 0.05f * r * r / small);
    st.wall_nu = rmin(1.0f * c * r, 0.25f * (r * r / small));
    st.held = 0;
    st.nonfinite = 0;
    most = max(most, st.substeps);
  }
  atomicMax(&most_s, most);
  __syncthreads();
  if (threadIdx.x == 0) {
    b.cur->max_sub = most_s;
    if (most_s > b.cur->slots) {
      b.ctl->flags |= kFlagSlots;
      b.ctl->abort = 1;
    }
  }
}

// Start of a substep slot: resets the bounds of the scenes that step in it
// and their report reductions; a dead slot scans an empty cell table.
__global__ __launch_bounds__(1024) void k_b_substep_begin(BatchArgs b) {
  if (b.ctl->abort) return;
  uint32_t k = b.cur->k;
  if (k >= b.cur->max_sub) {
    if (threadIdx.x == 0) b.ctl->volume = 0;
    return;
  }
  for (uint32_t s = threadIdx.x; s < b.scenes; s += blockDim.x) {
    SceneState& st = b.state[s];
    if (k >= st.substeps) continue;
    st.bmin[0] = st.bmin[1] = st.bmin[2] = 0x7fffffff;
    st.bmax[0] = st.bmax[1] = st.bmax[2] = (int32_t)0x80000000u;
    st.vmax_bits = 0;
    st.comp_bits = 0;
  }
}

template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_b_bounds(BatchArgs b, const float4* __restrict__ P,
                                                 const uint2* __restrict__ map) {
  if (slot_dead(b)) return;
  uint32_t s;
  bool live;
  uint32_t i = batch_index<BS>(b, map, s, live);
  SceneState& st = b.state[s];
  if (!scene_active(b, st)) return;
  // the extreme coordinates' cells are the extreme cells (k_bounds); a NaN
  // coordinate's cell is 0
  float fmn[3] = {INFINITY, INFINITY, INFINITY};
  float fmx[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool nan[3] = {false, false, false};
  if (live) {
    const float4 p = P[i];
    const float v[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      fmn[a] = fmx[a] = v[a];
      nan[a] = !(v[a] == v[a]);
    }
  }
  int32_t mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    for (int off = 32; off >= 1; off >>= 1) {
      fmn[a] = fminf(fmn[a], __shfl_xor(fmn[a], off, 64));
      fmx[a] = fmaxf(fmx[a], __shfl_xor(fmx[a], off, 64));
    }
    mn[a] = cell_coord(fmn[a], b.radius);
    mx[a] = cell_coord(fmx[a], b.radius);
    if (__any(nan[a])) {
      mn[a] = min(mn[a], 0);
      mx[a] = max(mx[a], 0);
    }
  }
  // the block's waves fold in LDS, then one lane per block updates the
  // scene (every lane of a block belongs to one scene)
  constexpr uint32_t kWaves = BS / 64;
  __shared__ int32_t part[kWaves][6];
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      part[w][a] = mn[a];
      part[w][3 + a] = mx[a];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    int a = threadIdx.x;
    int32_t v = part[0][a];
    for (uint32_t k = 1; k < kWaves; ++k) v = a < 3 ? min(v, part[k][a]) : max(v, part[k][a]);
    if (a < 3)
      atomicMin(&st.bmin[a], v);
    else
      atomicMax(&st.bmax[a - 3], v);
  }
}

// One block: every scene's cell box and the first slot of its range (scenes
// in order). A scene that does not step in this slot takes one slot holding
// all its particles, counted here. The total goes to the control block; more
// than `capacity` slots aborts the sequence for the host to grow the table.
__global__ __launch_bounds__(1024) void k_b_grid_setup(BatchArgs b, uint32_t* __restrict__ counts,
                                                       uint64_t capacity) {
  if (slot_dead(b)) return;
  __shared__ uint64_t sums[1024];
  __shared__ uint64_t carry_s;
  __shared__ int big_s;
  if (threadIdx.x == 0) {
    carry_s = 0;
    big_s = 0;
  }
  __syncthreads();
  const uint32_t k = b.cur->k;
  for (uint32_t base = 0; base < b.scenes; base += blockDim.x) {
    uint32_t s = base + threadIdx.x;
    uint64_t vol = 0;
    bool parked = false;
    if (s < b.scenes) {
      SceneState& st = b.state[s];
      if (b.desc[s].count == 0) {
        vol = 0;
      } else if (k >= st.substeps) {
        vol = 1;
        parked = true;
      } else {
        vol = 1;
        bool big = false;
        for (int a = 0; a < 3; ++a) {
          int64_t sp = (int64_t)st.bmax[a] - (int64_t)st.bmin[a] + 1;
          if (sp <= 0 || sp > 0xffffffffll) {
            big = true;
            sp = 1;
          }
          st.box.lo[a] = st.bmin[a];
          st.box.span[a] = (uint32_t)sp;
          uint64_t prod = vol * (uint64_t)sp;
          if (prod / (uint64_t)sp != vol || prod >= 0xffffffffull) big = true;
          vol = prod;
        }
        if (big) {
          big_s = 1;
          vol = 1;
        }
      }
    }
    sums[threadIdx.x] = vol;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
      uint64_t add = threadIdx.x >= off ? sums[threadIdx.x - off] : 0;
      __syncthreads();
      sums[threadIdx.x] += add;
      __syncthreads();
    }
    uint64_t carry = carry_s;
    uint64_t at = carry + sums[threadIdx.x] - vol;
    if (s < b.scenes) {
      b.state[s].slot_base = at < 0xffffffffull ? (uint32_t)at : 0xffffffffu;
      if (parked && at < capacity) counts[at] = b.desc[s].count;
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry_s = carry + sums[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint64_t total = carry_s;
    if (big_s || total > capacity || total >= 0xffffffffull) {
      b.ctl->flags |= kFlagOverflow;
      b.ctl->abort = 1;
      b.ctl->need_volume = big_s ? 0xffffffffffffffffull : total;
    } else {
      b.ctl->volume = (uint32_t)total;
    }
  }
}

template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_b_count(BatchArgs b, const float4* __restrict__ P,
                                                uint32_t* __restrict__ counts, uint2* __restrict__ slot_rank,
                                                const uint2* __restrict__ map) {
  if (slot_dead(b)) return;
  uint32_t s;
  bool live;
  uint32_t i = batch_index<BS>(b, map, s, live);
  const SceneState& st = b.state[s];
  if (!scene_active(b, st)) return;  // the same for the whole block
  const uint32_t slot = live ? scene_slot(st, P[i], b.radius) : 0xffffffffu;
  const uint32_t rank = wave_rank(counts, slot, live);
  if (live) slot_rank[i] = make_uint2(slot, rank);
}

template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_b_scatter(BatchArgs b, const float4* __restrict__ P,
                                                  const uint2* __restrict__ slot_rank,
                                                  const uint32_t* __restrict__ start, uint2* __restrict__ pairs,
                                                  const uint2* __restrict__ map) {
  if (slot_dead(b)) return;
  uint32_t s;
  bool live;
  uint32_t i = batch_index<BS>(b, map, s, live);
  if (!live || !scene_active(b, b.state[s])) return;
  uint2 sr = slot_rank[i];
  pairs[start[sr.x] + sr.y] = make_uint2(i, id_of(P[i]));
}

// Sorted slot = cell start + number of cell-mates with a smaller id.
template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_b_gather(BatchArgs b, const float4* __restrict__ PA,
                                                 const float4* __restrict__ VA, const uint2* __restrict__ pairs,
                                                 const uint32_t* __restrict__ start, float4* __restrict__ PB,
                                                 float4* __restrict__ VB, const uint2* __restrict__ map) {
  if (slot_dead(b)) return;
  uint32_t s;
  bool live;
  uint32_t q = batch_index<BS>(b, map, s, live);
  const SceneState& st = b.state[s];
  if (!live || !scene_active(b, st)) return;
  uint2 me = pairs[q];
  float4 p = PA[me.x];
  float4 v = VA[me.x];
  uint32_t c = scene_slot(st, p, b.radius);
  uint32_t s0 = start[c], e0 = start[c + 1];
  uint32_t rank = 0;
  for (uint32_t k = s0; k < e0; ++k) rank += pairs[k].y < me.y ? 1u : 0u;
  PB[s0 + rank] = p;
  VB[s0 + rank] = v;
}

// The density pass of the single world (k_wcsph_density) for the scenes
// that step in this slot.
template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_b_density(BatchArgs b, const float4* __restrict__ P, float4* __restrict__ V,
                                                  float* __restrict__ scaled, const uint2* __restrict__ map) {
  if (slot_dead(b)) return;
  uint32_t s;
  bool live;
  uint32_t i = batch_index<BS>(b, map, s, live);
  SceneState& st = b.state[s];
  if (!scene_active(b, st)) return;
  float comp = 0.0f;
  if (live) {
    const StepParams prm = scene_params(b, st, b.desc[s]);
    int32_t lo[3] = {st.box.lo[0], st.box.lo[1], st.box.lo[2]};
    uint32_t span[3] = {st.box.span[0], st.box.span[1], st.box.span[2]};
    float4 at = P[i];
    float sum = 0.0f;
    visit_dense(P, prm.cell_start, lo, span, at, prm.radius, prm.reach,
                [&](uint32_t, float4, float, float, float, float sq) {
                  sum += kernel_value(sqrtf(sq), prm.radius, prm.rinv, prm.scale);
                });
    float wsum = 0.0f;
    if (prm.wall_cells)  // a batch's walls are dense (searched ones are refused)
      visit_dense(prm.walls, prm.wall_start, prm.wall_box.lo, prm.wall_box.span, at, prm.radius, prm.reach,
                  [&](uint32_t, float4, float, float, float, float sq) {
                    wsum += kernel_value(sqrtf(sq), prm.radius, prm.rinv, prm.scale);
                  });
    float density = prm.mass * (sum + wsum);
    float pressure = tait_pressure(density, prm.rest, prm.stiffness);
    V[i].w = density;
    scaled[i] = pressure / (density * density);
    comp = density / prm.rest - 1.0f;
  }
  wave_max_to(&st.comp_bits, comp);
}

// Queue slots a lane of the force pass's two-phase walk.
constexpr uint32_t kBatchQueue = 16;

// The fused force, velocity and advance pass of the single world
// (k_wcsph_force) for the scenes that step in this slot.
template <uint32_t BS>
__global__ __launch_bounds__(BS, 8) void k_b_force(BatchArgs b, const float4* __restrict__ PB,
                                                const float4* __restrict__ VB, const float* __restrict__ scaled,
                                                float4* __restrict__ PA, float4* __restrict__ VA,
                                                const uint2* __restrict__ map) {
  if (slot_dead(b)) return;
  uint32_t s;
  bool live;
  uint32_t i = batch_index<BS>(b, map, s, live);
  SceneState& st = b.state[s];
  if (!scene_active(b, st)) return;
  __shared__ uint32_t queue[kBatchQueue * BS];
  float speed = 0.0f;
  uint32_t held = 0;
  if (live) {
    const StepParams prm = scene_params(b, st, b.desc[s]);
    int32_t lo[3] = {st.box.lo[0], st.box.lo[1], st.box.lo[2]};
    uint32_t span[3] = {st.box.span[0], st.box.span[1], st.box.span[2]};
    float4 at = PB[i];
    float4 vi = VB[i];
    float density = vi.w;
    float own = scaled[i];
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    // the single world's two-phase walk (visit_dense_queued): the launch
    // covers every scene, so it is as full as a large world's
    visit_dense_queued<kBatchQueue>(
        PB, prm.cell_start, lo, span, at, prm.radius, prm.reach, queue + threadIdx.x, BS,
        // gathered inside the term: the kernel's 64 registers hold no more
        [](uint32_t) { return Nothing{}; },
        [&](uint32_t j, float4, float ox, float oy, float oz, float sq, Nothing) {
          if (sq <= 0.0f) return;
          float distance = sqrtf(sq);
          float f = kernel_slope(distance, prm.radius, prm.rinv, prm.scale) / distance;
          float gx = ox * f, gy = oy * f, gz = oz * f;
          const float4 vj = VB[j];
          float toward = rmin(dot3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, ox, oy, oz), 0.0f);
          float damp = -prm.nu * toward / (sq + prm.soften) * 2.0f / (density + vj.w);
          float kk = prm.mass * (own + scaled[j] + damp);
          px = px - gx * kk;
          py = py - gy * kk;
          pz = pz - gz * kk;
        });
    float wx = 0.0f, wy = 0.0f, wz = 0.0f;
    if (prm.wall_cells) {
      float pressure = tait_pressure(density, prm.rest, prm.stiffness);
      float mirrored = own + pressure / (prm.rest * prm.rest);
      float awx = 0.0f, awy = 0.0f, awz = 0.0f;
      auto walls = [&](auto&& f) {
        visit_dense(prm.walls, prm.wall_start, prm.wall_box.lo, prm.wall_box.span, at, prm.radius, prm.reach, f);
      };
      walls([&](uint32_t, float4, float ox, float oy, float oz, float sq) {
        float gx, gy, gz;
        kernel_gradient(ox, oy, oz, sq, prm.radius, prm.rinv, prm.scale, gx, gy, gz);
        float kk = prm.mass * mirrored;
        px = px - gx * kk;
        py = py - gy * kk;
        pz = pz - gz * kk;
        awx = awx + -gx;
        awy = awy + -gy;
        awz = awz + -gz;
      });
      wall_viscous(prm, at, vi.x, vi.y, vi.z, density, awx, awy, awz, walls, wx, wy, wz);
    }
    float ax = px + wx + prm.gx;
    float ay = py + wy + prm.gy;
    float az = pz + wz + prm.gz;
    float vx = vi.x + ax * prm.dt;
    float vy = vi.y + ay * prm.dt;
    float vz = vi.z + az * prm.dt;
    float nx, ny, nz;
    bool ok = advance_one(prm, at, vx, vy, vz, nx, ny, nz, held);
    if (!ok) {
      atomicOr(&b.ctl->flags, kFlagNonFinite);
      st.nonfinite = 1;
    }
    PA[i] = make_float4(nx, ny, nz, at.w);
    VA[i] = make_float4(vx, vy, vz, density);
    speed = sqrtf(dot3(vx, vy, vz, vx, vy, vz));
  }
  wave_max_to(&st.vmax_bits, speed);
  wave_sum_to(&st.held, held);
}

__global__ void k_b_substep_end(BatchArgs b) {
  if (threadIdx.x != 0) return;
  if (slot_dead(b)) return;
  b.cur->k += 1;
}

// End of a step: every scene's substep total, then the cursor moves to the
// next step. The sequence stops at the target or after a non-finite step.
__global__ __launch_bounds__(1024) void k_b_step_end(BatchArgs b) {
  if (b.ctl->abort) return;
  for (uint32_t s = threadIdx.x; s < b.scenes; s += blockDim.x) {
    SceneState& st = b.state[s];
    uint64_t t = ((uint64_t)st.substeps_total_hi << 32 | st.substeps_total_lo) + st.substeps;
    st.substeps_total_lo = (uint32_t)t;
    st.substeps_total_hi = (uint32_t)(t >> 32);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    BatchCursor* c = b.cur;
    c->k = 0;
    c->step += 1;
    if ((b.ctl->flags & kFlagNonFinite) || c->step >= c->target) b.ctl->abort = 1;
  }
}

// ---- launchers -------------------------------------------------------------

#define SFL_BATCH_LAUNCH(KERNEL, block, blocks, stream, ...)                                   \
  do {                                                                                         \
    if ((blocks) == 0) break;                                                                  \
    switch (block) {                                                                           \
      case 64: hipLaunchKernelGGL(KERNEL<64>, blocks, 64, 0, stream, __VA_ARGS__); break;       \
      case 128: hipLaunchKernelGGL(KERNEL<128>, blocks, 128, 0, stream, __VA_ARGS__); break;    \
      case 512: hipLaunchKernelGGL(KERNEL<512>, blocks, 512, 0, stream, __VA_ARGS__); break;    \
      case 1024: hipLaunchKernelGGL(KERNEL<1024>, blocks, 1024, 0, stream, __VA_ARGS__); break; \
      default: hipLaunchKernelGGL(KERNEL<256>, blocks, 256, 0, stream, __VA_ARGS__); break;     \
    }                                                                                          \
  } while (0)

void launch_batch_fastest(hipStream_t s, const BatchArgs& b, const float4* V, const uint2* map, uint32_t blocks,
                          uint32_t block) {
  SFL_BATCH_LAUNCH(k_b_fastest, block, blocks, s, b, V, map);
}
void launch_batch_reset(hipStream_t s, const BatchArgs& b, int fresh, uint32_t steps, float dt, uint32_t slots) {
  hipLaunchKernelGGL(k_b_reset, 1, 64, 0, s, b, fresh, steps, dt, slots);
}
void launch_batch_plan(hipStream_t s, const BatchArgs& b) { hipLaunchKernelGGL(k_b_plan, 1, 1024, 0, s, b); }
void launch_batch_substep_begin(hipStream_t s, const BatchArgs& b) {
  hipLaunchKernelGGL(k_b_substep_begin, 1, 1024, 0, s, b);
}
void launch_batch_bounds(hipStream_t s, const BatchArgs& b, const float4* P, const uint2* map, uint32_t blocks,
                         uint32_t block) {
  SFL_BATCH_LAUNCH(k_b_bounds, block, blocks, s, b, P, map);
}
void launch_batch_grid_setup(hipStream_t s, const BatchArgs& b, uint32_t* counts, uint64_t capacity) {
  hipLaunchKernelGGL(k_b_grid_setup, 1, 1024, 0, s, b, counts, capacity);
}
void launch_batch_count(hipStream_t s, const BatchArgs& b, const float4* P, uint32_t* counts, uint2* slot_rank,
                        const uint2* map, uint32_t blocks, uint32_t block) {
  SFL_BATCH_LAUNCH(k_b_count, block, blocks, s, b, P, counts, slot_rank, map);
}
void launch_batch_scatter(hipStream_t s, const BatchArgs& b, const float4* P, const uint2* slot_rank,
                          const uint32_t* start, uint2* pairs, const uint2* map, uint32_t blocks, uint32_t block) {
  SFL_BATCH_LAUNCH(k_b_scatter, block, blocks, s, b, P, slot_rank, start, pairs, map);
}
void launch_batch_gather(hipStream_t s, const BatchArgs& b, const float4* PA, const float4* VA, const uint2* pairs,
                         const uint32_t* start, float4* PB, float4* VB, const uint2* map, uint32_t blocks,
                         uint32_t block) {
  SFL_BATCH_LAUNCH(k_b_gather, block, blocks, s, b, PA, VA, pairs, start, PB, VB, map);
}
void launch_batch_density(hipStream_t s, const BatchArgs& b, const float4* P, float4* V, float* scaled,
                          const uint2* map, uint32_t blocks, uint32_t block) {
  SFL_BATCH_LAUNCH(k_b_density, block, blocks, s, b, P, V, scaled, map);
}
void launch_batch_force(hipStream_t s, const BatchArgs& b, const float4* PB, const float4* VB, const float* scaled,
                        float4* PA, float4* VA, const uint2* map, uint32_t blocks, uint32_t block) {
  SFL_BATCH_LAUNCH(k_b_force, block, blocks, s, b, PB, VB, scaled, PA, VA, map);
}
void launch_batch_substep_end(hipStream_t s, const BatchArgs& b) {
  hipLaunchKernelGGL(k_b_substep_end, 1, 64, 0, s, b);
}
void launch_batch_step_end(hipStream_t s, const BatchArgs& b) { hipLaunchKernelGGL(k_b_step_end, 1, 1024, 0, s, b); }

}  // namespace sfl
