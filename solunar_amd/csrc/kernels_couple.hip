// Coupled bodies on gfx950 (src/fluid/couple.rs, src/fluid/lift.rs): the
// body points placed where their bodies stand and carried through the
// substeps, sorted by (cell, owner, previous place) with a stable radix sort;
// each 512-particle chunk's damping on each body summed pair by pair in the
// reference's order and added to the bodies in chunk order; the fastest body
// near water; the lift each body reads from the water around it, and what a
// dynamic body's lift takes back out of that water.
//
// Arithmetic follows the reference's f32 operations in order, as the rest of
// the library does (device.hpp). Sums the reference folds in order are
// folded in order by one lane; only maxima, minima and counts use atomics.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>

#include "couple.hpp"
#include "device.hpp"
#include "launch.hpp"

namespace sfl {

namespace {

__device__ __forceinline__ void cross3(float ax, float ay, float az, float bx, float by, float bz, float& cx,
                                       float& cy, float& cz) {
  cx = ay * bz - by * az;
  cy = az * bx - bz * ax;
  cz = ax * by - bx * ay;
}

// Quat::rotate_vec3, taken as glam's v (w^2 - |u|^2) + u (2 (v.u)) +
// (u x v)(2 w), u = (x, y, z)
__device__ __forceinline__ void rotate(const float q[4], float vx, float vy, float vz, float& ox, float& oy,
                                       float& oz) {
  float ux = q[0], uy = q[1], uz = q[2], w = q[3];
  float b2 = dot3(ux, uy, uz, ux, uy, uz);
  float s = w * w - b2;
  float d = dot3(vx, vy, vz, ux, uy, uz) * 2.0f;
  float cx, cy, cz;
  cross3(ux, uy, uz, vx, vy, vz, cx, cy, cz);
  float w2 = w * 2.0f;
  ox = vx * s + ux * d + cx * w2;
  oy = vy * s + uy * d + cy * w2;
  oz = vz * s + uz * d + cz * w2;
}

// a point's arm and velocity from its body's pose: arm = R local,
// velocity = linear + angular x arm (couple.rs:271-274)
__device__ __forceinline__ void place(const BodyState& b, float4 local, float& ax, float& ay, float& az, float& vx,
                                      float& vy, float& vz) {
  rotate(b.rotation, local.x, local.y, local.z, ax, ay, az);
  float cx, cy, cz;
  cross3(b.angular[0], b.angular[1], b.angular[2], ax, ay, az, cx, cy, cz);
  vx = b.linear[0] + cx;
  vy = b.linear[1] + cy;
  vz = b.linear[2] + cz;
}

// order-preserving int keys of floats for atomic min / max
__device__ __forceinline__ int32_t ordered(float v) {
  int32_t b = __float_as_int(v);
  return b >= 0 ? b : b ^ 0x7fffffff;
}
__device__ __forceinline__ float unordered(int32_t k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

// f32::total_cmp as an unsigned key
__device__ __forceinline__ uint32_t total_key(float v) {
  uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float total_value(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ uint32_t owner_of(float4 p) { return __float_as_uint(p.w); }

}  // namespace

// ---- placement and carry (couple.rs:258-319) -----------------------------

__global__ __launch_bounds__(256) void k_movers_place(const float4* __restrict__ local, uint32_t m,
                                                      const BodyState* __restrict__ bodies, float4* __restrict__ MS,
                                                      float4* __restrict__ MV) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float4 p = local[i];
  const BodyState& b = bodies[owner_of(p)];
  float ax, ay, az, vx, vy, vz;
  place(b, p, ax, ay, az, vx, vy, vz);
  MS[i] = make_float4(b.center[0] + ax, b.center[1] + ay, b.center[2] + az, p.w);
  MV[i] = make_float4(vx, vy, vz, 0.0f);
}

__global__ void k_movers_box_reset(MoverGrid* g, const Control* ctl) {
  if (threadIdx.x != 0 || ctl->abort) return;
  for (int a = 0; a < 3; ++a) {
    g->lo[a] = 0x7fffffff;
    g->hi[a] = (int32_t)0x80000000u;
  }
  g->bad = 0;
  g->dense = 0;
  g->volume = 0;
}

__device__ __forceinline__ void carried(const float4* MS, const float4* MV, uint32_t i, float elapsed, float& x,
                                        float& y, float& z) {
  float4 s = MS[i], v = MV[i];
  x = s.x + v.x * elapsed;
  y = s.y + v.y * elapsed;
  z = s.z + v.z * elapsed;
}

__global__ __launch_bounds__(256) void k_movers_cells(const float4* __restrict__ MS, const float4* __restrict__ MV,
                                                      uint32_t m, float elapsed, float radius, MoverGrid* g,
                                                      const Control* ctl) {
  if (ctl->abort) return;
  int32_t lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int32_t hi[3] = {(int32_t)0x80000000u, (int32_t)0x80000000u, (int32_t)0x80000000u};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    float x, y, z;
    carried(MS, MV, i, elapsed, x, y, z);
    int32_t c[3] = {cell_coord(x, radius), cell_coord(y, radius), cell_coord(z, radius)};
    for (int a = 0; a < 3; ++a) {
      lo[a] = min(lo[a], c[a]);
      hi[a] = max(hi[a], c[a]);
    }
  }
  for (int a = 0; a < 3; ++a) {
    for (int off = 32; off >= 1; off >>= 1) {
      lo[a] = min(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = max(hi[a], __shfl_xor(hi[a], off, 64));
    }
  }
  // the block's box, then one set of atomics a block
  __shared__ int32_t wlo[4][3], whi[4][3];
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      wlo[wave][a] = lo[a];
      whi[wave][a] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x == 0 && blockIdx.x * blockDim.x < m) {
    for (int a = 0; a < 3; ++a) {
      int32_t l = wlo[0][a], h = whi[0][a];
      for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
        l = min(l, wlo[w][a]);
        h = max(h, whi[w][a]);
      }
      atomicMin(&g->lo[a], l);
      atomicMax(&g->hi[a], h);
    }
  }
}

__global__ __launch_bounds__(256) void k_movers_keys(const float4* __restrict__ MS, const float4* __restrict__ MV,
                                                     uint32_t m, float elapsed, float radius, uint32_t owner_bits,
                                                     MoverGrid* g, uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ perm, Control* ctl) {
  if (ctl->abort) return;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t span[3];
  for (int a = 0; a < 3; ++a) span[a] = (uint64_t)((int64_t)g->hi[a] - (int64_t)g->lo[a] + 1);
  // the slots fit beside the owner bits when volume <= (2^64 - 1) >> owner_bits
  const uint64_t room = ~uint64_t(0) >> owner_bits;
  bool fits = span[0] <= room && span[1] <= room / span[0] && span[2] <= room / (span[0] * span[1]);
  if (i == 0) {
    for (int a = 0; a < 3; ++a) g->span[a] = span[a];
    g->volume = fits ? span[0] * span[1] * span[2] : 0;
    if (!fits) {
      g->bad = 1;
      atomicOr(&ctl->flags, kFlagMoverBox);
    }
  }
  if (i >= m) return;
  float x, y, z;
  carried(MS, MV, i, elapsed, x, y, z);
  uint64_t key = 0;
  if (fits) {
    uint64_t cx = (uint64_t)((int64_t)cell_coord(x, radius) - g->lo[0]);
    uint64_t cy = (uint64_t)((int64_t)cell_coord(y, radius) - g->lo[1]);
    uint64_t cz = (uint64_t)((int64_t)cell_coord(z, radius) - g->lo[2]);
    key = (((cx * span[1] + cy) * span[2] + cz) << owner_bits) | owner_of(MS[i]);
  }
  keys[i] = key;
  perm[i] = i;
}

// The sorted order's start, velocity and carried position; each body's
// `moved` = linear velocity * elapsed (couple.rs:315-318).
__global__ __launch_bounds__(256) void k_movers_gather(const float4* __restrict__ MS, const float4* __restrict__ MV,
                                                       const uint32_t* __restrict__ perm, uint32_t m, float elapsed,
                                                       float4* __restrict__ MS2, float4* __restrict__ MV2,
                                                       float4* __restrict__ MP, BodyState* __restrict__ bodies,
                                                       uint32_t nbodies, const Control* ctl) {
  if (ctl->abort) return;
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nbodies) {
    BodyState& b = bodies[k];
    for (int a = 0; a < 3; ++a) b.moved[a] = b.linear[a] * elapsed;
  }
  if (k >= m) return;
  uint32_t i = perm[k];
  float4 s = MS[i], v = MV[i];
  MS2[k] = s;
  MV2[k] = v;
  MP[k] = make_float4(s.x + v.x * elapsed, s.y + v.y * elapsed, s.z + v.z * elapsed, s.w);
}

// Slot c of the points' dense cell table: the first point of slot c or later.
__global__ __launch_bounds__(256) void k_mover_table(const uint64_t* __restrict__ keys, uint32_t m,
                                                     uint32_t owner_bits, MoverGrid* g, uint32_t* __restrict__ table,
                                                     uint64_t cap, const Control* ctl) {
  if (ctl->abort) return;
  const uint64_t volume = g->volume;
  const bool dense = !g->bad && volume + 1 <= cap;
  if (blockIdx.x == 0 && threadIdx.x == 0) g->dense = dense ? 1u : 0u;
  if (!dense) return;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= volume;
       c += (uint64_t)gridDim.x * blockDim.x)
    table[c] = c == volume ? m : lower_bound64(keys, m, c << owner_bits);
}

// ---- the damping each chunk gives each body (step.rs:173-221) ------------

// The pair records' room against what the density pass counted: past it the
// substep aborts before its force pass (the step runs it again with more).
// Slots past the count take the key of no body; values number the slots.
__global__ __launch_bounds__(256) void k_pairs_check(const uint32_t* __restrict__ total, uint32_t cap,
                                                     uint32_t none, uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ value, Control* ctl) {
  if (ctl->abort) return;
  const uint32_t used = *total;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&ctl->pairs_used, used);
  if (used > cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      atomicOr(&ctl->flags, kFlagPairRoom);
      ctl->need_pairs = used;
      ctl->abort = 1;
    }
    return;
  }
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cap; k += gridDim.x * blockDim.x) {
    if (k >= used) key[k] = none;
    value[k] = k;
  }
}

// One wave per body: its reaching pairs, in particle order (a stable sort by
// body kept it), summed per chunk from +0 and added to the body's impulse as
// force * dt chunk by chunk (step.rs:213-221). The wave stages 64 records at
// a time in LDS; lane 0 folds them in order.
__global__ __launch_bounds__(64) void k_body_fold(const uint32_t* __restrict__ key, const uint32_t* __restrict__ value,
                                                  uint32_t cap, const float4* __restrict__ force,
                                                  const float4* __restrict__ torque, BodyState* __restrict__ bodies,
                                                  float dt, const Control* ctl) {
  if (ctl->abort) return;
  const uint32_t b = blockIdx.x;
  __shared__ float4 fs[64], ts[64];
  uint32_t first = 0, last = 0;
  {
    uint32_t lo = 0, hi = cap;
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (key[mid] < b) lo = mid + 1; else hi = mid;
    }
    first = lo;
    hi = cap;
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (key[mid] < b + 1) lo = mid + 1; else hi = mid;
    }
    last = lo;
  }
  BodyState& s = bodies[b];
  float im[3] = {s.impulse[0], s.impulse[1], s.impulse[2]};
  float an[3] = {s.angular_impulse[0], s.angular_impulse[1], s.angular_impulse[2]};
  float af[3] = {0.0f, 0.0f, 0.0f}, at[3] = {0.0f, 0.0f, 0.0f};
  uint32_t chunk = 0xffffffffu;
  bool open = false;
  for (uint32_t base = first; base < last; base += 64) {
    uint32_t k = base + threadIdx.x;
    if (k < last) {
      uint32_t r = value[k];
      fs[threadIdx.x] = force[r];
      ts[threadIdx.x] = torque[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t n = min(64u, last - base);
      for (uint32_t j = 0; j < n; ++j) {
        float4 f = fs[j], t = ts[j];
        uint32_t c = __float_as_uint(f.w);
        if (!open || c != chunk) {
          if (open)
            for (int a = 0; a < 3; ++a) {
              im[a] = im[a] + af[a] * dt;
              an[a] = an[a] + at[a] * dt;
            }
          open = true;
          chunk = c;
          for (int a = 0; a < 3; ++a) af[a] = at[a] = 0.0f;
        }
        af[0] = af[0] + f.x;
        af[1] = af[1] + f.y;
        af[2] = af[2] + f.z;
        at[0] = at[0] + t.x;
        at[1] = at[1] + t.y;
        at[2] = at[2] + t.z;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (open)
    for (int a = 0; a < 3; ++a) {
      im[a] = im[a] + af[a] * dt;
      an[a] = an[a] + at[a] * dt;
    }
  for (int a = 0; a < 3; ++a) {
    s.impulse[a] = im[a];
    s.angular_impulse[a] = an[a];
  }
}

// ---- the fastest body near water (couple.rs:321-366) ----------------------

// One block per body over its points in placement order (its own
// contiguous range): max |v| and the box of the points where carry(0) puts
// them (start + v 0, the start up to the sign of a zero, which no cell of
// the box minus or plus a positive reach tells apart). Max and min are f32
// max and min, exact in any order (couple.rs:335-338).
__global__ __launch_bounds__(256) void k_extent_bodies(const float4* __restrict__ MS, const float4* __restrict__ MV,
                                                       const BodyState* __restrict__ bodies,
                                                       BodyExtent* __restrict__ e) {
  const BodyState& b = bodies[blockIdx.x];
  float speed = 0.0f;
  float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
  float hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
  for (uint32_t k = threadIdx.x; k < b.points; k += blockDim.x) {
    float4 p = MS[b.first_point + k], v = MV[b.first_point + k];
    speed = fmaxf(speed, sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z)));
    float c[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], c[a]);
      hi[a] = fmaxf(hi[a], c[a]);
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    speed = fmaxf(speed, __shfl_xor(speed, off, 64));
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
  }
  __shared__ float ws[4][7];
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ws[wave][0] = speed;
    for (int a = 0; a < 3; ++a) {
      ws[wave][1 + a] = lo[a];
      ws[wave][4 + a] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
    speed = fmaxf(speed, ws[w][0]);
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], ws[w][1 + a]);
      hi[a] = fmaxf(hi[a], ws[w][4 + a]);
    }
  }
  BodyExtent& x = e[blockIdx.x];
  x.speed_bits = __float_as_uint(speed);
  for (int a = 0; a < 3; ++a) {
    x.lo[a] = ordered(lo[a]);
    x.hi[a] = ordered(hi[a]);
  }
  x.wet = 0;
}

// Any occupied cell of the fluid's last grid inside each body's cell range
// (lo, hi from the host): the grid's table and box are the last sort's.
__global__ __launch_bounds__(256) void k_extent_wet(const uint32_t* __restrict__ start,
                                                    const uint64_t* __restrict__ keys, const Control* ctl,
                                                    const WetQuery* __restrict__ q, uint32_t nq,
                                                    BodyExtent* __restrict__ e) {
  uint32_t b = blockIdx.y;
  if (b >= nq) return;
  const WetQuery& w = q[b];
  const GridView g{{ctl->lo[0], ctl->lo[1], ctl->lo[2]}, {ctl->span[0], ctl->span[1], ctl->span[2]}, start, keys,
                   keys ? ctl->cells : 0u};
  int64_t lo[3], n[3];
  uint64_t volume = 1;
  for (int a = 0; a < 3; ++a) {
    int64_t t0 = ctl->lo[a], t1 = (int64_t)ctl->lo[a] + ctl->span[a] - 1;
    lo[a] = max((int64_t)w.lo[a], t0);
    int64_t hi = min((int64_t)w.hi[a], t1);
    n[a] = hi - lo[a] + 1;
    if (n[a] <= 0) return;
    volume *= (uint64_t)n[a];
  }
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < volume;
       k += (uint64_t)gridDim.x * blockDim.x) {
    int64_t z = (int64_t)(k % n[2]), rest = (int64_t)(k / n[2]);
    int64_t y = rest % n[1], x = rest / n[1];
    uint32_t s, f;
    if (grid_range(g, lo[0] + x, lo[1] + y, lo[2] + z, s, f) && s < f) {
      e[w.body].wet = 1;
      return;
    }
  }
}

// ---- lifts (src/fluid/lift.rs:25-152) -------------------------------------

// probes' arms and velocities, and the box round them grown by the reach
__global__ __launch_bounds__(256) void k_lift_probes(const float4* __restrict__ probes,
                                                     const BodyState* __restrict__ bodies, uint32_t nbodies,
                                                     float4* __restrict__ arm, float4* __restrict__ vel,
                                                     LiftState* __restrict__ ls) {
  uint32_t b = blockIdx.y;
  if (b >= nbodies) return;
  const BodyState& s = bodies[b];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < s.probes; k += gridDim.x * blockDim.x) {
    uint32_t i = s.first_probe + k;
    float ax, ay, az, vx, vy, vz;
    place(s, probes[i], ax, ay, az, vx, vy, vz);
    arm[i] = make_float4(ax, ay, az, 0.0f);
    vel[i] = make_float4(vx, vy, vz, 0.0f);
    float c[3] = {s.center[0] + ax, s.center[1] + ay, s.center[2] + az};
    for (int a = 0; a < 3; ++a) {
      atomicMin(&ls[b].box_lo[a], ordered(c[a]));
      atomicMax(&ls[b].box_hi[a], ordered(c[a]));
    }
  }
}

__global__ void k_lift_reset(LiftState* ls, uint32_t nbodies) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbodies) return;
  LiftState& l = ls[b];
  for (int a = 0; a < 3; ++a) {
    l.box_lo[a] = ordered(3.40282347e38f);
    l.box_hi[a] = ordered(-3.40282347e38f);
  }
  l.near = 0;
}

__device__ __forceinline__ bool in_box(const LiftState& l, float4 p) {
  // at.clamp(low, high) == at
  float lo[3] = {l.low[0], l.low[1], l.low[2]}, hi[3] = {l.high[0], l.high[1], l.high[2]};
  float c[3] = {p.x, p.y, p.z};
  for (int a = 0; a < 3; ++a)
    if (rmin(rmax(c[a], lo[a]), hi[a]) != c[a]) return false;
  return true;
}

// low = box - reach, high = box + reach (lift.rs:57-61), one lane a body
__global__ void k_lift_reach(LiftState* __restrict__ ls, uint32_t nbodies, float reach) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbodies) return;
  LiftState& l = ls[b];
  for (int a = 0; a < 3; ++a) {
    l.low[a] = unordered(l.box_lo[a]) - reach;
    l.high[a] = unordered(l.box_hi[a]) + reach;
  }
}

constexpr uint32_t kLiftChunk = 1024;

// The box of each 1024 particles of the water (its order after the step is
// the last sort's, so a chunk is a compact piece of it).
__global__ __launch_bounds__(256) void k_chunk_boxes(const float4* __restrict__ P, uint32_t n,
                                                     float4* __restrict__ boxes) {
  const uint32_t c = blockIdx.x;
  float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
  float hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
  for (uint32_t k = threadIdx.x; k < kLiftChunk; k += blockDim.x) {
    uint32_t i = c * kLiftChunk + k;
    if (i >= n) break;
    float4 p = P[i];
    float q[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], q[a]);
      hi[a] = fmaxf(hi[a], q[a]);
    }
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
  __shared__ float ws[4][6];
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      ws[wave][a] = lo[a];
      ws[wave][3 + a] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (uint32_t w = 1; w < 4; ++w)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], ws[w][a]);
      hi[a] = fmaxf(hi[a], ws[w][3 + a]);
    }
  boxes[2 * c] = make_float4(lo[0], lo[1], lo[2], 0.0f);
  boxes[2 * c + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

__device__ __forceinline__ bool chunk_meets(const LiftState& l, const float4* boxes, uint32_t c) {
  float4 lo = boxes[2 * c], hi = boxes[2 * c + 1];
  return !(hi.x < l.low[0] || lo.x > l.high[0] || hi.y < l.low[1] || lo.y > l.high[1] || hi.z < l.low[2] ||
           lo.z > l.high[2]);
}

// Block (body, chunk): the chunk's water inside the body's box, counted into
// counts[body * chunks + chunk].
__global__ __launch_bounds__(256) void k_lift_count(const float4* __restrict__ P, uint32_t n,
                                                    const LiftState* __restrict__ ls, const float4* __restrict__ boxes,
                                                    uint32_t chunks, uint32_t* __restrict__ counts) {
  const uint32_t b = blockIdx.x, c = blockIdx.y;
  const LiftState& l = ls[b];
  __shared__ uint32_t sum_s;
  if (!chunk_meets(l, boxes, c)) {
    if (threadIdx.x == 0) counts[(size_t)b * chunks + c] = 0;
    return;
  }
  if (threadIdx.x == 0) sum_s = 0;
  __syncthreads();
  uint32_t inside = 0;
  for (uint32_t k = threadIdx.x; k < kLiftChunk; k += blockDim.x) {
    uint32_t i = c * kLiftChunk + k;
    if (i < n && in_box(l, P[i])) ++inside;
  }
  for (int off = 32; off >= 1; off >>= 1) inside += __shfl_xor(inside, off, 64);
  if ((threadIdx.x & 63) == 0 && inside) atomicAdd(&sum_s, inside);
  __syncthreads();
  if (threadIdx.x == 0) counts[(size_t)b * chunks + c] = sum_s;
}

// Each body's water count and its offset in the near lists, from the scan
// of the (body, chunk) counts.
__global__ void k_lift_offsets(LiftState* __restrict__ ls, uint32_t nbodies, const uint32_t* __restrict__ offsets,
                               uint32_t chunks) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbodies) return;
  uint32_t first = offsets[(size_t)b * chunks], last = offsets[(size_t)(b + 1) * chunks];
  ls[b].near = last - first;
  ls[b].near_offset = first;
}

// Block (body, chunk) of 1024: the chunk's water inside the body's box in
// particle order at its offset, and each column's highest water
// (lift.rs:71-84) by total_cmp, as 1 << 32 | key in the column table.
__global__ __launch_bounds__(1024) void k_lift_near(const float4* __restrict__ P, uint32_t n,
                                                    const LiftState* __restrict__ ls, const float4* __restrict__ boxes,
                                                    uint32_t chunks, const uint32_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ near, unsigned long long* __restrict__ columns,
                                                    uint64_t* __restrict__ column_keys,
                                                    uint32_t* __restrict__ column_tops, float column, float upx,
                                                    float upy, float upz) {
  const uint32_t b = blockIdx.x, c = blockIdx.y;
  const LiftState& l = ls[b];
  const size_t at = (size_t)b * chunks + c;
  if (offsets[at + 1] == offsets[at]) return;
  __shared__ uint32_t wave_sums[16];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i = c * kLiftChunk + threadIdx.x;
  bool inside = false;
  float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i < n) {
    p = P[i];
    inside = in_box(l, p);
  }
  unsigned long long ballot = __ballot(inside);
  uint32_t before = __popcll(ballot & ((1ull << lane) - 1ull));
  if (lane == 0) wave_sums[wave] = __popcll(ballot);
  __syncthreads();
  if (!inside) return;
  uint32_t offset = offsets[at];
  for (uint32_t w = 0; w < wave; ++w) offset += wave_sums[w];
  near[offset + before] = i;
  int64_t kx = (int64_t)cell_coord(p.x, column) - l.col_lo[0];
  int64_t kz = (int64_t)cell_coord(p.z, column) - l.col_lo[1];
  float height = dot3(p.x, p.y, p.z, upx, upy, upz);
  if (l.sorted) {
    // the column as (x, z) in the reference's key order, for the sort
    column_keys[offset + before] = (uint64_t)(uint32_t)kx << 32 | (uint32_t)kz;
    column_tops[offset + before] = total_key(height);
    return;
  }
  unsigned long long* table = columns + l.column_offset;
  atomicMax(&table[kx * l.col_span[1] + kz], (1ull << 32) | total_key(height));
}

// One block per body: the columns' tops sorted (bitonic, total_cmp), the
// level from the mean of their middle half; then lanes 0-3 fold the flow of
// the water under the level and the probes' wetted share and middle, one
// component a lane, each component's sum in the reference's order
// (lift.rs:85-126).
__global__ __launch_bounds__(1024) void k_lift_level(const float4* __restrict__ P, const float4* __restrict__ V,
                                                     LiftState* __restrict__ ls, const BodyState* __restrict__ bodies,
                                                     const uint32_t* __restrict__ near,
                                                     const unsigned long long* __restrict__ columns,
                                                     uint32_t* __restrict__ sortbuf, const float4* __restrict__ arm,
                                                     const float4* __restrict__ vel, float half_spacing, float upx,
                                                     float upy, float upz) {
  LiftState& l = ls[blockIdx.x];
  if (l.near == 0) return;
  const BodyState& body = bodies[blockIdx.x];
  const unsigned long long* table = columns + l.column_offset;
  uint32_t* keys = sortbuf + l.sort_offset;
  const uint64_t cols = (uint64_t)l.col_span[0] * l.col_span[1];
  __shared__ uint32_t wave_sums[16];
  __shared__ uint32_t base_s;
  if (threadIdx.x == 0) base_s = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // a sorted body's tops are already in place
  for (uint64_t first = 0; !l.sorted && first < cols; first += 1024) {
    uint64_t c = first + threadIdx.x;
    unsigned long long t = c < cols ? table[c] : 0ull;
    bool full = t != 0ull;
    unsigned long long ballot = __ballot(full);
    uint32_t before = __popcll(ballot & ((1ull << lane) - 1ull));
    if (lane == 0) wave_sums[wave] = __popcll(ballot);
    __syncthreads();
    uint32_t offset = base_s;
    for (uint32_t w = 0; w < wave; ++w) offset += wave_sums[w];
    if (full) keys[offset + before] = (uint32_t)t;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t all = 0;
      for (uint32_t w = 0; w < 16; ++w) all += wave_sums[w];
      base_s += all;
    }
    __syncthreads();
  }
  const uint32_t count = l.sorted ? l.tops : base_s;
  uint32_t size = 1;  // a power of two, at most the buffer's sort_size
  while (size < count) size <<= 1;
  for (uint32_t k = count + threadIdx.x; k < size; k += 1024) keys[k] = 0xffffffffu;
  __syncthreads();
  for (uint32_t k = 2; k <= size; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < size; i += 1024) {
        uint32_t partner = i ^ j;
        if (partner > i) {
          uint32_t a = keys[i], b = keys[partner];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            keys[i] = b;
            keys[partner] = a;
          }
        }
      }
      __syncthreads();
    }
  // the mean of the middle half, plus half a spacing
  __shared__ float level_s;
  __shared__ float4 buf[1024];
  if (threadIdx.x == 0) {
    uint32_t quarter = count / 4;
    float sum = 0.0f;
    for (uint32_t k = quarter; k < count - quarter; ++k) sum = sum + total_value(keys[k]);
    level_s = sum / (float)(count - 2 * quarter) + half_spacing;
  }
  __syncthreads();
  const float level = level_s;
  // the flow of the water under the level, in particle order: the block
  // stages 1024 particles at a time; thread a < 3 folds component a, each in
  // particle order, the same sums one lane would make, side by side
  const uint32_t* mine = near + l.near_offset;
  const uint32_t who = threadIdx.x;
  float acc = 0.0f;
  uint32_t wet = 0;
  for (uint32_t base = 0; base < l.near; base += 1024) {
    uint32_t k = base + threadIdx.x;
    if (k < l.near) {
      uint32_t i = mine[k];
      float4 p = P[i];
      float4 v = V[i];
      buf[threadIdx.x] = make_float4(v.x, v.y, v.z, dot3(p.x, p.y, p.z, upx, upy, upz) <= level ? 1.0f : 0.0f);
    }
    __syncthreads();
    if (who < 4) {
      uint32_t n = min(1024u, l.near - base);
      for (uint32_t j = 0; j < n; ++j) {
        float4 b = buf[j];
        if (b.w != 0.0f) {
          acc = acc + (who == 0 ? b.x : (who == 1 ? b.y : b.z));
          wet += 1;
        }
      }
    }
    __syncthreads();
  }
  // the probes' shares under the level: wetted, middle, the waterline
  float spacing = body.probe_spacing;
  float part = 0.0f;  // lane 0: wetted; lanes 1..3: middle x, y, z
  uint32_t line = 0;
  for (uint32_t base = 0; base < body.probes; base += 1024) {
    uint32_t k = base + threadIdx.x;
    if (k < body.probes) {
      float4 a = arm[body.first_probe + k];
      float cx = body.center[0] + a.x, cy = body.center[1] + a.y, cz = body.center[2] + a.z;
      float depth = level - dot3(cx, cy, cz, upx, upy, upz);
      float share = depth / spacing + 0.5f;
      if (share < 0.0f) share = 0.0f;
      if (share > 1.0f) share = 1.0f;
      // a share that is not > 0 (0 or NaN) leaves the probe out
      buf[threadIdx.x] = share > 0.0f ? make_float4(a.x * share, a.y * share, a.z * share, share)
                                      : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    }
    __syncthreads();
    if (who < 4) {
      uint32_t n = min(1024u, body.probes - base);
      for (uint32_t j = 0; j < n; ++j) {
        float4 b = buf[j];
        if (!(b.w > 0.0f)) continue;
        part = part + (who == 0 ? b.w : (who == 1 ? b.x : (who == 2 ? b.y : b.z)));
        if (b.w < 1.0f) ++line;
      }
    }
    __syncthreads();
  }
  if (who >= 4) return;
  if (who < 3) l.flow[who] = acc / (float)(wet > 1 ? wet : 1);
  if (who == 0) {
    l.level = level;
    l.wetted = part;
    l.volume = part * (spacing * (spacing * spacing));
    l.line = line;
  } else {
    l.middle[who - 1] = part;
  }
}

// One block per body, after the host has set area = volume^(2/3) / wetted:
// buoyancy, then the drag on every probe under, then the waterplane's
// damping on every probe on the waterline, as impulses (lift.rs:127-150).
// The block works each probe's term out in parallel; lanes 0-5 sum them, one
// force or torque component a lane, each in the reference's order.
__global__ __launch_bounds__(256) void k_lift_force(LiftState* __restrict__ ls, const BodyState* __restrict__ bodies,
                                                    const float4* __restrict__ arm, const float4* __restrict__ vel,
                                                    float rest, float gravity, float upx, float upy, float upz,
                                                    float dt) {
  const uint32_t b = blockIdx.x;
  LiftState& l = ls[b];
  if (l.near == 0 || l.wetted <= 0.0f) {
    if (threadIdx.x == 0) l.lifted = 0;
    return;
  }
  const BodyState& body = bodies[b];
  const float spacing = body.probe_spacing;
  __shared__ float4 fs[256], ts[256];
  // lane a < 3 folds force component a, lane 3 + a torque component a, each
  // in probe order, from buoyancy and its torque
  const uint32_t lane = threadIdx.x;
  float acc = 0.0f;
  {
    float mx = l.middle[0] / l.wetted, my = l.middle[1] / l.wetted, mz = l.middle[2] / l.wetted;
    float lift = rest * gravity * l.volume;
    float bx = upx * lift, by = upy * lift, bz = upz * lift;
    float tx, ty, tz;
    cross3(mx, my, mz, bx, by, bz, tx, ty, tz);
    float start[6] = {bx, by, bz, tx, ty, tz};
    acc = 0.0f + start[lane < 6 ? lane : 0];
  }
  const float plane = (float)l.line * spacing * spacing;
  const float waves = rest * plane * sqrtf(gravity * sqrtf(plane));
  const float each = waves / (float)(l.line > 1 ? l.line : 1);
  for (int pass = 0; pass < 2; ++pass) {
    for (uint32_t base = 0; base < body.probes; base += 256) {
      uint32_t k = base + threadIdx.x;
      float4 fo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), to = fo;
      if (k < body.probes) {
        float4 a = arm[body.first_probe + k], v = vel[body.first_probe + k];
        float cx = body.center[0] + a.x, cy = body.center[1] + a.y, cz = body.center[2] + a.z;
        float depth = l.level - dot3(cx, cy, cz, upx, upy, upz);
        float share = depth / spacing + 0.5f;
        if (share < 0.0f) share = 0.0f;
        if (share > 1.0f) share = 1.0f;
        float dx, dy, dz;
        bool take;
        if (pass == 0) {
          // drag on the probe's share of the wetted area
          float px = l.flow[0] - v.x, py = l.flow[1] - v.y, pz = l.flow[2] - v.z;
          float kk = 0.5f * 1.0f * rest * l.area * share * sqrtf(dot3(px, py, pz, px, py, pz));
          dx = px * kk;
          dy = py * kk;
          dz = pz * kk;
          take = share > 0.0f;
        } else {
          // waves off the waterplane
          float kk = each * dot3(l.flow[0] - v.x, l.flow[1] - v.y, l.flow[2] - v.z, upx, upy, upz);
          dx = upx * kk;
          dy = upy * kk;
          dz = upz * kk;
          take = share > 0.0f && share < 1.0f;
        }
        if (take) {
          float qx, qy, qz;
          cross3(a.x, a.y, a.z, dx, dy, dz, qx, qy, qz);
          fo = make_float4(dx, dy, dz, 1.0f);
          to = make_float4(qx, qy, qz, 0.0f);
        }
      }
      fs[threadIdx.x] = fo;
      ts[threadIdx.x] = to;
      __syncthreads();
      if (lane < 6) {
        uint32_t n = min(256u, body.probes - base);
        for (uint32_t j = 0; j < n; ++j) {
          float4 a = fs[j], c = ts[j];
          if (a.w == 0.0f) continue;
          float x = lane == 0 ? a.x : lane == 1 ? a.y : lane == 2 ? a.z : lane == 3 ? c.x : lane == 4 ? c.y : c.z;
          acc = acc + x;
        }
      }
      __syncthreads();
    }
  }
  if (lane >= 6) return;
  if (lane < 3)
    l.impulse[lane] = acc * dt;
  else
    l.angular[lane - 3] = acc * dt;
  if (lane == 0) l.lifted = 1;
}

// give_back (lift.rs:154-166): the water read loses what the body got
__global__ __launch_bounds__(256) void k_give_back(const uint32_t* __restrict__ near, uint32_t count, float cx,
                                                   float cy, float cz, float4* __restrict__ V) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  uint32_t i = near[k];
  float4 v = V[i];
  V[i] = make_float4(v.x - cx, v.y - cy, v.z - cz, v.w);
}

// ---- launchers ------------------------------------------------------------

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

void launch_movers_place(hipStream_t s, const float4* local, uint32_t m, const BodyState* bodies, float4* MS,
                         float4* MV) {
  if (m == 0) return;
  hipLaunchKernelGGL(k_movers_place, blocks_for(m, 256), 256, 0, s, local, m, bodies, MS, MV);
}

hipError_t mover_sort_bytes(uint32_t m, size_t* bytes) {
  *bytes = 0;
  return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                   (const uint32_t*)nullptr, (uint32_t*)nullptr, m, 0, 64);
}

hipError_t launch_movers_carry(hipStream_t s, const MoverCarry& c, float elapsed, float radius, Control* ctl) {
  if (c.m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_movers_box_reset, 1, 64, 0, s, c.grid, ctl);
  const uint32_t cell_blocks = blocks_for(c.m, 256) < 256 ? blocks_for(c.m, 256) : 256;
  hipLaunchKernelGGL(k_movers_cells, cell_blocks, 256, 0, s, c.ms, c.mv, c.m, elapsed, radius, c.grid, ctl);
  hipLaunchKernelGGL(k_movers_keys, blocks_for(c.m, 256), 256, 0, s, c.ms, c.mv, c.m, elapsed, radius, c.owner_bits,
                     c.grid, c.keys_in, c.perm_in, ctl);
  size_t bytes = c.temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(c.temp, bytes, c.keys_in, c.keys, c.perm_in, c.perm, c.m, 0, 64, s);
  if (e != hipSuccess) return e;
  uint32_t threads = c.m > c.nbodies ? c.m : c.nbodies;
  hipLaunchKernelGGL(k_movers_gather, blocks_for(threads, 256), 256, 0, s, c.ms, c.mv, c.perm, c.m, elapsed, c.ms2,
                     c.mv2, c.mp, c.bodies, c.nbodies, ctl);
  if (c.table_cap)
    hipLaunchKernelGGL(k_mover_table, 1024, 256, 0, s, c.keys, c.m, c.owner_bits, c.grid, c.table, c.table_cap, ctl);
  return hipSuccess;
}

void launch_pairs_prepare(hipStream_t s, const PairRoom& room, uint32_t n, Control* ctl) {
  if (n == 0) return;
  launch_scan(s, room.counts, room.offsets, room.tiles, room.total, nullptr, n, 0, ctl);
  hipLaunchKernelGGL(k_pairs_check, 1024, 256, 0, s, room.total, room.cap, room.none, room.key, room.value, ctl);
}

hipError_t pair_sort_bytes(uint32_t cap, size_t* bytes) {
  *bytes = 0;
  return rocprim::radix_sort_pairs(nullptr, *bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (const uint32_t*)nullptr, (uint32_t*)nullptr, cap, 0, 32);
}

hipError_t launch_pairs_fold(hipStream_t s, const PairRoom& room, BodyState* bodies, uint32_t nbodies, float dt,
                             const Control* ctl) {
  if (nbodies == 0) return hipSuccess;
  size_t bytes = room.temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(room.temp, bytes, room.key, room.key_sorted, room.value,
                                           room.value_sorted, room.cap, 0, (int)room.key_bits, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_body_fold, nbodies, 64, 0, s, room.key_sorted, room.value_sorted, room.cap, room.force,
                     room.torque, bodies, dt, ctl);
  return hipSuccess;
}

void launch_extent(hipStream_t s, const float4* MS, const float4* MV, const BodyState* bodies, BodyExtent* e,
                   uint32_t nbodies) {
  if (nbodies == 0) return;
  hipLaunchKernelGGL(k_extent_bodies, nbodies, 256, 0, s, MS, MV, bodies, e);
}

void launch_extent_wet(hipStream_t s, const uint32_t* start, const uint64_t* keys, const Control* ctl,
                       const WetQuery* q, uint32_t nq, BodyExtent* e) {
  if (nq == 0) return;
  hipLaunchKernelGGL(k_extent_wet, dim3(64, nq), 256, 0, s, start, keys, ctl, q, nq, e);
}

void launch_lift_box(hipStream_t s, const float4* probes, const BodyState* bodies, uint32_t nbodies,
                     uint32_t most_probes, float4* arm, float4* vel, LiftState* ls, float reach) {
  if (nbodies == 0) return;
  hipLaunchKernelGGL(k_lift_reset, blocks_for(nbodies, 64), 64, 0, s, ls, nbodies);
  hipLaunchKernelGGL(k_lift_probes, dim3(blocks_for(most_probes, 256), nbodies), 256, 0, s, probes, bodies, nbodies,
                     arm, vel, ls);
  hipLaunchKernelGGL(k_lift_reach, blocks_for(nbodies, 64), 64, 0, s, ls, nbodies, reach);
}

void launch_lift_count(hipStream_t s, const float4* P, uint32_t n, LiftState* ls, uint32_t nbodies, float4* boxes,
                       uint32_t* counts, uint32_t* offsets, uint32_t* tiles, uint32_t* total) {
  if (nbodies == 0 || n == 0) return;
  const uint32_t chunks = blocks_for(n, kLiftChunk);
  hipLaunchKernelGGL(k_chunk_boxes, chunks, 256, 0, s, P, n, boxes);
  hipLaunchKernelGGL(k_lift_count, dim3(nbodies, chunks), 256, 0, s, P, n, ls, boxes, chunks, counts);
  launch_scan(s, counts, offsets, tiles, total, nullptr, (uint64_t)nbodies * chunks, 0, nullptr);
  hipLaunchKernelGGL(k_lift_offsets, blocks_for(nbodies, 64), 64, 0, s, ls, nbodies, offsets, chunks);
}

void launch_lift_near(hipStream_t s, const float4* P, uint32_t n, LiftState* ls, uint32_t nbodies,
                      const float4* boxes, const uint32_t* offsets, uint32_t* near, unsigned long long* columns,
                      uint64_t* column_keys, uint32_t* column_tops, float column, const float up[3]) {
  if (nbodies == 0 || n == 0) return;
  const uint32_t chunks = blocks_for(n, kLiftChunk);
  hipLaunchKernelGGL(k_lift_near, dim3(nbodies, chunks), 1024, 0, s, P, n, ls, boxes, chunks, offsets, near, columns,
                     column_keys, column_tops, column, up[0], up[1], up[2]);
}

hipError_t lift_sort_bytes(uint32_t most, size_t* bytes) {
  size_t a = 0, b = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, a, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, most, 0, 64);
  if (e != hipSuccess) return e;
  e = rocprim::reduce_by_key(nullptr, b, (const uint64_t*)nullptr, (const uint32_t*)nullptr, (size_t)most,
                             (uint64_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, rocprim::maximum<uint32_t>(),
                             rocprim::equal_to<uint64_t>());
  *bytes = std::max(a, b);
  return e;
}

// lift.rs:73-84 for one body: its water's (column, height) pairs sorted by
// column, then the highest of each column (total_cmp order is the order of
// total_key), in column order; the level pass sorts them by height.
hipError_t launch_lift_tops(hipStream_t s, const LiftSort& t, LiftState* ls, uint32_t b, const LiftState& l,
                            uint32_t* sortbuf) {
  if (l.near == 0) return hipSuccess;
  const size_t o = l.near_offset;
  size_t bytes = t.temp_bytes;
  hipError_t e = rocprim::radix_sort_pairs(t.temp, bytes, t.keys + o, t.keys_sorted + o, t.values + o,
                                           t.values_sorted + o, l.near, 0, 64, s);
  if (e != hipSuccess) return e;
  bytes = t.temp_bytes;
  return rocprim::reduce_by_key(t.temp, bytes, t.keys_sorted + o, t.values_sorted + o, (size_t)l.near, t.unique + o,
                                sortbuf + l.sort_offset, &ls[b].tops, rocprim::maximum<uint32_t>(),
                                rocprim::equal_to<uint64_t>(), s);
}

void launch_lift_level(hipStream_t s, const float4* P, const float4* V, LiftState* ls, const BodyState* bodies,
                       uint32_t nbodies, const uint32_t* near, const unsigned long long* columns, uint32_t* sortbuf,
                       const float4* arm, const float4* vel, float half_spacing, const float up[3]) {
  if (nbodies == 0) return;
  hipLaunchKernelGGL(k_lift_level, nbodies, 1024, 0, s, P, V, ls, bodies, near, columns, sortbuf, arm, vel,
                     half_spacing, up[0], up[1], up[2]);
}

void launch_lift_force(hipStream_t s, LiftState* ls, const BodyState* bodies, uint32_t nbodies, const float4* arm,
                       const float4* vel, float rest, float gravity, const float up[3], float dt) {
  if (nbodies == 0) return;
  hipLaunchKernelGGL(k_lift_force, nbodies, 256, 0, s, ls, bodies, arm, vel, rest, gravity, up[0], up[1], up[2], dt);
}

void launch_give_back(hipStream_t s, const uint32_t* near, uint32_t count, const float change[3], float4* V) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_give_back, blocks_for(count, 256), 256, 0, s, near, count, change[0], change[1], change[2], V);
}

}  // namespace sfl
