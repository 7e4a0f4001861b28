// WCSPH substep on gfx950 (src/fluid/step.rs:144-224): the density pass with
// the Tait pressure, then one fused pass of pressure force, artificial
// viscosity, wall pressure, wall viscosity, gravity, the velocity update and
// the symplectic advance. Both passes read the state gathered into sorted
// order (B) and the force pass writes the next state back to A, so a
// substep that aborted before the gather leaves A as it was.
#include "device.hpp"
#include "launch.hpp"

namespace sfl {

// Queue slots a lane for the force pass's two-phase neighbour walk, and the
// smallest launch that takes it: at 1M particles the force pass runs 362
// instead of 388 us (16 slots; 8, 12, 24 and 32 gave 377, 367, 377 and 435);
// at 128k and below the walk is bound by load latency and the queue's LDS
// round trips cost more than the pair terms it saves (113 against 109 us).
constexpr uint32_t kForceQueue = 16;
constexpr uint32_t kForceQueueFrom = 262144;

// compute_densities (src/fluid/mod.rs:621-675) for every particle, then the
// pressure over density squared of step.rs:151-172. rho goes to V.w, p/rho^2
// to `scaled`. Also folds max(rho/rest - 1, 0) for the report (step.rs:84-88).
// SEARCH: a table (water's or walls') may be searched rather than dense.
template <bool MOVERS, bool SEARCH>
__global__ __launch_bounds__(256) void k_wcsph_density(const float4* __restrict__ P, float4* __restrict__ V,
                                                       float* __restrict__ scaled, StepParams prm) {
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  bool live = i < prm.n;
  float comp = 0.0f;
  if (live) {
    float4 at = P[i];
    float sum = 0.0f;
    visit_fluid<SEARCH>(P, prm, at, [&](uint32_t, float4, float, float, float, float sq) {
      sum += kernel_value(sqrtf(sq), prm.radius, prm.rinv, prm.scale);
    });
    float wsum = 0.0f;
    if (prm.wall_cells)
      visit_walls<SEARCH>(prm, at, [&](uint32_t, float4, float, float, float, float sq) {
        wsum += kernel_value(sqrtf(sq), prm.radius, prm.rinv, prm.scale);
      });
    float density;
    if (MOVERS) {
      // the coupled bodies' points (mod.rs:648-656); their count sizes the
      // force pass's pair records
      float msum = 0.0f;
      uint32_t count = 0;
      visit_movers(prm.movers, at, prm.radius, prm.reach, [&](uint32_t, float4, float, float, float, float sq) {
        msum += kernel_value(sqrtf(sq), prm.radius, prm.rinv, prm.scale);
        ++count;
      });
      prm.movers.pairs.counts[i] = count;
      density = prm.mass * (sum + wsum + msum);
    } else {
      density = prm.mass * (sum + wsum);
    }
    if (prm.terrain.count) {
      // the ground's sites and slope (mod.rs:657-663)
      float ground, sx, sy, sz;
      ground_near(prm.terrain, at.x, at.y, at.z, ground, sx, sy, sz);
      density = density + ground;
      prm.terrain.slopes[i] = make_float4(sx, sy, sz, 0.0f);
    }
    float pressure = tait_pressure(density, prm.rest, prm.stiffness);
    V[i].w = density;
    scaled[i] = pressure / (density * density);
    comp = density / prm.rest - 1.0f;
  }
  wave_max_to(&stripe_of(ctl).comp_bits, comp);
}

// The push of wcsph_substep's force loop (step.rs:173-211): fluid pairs,
// walls, the coupled bodies' points (mover_push, src/fluid/couple.rs:
// 368-408), then + wall_viscous (step.rs:262-301) + gravity. With RECORD,
// each body-point pair's damping on its body goes to the pair records.
template <bool MOVERS, bool RECORD, bool SEARCH, uint32_t QUEUE = 0>
__device__ __forceinline__ void wcsph_push(const float4* __restrict__ PB, const float4* __restrict__ VB,
                                           const float* __restrict__ scaled, const StepParams& prm, uint32_t i,
                                           float4 at, float4 vi, float& ax, float& ay, float& az,
                                           uint32_t* queue = nullptr) {
  float density = vi.w;
  float own = scaled[i];
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  // one fluid pair's push, from the neighbour's velocity (density in w) and
  // p / rho^2
  auto pair = [&](float ox, float oy, float oz, float sq, float4 vj, float sj) {
    if (sq <= 0.0f) return;
    float distance = sqrtf(sq);
    float f = kernel_slope(distance, prm.radius, prm.rinv, prm.scale) / distance;
    float gx = ox * f, gy = oy * f, gz = oz * f;
    float toward = rmin(dot3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, ox, oy, oz), 0.0f);
    float damp = -prm.nu * toward / (sq + prm.soften) * 2.0f / (density + vj.w);
    float k = prm.mass * (own + sj + damp);
    px = px - gx * k;
    py = py - gy * k;
    pz = pz - gz * k;
  };
  if constexpr (QUEUE > 0 && !SEARCH) {
    const Control* ctl = prm.ctl;
    const int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
    const uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
    visit_dense_queued<QUEUE>(
        PB, prm.cell_start, lo, span, at, prm.radius, prm.reach, queue + threadIdx.x, blockDim.x,
        [&](uint32_t j) { return Neighbour{VB[j], scaled[j]}; },
        [&](uint32_t, float4, float ox, float oy, float oz, float sq, Neighbour nb) {
          pair(ox, oy, oz, sq, nb.v, nb.s);
        });
  } else {
    visit_fluid<SEARCH>(PB, prm, at, [&](uint32_t j, float4, float ox, float oy, float oz, float sq) {
      if (sq <= 0.0f) return;
      pair(ox, oy, oz, sq, VB[j], scaled[j]);
    });
  }
  float wx = 0.0f, wy = 0.0f, wz = 0.0f;
  float mirrored = 0.0f;
  if (MOVERS || prm.wall_cells || prm.terrain.count) {
    float pressure = tait_pressure(density, prm.rest, prm.stiffness);
    mirrored = own + pressure / (prm.rest * prm.rest);
  }
  auto walls = [&](auto&& f) {
    if (prm.wall_cells)
      visit_walls<SEARCH>(prm, at, f);
  };
  float awx = 0.0f, awy = 0.0f, awz = 0.0f;
  if (prm.wall_cells) {
    walls([&](uint32_t, float4, float ox, float oy, float oz, float sq) {
      float gx, gy, gz;
      kernel_gradient(ox, oy, oz, sq, prm.radius, prm.rinv, prm.scale, gx, gy, gz);
      float k = prm.mass * mirrored;
      px = px - gx * k;
      py = py - gy * k;
      pz = pz - gz * k;
      awx = awx + -gx;
      awy = awy + -gy;
      awz = awz + -gz;
    });
  }
  if (prm.terrain.count) {
    // push -= slope * mirrored (step.rs:203-205)
    float4 g = prm.terrain.slopes[i];
    px = px - g.x * mirrored;
    py = py - g.y * mirrored;
    pz = pz - g.z * mirrored;
  }
  if (MOVERS) {
    const MoverView& mv = prm.movers;
    uint32_t slot = RECORD ? mv.pairs.offsets[i] : 0;
    const float chunk = __uint_as_float(i / 512u);  // CHUNK, src/fluid/mod.rs:36
    visit_movers(mv, at, prm.radius, prm.reach, [&](uint32_t j, float4 p, float ox, float oy, float oz, float sq) {
      float4 v = mv.vel[j];
      float toward = dot3(vi.x - v.x, vi.y - v.y, vi.z - v.z, ox, oy, oz);
      float damp = toward < 0.0f ? -mv.damping * toward / (sq + prm.soften) / density : 0.0f;
      float gx, gy, gz;
      kernel_gradient(ox, oy, oz, sq, prm.radius, prm.rinv, prm.scale, gx, gy, gz);
      gx = gx * prm.mass;
      gy = gy * prm.mass;
      gz = gz * prm.mass;
      float k = mirrored + damp;
      px = px - gx * k;
      py = py - gy * k;
      pz = pz - gz * k;
      if (!RECORD) return;
      if (toward * toward < mv.settles * sq) {
        mv.pairs.key[slot++] = mv.pairs.none;
        return;
      }
      // force = gradient m (damp m), torque = arm x force, arm from the
      // body's centre moved on (couple.rs:400-407)
      float km = damp * prm.mass;
      float fx = gx * km, fy = gy * km, fz = gz * km;
      uint32_t owner = __float_as_uint(p.w);
      const BodyState& b = mv.bodies[owner];
      float rx = p.x - (b.center[0] + b.moved[0]);
      float ry = p.y - (b.center[1] + b.moved[1]);
      float rz = p.z - (b.center[2] + b.moved[2]);
      mv.pairs.key[slot] = owner;
      mv.pairs.force[slot] = make_float4(fx, fy, fz, chunk);
      mv.pairs.torque[slot] = make_float4(ry * fz - fy * rz, rz * fx - fz * rx, rx * fy - fx * ry, 0.0f);
      ++slot;
    });
  }
  if (prm.wall_cells || prm.terrain.count)
    wall_viscous(prm, at, vi.x, vi.y, vi.z, density, awx, awy, awz, walls, wx, wy, wz);
  // push + wall_viscous + gravity
  ax = px + wx + prm.gx;
  ay = py + wy + prm.gy;
  az = pz + wz + prm.gz;
}

// The force loop of wcsph_substep (step.rs:173-211) with wall_viscous
// (step.rs:262-301), v += push * dt (step.rs:212-222) and advance
// (step.rs:305-339), for every particle. Reads the sorted state PB/VB, writes
// PA/VA; `forces` (optional) keeps the push, `clamps` (optional) the
// clamped components of the advance (bits x, y, z). QUEUE: slots a lane of
// the two-phase walk (0: the one-phase walk).
template <bool MOVERS, bool SEARCH, uint32_t QUEUE>
__global__ __launch_bounds__(256) void k_wcsph_force(const float4* __restrict__ PB, const float4* __restrict__ VB,
                                                     const float* __restrict__ scaled, StepParams prm,
                                                     float4* __restrict__ PA, float4* __restrict__ VA,
                                                     float* __restrict__ forces, uint8_t* __restrict__ clamps) {
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  __shared__ uint32_t queue[QUEUE ? QUEUE * 256 : 1];
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  bool live = i < prm.n;
  float speed = 0.0f;
  uint32_t held = 0;
  if (live) {
    float4 at = PB[i];
    float4 vi = VB[i];
    float density = vi.w;
    float ax, ay, az;
    wcsph_push<MOVERS, MOVERS, SEARCH, QUEUE>(PB, VB, scaled, prm, i, at, vi, ax, ay, az, queue);
    if (forces) {
      forces[3 * i] = ax;
      forces[3 * i + 1] = ay;
      forces[3 * i + 2] = az;
    }
    float vx = vi.x + ax * prm.dt;
    float vy = vi.y + ay * prm.dt;
    float vz = vi.z + az * prm.dt;
    float nx, ny, nz;
    uint32_t clamped;
    bool ok = advance_one(prm, at, vx, vy, vz, nx, ny, nz, held, clamped);
    if (clamps) clamps[i] = (uint8_t)clamped;
    if (!ok) {
      atomicOr(&ctl->flags, kFlagNonFinite);
      atomicMin(&ctl->bad_index, i);
    }
    PA[i] = make_float4(nx, ny, nz, at.w);
    VA[i] = make_float4(vx, vy, vz, density);
    speed = sqrtf(dot3(vx, vy, vz, vx, vy, vz));
  }
  wave_max_to(&stripe_of(ctl).vmax_bits, speed);
  wave_sum_to(&stripe_of(ctl).held, held);
}

// Advances `PB` by v*dt into `PA` for particles after the first non-finite
// one, undoing what the fused pass wrote past it: the reference stops its
// advance there (step.rs:333-335), leaving later particles at their old
// position with the pushed, unclamped velocity. Recomputes that velocity.
template <bool MOVERS, bool SEARCH>
__global__ __launch_bounds__(256) void k_wcsph_repair(const float4* __restrict__ PB, const float4* __restrict__ VB,
                                                      const float* __restrict__ scaled, StepParams prm,
                                                      float4* __restrict__ PA, float4* __restrict__ VA) {
  Control* ctl = prm.ctl;
  uint32_t bad = ctl->bad_index;
  uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= prm.n || i < bad) return;
  float4 at = PB[i];
  if (i == bad) {
    PA[i] = at;  // velocity keeps its zeroed components, position stays
    return;
  }
  float4 vi = VB[i];
  float density = vi.w;
  float ax, ay, az;
  wcsph_push<MOVERS, false, SEARCH>(PB, VB, scaled, prm, i, at, vi, ax, ay, az);
  PA[i] = at;
  VA[i] = make_float4(vi.x + ax * prm.dt, vi.y + ay * prm.dt, vi.z + az * prm.dt, density);
}

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

// Whether a table of the substep is searched (the kernels' dense-only
// instantiations skip the check in the neighbour walk).
static inline bool searched(const StepParams& prm) { return prm.cell_keys || prm.wall_keys; }

void launch_wcsph_density(hipStream_t s, const float4* P, float4* V, float* scaled, const StepParams& prm) {
  if (prm.n == 0) return;
  const uint32_t blocks = blocks_for(prm.n, 256);
  const bool m = prm.movers.count != 0, k = searched(prm);
  if (m && k)
    hipLaunchKernelGGL((k_wcsph_density<true, true>), blocks, 256, 0, s, P, V, scaled, prm);
  else if (m)
    hipLaunchKernelGGL((k_wcsph_density<true, false>), blocks, 256, 0, s, P, V, scaled, prm);
  else if (k)
    hipLaunchKernelGGL((k_wcsph_density<false, true>), blocks, 256, 0, s, P, V, scaled, prm);
  else
    hipLaunchKernelGGL((k_wcsph_density<false, false>), blocks, 256, 0, s, P, V, scaled, prm);
}

void launch_wcsph_force(hipStream_t s, const float4* PB, const float4* VB, const float* scaled,
                        const StepParams& prm, float4* PA, float4* VA, float* forces, uint8_t* clamps) {
  if (prm.n == 0) return;
  const uint32_t blocks = blocks_for(prm.n, 256);
  const bool m = prm.movers.count != 0, k = searched(prm), q = !k && prm.n >= kForceQueueFrom;
  if (m && k)
    hipLaunchKernelGGL((k_wcsph_force<true, true, 0>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA, forces, clamps);
  else if (m && q)
    hipLaunchKernelGGL((k_wcsph_force<true, false, kForceQueue>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA,
                       forces, clamps);
  else if (m)
    hipLaunchKernelGGL((k_wcsph_force<true, false, 0>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA, forces,
                       clamps);
  else if (k)
    hipLaunchKernelGGL((k_wcsph_force<false, true, 0>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA, forces,
                       clamps);
  else if (q)
    hipLaunchKernelGGL((k_wcsph_force<false, false, kForceQueue>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA,
                       forces, clamps);
  else
    hipLaunchKernelGGL((k_wcsph_force<false, false, 0>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA, forces,
                       clamps);
}

void launch_wcsph_repair(hipStream_t s, const float4* PB, const float4* VB, const float* scaled,
                         const StepParams& prm, float4* PA, float4* VA) {
  if (prm.n == 0) return;
  const uint32_t blocks = blocks_for(prm.n, 256);
  const bool m = prm.movers.count != 0, k = searched(prm);
  if (m && k)
    hipLaunchKernelGGL((k_wcsph_repair<true, true>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA);
  else if (m)
    hipLaunchKernelGGL((k_wcsph_repair<true, false>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA);
  else if (k)
    hipLaunchKernelGGL((k_wcsph_repair<false, true>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA);
  else
    hipLaunchKernelGGL((k_wcsph_repair<false, false>), blocks, 256, 0, s, PB, VB, scaled, prm, PA, VA);
}

}  // namespace sfl
