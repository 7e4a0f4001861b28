// DFSPH substep on gfx950 (src/fluid/dfsph.rs:79-103): density and factors,
// the divergence and density Jacobi solves as one persistent cooperative
// launch each, viscosity with gravity, and the advance.
//
// The reference caches every pair's kernel gradient for the substep
// (dfsph.rs:24-41, 108-136). Here each pass recomputes it from the fixed
// positions with the same operations, which gives the same bits and trades
// 12 bytes of HBM per pair per pass for arithmetic.
//
// Stop test: the reference averages the per-particle errors with an f32 left
// fold in particle order (dfsph.rs:263-267), which no parallel sum can
// reproduce. The solve instead sums them in FP64 with a fixed shape that the
// CPU oracle replays exactly: lane l of the wave that owns chunk c (512
// particles) adds errors 512c + l + 64k, k = 0..7, in order; the wave folds
// lanes by halving (s_l += s_{l+h}, h = 32..1); chunk sums land in 256 slots,
// slot t adding chunks t, t+256, ... in order; slots fold by halving
// (h = 128..1). Every block computes that total itself after the barrier, so
// all blocks take the same decision without a second barrier.
#include "device.hpp"
#include "launch.hpp"

namespace sfl {

// compute_densities (src/fluid/mod.rs:621-675) fused with gradients + factors
// (dfsph.rs:108-163) and the neighbour count the divergence solve uses
// (dfsph.rs:249-255). rho goes to V.w; alpha to `alpha` with its sign bit set
// for particles with fewer than 20 fluid + wall neighbours.
__global__ __launch_bounds__(256) void k_dfsph_density(const float4* __restrict__ P, float4* __restrict__ V,
                                                       float* __restrict__ alpha, StepParams prm) {
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool live = i < prm.n;
  float comp = 0.0f;
  if (live) {
    int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
    uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
    float4 at = P[i];
    float sum = 0.0f, wsum = 0.0f, squares = 0.0f;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    uint32_t crowd = 0;
    visit_near(P, prm.cell_start, lo, span, at, prm.radius, prm.reach,
               [&](uint32_t, float4, float ox, float oy, float oz, float sq) {
                 float distance = sqrtf(sq);
                 sum += kernel_value(distance, prm.radius, prm.scale);
                 float gx, gy, gz;
                 kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
                 float qx = gx * prm.mass, qy = gy * prm.mass, qz = gz * prm.mass;
                 sx = sx + qx;
                 sy = sy + qy;
                 sz = sz + qz;
                 squares += dot3(qx, qy, qz, qx, qy, qz);
                 ++crowd;
               });
    if (prm.wall_cells)
      visit_near(prm.walls, prm.wall_start, prm.wall_box.lo, prm.wall_box.span, at, prm.radius, prm.reach,
                 [&](uint32_t, float4, float ox, float oy, float oz, float sq) {
                   wsum += kernel_value(sqrtf(sq), prm.radius, prm.scale);
                   float gx, gy, gz;
                   kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
                   sx = sx + gx * prm.mass;
                   sy = sy + gy * prm.mass;
                   sz = sz + gz * prm.mass;
                   ++crowd;
                 });
    float density = prm.mass * (sum + wsum);
    float total = dot3(sx, sy, sz, sx, sy, sz) + squares;
    float a = total > 0.0f ? 1.0f / total : 0.0f;
    alpha[i] = crowd < 20u ? -a : a;
    V[i].w = density;
    comp = density / prm.rest - 1.0f;
  }
  wave_max_to(&ctl->comp_bits, comp);
}

// ---- persistent Jacobi solve ------------------------------------------------

constexpr uint32_t kSpinLimit = 1u << 26;  // about 4 s of s_sleep(2)

// Grid-wide barrier of a cooperative launch (every block resident). Returns
// false if it waited past the spin limit, having flagged the world.
__device__ bool grid_barrier(Control* ctl) {
  __threadfence();
  __syncthreads();
  __shared__ int ok_s;
  if (threadIdx.x == 0) {
    int ok = 1;
    uint32_t gen = __hip_atomic_load(&ctl->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t arrived = __hip_atomic_fetch_add(&ctl->bar_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == gridDim.x - 1) {
      __hip_atomic_store(&ctl->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&ctl->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      uint32_t spins = 0;
      while (__hip_atomic_load(&ctl->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > kSpinLimit) {
          ok = 0;
          atomicOr(&ctl->flags, kFlagBarrier);
          break;
        }
      }
    }
    ok_s = ok;
  }
  __syncthreads();
  __threadfence();
  return ok_s != 0;
}

struct SolveArgs {
  StepParams prm;
  const float4* P;      // positions, fixed during the solve
  float4* V;            // velocities (updated in place); w = density
  const float* alpha;   // factors, sign bit = surface particle
  float* kappa;         // stiffness of the last rate pass
  double* partial;      // per chunk error sums
  uint32_t* pcount;     // per chunk count of errors > 0
  uint32_t chunks;
  int32_t which;        // 0 divergence, 1 density
  float tolerance;
  uint32_t min_iterations;
  uint32_t max_iterations;
};

// solve (src/fluid/dfsph.rs:222-275) with density_rate (dfsph.rs:166-181) and
// push_apart (dfsph.rs:184-216).
__global__ __launch_bounds__(256) void k_dfsph_solve(SolveArgs a) {
  const StepParams& prm = a.prm;
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  const uint32_t waves = gridDim.x * 4u;
  const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  __shared__ double slot[256];
  __shared__ uint32_t slot_count[256];
  int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
  uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
  const float dt = prm.dt;
  auto walls = [&](float4 at, auto&& f) {
    if (prm.wall_cells)
      visit_near(prm.walls, prm.wall_start, prm.wall_box.lo, prm.wall_box.span, at, prm.radius, prm.reach, f);
  };
  uint32_t it = 0;
  for (;;) {
    // rate pass: stiffness and error of every particle
    for (uint32_t c = wave; c < a.chunks; c += waves) {
      double acc = 0.0;
      uint32_t cnt = 0;
      for (uint32_t k = 0; k < 8; ++k) {
        uint32_t i = c * 512u + k * 64u + lane;
        if (i >= prm.n) continue;
        float4 at = a.P[i];
        float4 vi = a.V[i];
        float rate = 0.0f;
        visit_near(a.P, prm.cell_start, lo, span, at, prm.radius, prm.reach,
                   [&](uint32_t j, float4, float ox, float oy, float oz, float sq) {
                     float gx, gy, gz;
                     kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
                     float4 vj = a.V[j];
                     rate += dot3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, gx, gy, gz);
                   });
        walls(at, [&](uint32_t, float4, float ox, float oy, float oz, float sq) {
          float gx, gy, gz;
          kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
          rate += dot3(vi.x, vi.y, vi.z, gx, gy, gz);
        });
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
        a.kappa[i] = target * fabsf(fa);
        acc += (double)error;
        cnt += error > 0.0f ? 1u : 0u;
      }
      for (int h = 32; h >= 1; h >>= 1) {
        acc = acc + __shfl_down(acc, h, 64);
        cnt += (uint32_t)__shfl_down((int)cnt, h, 64);
      }
      if (lane == 0) {
        a.partial[c] = acc;
        a.pcount[c] = cnt;
      }
    }
    if (!grid_barrier(ctl)) return;
    // every block folds the chunk sums the same way
    {
      double s = 0.0;
      uint32_t cn = 0;
      for (uint32_t c = threadIdx.x; c < a.chunks; c += 256u) {
        s += a.partial[c];
        cn += a.pcount[c];
      }
      slot[threadIdx.x] = s;
      slot_count[threadIdx.x] = cn;
      __syncthreads();
      for (uint32_t h = 128; h >= 1; h >>= 1) {
        if (threadIdx.x < h) {
          slot[threadIdx.x] = slot[threadIdx.x] + slot[threadIdx.x + h];
          slot_count[threadIdx.x] += slot_count[threadIdx.x + h];
        }
        __syncthreads();
      }
    }
    double total = slot[0];
    uint32_t loaded = slot_count[0];
    __syncthreads();
    double average = total / (double)(loaded > 1u ? loaded : 1u);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
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
    for (uint32_t c = wave; c < a.chunks; c += waves) {
      for (uint32_t k = 0; k < 8; ++k) {
        uint32_t i = c * 512u + k * 64u + lane;
        if (i >= prm.n) continue;
        float4 at = a.P[i];
        float own = a.kappa[i];
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        visit_near(a.P, prm.cell_start, lo, span, at, prm.radius, prm.reach,
                   [&](uint32_t j, float4, float ox, float oy, float oz, float sq) {
                     float gx, gy, gz;
                     kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
                     float k2 = prm.mass * (own + a.kappa[j]);
                     cx = cx - gx * k2;
                     cy = cy - gy * k2;
                     cz = cz - gz * k2;
                   });
        walls(at, [&](uint32_t, float4, float ox, float oy, float oz, float sq) {
          float gx, gy, gz;
          kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
          float k2 = prm.mass * own;
          cx = cx - gx * k2;
          cy = cy - gy * k2;
          cz = cz - gz * k2;
        });
        float4 vi = a.V[i];
        vi.x = vi.x + cx * dt;
        vi.y = vi.y + cy * dt;
        vi.z = vi.z + cz * dt;
        a.V[i] = vi;
      }
    }
    if (!grid_barrier(ctl)) return;
    ++it;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->iterations[a.which] = it;
    ctl->iterations_step[a.which] += it;
    ctl->rates_step += it + 1;
  }
}

// The viscosity pass of dfsph_substep (dfsph.rs:90-99): viscous
// (src/fluid/step.rs:229-251) + gravity, v += push * dt. Reads VB, writes VA
// (density carried along).
__global__ __launch_bounds__(256) void k_dfsph_viscous(const float4* __restrict__ PB, const float4* __restrict__ VB,
                                                       StepParams prm, float4* __restrict__ VA,
                                                       float* __restrict__ forces) {
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= prm.n) return;
  int32_t lo[3] = {ctl->lo[0], ctl->lo[1], ctl->lo[2]};
  uint32_t span[3] = {ctl->span[0], ctl->span[1], ctl->span[2]};
  float4 at = PB[i];
  float4 vi = VB[i];
  float density = vi.w;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  visit_near(PB, prm.cell_start, lo, span, at, prm.radius, prm.reach,
             [&](uint32_t j, float4, float ox, float oy, float oz, float sq) {
               if (j == i) return;
               float4 vj = VB[j];
               float toward = dot3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, ox, oy, oz);
               if (toward < 0.0f) {
                 float squeeze = toward / (sq + prm.soften);
                 float damp = -prm.nu * squeeze * 2.0f / (density + vj.w);
                 float gx, gy, gz;
                 kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
                 float k = prm.mass * damp;
                 px = px - gx * k;
                 py = py - gy * k;
                 pz = pz - gz * k;
               }
             });
  float wx = 0.0f, wy = 0.0f, wz = 0.0f;
  if (prm.wall_cells) {
    auto walls = [&](auto&& f) {
      visit_near(prm.walls, prm.wall_start, prm.wall_box.lo, prm.wall_box.span, at, prm.radius, prm.reach, f);
    };
    float awx = 0.0f, awy = 0.0f, awz = 0.0f;
    walls([&](uint32_t, float4, float ox, float oy, float oz, float sq) {
      float gx, gy, gz;
      kernel_gradient(ox, oy, oz, sq, prm.radius, prm.scale, gx, gy, gz);
      awx = awx + -gx;
      awy = awy + -gy;
      awz = awz + -gz;
    });
    wall_viscous(prm, at, vi.x, vi.y, vi.z, density, awx, awy, awz, walls, wx, wy, wz);
  }
  float ax = px + wx + prm.gx;
  float ay = py + wy + prm.gy;
  float az = pz + wz + prm.gz;
  if (forces) {
    forces[3 * i] = ax;
    forces[3 * i + 1] = ay;
    forces[3 * i + 2] = az;
  }
  VA[i] = make_float4(vi.x + ax * prm.dt, vi.y + ay * prm.dt, vi.z + az * prm.dt, density);
}

// advance (src/fluid/step.rs:305-339): PB + VA*dt -> PA, VA. Particles past a
// non-finite one keep their position and velocity, as in the reference.
__global__ __launch_bounds__(256) void k_dfsph_advance(const float4* __restrict__ PB, StepParams prm,
                                                       float4* __restrict__ PA, float4* __restrict__ VA,
                                                       float4* __restrict__ VKEEP) {
  Control* ctl = prm.ctl;
  if (ctl->abort) return;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
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
  wave_max_to(&ctl->vmax_bits, speed);
  wave_sum_to(&ctl->held, held);
}

// After a non-finite advance: particles past the first bad one go back to
// their old position and pre-advance velocity (kept in VKEEP); the bad one
// keeps its old position and its zeroed velocity components.
__global__ __launch_bounds__(256) void k_dfsph_repair(const float4* __restrict__ PB, const float4* __restrict__ VKEEP,
                                                      StepParams prm, float4* __restrict__ PA,
                                                      float4* __restrict__ VA) {
  Control* ctl = prm.ctl;
  uint32_t bad = ctl->bad_index;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= prm.n || i < bad) return;
  PA[i] = PB[i];
  if (i > bad) VA[i] = VKEEP[i];
}

static inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

void launch_dfsph_density(hipStream_t s, const float4* P, float4* V, float* alpha, const StepParams& prm) {
  if (prm.n == 0) return;
  hipLaunchKernelGGL(k_dfsph_density, blocks_for(prm.n, 256), 256, 0, s, P, V, alpha, prm);
}

int solve_grid_limit() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_dfsph_solve), 256, 0) !=
      hipSuccess)
    return 0;
  return per * cus;
}

hipError_t launch_dfsph_solve(hipStream_t s, const StepParams& prm, const float4* P, float4* V, const float* alpha,
                              float* kappa, double* partial, uint32_t* pcount, int which, float tolerance,
                              uint32_t min_it, uint32_t max_it, int grid) {
  if (prm.n == 0) return hipSuccess;
  SolveArgs a;
  a.prm = prm;
  a.P = P;
  a.V = V;
  a.alpha = alpha;
  a.kappa = kappa;
  a.partial = partial;
  a.pcount = pcount;
  a.chunks = (prm.n + 511u) / 512u;
  a.which = which;
  a.tolerance = tolerance;
  a.min_iterations = min_it;
  a.max_iterations = max_it;
  uint32_t want = (a.chunks + 3u) / 4u;  // one wave per chunk at most
  uint32_t blocks = (uint32_t)grid < want ? (uint32_t)grid : want;
  if (blocks == 0) blocks = 1;
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_dfsph_solve), dim3(blocks), dim3(256), args, 0,
                                    s);
}

void launch_dfsph_viscous(hipStream_t s, const float4* PB, const float4* VB, const StepParams& prm, float4* VA,
                          float* forces) {
  if (prm.n == 0) return;
  hipLaunchKernelGGL(k_dfsph_viscous, blocks_for(prm.n, 256), 256, 0, s, PB, VB, prm, VA, forces);
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
