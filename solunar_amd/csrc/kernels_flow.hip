// Flows and the domain after a step (src/fluid/flow.rs:144-223,
// src/fluid/mod.rs:451-489) on gfx950.
//
// Drains and the domain take particles out and keep the rest in order: a
// keep flag per particle, an exclusive scan of the flags, and a gather of
// the kept particles to their new slots (keep_only). An emitter marks the
// lattice slots of its box that water is nearer to than 0.9 spacing (each
// particle checks the 27 slots around its nearest one), then the free
// slots, in the reference's y, z, x order, become new particles: the scan
// of the free flags gives each one its index and id, so the result does not
// depend on the order the marks landed in.
#include <algorithm>

#include "device.hpp"
#include "launch.hpp"

namespace sfl {

// keep[i]: outside the drain box (inclusive bounds), or inside the domain
// (at.clamp(low, high) == at: every component within its bounds, NaN out).
__global__ __launch_bounds__(256) void k_flow_keep(const float4* __restrict__ P, uint32_t n, FlowBox box, int domain,
                                                   uint32_t* __restrict__ keep) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 at = P[i];
  const bool within = at.x >= box.lo[0] && at.y >= box.lo[1] && at.z >= box.lo[2] && at.x <= box.hi[0] &&
                      at.y <= box.hi[1] && at.z <= box.hi[2];
  keep[i] = (domain ? within : !within) ? 1u : 0u;
}

// The kept particles to their slots start[i] of B.
__global__ __launch_bounds__(256) void k_flow_gather(const float4* __restrict__ PA, const float4* __restrict__ VA,
                                                     uint32_t n, const uint32_t* __restrict__ keep,
                                                     const uint32_t* __restrict__ start, float4* __restrict__ PB,
                                                     float4* __restrict__ VB) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  const uint32_t to = start[i];
  PB[to] = PA[i];
  VB[to] = VA[i];
}

__global__ __launch_bounds__(256) void k_emit_clear(uint32_t* __restrict__ free, uint64_t slots) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += (uint64_t)gridDim.x * blockDim.x)
    free[s] = 1u;
}

// `value.round() as i64` of the reference, NaN as 0, clamped far outside
// any slot range (the callers skip those)
__device__ __forceinline__ int64_t nearest_slot(float value) {
  const float r = roundf(value);
  if (!(r == r)) return 0;
  return (int64_t)fminf(fmaxf(r, -8192.0f), 8192.0f);
}

// Slot (x, y, z) of an emitter is taken when a particle is nearer its centre
// than 0.9 spacing (src/fluid/flow.rs:177-202).
__global__ __launch_bounds__(256) void k_emit_taken(const float4* __restrict__ P, uint32_t n, EmitBox e,
                                                    uint32_t* __restrict__ free) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 at = P[i];
  const float lx = (at.x - e.lo[0]) / e.spacing - 0.5f;
  const float ly = (at.y - e.lo[1]) / e.spacing - 0.5f;
  const float lz = (at.z - e.lo[2]) / e.spacing - 0.5f;
  const int64_t px = nearest_slot(lx), py = nearest_slot(ly), pz = nearest_slot(lz);
  if (px < -1 || px > e.cells[0] || py < -1 || py > e.cells[1] || pz < -1 || pz > e.cells[2]) return;
  for (int64_t y = py - 1; y <= py + 1; ++y)
    for (int64_t z = pz - 1; z <= pz + 1; ++z)
      for (int64_t x = px - 1; x <= px + 1; ++x) {
        if (x < 0 || x >= e.cells[0] || y < 0 || y >= e.cells[1] || z < 0 || z >= e.cells[2]) continue;
        const float cx = e.lo[0] + ((float)x + 0.5f) * e.spacing;
        const float cy = e.lo[1] + ((float)y + 0.5f) * e.spacing;
        const float cz = e.lo[2] + ((float)z + 0.5f) * e.spacing;
        const float dx = at.x - cx, dy = at.y - cy, dz = at.z - cz;
        if (dot3(dx, dy, dz, dx, dy, dz) < e.clear) free[(uint64_t)((y * e.cells[2] + z) * e.cells[0] + x)] = 0u;
      }
}

// Free slot s (rank start[s] among the free ones, below `limit`) becomes
// particle n + rank with id first_id + rank, at the slot's centre, moving at
// the emitter's velocity, density 0 (add_particle, src/fluid/mod.rs:412-426).
__global__ __launch_bounds__(256) void k_emit_append(const uint32_t* __restrict__ free,
                                                     const uint32_t* __restrict__ start, uint64_t slots, EmitBox e,
                                                     uint32_t n, uint32_t limit, uint32_t first_id,
                                                     float4* __restrict__ P, float4* __restrict__ V) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots;
       s += (uint64_t)gridDim.x * blockDim.x) {
    if (!free[s]) continue;
    const uint32_t rank = start[s];
    if (rank >= limit) continue;
    const int64_t x = (int64_t)(s % (uint64_t)e.cells[0]);
    const int64_t z = (int64_t)((s / (uint64_t)e.cells[0]) % (uint64_t)e.cells[2]);
    const int64_t y = (int64_t)(s / ((uint64_t)e.cells[0] * (uint64_t)e.cells[2]));
    const float ax = e.lo[0] + ((float)x + 0.5f) * e.spacing;
    const float ay = e.lo[1] + ((float)y + 0.5f) * e.spacing;
    const float az = e.lo[2] + ((float)z + 0.5f) * e.spacing;
    P[n + rank] = make_float4(ax, ay, az, __uint_as_float(first_id + rank));
    V[n + rank] = make_float4(e.vel[0], e.vel[1], e.vel[2], 0.0f);
  }
}

// The report's max |v| and max(rho / rest - 1, 0) over the particles after
// the flows (src/fluid/step.rs:83-88), into the stripes cleared first.
__global__ void k_flow_clear_totals(Control* ctl) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < kStripes; ++k) {
    ctl->stripes[k].vmax_bits = 0;
    ctl->stripes[k].comp_bits = 0;
  }
}

__global__ __launch_bounds__(256) void k_flow_totals(const float4* __restrict__ V, uint32_t n, float rest,
                                                     Control* ctl) {
  float most = 0.0f, comp = 0.0f;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 v = V[i];
    most = rmax(most, sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z)));
    comp = rmax(comp, v.w / rest - 1.0f);
  }
  wave_max_to(&stripe_of(ctl).vmax_bits, most);
  wave_max_to(&stripe_of(ctl).comp_bits, comp);
}

static inline uint32_t flow_blocks(uint64_t n) { return (uint32_t)((n + 255) / 256); }

void launch_flow_keep(hipStream_t s, const float4* P, uint32_t n, const FlowBox& box, int domain, uint32_t* keep) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_flow_keep, flow_blocks(n), 256, 0, s, P, n, box, domain, keep);
}

void launch_flow_gather(hipStream_t s, const float4* PA, const float4* VA, uint32_t n, const uint32_t* keep,
                        const uint32_t* start, float4* PB, float4* VB) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_flow_gather, flow_blocks(n), 256, 0, s, PA, VA, n, keep, start, PB, VB);
}

void launch_emit_mark(hipStream_t s, const float4* P, uint32_t n, const EmitBox& e, uint32_t* free,
                      uint64_t slots) {
  const uint32_t grid = (uint32_t)std::min<uint64_t>(flow_blocks(slots), 4096);
  hipLaunchKernelGGL(k_emit_clear, grid, 256, 0, s, free, slots);
  if (n) hipLaunchKernelGGL(k_emit_taken, flow_blocks(n), 256, 0, s, P, n, e, free);
}

void launch_emit_append(hipStream_t s, const uint32_t* free, const uint32_t* start, uint64_t slots, const EmitBox& e,
                        uint32_t n, uint32_t limit, uint32_t first_id, float4* P, float4* V) {
  const uint32_t grid = (uint32_t)std::min<uint64_t>(flow_blocks(slots), 4096);
  hipLaunchKernelGGL(k_emit_append, grid, 256, 0, s, free, start, slots, e, n, limit, first_id, P, V);
}

void launch_flow_totals(hipStream_t s, const float4* V, uint32_t n, float rest, Control* ctl) {
  hipLaunchKernelGGL(k_flow_clear_totals, 1, 64, 0, s, ctl);
  if (n) {
    const uint32_t grid = std::min<uint32_t>(flow_blocks(n), 2048);
    hipLaunchKernelGGL(k_flow_totals, grid, 256, 0, s, V, n, rest, ctl);
  }
}

}  // namespace sfl
