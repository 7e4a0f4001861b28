// Host-side launchers of the gfx950 kernels (defined next to the kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace sfl {

// neighbour build (kernels_grid.hip)
void launch_step_begin(hipStream_t s, Control* ctl);
void launch_substep_begin(hipStream_t s, Control* ctl);
void launch_fastest(hipStream_t s, const float4* V, uint32_t n, Control* ctl);
void launch_bounds(hipStream_t s, const float4* P, uint32_t n, float radius, Control* ctl);
void launch_grid_setup(hipStream_t s, Control* ctl, uint32_t n, uint64_t capacity);
// capacity kSearchedTable: the box goes to a searched table (any size whose
// slots fit 63 bits)
constexpr uint64_t kSearchedTable = 0xffffffffffffffffull;
// The buffers of a sort on a searched table (n particles; scan n + 1).
struct SearchedSort {
  uint64_t* keys_in;
  uint64_t* keys;
  uint32_t* values_in;
  uint32_t* values;
  uint32_t* heads;
  uint32_t* scan;
  uint32_t* tile_sums;  // the scan's, (n + 2048) / 2048 + 1
  uint32_t* total;
  uint64_t* cell_keys;
  uint32_t* start;  // n + 1
  void* temp;
  size_t temp_bytes;
};
hipError_t searched_sort_bytes(uint32_t n, size_t* bytes);
hipError_t launch_searched_sort(hipStream_t s, const SearchedSort& t, const float4* PA, const float4* VA, uint32_t n,
                                float radius, float4* PB, float4* VB, Control* ctl, uint32_t* src = nullptr);
void launch_count(hipStream_t s, const float4* P, uint32_t n, float radius, uint32_t* counts, uint2* slot_rank,
                  Control* ctl);
void launch_scan(hipStream_t s, uint32_t* in, uint32_t* out, uint32_t* tile_sums, uint32_t* total,
                 const uint32_t* len_dev, uint64_t len_max, int zero_input, const Control* ctl);
void launch_scatter(hipStream_t s, const float4* P, uint32_t n, const uint2* slot_rank, const uint32_t* start,
                    uint2* pairs, const Control* ctl);
void launch_rank_gather(hipStream_t s, const float4* PA, const float4* VA, uint32_t n, float radius,
                        const uint2* pairs, const uint32_t* start, float4* PB, float4* VB, const Control* ctl,
                        uint32_t* src = nullptr);
void launch_list_count(hipStream_t s, const float4* lp, const StepParams& prm, uint32_t* counts, int kind);
void launch_list_emit(hipStream_t s, const float4* lp, const StepParams& prm, const uint32_t* list_start,
                      uint32_t* items, int kind);
void launch_list_density(hipStream_t s, const float4* P, float4* V, const StepParams& prm, const uint32_t* fstart,
                         const uint32_t* fitems, const uint32_t* wstart, const uint32_t* witems);

// WCSPH (kernels_wcsph.hip)
void launch_wcsph_density(hipStream_t s, const float4* P, float4* V, float* scaled, const StepParams& prm);
void launch_wcsph_force(hipStream_t s, const float4* PB, const float4* VB, const float* scaled,
                        const StepParams& prm, float4* PA, float4* VA, float* forces, uint8_t* clamps = nullptr);
void launch_wcsph_repair(hipStream_t s, const float4* PB, const float4* VB, const float* scaled,
                         const StepParams& prm, float4* PA, float4* VA);

// Flows and the domain after a step (kernels_flow.hip)
void launch_flow_keep(hipStream_t s, const float4* P, uint32_t n, const FlowBox& box, int domain, uint32_t* keep);
void launch_flow_gather(hipStream_t s, const float4* PA, const float4* VA, uint32_t n, const uint32_t* keep,
                        const uint32_t* start, float4* PB, float4* VB);
void launch_emit_mark(hipStream_t s, const float4* P, uint32_t n, const EmitBox& e, uint32_t* free, uint64_t slots);
void launch_emit_append(hipStream_t s, const uint32_t* free, const uint32_t* start, uint64_t slots, const EmitBox& e,
                        uint32_t n, uint32_t limit, uint32_t first_id, float4* P, float4* V);
void launch_flow_totals(hipStream_t s, const float4* V, uint32_t n, float rest, Control* ctl);

// DFSPH (kernels_dfsph.hip)
// The flow solve's per-substep tile bookkeeping: both solves' tile sums and
// counts (tile_cap each), pass counters, weights and list.
struct FlowSetup {
  double* tile_sum = nullptr;
  uint32_t* tile_count = nullptr;
  uint32_t tile_cap = 0;
  uint32_t* tile_done = nullptr;
  uint32_t* tile_work = nullptr;
  uint32_t* tile_list = nullptr;
};
// The DFSPH density pass; with a set-up (tile_done set), one more block of
// the launch sets the flow solve's tiles up for the substep beside it.
hipError_t launch_dfsph_density(hipStream_t s, const float4* P, float4* V, float* alpha, const StepParams& prm,
                                const FlowSetup& setup = FlowSetup());
int solve_grid_limit(int blocks_per_cu, bool tiles, uint32_t* tile_bytes);
// The flow solve's buffers: a second velocity buffer (n), and per solve tile
// its passes finished (2 * tile_cap: a half for each solve of the substep),
// its weight and the list of non-empty tiles (tile_cap each; tile_sum and
// tile_count then hold 2 * tile_cap). All null: the tiled solve's passes
// meet at grid barriers instead.
struct FlowBuffers {
  float4* v2 = nullptr;
  uint32_t* tile_done = nullptr;
  uint32_t* tile_work = nullptr;
  uint32_t* tile_list = nullptr;
};
hipError_t launch_dfsph_solve(hipStream_t s, const StepParams& prm, const float4* P, float4* V, const float* alpha,
                              float4* kappa, float* errors, double* tile_sum, uint32_t* tile_count,
                              uint32_t tile_cap, int which, float tolerance, uint32_t min_it, uint32_t max_it,
                              int grid, uint32_t tile_bytes, const float4* vin = nullptr, float* forces = nullptr,
                              const FlowBuffers& flow = FlowBuffers());
// Whether the solve runs on LDS tiles (then its density solve also runs the
// viscosity pass, given the velocities before it).
bool dfsph_solve_tiled(const StepParams& prm, uint32_t tile_bytes);
void launch_dfsph_viscous(hipStream_t s, const float4* PB, const float4* VB, const StepParams& prm, float4* VA,
                          float* forces);
void launch_dfsph_advance(hipStream_t s, const float4* PB, const StepParams& prm, float4* PA, float4* VA,
                          float4* VKEEP);
void launch_dfsph_repair(hipStream_t s, const float4* PB, const float4* VKEEP, const StepParams& prm, float4* PA,
                         float4* VA);

}  // namespace sfl
