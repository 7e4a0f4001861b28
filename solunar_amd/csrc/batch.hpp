// Batched WCSPH: many independent fluid worlds ("scenes") stepped together.
//
// The scenes' particles share one set of device arrays: scene s owns the
// index range [first_s, first_s + count_s) for the life of the batch, and the
// per-substep sort orders particles by (scene, cell, id), which never moves a
// particle out of its scene's range. Every scene keeps its own cell box, wall
// table, container, substep count and substep length, so each scene's
// arithmetic is that of a World stepping alone, bit for bit.
//
// A launch over particles maps each block to one scene (blocks are cut per
// scene, the last block of a scene partly idle), so the scene's parameters
// are wave-uniform. A step of the whole batch is a fixed sequence of launches
// steered by a device cursor (step, substep), which the host captures once
// in a HIP graph for a run of many steps and launches with one call.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/solunar_fluid.h"
#include "device.hpp"
#include "world.hpp"

namespace sfl {

constexpr uint32_t kFlagSlots = 16u;  // a scene wants more substeps than the sequence has slots

// Where the launch sequence stands. The reset kernel sets dt, target and
// slots before a run of launches; the kernels advance the rest.
struct BatchCursor {
  uint32_t step;       // steps the batch has finished
  uint32_t target;     // the sequence stops when `step` reaches it
  uint32_t k;          // substep of the current step
  uint32_t max_sub;    // most substeps any scene takes this step
  float dt;            // length of each step
  uint32_t slots;      // substep slots per step in the launch sequence
  uint32_t pad[2];
};

// Static description of scene s.
struct SceneDesc {
  uint32_t first, count;  // particle range
  uint32_t wall_base;     // first entry of its wall-cell table in the shared one
  uint32_t wall_cells;    // 0: no walls
  CellBox wall_box;
  int32_t has_bounds;
  float blo[3], bhi[3];
  uint32_t pad;
};

// Per-scene state the kernels keep.
struct SceneState {
  CellBox box;          // this substep's cell box
  uint32_t slot_base;   // first slot of the box in the shared cell table
  int32_t bmin[3], bmax[3];
  uint32_t substeps;    // of the current (or last) step
  uint32_t capped;
  float dt, nu, wall_nu;
  uint32_t vmax_bits;   // max |v| after the last advance
  uint32_t comp_bits;   // max(rho/rest - 1, 0) of the last density pass
  uint32_t held;        // particles held back, current (or last) step
  uint32_t nonfinite;   // an advance of this scene went non-finite
  uint32_t substeps_total_lo, substeps_total_hi;
  uint32_t pad[3];
};

// Everything a batch kernel reads besides the particle arrays.
struct BatchArgs {
  uint32_t n, scenes;
  float radius, rinv, reach, scale, mass, rest, stiffness, soften, half_spacing;
  float gx, gy, gz;
  float sound_speed, viscosity, courant;
  uint32_t max_substeps;
  const SceneDesc* desc;
  SceneState* state;
  const float4* walls;
  const uint32_t* wall_start;
  const uint32_t* cell_start;
  Control* ctl;
  BatchCursor* cur;
};

// Block sizes of the batch kernels (threads per block); the autotuner picks
// them from kBatchBlocks.
struct BatchShape {
  uint32_t sort = 256;     // bounds, count, scatter, gather
  uint32_t density = 256;
  uint32_t force = 256;
  bool operator<(const BatchShape& o) const {
    return sort != o.sort ? sort < o.sort : density != o.density ? density < o.density : force < o.force;
  }
  bool operator==(const BatchShape& o) const { return sort == o.sort && density == o.density && force == o.force; }
};
constexpr uint32_t kBatchBlocks[] = {64, 128, 256, 512, 1024};

// Launchers (kernels_batch.hip). `map` is the block -> (scene, block of
// scene) table for that block size, `blocks` its length.
void launch_batch_fastest(hipStream_t s, const BatchArgs& b, const float4* V, const uint2* map, uint32_t blocks,
                          uint32_t block);
// Before a run of launches: clears the stop word and flags; `fresh` starts a
// new call (substep 0, target = finished steps + steps).
void launch_batch_reset(hipStream_t s, const BatchArgs& b, int fresh, uint32_t steps, float dt, uint32_t slots);
void launch_batch_plan(hipStream_t s, const BatchArgs& b);
void launch_batch_substep_begin(hipStream_t s, const BatchArgs& b);
void launch_batch_bounds(hipStream_t s, const BatchArgs& b, const float4* P, const uint2* map, uint32_t blocks,
                         uint32_t block);
void launch_batch_grid_setup(hipStream_t s, const BatchArgs& b, uint32_t* counts, uint64_t capacity);
void launch_batch_count(hipStream_t s, const BatchArgs& b, const float4* P, uint32_t* counts, uint2* slot_rank,
                        const uint2* map, uint32_t blocks, uint32_t block);
void launch_batch_scatter(hipStream_t s, const BatchArgs& b, const float4* P, const uint2* slot_rank,
                          const uint32_t* start, uint2* pairs, const uint2* map, uint32_t blocks, uint32_t block);
void launch_batch_gather(hipStream_t s, const BatchArgs& b, const float4* PA, const float4* VA, const uint2* pairs,
                         const uint32_t* start, float4* PB, float4* VB, const uint2* map, uint32_t blocks,
                         uint32_t block);
void launch_batch_density(hipStream_t s, const BatchArgs& b, const float4* P, float4* V, float* scaled,
                          const uint2* map, uint32_t blocks, uint32_t block);
void launch_batch_force(hipStream_t s, const BatchArgs& b, const float4* PB, const float4* VB, const float* scaled,
                        float4* PA, float4* VA, const uint2* map, uint32_t blocks, uint32_t block);
void launch_batch_substep_end(hipStream_t s, const BatchArgs& b);
void launch_batch_step_end(hipStream_t s, const BatchArgs& b);

class Batch {
 public:
  static int32_t create(World* const* worlds, size_t count, Batch** out, std::string& why);
  ~Batch();

  int32_t step(float dt, uint32_t steps, sfl_report* reports);
  int32_t prepare(uint32_t steps);
  int32_t autotune(float dt, uint32_t repeats, sfl_batch_tuning* out);
  int32_t set_tuning(const sfl_batch_tuning& t);
  int32_t get_tuning(sfl_batch_tuning* out) const;
  int32_t use_graph(bool on);
  int32_t read_state(size_t scene, size_t cap, uint32_t* ids, float* xyz, float* vxyz, float* rho);
  int32_t state_hash(size_t scene, uint64_t* out);
  int32_t store(World* const* worlds, size_t count);
  int32_t scene_failed(size_t scene, int32_t* failed) const;
  int32_t synchronize();

  size_t scenes() const { return desc_.size(); }
  size_t len(size_t scene) const { return scene < desc_.size() ? desc_[scene].count : 0; }
  size_t total() const { return n_; }
  size_t device_bytes() const;
  uint64_t graph_nodes() const { return graph_nodes_; }
  const char* detail() const { return detail_.c_str(); }
  void set_detail(const std::string& d) { detail_ = d; }

 private:
  Batch() = default;
  int32_t init(World* const* worlds, size_t count);
  int32_t device_error(const char* what, hipError_t e);
  BatchArgs args() const;
  int32_t map_for(uint32_t block, const uint2** map, uint32_t* blocks);
  int32_t ensure_cells(uint64_t volume);
  void enqueue_substep(const BatchShape& shape);
  void enqueue_steps(uint32_t steps, uint32_t slots, const BatchShape& shape);
  int32_t prepare_maps(const BatchShape& shape);
  uint32_t chunk_for(uint32_t steps) const;
  int32_t graph_for(uint32_t steps, uint32_t slots, hipGraphExec_t* exec);
  void drop_graph();
  int32_t read_back(bool states);

  sfl_settings settings_{};
  float radius_ = 0, reach_ = 0, scale_ = 0, mass_ = 0, stiffness_ = 0;
  size_t n_ = 0;
  std::vector<SceneDesc> desc_;
  std::vector<SceneState> state_host_;
  std::string detail_;

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  Control* ctl_ = nullptr;
  Control* ctl_host_ = nullptr;      // pinned
  BatchCursor* cur_ = nullptr;
  BatchCursor* cur_host_ = nullptr;  // pinned
  DeviceBuffer<float4> pa_, va_, pb_, vb_, walls_;
  DeviceBuffer<float> scaled_;
  DeviceBuffer<uint2> slot_rank_, scatter_;
  DeviceBuffer<uint32_t> counts_, cell_start_, tile_sums_, scan_total_, wall_start_;
  DeviceBuffer<SceneDesc> desc_d_;
  DeviceBuffer<SceneState> state_d_;
  uint64_t cell_cap_ = 0;
  std::map<uint32_t, std::pair<DeviceBuffer<uint2>*, uint32_t>> maps_;

  BatchShape shape_;
  sfl_batch_tuning tuning_{};
  bool graphs_enabled_ = true;
  // captured sequences by (steps, slots); all of them hold the current
  // shape's kernels and cell table, so a change of either drops them
  struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t nodes = 0;
  };
  std::map<std::pair<uint32_t, uint32_t>, Graph> graphs_;
  uint64_t graph_nodes_ = 0;
  uint32_t slots_ = 1;
};

}  // namespace sfl
