// Host side of the device fluid world: the C++ counterpart of the
// reference's FluidWorld (src/fluid/mod.rs:310-339) whose particle state
// lives in HBM. It validates input, keeps the walls, sizes device buffers,
// orders the launches of each step on the world's stream, and turns the
// control block into FluidReport fields and PhysicsError codes.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/solunar_fluid.h"
#include "couple.hpp"
#include "device.hpp"
#include "launch.hpp"

namespace sfl {

// Set while an allocation-free step runs on this thread: a buffer that would
// have to grow then fails instead of calling hipMalloc.
inline thread_local bool t_no_alloc = false;

// The allocation-free step's scratch (no_alloc.hip): its size, the device
// address of one GPU's copy, and a kernel that clears part of it.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
hipError_t no_alloc_scratch(int device, unsigned char** out);
void launch_fill_zero(hipStream_t s, void* p, size_t bytes);

template <class T>
struct DeviceBuffer {
  T* ptr = nullptr;
  size_t cap = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
  // Grows to at least `n` elements; keeps the first `keep` elements.
  hipError_t reserve(size_t n, size_t keep, hipStream_t s);
  void swap(DeviceBuffer& o) {
    std::swap(ptr, o.ptr);
    std::swap(cap, o.cap);
  }
};

template <class T>
hipError_t DeviceBuffer<T>::reserve(size_t n, size_t keep, hipStream_t s) {
  if (n <= cap) return hipSuccess;
  if (t_no_alloc) return hipErrorOutOfMemory;
  size_t want = std::max(n, cap + cap / 2);
  T* fresh = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&fresh), want * sizeof(T));
  if (e != hipSuccess) return e;
  if (keep && ptr) {
    e = hipMemcpyAsync(fresh, ptr, keep * sizeof(T), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) {
      (void)hipFree(fresh);
      return e;
    }
  }
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(fresh);
    return e;
  }
  release();
  ptr = fresh;
  cap = want;
  return hipSuccess;
}

// FNV-1a 64 of (id, x, v) in id order (src/fluid/mod.rs:678-691,
// src/snapshot/bytes.rs:6-13) over float4 pairs (x, id bits) and (v, rho).
uint64_t hash_particles(const float4* p, const float4* v, size_t n);

class Batch;
class Tape;

// Why a snapshot was refused (sfl_snapshot_error).
struct SnapshotFailure {
  int32_t kind;
  uint32_t version;
  const char* what;
};

class World {
  friend class Batch;
  friend class Tape;

 public:
  static int32_t create(const sfl_settings& s, World** out, std::string& why);
  static bool valid_settings(const sfl_settings& s);
  // snapshots (src/fluid/save.rs), snapshot.cpp
  size_t snapshot_size() const;
  int32_t save_snapshot(uint8_t* bytes);
  static int32_t load_snapshot(const uint8_t* data, size_t size, const sfl_settings& base, World** out,
                               SnapshotFailure& fail, std::string& why);
  ~World();

  int32_t add_particles(size_t count, const float* xyz, const float* vxyz, uint32_t* ids);
  int32_t fill_box(sfl_vec3 low, sfl_vec3 high, float jitter, uint32_t seed, size_t* count);
  int32_t set_domain(const sfl_vec3* low, const sfl_vec3* high);
  int32_t add_emitter(sfl_vec3 low, sfl_vec3 high, sfl_vec3 velocity, uint32_t* id);
  int32_t add_drain(sfl_vec3 low, sfl_vec3 high, uint32_t* id);
  int32_t remove_flow(uint32_t id, int32_t* removed);
  int32_t add_box_container(sfl_vec3 low, sfl_vec3 high, uint32_t* id);
  int32_t add_wall(sfl_vec3 low, sfl_vec3 high, uint32_t* id);
  int32_t add_boundary_points(size_t count, const float* xyz, uint32_t* id);
  int32_t remove_boundary(uint32_t id, int32_t* removed);

  int32_t step(float dt, sfl_report* report);
  // step of a WCSPH world with its control block and cell table in the
  // shared scratch (no_alloc.hip): calls no hipMalloc or hipFree
  int32_t step_no_alloc(float dt, sfl_report* report);
  // coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs), couple.cpp
  int32_t couple(uint32_t index, uint32_t generation, const sfl_shape& shape);
  int32_t uncouple(uint32_t index, uint32_t generation, int32_t* was);
  size_t coupled_len() const { return coupled_.size(); }
  int32_t step_with_bodies(float dt, const sfl_body_pose* poses, size_t count, sfl_report* report,
                           sfl_body_push* pushes, size_t cap, size_t* out_count);
  // terrain ground (src/fluid/terrain.rs), terrain.cpp
  void terrain_nodes(sfl_vec3 low, sfl_vec3 high, uint64_t nodes[3]) const;
  int32_t add_terrain(sfl_vec3 low, sfl_vec3 high, const float* distances, size_t count, uint32_t* id);
  size_t terrain_len() const { return terrains_.size(); }
  int32_t update_neighbors();
  int32_t compute_densities();

  int32_t read_ids(size_t cap, uint32_t* out);
  int32_t read_positions(size_t cap, float* out);
  int32_t read_velocities(size_t cap, float* out);
  int32_t read_densities(size_t cap, float* out);
  int32_t read_forces(size_t cap, float* out);
  int32_t read_neighbors(int kind, size_t start_cap, uint32_t* start, size_t items_cap, uint32_t* items,
                         size_t* items_len);
  int32_t read_walls(size_t cap, float* out) const;
  int32_t state_hash(uint64_t* out);
  int32_t keep_forces(bool keep);
  int32_t synchronize();

  int32_t profile(bool enable);
  int32_t profile_read(double* ms, uint64_t* launches, bool reset);
  int32_t work_read(uint64_t* substeps, uint64_t* pushes, uint64_t* rates, bool reset);
  int32_t read_solve_log(size_t cap, double* average, uint32_t* loaded, size_t* len);

  const sfl_settings& settings() const { return settings_; }
  // domain (src/fluid/mod.rs:446-448): false when none is set
  bool domain(sfl_vec3* low, sfl_vec3* high) const;
  // kernel (mod.rs:376-378) with Kernel::value / slope / gradient
  // (mod.rs:140-164), in the device's arithmetic
  float kernel_value_at(float distance) const { return kernel_value(distance, radius_, scale_); }
  float kernel_slope_at(float distance) const { return kernel_slope(distance, radius_, scale_); }
  sfl_vec3 kernel_gradient_at(sfl_vec3 offset) const;
  // neighbors_of (mod.rs:612-614): particle `index`'s fluid list of the last
  // build
  int32_t neighbors_of(size_t index, size_t cap, uint32_t* out, size_t* len);
  float mass() const { return mass_; }
  size_t len() const { return n_; }
  size_t wall_len() const { return wall_owner_.size(); }
  int solve_grid() const { return solve_grid_; }
  size_t device_bytes() const;
  const char* detail() const { return detail_.c_str(); }
  void set_detail(const std::string& d) { detail_ = d; }

 private:
  World() = default;
  int32_t device_error(const char* what, hipError_t e);
  int32_t init_device();
  StepParams params(float dt) const;
  int32_t ensure_particles(size_t n);
  int32_t ensure_cells(uint64_t volume);
  int32_t ensure_cells_for_bounds();
  int32_t read_control();
  struct Totals {
    float vmax, comp;
    uint32_t held, pairs_max;
  };
  Totals totals() const;
  int32_t ensure_pairs(uint32_t cap);
  int32_t settle_walls();
  int32_t wcsph(float dt, sfl_report& r);
  int32_t dfsph(float dt, sfl_report& r);
  int32_t enqueue_sort();
  int32_t enqueue_wcsph_substep(const StepParams& prm, uint32_t index, float small);
  int32_t enqueue_dfsph_substep(const StepParams& prm);
  int32_t emit_lists();
  int32_t ensure_fastest();
  void invalidate_lists() {
    lists_valid_ = false;
    lists_emitted_ = false;
  }
  int32_t add_wall_points(const std::vector<float>& pts, uint32_t* id);
  int32_t add_flow(sfl_vec3 low, sfl_vec3 high, bool drain, sfl_vec3 velocity, uint32_t* id);
  int32_t restore(size_t count, const uint32_t* ids, const float* xyz, const float* vxyz, uint32_t next_id);
  int32_t run_flows(uint32_t& emitted, uint32_t& drained, bool& changed);
  int32_t keep_only(const FlowBox& box, bool domain, uint32_t& gone);
  int32_t emit(const sfl_vec3& low, const sfl_vec3& high, const sfl_vec3& velocity, uint32_t& added);
  int32_t read_count(const uint32_t* dev, uint32_t* out);
  size_t coupled_points() const;
  int32_t place_movers(const sfl_body_pose* poses, size_t count, float dt);
  int32_t carry_movers(float elapsed);
  int32_t replay_movers(uint32_t done, float small);
  MoverView mover_view(float dt) const;
  int32_t fastest_near_water(float dt, float slow, float& most);
  int32_t lifts(float dt, const sfl_body_pose* poses, size_t count, sfl_body_push* pushes);
  void phase_begin(int phase);
  void phase_end(int phase);
  void collect_profile();

  sfl_settings settings_{};
  float radius_ = 0, reach_ = 0, scale_ = 0, mass_ = 0;
  size_t n_ = 0;
  uint32_t next_id_ = 0;
  uint64_t version_ = 0;  // bumped by every call that may change the particles
  float vmax_ = 0;
  bool vmax_valid_ = true;
  bool keep_forces_ = false;
  bool forces_valid_ = false;
  std::string detail_;

  // walls, host copy in settled order (src/fluid/boundary.rs:23-33)
  std::vector<float> wall_pos_;
  std::vector<uint32_t> wall_owner_;
  uint32_t wall_next_ = 0;
  struct Container {
    uint32_t id;
    float lo[3], hi[3];
  };
  std::vector<Container> containers_;
  // flows (src/fluid/flow.rs) and the domain (src/fluid/mod.rs:428-465)
  struct Flow {
    uint32_t id;
    sfl_vec3 low, high;
    bool drain;
    sfl_vec3 velocity;
  };
  std::vector<Flow> flows_;
  uint32_t flow_next_ = 0;
  bool has_domain_ = false;
  FlowBox domain_{};
  bool has_bounds_ = false;
  float blo_[3] = {0, 0, 0}, bhi_[3] = {0, 0, 0};

  // device
  int device_ = 0;
  hipStream_t stream_ = nullptr;  // the world's only stream
  Control* ctl_ = nullptr;       // device
  Control* ctl_host_ = nullptr;  // pinned host mirror
  int solve_grid_ = 0;
  uint32_t solve_lds_ = 0;  // dynamic LDS of the list solve's tiles
  DeviceBuffer<float4> pa_, va_, pb_, vb_;
  DeviceBuffer<float> aux_, errors_, forces_;
  DeviceBuffer<float4> kappa_;  // (x, stiffness) per particle during a solve
  DeviceBuffer<uint2> slot_rank_, scatter_;
  DeviceBuffer<uint32_t> counts_, cell_start_, tile_sums_, scan_total_;
  uint64_t cell_cap_ = 0;
  // The searched fluid table (the reference's Grid without its dense index,
  // src/fluid/mod.rs:214-236), used while the particles' cell box is more
  // than dense_limit() cells: each sort radix-sorts the cells' 64-bit slots
  // and keeps the occupied ones (cell_keys_) with their ranges (sstart_).
  bool searched_ = false;
  uint64_t last_box_ = 0;  // cells of the box of the last searched sort
  DeviceBuffer<uint64_t> skeys_in_, skeys_, cell_keys_;
  DeviceBuffer<uint32_t> svals_in_, svals_, heads_, head_scan_, sstart_;
  DeviceBuffer<uint8_t> ssort_temp_;
  uint64_t dense_limit() const;
  int32_t use_searched();
  DeviceBuffer<uint2> pairs_;
  DeviceBuffer<uint32_t> pair_counts_;
  uint32_t pair_cap_ = 0, pair_stride_ = 0;
  // DFSPH neighbour format: compact lists (6 bytes a fluid entry, read by
  // the tiled solve; the default when the solve has its LDS tiles), full
  // lists (8 bytes), or masks (27 per particle, 16 or 32 bits, with each
  // particle's own cell). Each widens on demand: masks 16 -> 32 -> lists,
  // compact -> lists.
  DeviceBuffer<uint32_t> masks_, own_slot_;
  DeviceBuffer<uint16_t> near_;  // compact format: fluid neighbours
  DeviceBuffer<float> factor_;
  int format_ = kLists;
  // the DFSPH stop sum's per solve tile sums and counts, sized with the
  // cell table
  // keep flags and their scan (drains and the domain), free-slot flags and
  // their scan (emitters)
  DeviceBuffer<uint32_t> flow_flags_, flow_start_, flow_tiles_, flow_total_;
  // (two sets of stop_cap(): the flow solve keeps one per iteration parity)
  DeviceBuffer<double> stop_sum_;
  DeviceBuffer<uint32_t> stop_count_;
  uint32_t stop_cap() const {
    return (uint32_t)std::min<size_t>(std::min(stop_sum_.cap, stop_count_.cap) / 2, 0xffffffffu);
  }
  // the flow solve's per tile passes (stop_cap() for each of the two
  // solves), weights and list of non-empty tiles (stop_cap() each); flow_ is
  // false when SOLUNAR_SOLVE_FLOW=0 makes
  // the tiled solve's passes meet at grid barriers
  DeviceBuffer<uint32_t> flow_done_, flow_work_, flow_list_;
  bool flow_ = true;
  FlowBuffers flow_buffers(float4* v2) const {
    FlowBuffers f;
    if (flow_ && flow_done_.cap >= 2 * (size_t)stop_cap() && flow_work_.cap >= stop_cap() &&
        flow_list_.cap >= stop_cap()) {
      f.v2 = v2;
      f.tile_done = flow_done_.ptr;
      f.tile_work = flow_work_.ptr;
      f.tile_list = flow_list_.ptr;
    }
    return f;
  }
  DeviceBuffer<float4> walls_d_;
  DeviceBuffer<uint32_t> wall_start_d_;
  DeviceBuffer<uint64_t> wall_keys_d_;  // searched wall table: occupied cells' slots
  uint32_t wall_keyed_ = 0;             // occupied wall cells when searched, else 0
  CellBox wall_box_{};
  uint32_t wall_cells_ = 0;

  // neighbour lists of the last build (reference `neighbors`, `wall_neighbors`)
  bool lists_valid_ = false;
  bool lists_emitted_ = false;
  bool lists_on_b_ = false;  // true: the build's positions are PB (after a step)
  DeviceBuffer<uint32_t> fstart_, fitems_, wstart_, witems_, list_counts_;
  size_t fitems_len_ = 0, witems_len_ = 0;

  // coupled bodies (src/fluid/couple.rs:15-47), in handle order: wall
  // points and probes in body axes
  struct Coupled {
    uint32_t index, generation;
    std::vector<float> points, probes;
    float probe_spacing;
  };
  std::vector<float> body_reach_;  // each body's farthest point from its centre
  // terrain regions (TerrainField, src/fluid/terrain.rs:35-42) and the
  // world's GroundSites (terrain.rs:173-247), built with the first region
  struct Terrain {
    uint32_t owner;
    float low[3];
    uint32_t nodes[3];
    float step;
    std::vector<float> distances;
  };
  std::vector<Terrain> terrains_;
  bool terrain_dirty_ = false;
  std::vector<float4> sites_h_;
  std::vector<float> far_h_, damping_h_;
  float ground_low_ = 0, ground_step_ = 0, ground_reach_ = 0;
  DeviceBuffer<TerrainDesc> terrain_desc_;
  DeviceBuffer<float> terrain_dist_, far_d_, damping_d_;
  DeviceBuffer<float4> sites_d_, slopes_;
  int32_t upload_terrains();
  void build_ground_sites();
  TerrainView terrain_view() const;
  std::vector<Coupled> coupled_;
  bool coupled_dirty_ = true;  // the device copies of the points are stale
  uint32_t most_probes_ = 0;
  std::vector<BodyState> bodies_h_;
  uint32_t movers_m_ = 0;  // body points placed for this step; 0 outside step_with_bodies
  uint32_t owner_bits_ = 0;
  int mcur_ = 0;           // which of ms_/mv_ holds the current order
  bool grid_any_ = false;  // the water's last grid (update_neighbors) holds a cell
  // an allocation-free step left its last grid in the shared scratch, where
  // the world could not keep it: the lists are stale and the fast-body rule
  // has no grid to read until the next sort
  bool grid_dropped_ = false;
  bool no_alloc_ = false;  // inside step_no_alloc
  uint64_t sorts_ = 0;     // sorts enqueued so far
  DeviceBuffer<float4> body_local_, probe_local_;
  DeviceBuffer<BodyState> bodies_d_;
  DeviceBuffer<float4> ms_[2], mv_[2], mp_;
  DeviceBuffer<uint64_t> mkeys_in_, mkeys_;
  DeviceBuffer<uint32_t> mperm_in_, mperm_;
  DeviceBuffer<uint8_t> msort_temp_;
  DeviceBuffer<MoverGrid> mgrid_;
  DeviceBuffer<uint32_t> mtable_;  // the points' dense cell table
  // the force pass's body-point pair records and their sort by body
  DeviceBuffer<uint32_t> pair_counts_b_, pair_offsets_b_, pair_tiles_b_, pair_total_b_;
  DeviceBuffer<uint32_t> pair_key_[2], pair_value_[2];
  DeviceBuffer<float4> pair_force_, pair_torque_;
  DeviceBuffer<uint8_t> pair_temp_;
  uint32_t body_pair_cap_ = 0;  // records the buffers hold
  uint32_t body_pair_use_ = 0;  // records a substep sorts: the last step's most, and a quarter
  uint32_t pair_slack_ = 65536;  // smallest room, and the slack over the last step's most
  int32_t ensure_pair_room(uint64_t cap);
  PairRoom pair_room() const;
  DeviceBuffer<BodyExtent> extent_;
  DeviceBuffer<WetQuery> wet_;
  DeviceBuffer<LiftState> lift_;
  DeviceBuffer<float4> probe_arm_, probe_vel_;
  DeviceBuffer<uint32_t> lift_near_, lift_sort_, lift_counts_, lift_offsets_, lift_tiles_;
  DeviceBuffer<float4> lift_boxes_;
  DeviceBuffer<unsigned long long> lift_columns_;
  // the sorted column path (bodies whose column box is too sparse for the table)
  DeviceBuffer<uint64_t> lift_ckeys_, lift_ckeys2_, lift_ukeys_;
  DeviceBuffer<uint32_t> lift_cvals_, lift_cvals2_;
  DeviceBuffer<unsigned char> lift_sort_temp_;
  bool lift_tables_ = true;  // false: every body takes the sorted path (tests)

  // measurement
  bool profiling_ = false;
  struct EventPair {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<EventPair> pending_;
  std::vector<hipEvent_t> free_events_;
  hipEvent_t open_ = nullptr;
  double phase_ms_[SFL_PHASE_COUNT] = {};
  uint64_t phase_launches_[SFL_PHASE_COUNT] = {};
  uint64_t work_substeps_ = 0, work_pushes_ = 0, work_rates_ = 0;
};

}  // namespace sfl
