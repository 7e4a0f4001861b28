// Reverse-mode differentiation of a WCSPH rollout: the recording forward
// (the world's own step, with checkpoints of (x, v) every `every` steps) and
// the reverse pass (each segment replayed from its checkpoint with the same
// kernels, keeping per substep the sorted state, cell table, sort
// permutation and clamp flags, then the hand-written adjoint kernels run
// back through it). The replay of the segment before runs on a stream of
// its own beside the reverse kernels of the current one.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/solunar_fluid.h"
#include "device.hpp"
#include "world.hpp"

namespace sfl {

// One recorded substep, as the reverse kernels read it.
struct AdjointArgs {
  uint32_t n;
  float radius, reach, scale, mass, rest, stiffness, soften;
  float dt, nu, wall_nu;
  float dnu_dalpha;  // d nu / d (viscosity setting) at this substep, 0 where the cap binds
  const float4* P;   // sorted positions (x, id bits)
  const float4* V;   // sorted velocities, densities in .w
  const float* scaled;
  const uint8_t* clamps;
  const uint32_t* src;
  const uint32_t* cell_start;
  const CellBox* box;
  const float4* walls;
  const uint32_t* wall_start;
  const uint64_t* wall_keys;  // searched wall table, or null
  uint32_t wall_keyed;
  CellBox wall_box;
  uint32_t wall_cells;
};

void launch_adj_load(hipStream_t s, uint32_t n, const float4* P, const float* xbar, const float* vbar, float4* Xb,
                     float4* Vb);
void launch_adj_store(hipStream_t s, uint32_t n, const float4* P, const float4* Xb, const float4* Vb, float* xbar,
                      float* vbar);
void launch_adj_advance(hipStream_t s, const AdjointArgs& a, const float4* XbOut, const float4* VbOut, float4* Xb,
                        float4* Vb, float4* Ab);
uint32_t adj_force_blocks(uint32_t n);
void launch_adj_force(hipStream_t s, const AdjointArgs& a, const float4* Ab, float4* Xb, float4* Vb, float* Rb,
                      double* nu_acc);
void launch_adj_density(hipStream_t s, const AdjointArgs& a, const float* Rb, const float4* Xb, const float4* Vb,
                        float4* XbPrev, float4* VbPrev);
void launch_adj_keep_box(hipStream_t s, const Control* ctl, CellBox* box);
void launch_adj_compare(hipStream_t s, uint32_t n, const float4* P, const float4* V, const float4* P2,
                        const float4* V2, uint32_t* count);

class Tape {
 public:
  static int32_t record(World* w, float dt, uint32_t steps, uint32_t every, Tape** out, std::string& why);
  ~Tape();
  int32_t adjoint(const float* xbar, const float* vbar, float* xbar0, float* vbar0, float* viscosity_bar);
  size_t device_bytes() const;
  uint64_t substeps() const;
  const char* detail() const { return detail_.c_str(); }
  void set_detail(const std::string& d) { detail_ = d; }

 private:
  Tape() = default;
  int32_t device_error(const char* what, hipError_t e);
  int32_t replay(uint32_t segment, uint32_t set, uint32_t* count);
  int32_t ensure_slots(uint32_t set, uint32_t count);
  int32_t reverse(uint32_t segment, uint32_t set, uint32_t count);
  int32_t ensure_streams();

  struct Slot {
    DeviceBuffer<float4> P, V;
    DeviceBuffer<float> scaled;
    DeviceBuffer<uint8_t> clamps;
    DeviceBuffer<uint32_t> src, cells;
  };

  World* w_ = nullptr;
  uint64_t version_ = 0;
  float dt_ = 0;
  uint32_t steps_ = 0, every_ = 1;
  size_t n_ = 0;
  std::vector<uint32_t> substeps_;  // per step
  std::vector<float> small_;        // substep length per step
  std::vector<std::unique_ptr<DeviceBuffer<float4>>> ckp_, ckv_;  // checkpoints (x, v) per segment
  // Two slot sets: segment c replays into set c % 2 on the replay stream
  // while the reverse kernels read the other set on the world's stream.
  std::vector<std::unique_ptr<Slot>> slots_[2];
  DeviceBuffer<CellBox> boxes_[2];
  hipStream_t replay_stream_ = nullptr;
  hipEvent_t ready_ = nullptr;        // world stream: set-up done / replay stream: a replay done
  hipEvent_t reversed_[2] = {};       // reverse kernels done reading set k
  DeviceBuffer<float4> ra_p_, ra_v_;            // replay state
  DeviceBuffer<float4> xb_, vb_, xb2_, vb2_, ab_;  // adjoints
  DeviceBuffer<float> rb_, bar_in_, bar_out_;
  DeviceBuffer<double> nu_acc_;
  DeviceBuffer<uint32_t> counts_, tile_sums_, scan_total_, mismatch_;
  DeviceBuffer<uint2> slot_rank_, scatter_;
  uint64_t cell_cap_ = 0;
  Control* ctl_ = nullptr;
  Control* ctl_host_ = nullptr;
  std::string detail_;
};

}  // namespace sfl
