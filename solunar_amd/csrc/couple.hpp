// Coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs): the records the
// coupling kernels (kernels_couple.hip) share with the host side.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace sfl {

// The pair records of one step and their sort (PairRecords, device.hpp).
struct PairRoom {
  uint32_t* counts;  // n
  uint32_t* offsets;  // n + 1
  uint32_t* tiles;
  uint32_t* total;
  uint32_t cap;
  uint32_t none;
  uint32_t key_bits;
  uint32_t* key;
  uint32_t* value;
  uint32_t* key_sorted;
  uint32_t* value_sorted;
  float4* force;
  float4* torque;
  void* temp;
  size_t temp_bytes;
};

// The fastest point and the box of each body (couple.rs:321-339), and
// whether the water's last grid has a cell in reach of it.
struct BodyExtent {
  uint32_t speed_bits;
  int32_t lo[3], hi[3];  // order-preserving keys of the box
  uint32_t wet;
};

// A body's cell range (from the host) to test against the water's grid.
struct WetQuery {
  uint32_t body;
  int32_t lo[3], hi[3];
};

// What the lift of one body reads from the water (lift.rs:25-152).
struct LiftState {
  int32_t box_lo[3], box_hi[3];  // order-preserving keys of the probes' box
  float low[3], high[3];         // the box grown by READ_REACH radii
  uint32_t near;                 // water particles inside
  uint32_t near_offset;          // into the near-list buffer (set by the host)
  int32_t col_lo[2];             // first column (x, z) of the box
  uint32_t col_span[2];
  uint32_t column_offset;  // into the column table
  uint32_t sorted;         // 1: the tops come from a sort of the water's column keys, not the table
  uint32_t tops;           // sorted: the number of columns (written on the device)
  uint32_t sort_offset, sort_size;
  float level;
  float flow[3];
  float wetted, volume;
  float middle[3];  // sum of arm * share, not yet divided
  uint32_t line;    // probes on the waterline
  float area;       // volume^(2/3) / wetted, set by the host
  float impulse[3], angular[3];
  uint32_t lifted;
};

// The buffers of one carry (couple.rs:293-319).
struct MoverCarry {
  uint32_t m, nbodies, owner_bits;
  const float4* ms;  // start (x, y, z, owner bits), current order
  const float4* mv;  // velocity
  float4* ms2;       // sorted start, velocity and carried position
  float4* mv2;
  float4* mp;
  MoverGrid* grid;
  uint64_t* keys_in;
  uint32_t* perm_in;
  uint64_t* keys;  // sorted keys, read by the WCSPH passes
  uint32_t* perm;
  void* temp;
  size_t temp_bytes;
  BodyState* bodies;
  uint32_t* table;  // dense cell table, table_cap slots (0: none)
  uint64_t table_cap;
};

void launch_movers_place(hipStream_t s, const float4* local, uint32_t m, const BodyState* bodies, float4* MS,
                         float4* MV);
hipError_t mover_sort_bytes(uint32_t m, size_t* bytes);
hipError_t launch_movers_carry(hipStream_t s, const MoverCarry& c, float elapsed, float radius, Control* ctl);
void launch_pairs_prepare(hipStream_t s, const PairRoom& room, uint32_t n, Control* ctl);
hipError_t pair_sort_bytes(uint32_t cap, size_t* bytes);
hipError_t launch_pairs_fold(hipStream_t s, const PairRoom& room, BodyState* bodies, uint32_t nbodies, float dt,
                             const Control* ctl);
void launch_extent(hipStream_t s, const float4* MS, const float4* MV, const BodyState* bodies, BodyExtent* e,
                   uint32_t nbodies);
void launch_extent_wet(hipStream_t s, const uint32_t* start, const uint64_t* keys, const Control* ctl,
                       const WetQuery* q, uint32_t nq,
                       BodyExtent* e);
void launch_lift_box(hipStream_t s, const float4* probes, const BodyState* bodies, uint32_t nbodies,
                     uint32_t most_probes, float4* arm, float4* vel, LiftState* ls, float reach);
void launch_lift_count(hipStream_t s, const float4* P, uint32_t n, LiftState* ls, uint32_t nbodies, float4* boxes,
                       uint32_t* counts, uint32_t* offsets, uint32_t* tiles, uint32_t* total);
void launch_lift_near(hipStream_t s, const float4* P, uint32_t n, LiftState* ls, uint32_t nbodies,
                      const float4* boxes, const uint32_t* offsets, uint32_t* near, unsigned long long* columns,
                      uint64_t* column_keys, uint32_t* column_tops, float column, const float up[3]);
// The buffers of the sorted column path (bodies whose column box is too
// sparse for a table): the near water's column keys and heights, their
// sorted copies, the unique keys, and rocPRIM's temporary storage.
struct LiftSort {
  uint64_t* keys;
  uint32_t* values;
  uint64_t* keys_sorted;
  uint32_t* values_sorted;
  uint64_t* unique;
  void* temp;
  size_t temp_bytes;
};
hipError_t lift_sort_bytes(uint32_t most, size_t* bytes);
// Body b's column tops into sortbuf + sort_offset and their count into its
// LiftState (the host's copy `l` gives the offsets and the count of water).
hipError_t launch_lift_tops(hipStream_t s, const LiftSort& t, LiftState* ls, uint32_t b, const LiftState& l,
                            uint32_t* sortbuf);
void launch_lift_level(hipStream_t s, const float4* P, const float4* V, LiftState* ls, const BodyState* bodies,
                       uint32_t nbodies, const uint32_t* near, const unsigned long long* columns, uint32_t* sortbuf,
                       const float4* arm, const float4* vel, float half_spacing, const float up[3]);
void launch_lift_force(hipStream_t s, LiftState* ls, const BodyState* bodies, uint32_t nbodies, const float4* arm,
                       const float4* vel, float rest, float gravity, const float up[3], float dt);
void launch_give_back(hipStream_t s, const uint32_t* near, uint32_t count, const float change[3], float4* V);

}  // namespace sfl
