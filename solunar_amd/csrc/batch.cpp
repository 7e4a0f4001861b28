// Host side of the batched WCSPH path (batch.hpp): builds the shared arrays
// from a set of worlds, captures and launches the step sequence, grows the
// cell table or the substep slots when the device asks for it, times the
// block-size candidates, and reads scenes back.
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>

#include "launch.hpp"

namespace sfl {

namespace {

constexpr uint64_t kMaxCells = uint64_t(1) << 31;
constexpr uint32_t kMaxGraphSlots = 64;      // more substeps per step: plain launches
constexpr uint64_t kGraphLaunches = 16384;   // kernel nodes per captured graph, about
constexpr uint32_t kSlotLaunches = 12;       // launches of one substep slot

struct DeviceGuard {
  explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

int32_t cell_coord_host(float v, float radius) {
  float c = std::floor(v / radius);
  if (std::isnan(c)) return 0;
  if (c >= 2147483648.0f) return std::numeric_limits<int32_t>::max();
  if (c <= -2147483648.0f) return std::numeric_limits<int32_t>::min();
  return (int32_t)c;
}

// Settings that change a scene's arithmetic must be equal across a batch.
bool same_physics(const sfl_settings& a, const sfl_settings& b) {
  return a.solver == b.solver && a.spacing == b.spacing && a.rest_density == b.rest_density &&
         a.support == b.support && a.sound_speed == b.sound_speed && a.viscosity == b.viscosity &&
         a.gravity.x == b.gravity.x && a.gravity.y == b.gravity.y && a.gravity.z == b.gravity.z &&
         a.courant == b.courant && a.max_substeps == b.max_substeps && a.device == b.device;
}

}  // namespace

int32_t Batch::device_error(const char* what, hipError_t e) {
  detail_ = std::string(what) + ": " + hipGetErrorString(e);
  return SFL_DEVICE_ERROR;
}

#define SFL_HIP(call, what)                                \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return device_error(what, e_);   \
  } while (0)

int32_t Batch::create(World* const* worlds, size_t count, Batch** out, std::string& why) {
  *out = nullptr;
  if (!worlds || count == 0) {
    why = "a batch needs at least one world";
    return SFL_OUT_OF_RANGE;
  }
  if (count > (size_t(1) << 20)) {
    why = "a batch holds at most 2^20 scenes";
    return SFL_OUT_OF_RANGE;
  }
  for (size_t s = 0; s < count; ++s)
    if (!worlds[s]) return SFL_INVALID_HANDLE;
  const World& w0 = *worlds[0];
  uint64_t total = 0;
  for (size_t s = 0; s < count; ++s) {
    const World& w = *worlds[s];
    if (w.settings_.solver != SFL_WCSPH) {
      why = "a batch steps WCSPH worlds";
      return SFL_INVALID_SETTINGS;
    }
    if (!same_physics(w.settings_, w0.settings_)) {
      why = "the worlds of a batch need equal settings";
      return SFL_INVALID_SETTINGS;
    }
    if (!w.flows_.empty() || w.has_domain_ || !w.terrains_.empty()) {
      why = "a batch steps worlds without flows, a domain or terrain";
      return SFL_UNSUPPORTED;
    }
    if (w.wall_keyed_ || w.searched_) {
      why = "a batch keeps dense cell tables: a world's walls or water are too spread out for one";
      return SFL_UNSUPPORTED;
    }
    total += w.n_;
  }
  if (total >= (uint64_t(1) << 31)) {
    why = "a batch holds fewer than 2^31 particles";
    return SFL_OUT_OF_RANGE;
  }
  std::unique_ptr<Batch> b(new Batch());
  b->settings_ = w0.settings_;
  b->radius_ = w0.radius_;
  b->reach_ = w0.reach_;
  b->scale_ = w0.scale_;
  b->mass_ = w0.mass_;
  float c = w0.settings_.sound_speed;
  b->stiffness_ = w0.settings_.rest_density * (c * c) / 7.0f;  // step.rs:149
  b->n_ = (size_t)total;
  b->device_ = w0.device_;
  int32_t st = b->init(worlds, count);
  if (st != SFL_OK) {
    why = b->detail_;
    return st;
  }
  *out = b.release();
  return SFL_OK;
}

int32_t Batch::init(World* const* worlds, size_t count) {
  DeviceGuard g(device_);
  SFL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  SFL_HIP(hipMalloc(reinterpret_cast<void**>(&ctl_), sizeof(Control)), "hipMalloc(control)");
  SFL_HIP(hipMemset(ctl_, 0, sizeof(Control)), "hipMemset(control)");
  SFL_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctl_host_), sizeof(Control), hipHostMallocDefault),
          "hipHostMalloc(control)");
  std::memset(ctl_host_, 0, sizeof(Control));
  SFL_HIP(hipMalloc(reinterpret_cast<void**>(&cur_), sizeof(BatchCursor)), "hipMalloc(cursor)");
  SFL_HIP(hipMemset(cur_, 0, sizeof(BatchCursor)), "hipMemset(cursor)");
  SFL_HIP(hipHostMalloc(reinterpret_cast<void**>(&cur_host_), sizeof(BatchCursor), hipHostMallocDefault),
          "hipHostMalloc(cursor)");
  std::memset(cur_host_, 0, sizeof(BatchCursor));

  size_t n = std::max<size_t>(n_, 1);
  SFL_HIP(pa_.reserve(n, 0, stream_), "hipMalloc(positions)");
  SFL_HIP(va_.reserve(n, 0, stream_), "hipMalloc(velocities)");
  SFL_HIP(pb_.reserve(n, 0, stream_), "hipMalloc(sorted positions)");
  SFL_HIP(vb_.reserve(n, 0, stream_), "hipMalloc(sorted velocities)");
  SFL_HIP(scaled_.reserve(n, 0, stream_), "hipMalloc(pressure terms)");
  SFL_HIP(slot_rank_.reserve(n, 0, stream_), "hipMalloc(slot ranks)");
  SFL_HIP(scatter_.reserve(n, 0, stream_), "hipMalloc(sort pairs)");
  SFL_HIP(scan_total_.reserve(1, 0, stream_), "hipMalloc(scan total)");

  uint64_t wall_total = 0, wall_cells_total = 0;
  for (size_t s = 0; s < count; ++s) {
    const World& w = *worlds[s];
    if (!w.wall_owner_.empty()) {
      wall_total += w.wall_owner_.size();
      wall_cells_total += (uint64_t)w.wall_cells_ + 1;
    }
  }
  if (wall_total >= (uint64_t(1) << 32) || wall_cells_total >= (uint64_t(1) << 32)) {
    detail_ = "the batch's walls outgrow 32-bit indices";
    return SFL_OUT_OF_RANGE;
  }
  SFL_HIP(walls_.reserve(std::max<uint64_t>(wall_total, 1), 0, stream_), "hipMalloc(walls)");
  SFL_HIP(wall_start_.reserve(std::max<uint64_t>(wall_cells_total, 1), 0, stream_), "hipMalloc(wall tables)");

  desc_.resize(count);
  uint32_t first = 0, woff = 0, cbase = 0;
  uint64_t cells = 0;
  std::vector<uint32_t> table;
  for (size_t s = 0; s < count; ++s) {
    World& w = *worlds[s];
    int32_t st = w.synchronize();
    if (st != SFL_OK) {
      detail_ = w.detail_;
      return st;
    }
    SceneDesc& d = desc_[s];
    std::memset(&d, 0, sizeof(d));
    d.first = first;
    d.count = (uint32_t)w.n_;
    if (w.n_) {
      SFL_HIP(hipMemcpyAsync(pa_.ptr + first, w.pa_.ptr, w.n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_),
              "copy positions");
      SFL_HIP(hipMemcpyAsync(va_.ptr + first, w.va_.ptr, w.n_ * sizeof(float4), hipMemcpyDeviceToDevice, stream_),
              "copy velocities");
    }
    if (!w.wall_owner_.empty()) {
      uint32_t nw = (uint32_t)w.wall_owner_.size();
      SFL_HIP(hipMemcpyAsync(walls_.ptr + woff, w.walls_d_.ptr, nw * sizeof(float4), hipMemcpyDeviceToDevice,
                             stream_),
              "copy walls");
      table.resize((size_t)w.wall_cells_ + 1);
      SFL_HIP(hipMemcpy(table.data(), w.wall_start_d_.ptr, table.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
              "read wall table");
      for (auto& v : table) v += woff;
      SFL_HIP(hipMemcpy(wall_start_.ptr + cbase, table.data(), table.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice),
              "copy wall table");
      d.wall_base = cbase;
      d.wall_cells = w.wall_cells_;
      d.wall_box = w.wall_box_;
      woff += nw;
      cbase += (uint32_t)table.size();
    }
    d.has_bounds = w.has_bounds_ ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
      d.blo[a] = w.blo_[a];
      d.bhi[a] = w.bhi_[a];
    }
    // the scene's cell box can only cover its container (advance clamps
    // inside it); without one, a start the overflow path grows from
    uint64_t want = std::max<uint64_t>(64, 2 * (uint64_t)w.n_);
    if (w.has_bounds_) {
      uint64_t v = 1;
      for (int a = 0; a < 3; ++a) {
        int64_t lo = (int64_t)cell_coord_host(w.blo_[a], radius_) - 1;
        int64_t hi = (int64_t)cell_coord_host(w.bhi_[a], radius_) + 1;
        v = std::min<uint64_t>(v * (uint64_t)std::max<int64_t>(hi - lo + 1, 1), kMaxCells);
      }
      want = v;
    }
    cells = std::min<uint64_t>(cells + want, kMaxCells);
    first += (uint32_t)w.n_;
  }
  SFL_HIP(desc_d_.reserve(count, 0, stream_), "hipMalloc(scenes)");
  SFL_HIP(state_d_.reserve(count, 0, stream_), "hipMalloc(scene state)");
  SFL_HIP(hipMemcpyAsync(desc_d_.ptr, desc_.data(), count * sizeof(SceneDesc), hipMemcpyHostToDevice, stream_),
          "copy scenes");
  SFL_HIP(hipMemsetAsync(state_d_.ptr, 0, count * sizeof(SceneState), stream_), "clear scene state");
  state_host_.assign(count, SceneState{});
  int32_t st = ensure_cells(cells);
  if (st != SFL_OK) return st;
  // fastest (step.rs:92-96) of each scene for the first step's plan
  const uint2* map = nullptr;
  uint32_t blocks = 0;
  st = map_for(shape_.sort, &map, &blocks);
  if (st != SFL_OK) return st;
  launch_batch_fastest(stream_, args(), va_.ptr, map, blocks, shape_.sort);
  SFL_HIP(hipGetLastError(), "launch");
  SFL_HIP(hipStreamSynchronize(stream_), "batch set-up");
  return SFL_OK;
}

Batch::~Batch() {
  DeviceGuard g(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  drop_graph();
  for (auto& m : maps_) delete m.second.first;
  if (ctl_) (void)hipFree(ctl_);
  if (cur_) (void)hipFree(cur_);
  if (ctl_host_) (void)hipHostFree(ctl_host_);
  if (cur_host_) (void)hipHostFree(cur_host_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

BatchArgs Batch::args() const {
  const sfl_settings& s = settings_;
  BatchArgs a{};
  a.n = (uint32_t)n_;
  a.scenes = (uint32_t)desc_.size();
  a.radius = radius_;
  a.rinv = 1.0f / radius_;
  a.reach = reach_;
  a.scale = scale_;
  a.mass = mass_;
  a.rest = s.rest_density;
  a.stiffness = stiffness_;
  a.soften = 0.01f * radius_ * radius_;  // step.rs:165
  a.half_spacing = s.spacing / 2.0f;     // step.rs:308
  a.gx = s.gravity.x;
  a.gy = s.gravity.y;
  a.gz = s.gravity.z;
  a.sound_speed = s.sound_speed;
  a.viscosity = s.viscosity;
  a.courant = s.courant;
  a.max_substeps = s.max_substeps;
  a.desc = desc_d_.ptr;
  a.state = state_d_.ptr;
  a.walls = walls_.ptr;
  a.wall_start = wall_start_.ptr;
  a.cell_start = cell_start_.ptr;
  a.ctl = ctl_;
  a.cur = cur_;
  return a;
}

// Block table for `block` threads per block: entry b = (scene, block of
// the scene), scenes in order, ceil(count / block) blocks each.
int32_t Batch::map_for(uint32_t block, const uint2** map, uint32_t* blocks) {
  auto it = maps_.find(block);
  if (it == maps_.end()) {
    std::vector<uint2> m;
    for (size_t s = 0; s < desc_.size(); ++s) {
      uint32_t nb = (desc_[s].count + block - 1) / block;
      for (uint32_t j = 0; j < nb; ++j) m.push_back(make_uint2((uint32_t)s, j));
    }
    std::unique_ptr<DeviceBuffer<uint2>> buf(new DeviceBuffer<uint2>());
    SFL_HIP(buf->reserve(std::max<size_t>(m.size(), 1), 0, stream_), "hipMalloc(block table)");
    if (!m.empty())
      SFL_HIP(hipMemcpy(buf->ptr, m.data(), m.size() * sizeof(uint2), hipMemcpyHostToDevice), "copy block table");
    it = maps_.emplace(block, std::make_pair(buf.release(), (uint32_t)m.size())).first;
  }
  *map = it->second.first->ptr;
  *blocks = it->second.second;
  return SFL_OK;
}

int32_t Batch::prepare_maps(const BatchShape& shape) {
  const uint2* m;
  uint32_t nb;
  for (uint32_t b : {shape.sort, shape.density, shape.force}) {
    int32_t st = map_for(b, &m, &nb);
    if (st != SFL_OK) return st;
  }
  return SFL_OK;
}

int32_t Batch::ensure_cells(uint64_t volume) {
  if (volume <= cell_cap_ && counts_.ptr) return SFL_OK;
  if (volume > kMaxCells) {
    detail_ = "the scenes spread over more cells than the device cell table holds";
    return SFL_OUT_OF_RANGE;
  }
  uint64_t cap = std::min<uint64_t>(volume + volume / 4 + 4096, kMaxCells);
  drop_graph();
  SFL_HIP(hipStreamSynchronize(stream_), "sync");
  counts_.release();
  cell_start_.release();
  SFL_HIP(counts_.reserve(cap, 0, stream_), "hipMalloc(cell counts)");
  SFL_HIP(hipMemsetAsync(counts_.ptr, 0, cap * sizeof(uint32_t), stream_), "clear cell counts");
  SFL_HIP(cell_start_.reserve(cap + 1, 0, stream_), "hipMalloc(cell table)");
  SFL_HIP(tile_sums_.reserve((cap + 2047) / 2048 + 1, 0, stream_), "hipMalloc(tile sums)");
  SFL_HIP(hipStreamSynchronize(stream_), "sync");
  cell_cap_ = cap;
  return SFL_OK;
}

// One substep slot (wcsph_substep, src/fluid/step.rs:144-224, after the
// update_neighbors sort of src/fluid/mod.rs:540-566).
void Batch::enqueue_substep(const BatchShape& shape) {
  BatchArgs a = args();
  const uint2 *ms, *md, *mf;
  uint32_t ns, nd, nf;
  map_for(shape.sort, &ms, &ns);
  map_for(shape.density, &md, &nd);
  map_for(shape.force, &mf, &nf);
  launch_batch_substep_begin(stream_, a);
  launch_batch_bounds(stream_, a, pa_.ptr, ms, ns, shape.sort);
  launch_batch_grid_setup(stream_, a, counts_.ptr, cell_cap_);
  launch_batch_count(stream_, a, pa_.ptr, counts_.ptr, slot_rank_.ptr, ms, ns, shape.sort);
  launch_scan(stream_, counts_.ptr, cell_start_.ptr, tile_sums_.ptr, scan_total_.ptr, &ctl_->volume, cell_cap_, 1,
              ctl_);
  launch_batch_scatter(stream_, a, pa_.ptr, slot_rank_.ptr, cell_start_.ptr, scatter_.ptr, ms, ns, shape.sort);
  launch_batch_gather(stream_, a, pa_.ptr, va_.ptr, scatter_.ptr, cell_start_.ptr, pb_.ptr, vb_.ptr, ms, ns,
                      shape.sort);
  launch_batch_density(stream_, a, pb_.ptr, vb_.ptr, scaled_.ptr, md, nd, shape.density);
  launch_batch_force(stream_, a, pb_.ptr, vb_.ptr, scaled_.ptr, pa_.ptr, va_.ptr, mf, nf, shape.force);
  launch_batch_substep_end(stream_, a);
}

void Batch::enqueue_steps(uint32_t steps, uint32_t slots, const BatchShape& shape) {
  BatchArgs a = args();
  for (uint32_t t = 0; t < steps; ++t) {
    launch_batch_plan(stream_, a);
    for (uint32_t k = 0; k < slots; ++k) enqueue_substep(shape);
    launch_batch_step_end(stream_, a);
  }
}

void Batch::drop_graph() {
  for (auto& g : graphs_) {
    if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
    if (g.second.graph) (void)hipGraphDestroy(g.second.graph);
  }
  graphs_.clear();
  graph_nodes_ = 0;
}

// Steps per captured sequence: the whole call, or fewer when that would
// make a very large graph (the call then launches it several times).
uint32_t Batch::chunk_for(uint32_t steps) const {
  uint64_t per_step = 2 + (uint64_t)slots_ * kSlotLaunches;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(steps, kGraphLaunches / per_step));
}

int32_t Batch::graph_for(uint32_t steps, uint32_t slots, hipGraphExec_t* exec) {
  auto key = std::make_pair(steps, slots);
  auto it = graphs_.find(key);
  if (it == graphs_.end()) {
    int32_t st = prepare_maps(shape_);
    if (st != SFL_OK) return st;
    Graph g;
    SFL_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    enqueue_steps(steps, slots, shape_);
    hipError_t e = hipStreamEndCapture(stream_, &g.graph);
    if (e != hipSuccess) return device_error("hipStreamEndCapture", e);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    size_t nodes = 0;
    if (e == hipSuccess) e = hipGraphGetNodes(g.graph, nullptr, &nodes);
    if (e != hipSuccess) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      (void)hipGraphDestroy(g.graph);
      return device_error("graph capture", e);
    }
    g.nodes = nodes;
    it = graphs_.emplace(key, g).first;
  }
  graph_nodes_ = it->second.nodes;
  *exec = it->second.exec;
  return SFL_OK;
}

// Captures the sequence a call of `steps` steps will launch, so that the
// call itself only launches it.
int32_t Batch::prepare(uint32_t steps) {
  DeviceGuard g(device_);
  if (!graphs_enabled_ || steps == 0 || n_ == 0 || slots_ > kMaxGraphSlots) return SFL_OK;
  hipGraphExec_t exec;
  return graph_for(chunk_for(steps), slots_, &exec);
}

int32_t Batch::read_back(bool states) {
  SFL_HIP(hipMemcpyAsync(ctl_host_, ctl_, offsetof(Control, bar_count), hipMemcpyDeviceToHost, stream_),
          "read control");
  SFL_HIP(hipMemcpyAsync(cur_host_, cur_, sizeof(BatchCursor), hipMemcpyDeviceToHost, stream_), "read cursor");
  if (states && !desc_.empty())
    SFL_HIP(hipMemcpyAsync(state_host_.data(), state_d_.ptr, desc_.size() * sizeof(SceneState),
                           hipMemcpyDeviceToHost, stream_),
            "read scene state");
  SFL_HIP(hipStreamSynchronize(stream_), "batch step");
  return SFL_OK;
}

// `steps` calls of wcsph (src/fluid/step.rs:102-142) for every scene.
int32_t Batch::step(float dt, uint32_t steps, sfl_report* reports) {
  DeviceGuard g(device_);
  size_t S = desc_.size();
  if (!std::isfinite(dt) || dt < 0.0f) {
    detail_ = "fluid step";
    return SFL_NON_FINITE;
  }
  if (dt == 0.0f || steps == 0 || n_ == 0) {
    if (reports) std::memset(reports, 0, S * sizeof(sfl_report));
    return SFL_OK;
  }
  int32_t st = prepare_maps(shape_);
  if (st != SFL_OK) return st;
  bool fresh = true;
  uint32_t left = steps, progress = 0xffffffffu;
  int stalls = 0;
  for (;;) {
    bool graph = graphs_enabled_ && slots_ <= kMaxGraphSlots;
    uint32_t chunk = 0;
    hipGraphExec_t exec = nullptr;
    if (graph) {
      chunk = chunk_for(steps);
      st = graph_for(chunk, slots_, &exec);
      if (st != SFL_OK) return st;
    }
    launch_batch_reset(stream_, args(), fresh ? 1 : 0, steps, dt, slots_);
    fresh = false;
    if (graph) {
      for (uint32_t l = 0; l < (left + chunk - 1) / chunk; ++l)
        SFL_HIP(hipGraphLaunch(exec, stream_), "hipGraphLaunch");
    } else {
      enqueue_steps(left, slots_, shape_);
    }
    SFL_HIP(hipGetLastError(), "launch");
    st = read_back(false);
    if (st != SFL_OK) return st;
    uint32_t flags = ctl_host_->flags;
    uint32_t now = cur_host_->step, target = cur_host_->target;
    if (flags & kFlagOverflow) {
      st = ensure_cells(ctl_host_->need_volume);
      if (st != SFL_OK) return st;
    } else if (flags & kFlagSlots) {
      slots_ = std::max(slots_, cur_host_->max_sub);
    } else if ((flags & kFlagNonFinite) || now >= target) {
      break;
    } else if (now == progress && ++stalls > 2) {
      detail_ = "the batch's launch sequence made no progress";
      return SFL_DEVICE_ERROR;
    }
    progress = now;
    left = target > now ? target - now : 0;
  }
  st = read_back(true);
  if (st != SFL_OK) return st;
  if (reports)
    for (size_t s = 0; s < S; ++s) {
      sfl_report& r = reports[s];
      std::memset(&r, 0, sizeof(r));
      if (desc_[s].count == 0) continue;
      const SceneState& x = state_host_[s];
      r.substeps = x.substeps;
      r.capped = x.capped;
      std::memcpy(&r.max_compression, &x.comp_bits, 4);
      std::memcpy(&r.max_speed, &x.vmax_bits, 4);
      r.held_back = x.held;
      r.sound_speed = settings_.sound_speed;
    }
  if (ctl_host_->flags & kFlagNonFinite) {
    detail_ = "fluid particle";
    return SFL_NON_FINITE;
  }
  return SFL_OK;
}

int32_t Batch::autotune(float dt, uint32_t repeats, sfl_batch_tuning* out) {
  DeviceGuard g(device_);
  if (!std::isfinite(dt) || dt <= 0.0f) {
    detail_ = "fluid step";
    return SFL_NON_FINITE;
  }
  repeats = std::max<uint32_t>(repeats, 1);
  sfl_batch_tuning t{};
  t.candidates = (uint32_t)(sizeof(kBatchBlocks) / sizeof(kBatchBlocks[0]));
  for (uint32_t c = 0; c < t.candidates; ++c) t.blocks[c] = kBatchBlocks[c];
  if (n_ == 0) {
    t.sort_block = shape_.sort;
    t.density_block = shape_.density;
    t.force_block = shape_.force;
    tuning_ = t;
    if (out) *out = t;
    return SFL_OK;
  }
  size_t S = desc_.size();
  // the persistent state the timed launches touch: scene state and cursor
  DeviceBuffer<SceneState> saved;
  DeviceBuffer<BatchCursor> saved_cur;
  DeviceBuffer<float4> tpa, tva;
  SFL_HIP(saved.reserve(S, 0, stream_), "hipMalloc(tuning)");
  SFL_HIP(saved_cur.reserve(1, 0, stream_), "hipMalloc(tuning)");
  SFL_HIP(tpa.reserve(n_, 0, stream_), "hipMalloc(tuning)");
  SFL_HIP(tva.reserve(n_, 0, stream_), "hipMalloc(tuning)");
  SFL_HIP(hipMemcpyAsync(saved.ptr, state_d_.ptr, S * sizeof(SceneState), hipMemcpyDeviceToDevice, stream_), "save");
  SFL_HIP(hipMemcpyAsync(saved_cur.ptr, cur_, sizeof(BatchCursor), hipMemcpyDeviceToDevice, stream_), "save");
  for (uint32_t b : kBatchBlocks) {
    const uint2* m;
    uint32_t nb;
    int32_t st = map_for(b, &m, &nb);
    if (st != SFL_OK) return st;
  }
  hipEvent_t e0, e1;
  SFL_HIP(hipEventCreate(&e0), "hipEventCreate");
  SFL_HIP(hipEventCreate(&e1), "hipEventCreate");
  struct Events {
    hipEvent_t a, b;
    ~Events() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } guard{e0, e1};
  BatchArgs a = args();
  auto sort_group = [&](uint32_t block) {
    const uint2* m;
    uint32_t nb;
    map_for(block, &m, &nb);
    launch_batch_substep_begin(stream_, a);
    launch_batch_bounds(stream_, a, pa_.ptr, m, nb, block);
    launch_batch_grid_setup(stream_, a, counts_.ptr, cell_cap_);
    launch_batch_count(stream_, a, pa_.ptr, counts_.ptr, slot_rank_.ptr, m, nb, block);
    launch_scan(stream_, counts_.ptr, cell_start_.ptr, tile_sums_.ptr, scan_total_.ptr, &ctl_->volume, cell_cap_, 1,
                ctl_);
    launch_batch_scatter(stream_, a, pa_.ptr, slot_rank_.ptr, cell_start_.ptr, scatter_.ptr, m, nb, block);
    launch_batch_gather(stream_, a, pa_.ptr, va_.ptr, scatter_.ptr, cell_start_.ptr, pb_.ptr, vb_.ptr, m, nb, block);
  };
  auto density = [&](uint32_t block) {
    const uint2* m;
    uint32_t nb;
    map_for(block, &m, &nb);
    launch_batch_density(stream_, a, pb_.ptr, vb_.ptr, scaled_.ptr, m, nb, block);
  };
  auto force = [&](uint32_t block) {
    const uint2* m;
    uint32_t nb;
    map_for(block, &m, &nb);
    launch_batch_force(stream_, a, pb_.ptr, vb_.ptr, scaled_.ptr, tpa.ptr, tva.ptr, m, nb, block);
  };
  auto best_of = [&](auto&& run, uint32_t block, float* ms) -> int32_t {
    float best = std::numeric_limits<float>::infinity();
    for (uint32_t r = 0; r < repeats; ++r) {
      SFL_HIP(hipEventRecord(e0, stream_), "hipEventRecord");
      run(block);
      SFL_HIP(hipEventRecord(e1, stream_), "hipEventRecord");
      SFL_HIP(hipEventSynchronize(e1), "tuning run");
      float x = 0.0f;
      SFL_HIP(hipEventElapsedTime(&x, e0, e1), "hipEventElapsedTime");
      best = std::min(best, x);
    }
    *ms = best;
    return SFL_OK;
  };
  // one planned substep with a live slot; the cell table grown to fit
  for (int attempt = 0;; ++attempt) {
    launch_batch_reset(stream_, a, 1, 1, dt, 0xffffffffu);
    launch_batch_plan(stream_, a);
    sort_group(shape_.sort);
    int32_t st = read_back(false);
    if (st != SFL_OK) return st;
    if (!(ctl_host_->flags & kFlagOverflow)) break;
    if (attempt > 2) {
      detail_ = "the cell table did not settle while tuning";
      return SFL_DEVICE_ERROR;
    }
    SFL_HIP(hipMemcpyAsync(cur_, saved_cur.ptr, sizeof(BatchCursor), hipMemcpyDeviceToDevice, stream_), "restore");
    st = ensure_cells(ctl_host_->need_volume);
    if (st != SFL_OK) return st;
    a = args();
  }
  BatchShape best = shape_;
  float bs = 0, bd = 0, bf = 0;
  for (uint32_t c = 0; c < t.candidates; ++c) {
    uint32_t block = kBatchBlocks[c];
    int32_t st = best_of(sort_group, block, &t.sort_ms[c]);
    if (st == SFL_OK) st = best_of(density, block, &t.density_ms[c]);
    if (st == SFL_OK) st = best_of(force, block, &t.force_ms[c]);
    if (st != SFL_OK) return st;
    if (c == 0 || t.sort_ms[c] < bs) bs = t.sort_ms[c], best.sort = block;
    if (c == 0 || t.density_ms[c] < bd) bd = t.density_ms[c], best.density = block;
    if (c == 0 || t.force_ms[c] < bf) bf = t.force_ms[c], best.force = block;
  }
  SFL_HIP(hipGetLastError(), "tuning launch");
  SFL_HIP(hipMemcpyAsync(state_d_.ptr, saved.ptr, S * sizeof(SceneState), hipMemcpyDeviceToDevice, stream_),
          "restore");
  SFL_HIP(hipMemcpyAsync(cur_, saved_cur.ptr, sizeof(BatchCursor), hipMemcpyDeviceToDevice, stream_), "restore");
  SFL_HIP(hipStreamSynchronize(stream_), "tuning");
  t.sort_block = best.sort;
  t.density_block = best.density;
  t.force_block = best.force;
  if (!(best == shape_)) drop_graph();
  shape_ = best;
  tuning_ = t;
  if (out) *out = t;
  return SFL_OK;
}

int32_t Batch::set_tuning(const sfl_batch_tuning& t) {
  auto ok = [](uint32_t b) {
    for (uint32_t c : kBatchBlocks)
      if (c == b) return true;
    return false;
  };
  if (!ok(t.sort_block) || !ok(t.density_block) || !ok(t.force_block)) return SFL_OUT_OF_RANGE;
  BatchShape s;
  s.sort = t.sort_block;
  s.density = t.density_block;
  s.force = t.force_block;
  if (!(s == shape_)) drop_graph();
  shape_ = s;
  tuning_.sort_block = s.sort;
  tuning_.density_block = s.density;
  tuning_.force_block = s.force;
  return SFL_OK;
}

int32_t Batch::get_tuning(sfl_batch_tuning* out) const {
  *out = tuning_;
  out->sort_block = shape_.sort;
  out->density_block = shape_.density;
  out->force_block = shape_.force;
  return SFL_OK;
}

int32_t Batch::use_graph(bool on) {
  graphs_enabled_ = on;
  if (!on) drop_graph();
  return SFL_OK;
}

int32_t Batch::read_state(size_t scene, size_t cap, uint32_t* ids, float* xyz, float* vxyz, float* rho) {
  DeviceGuard g(device_);
  if (scene >= desc_.size()) return SFL_OUT_OF_RANGE;
  const SceneDesc& d = desc_[scene];
  size_t n = std::min<size_t>(cap, d.count);
  if (!n) return SFL_OK;
  std::vector<float4> p(n), v(n);
  SFL_HIP(hipMemcpyAsync(p.data(), pa_.ptr + d.first, n * sizeof(float4), hipMemcpyDeviceToHost, stream_),
          "read scene");
  SFL_HIP(hipMemcpyAsync(v.data(), va_.ptr + d.first, n * sizeof(float4), hipMemcpyDeviceToHost, stream_),
          "read scene");
  SFL_HIP(hipStreamSynchronize(stream_), "read scene");
  for (size_t i = 0; i < n; ++i) {
    if (ids) std::memcpy(ids + i, &p[i].w, 4);
    if (xyz) {
      xyz[3 * i] = p[i].x;
      xyz[3 * i + 1] = p[i].y;
      xyz[3 * i + 2] = p[i].z;
    }
    if (vxyz) {
      vxyz[3 * i] = v[i].x;
      vxyz[3 * i + 1] = v[i].y;
      vxyz[3 * i + 2] = v[i].z;
    }
    if (rho) rho[i] = v[i].w;
  }
  return SFL_OK;
}

int32_t Batch::state_hash(size_t scene, uint64_t* out) {
  DeviceGuard g(device_);
  if (scene >= desc_.size()) return SFL_OUT_OF_RANGE;
  const SceneDesc& d = desc_[scene];
  std::vector<float4> p(d.count), v(d.count);
  if (d.count) {
    SFL_HIP(hipMemcpyAsync(p.data(), pa_.ptr + d.first, d.count * sizeof(float4), hipMemcpyDeviceToHost, stream_),
            "read scene");
    SFL_HIP(hipMemcpyAsync(v.data(), va_.ptr + d.first, d.count * sizeof(float4), hipMemcpyDeviceToHost, stream_),
            "read scene");
    SFL_HIP(hipStreamSynchronize(stream_), "read scene");
  }
  *out = hash_particles(p.data(), v.data(), d.count);
  return SFL_OK;
}

int32_t Batch::scene_failed(size_t scene, int32_t* failed) const {
  if (scene >= desc_.size()) return SFL_OUT_OF_RANGE;
  *failed = state_host_[scene].nonfinite ? 1 : 0;
  return SFL_OK;
}

// Hands every scene back to a world: positions, velocities and densities
// in the scene's current order, and the fastest speed for the next plan.
int32_t Batch::store(World* const* worlds, size_t count) {
  DeviceGuard g(device_);
  if (!worlds || count != desc_.size()) return SFL_OUT_OF_RANGE;
  for (size_t s = 0; s < count; ++s) {
    if (!worlds[s]) return SFL_INVALID_HANDLE;
    if (worlds[s]->n_ != desc_[s].count) return SFL_OUT_OF_RANGE;
    if (worlds[s]->settings_.solver != SFL_WCSPH || !same_physics(worlds[s]->settings_, settings_))
      return SFL_INVALID_SETTINGS;
  }
  int32_t st = read_back(true);
  if (st != SFL_OK) return st;
  for (size_t s = 0; s < count; ++s) {
    World& w = *worlds[s];
    const SceneDesc& d = desc_[s];
    st = w.synchronize();
    if (st != SFL_OK) return st;
    if (d.count) {
      SFL_HIP(hipMemcpyAsync(w.pa_.ptr, pa_.ptr + d.first, d.count * sizeof(float4), hipMemcpyDeviceToDevice,
                             stream_),
              "store positions");
      SFL_HIP(hipMemcpyAsync(w.va_.ptr, va_.ptr + d.first, d.count * sizeof(float4), hipMemcpyDeviceToDevice,
                             stream_),
              "store velocities");
    }
    std::memcpy(&w.vmax_, &state_host_[s].vmax_bits, 4);
    w.vmax_valid_ = true;
    w.forces_valid_ = false;
    ++w.version_;
    w.invalidate_lists();
  }
  SFL_HIP(hipStreamSynchronize(stream_), "store");
  return SFL_OK;
}

int32_t Batch::synchronize() {
  DeviceGuard g(device_);
  SFL_HIP(hipStreamSynchronize(stream_), "synchronize");
  return SFL_OK;
}

size_t Batch::device_bytes() const {
  auto b = [](const auto& buf) { return buf.cap * sizeof(*buf.ptr); };
  size_t maps = 0;
  for (auto& m : maps_) maps += b(*m.second.first);
  return sizeof(Control) + sizeof(BatchCursor) + b(pa_) + b(va_) + b(pb_) + b(vb_) + b(walls_) + b(scaled_) +
         b(slot_rank_) + b(scatter_) + b(counts_) + b(cell_start_) + b(tile_sums_) + b(scan_total_) +
         b(wall_start_) + b(desc_d_) + b(state_d_) + maps;
}

}  // namespace sfl
