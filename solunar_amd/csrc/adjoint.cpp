// Host side of the WCSPH adjoint (adjoint.hpp).
#include "adjoint.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "launch.hpp"

namespace sfl {

namespace {

struct DeviceGuard {
  explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

}  // namespace

int32_t Tape::device_error(const char* what, hipError_t e) {
  detail_ = std::string(what) + ": " + hipGetErrorString(e);
  return SFL_DEVICE_ERROR;
}

#define SFL_HIP(call, what)                                \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return device_error(what, e_);   \
  } while (0)

Tape::~Tape() {
  if (w_) {
    DeviceGuard g(w_->device_);
    if (w_->stream_) (void)hipStreamSynchronize(w_->stream_);
    if (replay_stream_) (void)hipStreamSynchronize(replay_stream_);
  }
  if (ctl_) (void)hipFree(ctl_);
  if (ctl_host_) (void)hipHostFree(ctl_host_);
  if (ready_) (void)hipEventDestroy(ready_);
  for (hipEvent_t e : reversed_)
    if (e) (void)hipEventDestroy(e);
  if (replay_stream_) (void)hipStreamDestroy(replay_stream_);
}

// The replay stream (default priority, like the world's) and the events that
// order it against the world's stream.
int32_t Tape::ensure_streams() {
  if (!replay_stream_) SFL_HIP(hipStreamCreateWithFlags(&replay_stream_, hipStreamNonBlocking), "replay stream");
  if (!ready_) SFL_HIP(hipEventCreateWithFlags(&ready_, hipEventDisableTiming), "replay event");
  for (hipEvent_t& e : reversed_)
    if (!e) SFL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "replay event");
  return SFL_OK;
}

// `steps` calls of FluidWorld::step (src/fluid/step.rs:58-90) on the world,
// keeping (x, v) before steps 0, every, 2 every, ...
int32_t Tape::record(World* w, float dt, uint32_t steps, uint32_t every, Tape** out, std::string& why) {
  *out = nullptr;
  if (w->settings_.solver != SFL_WCSPH) {
    why = "the adjoint differentiates WCSPH steps";
    return SFL_INVALID_SETTINGS;
  }
  if (!std::isfinite(dt) || dt <= 0.0f) {
    why = "fluid step";
    return SFL_NON_FINITE;
  }
  if (steps == 0 || every == 0) {
    why = "a tape needs at least one step and a checkpoint interval of at least one";
    return SFL_OUT_OF_RANGE;
  }
  if (w->n_ == 0 || w->next_id_ != w->n_) {
    why = "the adjoint needs particles with ids 0..len-1";
    return SFL_OUT_OF_RANGE;
  }
  if (!w->flows_.empty() || w->has_domain_) {
    why = "the adjoint differentiates steps that keep the particles: no flows or domain";
    return SFL_UNSUPPORTED;
  }
  if (!w->terrains_.empty()) {
    why = "the adjoint's reverse pass has no ground terms";
    return SFL_UNSUPPORTED;
  }
  std::unique_ptr<Tape> t(new Tape());
  t->w_ = w;
  t->dt_ = dt;
  t->steps_ = steps;
  t->every_ = every;
  t->n_ = w->n_;
  DeviceGuard g(w->device_);
  hipStream_t s = w->stream_;
  uint32_t segments = (steps + every - 1) / every;
  for (uint32_t c = 0; c < segments; ++c) {
    t->ckp_.emplace_back(new DeviceBuffer<float4>());
    t->ckv_.emplace_back(new DeviceBuffer<float4>());
    hipError_t e = t->ckp_.back()->reserve(t->n_, 0, s);
    if (e == hipSuccess) e = t->ckv_.back()->reserve(t->n_, 0, s);
    if (e != hipSuccess) {
      why = std::string("hipMalloc(checkpoints): ") + hipGetErrorString(e);
      return SFL_DEVICE_ERROR;
    }
  }
  for (uint32_t k = 0; k < steps; ++k) {
    if (k % every == 0) {
      uint32_t c = k / every;
      hipError_t e = hipMemcpyAsync(t->ckp_[c]->ptr, w->pa_.ptr, t->n_ * sizeof(float4), hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess)
        e = hipMemcpyAsync(t->ckv_[c]->ptr, w->va_.ptr, t->n_ * sizeof(float4), hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) {
        why = std::string("checkpoint copy: ") + hipGetErrorString(e);
        return SFL_DEVICE_ERROR;
      }
    }
    sfl_report r{};
    int32_t st = w->step(dt, &r);
    if (st != SFL_OK) {
      why = w->detail_;
      return st;
    }
    if (w->searched_) {
      why = "the tape replays its segments on dense cell tables; the water spread onto a searched one";
      return SFL_UNSUPPORTED;
    }
    t->substeps_.push_back(r.substeps);
    t->small_.push_back(dt / (float)r.substeps);  // wcsph_steps (step.rs:131)
  }
  t->version_ = w->version_;
  *out = t.release();
  return SFL_OK;
}

uint64_t Tape::substeps() const {
  uint64_t total = 0;
  for (uint32_t k : substeps_) total += k;
  return total;
}

size_t Tape::device_bytes() const {
  auto b = [](const auto& buf) { return buf.cap * sizeof(*buf.ptr); };
  size_t total = sizeof(Control);
  for (auto& c : ckp_) total += b(*c);
  for (auto& c : ckv_) total += b(*c);
  for (const auto& set : slots_)
    for (auto& sl : set) total += b(sl->P) + b(sl->V) + b(sl->scaled) + b(sl->clamps) + b(sl->src) + b(sl->cells);
  return total + b(boxes_[0]) + b(boxes_[1]) + b(ra_p_) + b(ra_v_) + b(xb_) + b(vb_) + b(xb2_) + b(vb2_) + b(ab_) + b(rb_) +
         b(bar_in_) + b(bar_out_) + b(nu_acc_) + b(counts_) + b(tile_sums_) + b(scan_total_) + b(slot_rank_) +
         b(scatter_);
}

// Slot set `set` (written only by the replay stream) for `count` substeps.
// Growing a buffer frees the old one at once, on the host's clock, while the
// reverse kernels of an earlier segment may still read this set: growth
// waits for the world's stream first.
int32_t Tape::ensure_slots(uint32_t set, uint32_t count) {
  hipStream_t s = replay_stream_;
  std::vector<std::unique_ptr<Slot>>& slots = slots_[set];
  bool grows = boxes_[set].ptr && boxes_[set].cap < count;
  for (size_t k = 0; k < slots.size() && k < count; ++k) grows |= slots[k]->cells.ptr && slots[k]->cells.cap < cell_cap_ + 1;
  if (grows) SFL_HIP(hipStreamSynchronize(w_->stream_), "replay");
  while (slots.size() < count) slots.emplace_back(new Slot());
  for (uint32_t k = 0; k < count; ++k) {
    Slot& sl = *slots[k];
    SFL_HIP(sl.P.reserve(n_, 0, s), "hipMalloc(tape)");
    SFL_HIP(sl.V.reserve(n_, 0, s), "hipMalloc(tape)");
    SFL_HIP(sl.scaled.reserve(n_, 0, s), "hipMalloc(tape)");
    SFL_HIP(sl.clamps.reserve(n_, 0, s), "hipMalloc(tape)");
    SFL_HIP(sl.src.reserve(n_, 0, s), "hipMalloc(tape)");
    if (sl.cells.cap < cell_cap_ + 1) {
      sl.cells.release();
      SFL_HIP(sl.cells.reserve(cell_cap_ + 1, 0, s), "hipMalloc(tape cells)");
    }
  }
  SFL_HIP(boxes_[set].reserve(std::max<uint32_t>(count, 1), 0, s), "hipMalloc(tape boxes)");
  return SFL_OK;
}

// Replays segment c from its checkpoint with the world's kernels on the
// replay stream, keeping every substep in a slot of set `set`; `count` gets
// the number of substeps. Waits for the replay (its check reads the result).
int32_t Tape::replay(uint32_t c, uint32_t set, uint32_t* count) {
  World& w = *w_;
  hipStream_t s = replay_stream_;
  uint32_t first = c * every_, last = std::min(steps_, first + every_);
  uint32_t total = 0;
  for (uint32_t k = first; k < last; ++k) total += substeps_[k];
  *count = total;
  uint32_t n = (uint32_t)n_;
  for (int attempt = 0;; ++attempt) {
    int32_t st = ensure_slots(set, total);
    if (st != SFL_OK) return st;
    SFL_HIP(hipMemcpyAsync(ra_p_.ptr, ckp_[c]->ptr, n_ * sizeof(float4), hipMemcpyDeviceToDevice, s), "replay");
    SFL_HIP(hipMemcpyAsync(ra_v_.ptr, ckv_[c]->ptr, n_ * sizeof(float4), hipMemcpyDeviceToDevice, s), "replay");
    launch_step_begin(s, ctl_);
    uint32_t slot = 0;
    for (uint32_t k = first; k < last; ++k) {
      StepParams prm = w.params(small_[k]);
      prm.ctl = ctl_;
      for (uint32_t j = 0; j < substeps_[k]; ++j, ++slot) {
        Slot& sl = *slots_[set][slot];
        prm.cell_start = sl.cells.ptr;
        prm.cell_keys = nullptr;  // the tape's tables are dense
        launch_substep_begin(s, ctl_);
        launch_bounds(s, ra_p_.ptr, n, w.radius_, ctl_);
        launch_grid_setup(s, ctl_, n, cell_cap_);
        launch_count(s, ra_p_.ptr, n, w.radius_, counts_.ptr, slot_rank_.ptr, ctl_);
        launch_scan(s, counts_.ptr, sl.cells.ptr, tile_sums_.ptr, scan_total_.ptr, &ctl_->volume, cell_cap_, 1, ctl_);
        launch_scatter(s, ra_p_.ptr, n, slot_rank_.ptr, sl.cells.ptr, scatter_.ptr, ctl_);
        launch_rank_gather(s, ra_p_.ptr, ra_v_.ptr, n, w.radius_, scatter_.ptr, sl.cells.ptr, sl.P.ptr, sl.V.ptr, ctl_,
                           sl.src.ptr);
        launch_adj_keep_box(s, ctl_, boxes_[set].ptr + slot);
        launch_wcsph_density(s, sl.P.ptr, sl.V.ptr, sl.scaled.ptr, prm);
        launch_wcsph_force(s, sl.P.ptr, sl.V.ptr, sl.scaled.ptr, prm, ra_p_.ptr, ra_v_.ptr, nullptr, sl.clamps.ptr);
      }
    }
    // the replay must end where the recorded run was at the next checkpoint
    // (or at its end): bit for bit
    bool at_end = c + 1 == ckp_.size();
    SFL_HIP(hipMemsetAsync(mismatch_.ptr, 0, sizeof(uint32_t), s), "clear");
    launch_adj_compare(s, n, ra_p_.ptr, ra_v_.ptr, at_end ? w.pa_.ptr : ckp_[c + 1]->ptr,
                       at_end ? w.va_.ptr : ckv_[c + 1]->ptr, mismatch_.ptr);
    SFL_HIP(hipGetLastError(), "replay launch");
    uint32_t mismatched = 0;
    SFL_HIP(hipMemcpyAsync(ctl_host_, ctl_, offsetof(Control, bar_count), hipMemcpyDeviceToHost, s), "replay");
    SFL_HIP(hipMemcpyAsync(&mismatched, mismatch_.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "replay");
    SFL_HIP(hipStreamSynchronize(s), "replay");
    uint32_t flags = ctl_host_->flags;
    if (flags & kFlagNonFinite) {
      detail_ = "fluid particle";
      return SFL_NON_FINITE;
    }
    if (!(flags & kFlagOverflow)) {
      if (mismatched) {
        detail_ = "the replay of a tape segment differs from the recorded run";
        return SFL_DEVICE_ERROR;
      }
      return SFL_OK;
    }
    if (attempt > 4 || ctl_host_->need_volume > (uint64_t(1) << 31)) {
      detail_ = "the replay's cell table did not settle";
      return SFL_OUT_OF_RANGE;
    }
    uint64_t want = ctl_host_->need_volume;
    cell_cap_ = want + want / 4 + 4096;
    counts_.release();
    SFL_HIP(counts_.reserve(cell_cap_, 0, s), "hipMalloc(cell counts)");
    SFL_HIP(hipMemsetAsync(counts_.ptr, 0, cell_cap_ * sizeof(uint32_t), s), "clear cell counts");
    SFL_HIP(tile_sums_.reserve((cell_cap_ + 2047) / 2048 + 1, 0, s), "hipMalloc(tile sums)");
  }
}

// The adjoint kernels over segment c's substeps (slot set `which`), last to
// first, on the world's stream. On entry xb_/vb_ hold the adjoints of the
// segment's final state (in its order); on exit, of its first state.
int32_t Tape::reverse(uint32_t c, uint32_t which, uint32_t count) {
  World& w = *w_;
  hipStream_t s = w.stream_;
  const sfl_settings& set = w.settings_;
  uint32_t first = c * every_, last = std::min(steps_, first + every_);
  // the step of each slot
  std::vector<uint32_t> step_of;
  for (uint32_t k = first; k < last; ++k)
    for (uint32_t j = 0; j < substeps_[k]; ++j) step_of.push_back(k);
  for (int slot = (int)count - 1; slot >= 0; --slot) {
    Slot& sl = *slots_[which][slot];
    float small = small_[step_of[slot]];
    StepParams prm = w.params(small);
    float r = w.radius_, cs = set.sound_speed;
    float free = set.viscosity * cs * r, cap = 0.05f * r * r / small;  // viscosity (step.rs:254-259)
    AdjointArgs a{};
    a.n = (uint32_t)n_;
    a.radius = r;
    a.reach = w.reach_;
    a.scale = w.scale_;
    a.mass = w.mass_;
    a.rest = set.rest_density;
    a.stiffness = prm.stiffness;
    a.soften = prm.soften;
    a.dt = small;
    a.nu = prm.nu;
    a.wall_nu = prm.wall_nu;
    a.dnu_dalpha = free < cap ? cs * r : 0.0f;  // fmin picks `free` strictly below the cap
    a.P = sl.P.ptr;
    a.V = sl.V.ptr;
    a.scaled = sl.scaled.ptr;
    a.clamps = sl.clamps.ptr;
    a.src = sl.src.ptr;
    a.cell_start = sl.cells.ptr;
    a.box = boxes_[which].ptr + slot;
    a.walls = prm.walls;
    a.wall_start = prm.wall_start;
    a.wall_keys = prm.wall_keys;
    a.wall_keyed = prm.wall_keyed;
    a.wall_box = prm.wall_box;
    a.wall_cells = prm.wall_cells;
    launch_adj_advance(s, a, xb_.ptr, vb_.ptr, xb2_.ptr, vb2_.ptr, ab_.ptr);
    launch_adj_force(s, a, ab_.ptr, xb2_.ptr, vb2_.ptr, rb_.ptr, nu_acc_.ptr);
    launch_adj_density(s, a, rb_.ptr, xb2_.ptr, vb2_.ptr, xb_.ptr, vb_.ptr);
  }
  SFL_HIP(hipGetLastError(), "adjoint launch");
  return SFL_OK;
}

// Reverse pass: adjoints of the final x, v (per id) in; adjoints of the
// recorded initial x, v (per id) and of the viscosity setting out.
int32_t Tape::adjoint(const float* xbar, const float* vbar, float* xbar0, float* vbar0, float* viscosity_bar) {
  World& w = *w_;
  DeviceGuard g(w.device_);
  if (w.version_ != version_ || w.n_ != n_) {
    detail_ = "the world changed after the tape was recorded";
    return SFL_OUT_OF_RANGE;
  }
  hipStream_t s = w.stream_;
  if (!ctl_) {
    SFL_HIP(hipMalloc(reinterpret_cast<void**>(&ctl_), sizeof(Control)), "hipMalloc(control)");
    SFL_HIP(hipMemset(ctl_, 0, sizeof(Control)), "hipMemset(control)");
    SFL_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctl_host_), sizeof(Control), hipHostMallocDefault),
            "hipHostMalloc(control)");
  }
  if (!counts_.ptr) {
    cell_cap_ = std::max<uint64_t>(w.cell_cap_, 4096);
    SFL_HIP(counts_.reserve(cell_cap_, 0, s), "hipMalloc(cell counts)");
    SFL_HIP(hipMemsetAsync(counts_.ptr, 0, cell_cap_ * sizeof(uint32_t), s), "clear cell counts");
    SFL_HIP(tile_sums_.reserve((cell_cap_ + 2047) / 2048 + 1, 0, s), "hipMalloc(tile sums)");
    SFL_HIP(scan_total_.reserve(1, 0, s), "hipMalloc(scan total)");
  }
  SFL_HIP(mismatch_.reserve(1, 0, s), "hipMalloc(replay)");
  SFL_HIP(slot_rank_.reserve(n_, 0, s), "hipMalloc(replay)");
  SFL_HIP(scatter_.reserve(n_, 0, s), "hipMalloc(replay)");
  SFL_HIP(ra_p_.reserve(n_, 0, s), "hipMalloc(replay)");
  SFL_HIP(ra_v_.reserve(n_, 0, s), "hipMalloc(replay)");
  for (auto* b : {&xb_, &vb_, &xb2_, &vb2_, &ab_}) SFL_HIP(b->reserve(n_, 0, s), "hipMalloc(adjoints)");
  SFL_HIP(rb_.reserve(n_, 0, s), "hipMalloc(adjoints)");
  SFL_HIP(bar_in_.reserve(6 * n_, 0, s), "hipMalloc(adjoints)");
  SFL_HIP(bar_out_.reserve(6 * n_, 0, s), "hipMalloc(adjoints)");
  uint32_t blocks = adj_force_blocks((uint32_t)n_);
  SFL_HIP(nu_acc_.reserve(blocks, 0, s), "hipMalloc(adjoints)");
  SFL_HIP(hipMemsetAsync(nu_acc_.ptr, 0, blocks * sizeof(double), s), "clear");
  // final adjoints into the world's order
  SFL_HIP(hipMemcpyAsync(bar_in_.ptr, xbar, 3 * n_ * sizeof(float), hipMemcpyHostToDevice, s), "upload");
  SFL_HIP(hipMemcpyAsync(bar_in_.ptr + 3 * n_, vbar, 3 * n_ * sizeof(float), hipMemcpyHostToDevice, s), "upload");
  launch_adj_load(s, (uint32_t)n_, w.pa_.ptr, bar_in_.ptr, bar_in_.ptr + 3 * n_, xb_.ptr, vb_.ptr);
  // Segment c's reverse kernels (world stream, set c % 2) run beside the
  // replay of segment c - 1 (replay stream, the other set). A replay waits
  // for the set-up above and for the reverse that last read its set; a
  // reverse waits for its replay. On an error, both streams drain first.
  int32_t st = ensure_streams();
  if (st != SFL_OK) return st;
  hipStream_t r = replay_stream_;
  auto fail = [&](int32_t code) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamSynchronize(r);
    return code;
  };
  SFL_HIP(hipEventRecord(ready_, s), "replay event");
  SFL_HIP(hipStreamWaitEvent(r, ready_, 0), "replay event");
  const uint32_t segments = (uint32_t)ckp_.size();
  uint32_t count[2] = {0, 0};
  uint32_t c = segments - 1;
  st = replay(c, c % 2, &count[c % 2]);
  if (st != SFL_OK) return fail(st);
  for (;;) {
    const uint32_t set = c % 2;
    SFL_HIP(hipEventRecord(ready_, r), "replay event");
    SFL_HIP(hipStreamWaitEvent(s, ready_, 0), "replay event");
    st = reverse(c, set, count[set]);
    if (st != SFL_OK) return fail(st);
    SFL_HIP(hipEventRecord(reversed_[set], s), "replay event");
    if (c == 0) break;
    --c;
    // set c % 2 was last read by the reverse of segment c + 2
    SFL_HIP(hipStreamWaitEvent(r, reversed_[c % 2], 0), "replay event");
    st = replay(c, c % 2, &count[c % 2]);
    if (st != SFL_OK) return fail(st);
  }
  launch_adj_store(s, (uint32_t)n_, ckp_[0]->ptr, xb_.ptr, vb_.ptr, bar_out_.ptr, bar_out_.ptr + 3 * n_);
  std::vector<double> acc(blocks);
  SFL_HIP(hipMemcpyAsync(xbar0, bar_out_.ptr, 3 * n_ * sizeof(float), hipMemcpyDeviceToHost, s), "download");
  SFL_HIP(hipMemcpyAsync(vbar0, bar_out_.ptr + 3 * n_, 3 * n_ * sizeof(float), hipMemcpyDeviceToHost, s), "download");
  SFL_HIP(hipMemcpyAsync(acc.data(), nu_acc_.ptr, blocks * sizeof(double), hipMemcpyDeviceToHost, s), "download");
  SFL_HIP(hipStreamSynchronize(s), "adjoint");
  double total = 0.0;
  for (double v : acc) total += v;
  if (viscosity_bar) *viscosity_bar = (float)total;
  return SFL_OK;
}

}  // namespace sfl
