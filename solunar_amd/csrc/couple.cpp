// Host side of coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs): the
// shapes sampled into body points and probes, the poses a coupled step
// takes, the launches that place and carry the points through the substeps,
// the fastest body near water, and the lifts. Scalar arithmetic is the
// reference's f32 arithmetic in its order (compiled with -ffp-contract=off);
// the one libm call, the lift's volume^(2/3), is the same `powf` the
// reference's f32::powf calls.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <tuple>

#include "launch.hpp"
#include "world.hpp"

namespace sfl {

namespace {

struct DeviceGuard {
  explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

constexpr size_t kMaxProbes = 4096;  // MAX_PROBES, src/fluid/lift.rs:8
constexpr float kReadReach = 3.0f;    // READ_REACH, src/fluid/lift.rs:12
// Column-table slots all lifts of one step may use.
constexpr uint64_t kMaxColumns = uint64_t(1) << 26;

float fmin_r(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }
float length3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

// `x as i32` of floor(v / radius) (mod.rs:526-533)
int32_t cell_of(float v, float radius) {
  float c = std::floor(v / radius);
  if (std::isnan(c)) return 0;
  if (c >= 2147483648.0f) return std::numeric_limits<int32_t>::max();
  if (c <= -2147483648.0f) return std::numeric_limits<int32_t>::min();
  return (int32_t)c;
}

float from_ordered(int32_t k) {
  int32_t b = k >= 0 ? k : k ^ 0x7fffffff;
  float v;
  std::memcpy(&v, &b, 4);
  return v;
}

// shape_points (src/fluid/couple.rs:443-502): lattice points `spacing`
// apart inside the shape, no deeper than `reach`, centred on it, loops y, z,
// x. A capsule with a negative or non-finite half height is refused as
// Shape::validate refuses it (src/shape.rs:59-62).
int32_t shape_points(const sfl_shape& s, float spacing, float reach, uint64_t limit, std::vector<float>& out) {
  out.clear();
  float half[3];
  switch (s.kind) {
    case SFL_SHAPE_CUBOID:
      half[0] = s.a, half[1] = s.b, half[2] = s.c;
      break;
    case SFL_SHAPE_SPHERE:
      half[0] = half[1] = half[2] = s.a;
      break;
    case SFL_SHAPE_CAPSULE:
      if (!std::isfinite(s.b) || s.b < 0.0f) return SFL_INVALID_SHAPE;
      half[0] = s.a, half[1] = s.b + s.a, half[2] = s.a;
      break;
    default:
      return SFL_INVALID_SHAPE;
  }
  for (float h : half)
    if (!std::isfinite(h)) return SFL_INVALID_SHAPE;
  if (fmin_r(half[0], fmin_r(half[1], half[2])) <= 0.0f) return SFL_INVALID_SHAPE;
  int64_t cells[3];
  for (int a = 0; a < 3; ++a) {
    float f = std::floor(2.0f * half[a] / spacing);
    int64_t c = std::isnan(f) ? 0 : (f >= 9.2e18f ? INT64_MAX : (f <= -9.2e18f ? INT64_MIN : (int64_t)f));
    cells[a] = std::min<int64_t>(std::max<int64_t>(c, 0), 1 << 12);
  }
  if (cells[0] * cells[1] * cells[2] > (int64_t)limit) return SFL_FLUID_FULL;
  float origin[3];
  for (int a = 0; a < 3; ++a) origin[a] = -((float)cells[a] - 1.0f) * spacing / 2.0f;
  for (int64_t y = 0; y < cells[1]; ++y)
    for (int64_t z = 0; z < cells[2]; ++z)
      for (int64_t x = 0; x < cells[0]; ++x) {
        float at[3] = {origin[0] + (float)x * spacing, origin[1] + (float)y * spacing,
                       origin[2] + (float)z * spacing};
        float inside;
        if (s.kind == SFL_SHAPE_CUBOID) {
          inside = fmin_r(fmin_r(half[0] - std::fabs(at[0]), half[1] - std::fabs(at[1])), half[2] - std::fabs(at[2]));
        } else if (s.kind == SFL_SHAPE_SPHERE) {
          inside = s.a - length3(at[0], at[1], at[2]);
        } else {
          float yy = at[1];  // f32::clamp
          if (yy < -s.b) yy = -s.b;
          if (yy > s.b) yy = s.b;
          inside = s.a - length3(at[0], at[1] - yy, at[2]);
        }
        if (inside >= 0.0f && inside <= reach) out.insert(out.end(), at, at + 3);
      }
  return SFL_OK;
}

}  // namespace

#define SFL_HIP(call, what)                          \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return device_error(what, e_); \
  } while (0)

size_t World::coupled_points() const {
  size_t n = 0;
  for (const Coupled& k : coupled_) n += k.points.size() / 3;
  return n;
}

// couple (src/fluid/couple.rs:165-198)
int32_t World::couple(uint32_t index, uint32_t generation, const sfl_shape& shape) {
  std::vector<float> points;
  int32_t st = shape_points(shape, settings_.spacing, radius_, settings_.max_particles, points);
  if (st != SFL_OK) return st;
  if (points.empty()) return SFL_INVALID_SHAPE;
  size_t used = coupled_points();
  size_t room = settings_.max_particles > used ? (size_t)settings_.max_particles - used : 0;
  if (points.size() / 3 > room) return SFL_FLUID_FULL;
  float probe_spacing = settings_.spacing;
  std::vector<float> probes;
  for (;;) {
    st = shape_points(shape, probe_spacing, std::numeric_limits<float>::infinity(), kMaxProbes, probes);
    if (st == SFL_FLUID_FULL) {
      probe_spacing *= 1.25f;
      continue;
    }
    if (st != SFL_OK) return st;
    if (probes.empty()) {
      probe_spacing = settings_.spacing;
      probes = points;
    }
    break;
  }
  ++version_;
  uncouple(index, generation, nullptr);
  coupled_.push_back({index, generation, std::move(points), std::move(probes), probe_spacing});
  std::stable_sort(coupled_.begin(), coupled_.end(), [](const Coupled& a, const Coupled& b) {
    return std::tie(a.index, a.generation) < std::tie(b.index, b.generation);
  });
  coupled_dirty_ = true;
  body_pair_use_ = 0;  // the last step's pair count says nothing of the new set
  return SFL_OK;
}

// uncouple (src/fluid/couple.rs:200-205)
int32_t World::uncouple(uint32_t index, uint32_t generation, int32_t* was) {
  size_t before = coupled_.size();
  coupled_.erase(std::remove_if(coupled_.begin(), coupled_.end(),
                                [&](const Coupled& k) { return k.index == index && k.generation == generation; }),
                 coupled_.end());
  bool gone = coupled_.size() != before;
  if (gone) {
    coupled_dirty_ = true;
    body_pair_use_ = 0;
  }
  if (was) *was = gone ? 1 : 0;
  return SFL_OK;
}

namespace {
const sfl_body_pose* pose_of(const sfl_body_pose* poses, size_t count, uint32_t index, uint32_t generation) {
  for (size_t k = 0; k < count; ++k)
    if (poses[k].index == index && poses[k].generation == generation) return &poses[k];
  return nullptr;
}
}  // namespace

// place_movers_at (src/fluid/couple.rs:258-291): bodies no pose names are
// uncoupled; the others' points placed where their poses put them, then
// carried to 0 (sorted).
int32_t World::place_movers(const sfl_body_pose* poses, size_t count, float dt) {
  size_t before = coupled_.size();
  coupled_.erase(std::remove_if(coupled_.begin(), coupled_.end(),
                                [&](const Coupled& k) { return !pose_of(poses, count, k.index, k.generation); }),
                 coupled_.end());
  if (coupled_.size() != before) {
    coupled_dirty_ = true;
    body_pair_use_ = 0;
  }
  const uint32_t nb = (uint32_t)coupled_.size();
  movers_m_ = 0;
  if (nb == 0) return SFL_OK;
  if (coupled_dirty_) {
    std::vector<float4> local, probes;
    most_probes_ = 0;
    body_reach_.assign(nb, 0.0f);
    for (uint32_t b = 0; b < nb; ++b) {
      const Coupled& k = coupled_[b];
      float owner;
      std::memcpy(&owner, &b, 4);
      for (size_t i = 0; i < k.points.size(); i += 3) {
        local.push_back(make_float4(k.points[i], k.points[i + 1], k.points[i + 2], owner));
        body_reach_[b] = std::max(body_reach_[b], length3(k.points[i], k.points[i + 1], k.points[i + 2]));
      }
      for (size_t i = 0; i < k.probes.size(); i += 3)
        probes.push_back(make_float4(k.probes[i], k.probes[i + 1], k.probes[i + 2], 0.0f));
      most_probes_ = std::max<uint32_t>(most_probes_, (uint32_t)(k.probes.size() / 3));
    }
    SFL_HIP(body_local_.reserve(local.size(), 0, stream_), "hipMalloc(body points)");
    SFL_HIP(probe_local_.reserve(probes.size(), 0, stream_), "hipMalloc(body probes)");
    SFL_HIP(hipMemcpyAsync(body_local_.ptr, local.data(), local.size() * sizeof(float4), hipMemcpyHostToDevice,
                           stream_),
            "upload body points");
    SFL_HIP(hipMemcpyAsync(probe_local_.ptr, probes.data(), probes.size() * sizeof(float4), hipMemcpyHostToDevice,
                           stream_),
            "upload body probes");
    SFL_HIP(hipStreamSynchronize(stream_), "upload body points");
    coupled_dirty_ = false;
  }
  bodies_h_.assign(nb, BodyState{});
  uint32_t points = 0, probes = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    const Coupled& k = coupled_[b];
    const sfl_body_pose& p = *pose_of(poses, count, k.index, k.generation);
    BodyState& s = bodies_h_[b];
    for (int a = 0; a < 3; ++a) {
      s.center[a] = (float)p.position[a];  // position.to_dvec3().as_vec3()
      s.linear[a] = p.linear_velocity[a];
      s.angular[a] = p.angular_velocity[a];
    }
    for (int a = 0; a < 4; ++a) s.rotation[a] = p.rotation[a];
    s.first_point = points;
    s.points = (uint32_t)(k.points.size() / 3);
    s.first_probe = probes;
    s.probes = (uint32_t)(k.probes.size() / 3);
    s.probe_spacing = k.probe_spacing;
    points += s.points;
    probes += s.probes;
  }
  SFL_HIP(bodies_d_.reserve(nb, 0, stream_), "hipMalloc(bodies)");
  // bodies_h_ stays as it is until the next step, which reads back (and so
  // waits on) this stream before it writes it again
  SFL_HIP(hipMemcpyAsync(bodies_d_.ptr, bodies_h_.data(), nb * sizeof(BodyState), hipMemcpyHostToDevice, stream_),
          "upload bodies");
  const uint32_t m = points;
  owner_bits_ = 0;
  while ((uint64_t(1) << owner_bits_) < nb) ++owner_bits_;
  for (int k = 0; k < 2; ++k) {
    SFL_HIP(ms_[k].reserve(m, 0, stream_), "hipMalloc(body point starts)");
    SFL_HIP(mv_[k].reserve(m, 0, stream_), "hipMalloc(body point velocities)");
  }
  SFL_HIP(mp_.reserve(m, 0, stream_), "hipMalloc(body point positions)");
  SFL_HIP(mkeys_in_.reserve(m, 0, stream_), "hipMalloc(body point keys)");
  SFL_HIP(mkeys_.reserve(m, 0, stream_), "hipMalloc(body point keys)");
  SFL_HIP(mperm_in_.reserve(m, 0, stream_), "hipMalloc(body point order)");
  SFL_HIP(mperm_.reserve(m, 0, stream_), "hipMalloc(body point order)");
  SFL_HIP(mgrid_.reserve(1, 0, stream_), "hipMalloc(body point grid)");
  // the dense cell table over the box the points may reach this step: each
  // body's centre, give or take its reach, what it moves in dt and a cell
  {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t b = 0; b < nb; ++b) {
      const BodyState& st = bodies_h_[b];
      double v = std::sqrt((double)st.linear[0] * st.linear[0] + (double)st.linear[1] * st.linear[1] +
                           (double)st.linear[2] * st.linear[2]);
      double w = std::sqrt((double)st.angular[0] * st.angular[0] + (double)st.angular[1] * st.angular[1] +
                           (double)st.angular[2] * st.angular[2]);
      double reach = body_reach_[b] * 1.001 + (v + w * body_reach_[b]) * std::fabs((double)dt) + 2.0 * radius_;
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], (double)st.center[a] - reach);
        hi[a] = std::max(hi[a], (double)st.center[a] + reach);
      }
    }
    double volume = 1.0;
    for (int a = 0; a < 3; ++a) volume *= std::floor(hi[a] / radius_) - std::floor(lo[a] / radius_) + 3.0;
    if (std::isfinite(volume) && volume < (double)(uint64_t(1) << 27))
      SFL_HIP(mtable_.reserve((size_t)volume + 2, 0, stream_), "hipMalloc(body point table)");
  }
  size_t temp = 0;
  SFL_HIP(mover_sort_bytes(m, &temp), "radix sort size");
  SFL_HIP(msort_temp_.reserve(std::max<size_t>(temp, 1), 0, stream_), "hipMalloc(radix sort)");
  mcur_ = 0;
  movers_m_ = m;
  launch_movers_place(stream_, body_local_.ptr, m, bodies_d_.ptr, ms_[0].ptr, mv_[0].ptr);
  // each body's fastest point and box, for fastest_near_water
  SFL_HIP(extent_.reserve(nb, 0, stream_), "hipMalloc(body extents)");
  launch_extent(stream_, ms_[0].ptr, mv_[0].ptr, bodies_d_.ptr, extent_.ptr, nb);
  return carry_movers(0.0f);
}

// carry_movers (src/fluid/couple.rs:293-319) on the stream
int32_t World::carry_movers(float elapsed) {
  MoverCarry c{};
  c.m = movers_m_;
  c.nbodies = (uint32_t)coupled_.size();
  c.owner_bits = owner_bits_;
  c.ms = ms_[mcur_].ptr;
  c.mv = mv_[mcur_].ptr;
  c.ms2 = ms_[1 - mcur_].ptr;
  c.mv2 = mv_[1 - mcur_].ptr;
  c.mp = mp_.ptr;
  c.grid = mgrid_.ptr;
  c.keys_in = mkeys_in_.ptr;
  c.perm_in = mperm_in_.ptr;
  c.keys = mkeys_.ptr;
  c.perm = mperm_.ptr;
  c.temp = msort_temp_.ptr;
  c.temp_bytes = msort_temp_.cap;
  c.bodies = bodies_d_.ptr;
  c.table = mtable_.ptr;
  c.table_cap = mtable_.cap;
  SFL_HIP(launch_movers_carry(stream_, c, elapsed, radius_, ctl_), "radix sort (body points)");
  mcur_ = 1 - mcur_;
  return SFL_OK;
}

// The points' order after `done` substeps of a step whose later launches
// were aborted: placed again and carried through the same elapsed times.
int32_t World::replay_movers(uint32_t done, float small) {
  launch_movers_place(stream_, body_local_.ptr, movers_m_, bodies_d_.ptr, ms_[0].ptr, mv_[0].ptr);
  mcur_ = 0;
  int32_t st = carry_movers(0.0f);
  for (uint32_t k = 0; st == SFL_OK && k < done; ++k) st = carry_movers(small * (float)k);
  return st;
}

MoverView World::mover_view(float dt) const {
  MoverView v{};
  v.pos = mp_.ptr;
  v.vel = mv_[mcur_].ptr;
  v.keys = mkeys_.ptr;
  v.count = movers_m_;
  v.owner_bits = owner_bits_;
  v.grid = mgrid_.ptr;
  v.table = mtable_.ptr;
  v.bodies = bodies_d_.ptr;
  const float c = settings_.sound_speed, r = radius_;
  v.damping = fmin_r(1.0f * c * r, 0.25f * r * r / dt);  // step.rs:163-164
  const sfl_vec3& g = settings_.gravity;
  v.settles = length3(g.x, g.y, g.z) * settings_.spacing;  // couple.rs:387
  v.pairs.counts = pair_counts_b_.ptr;
  v.pairs.offsets = pair_offsets_b_.ptr;
  v.pairs.key = pair_key_[0].ptr;
  v.pairs.force = pair_force_.ptr;
  v.pairs.torque = pair_torque_.ptr;
  v.pairs.none = (uint32_t)coupled_.size();
  return v;
}

int32_t World::ensure_pair_room(uint64_t cap) {
  cap = std::min<uint64_t>(cap, 0xffffffffu);
  if (cap <= body_pair_cap_) return SFL_OK;
  for (int k = 0; k < 2; ++k) {
    SFL_HIP(pair_key_[k].reserve(cap, 0, stream_), "hipMalloc(body pair keys)");
    SFL_HIP(pair_value_[k].reserve(cap, 0, stream_), "hipMalloc(body pair order)");
  }
  SFL_HIP(pair_force_.reserve(cap, 0, stream_), "hipMalloc(body pair forces)");
  SFL_HIP(pair_torque_.reserve(cap, 0, stream_), "hipMalloc(body pair torques)");
  size_t temp = 0;
  SFL_HIP(pair_sort_bytes((uint32_t)cap, &temp), "radix sort size");
  SFL_HIP(pair_temp_.reserve(std::max<size_t>(temp, 1), 0, stream_), "hipMalloc(radix sort)");
  body_pair_cap_ = (uint32_t)cap;
  return SFL_OK;
}

PairRoom World::pair_room() const {
  PairRoom r{};
  r.counts = pair_counts_b_.ptr;
  r.offsets = pair_offsets_b_.ptr;
  r.tiles = pair_tiles_b_.ptr;
  r.total = pair_total_b_.ptr;
  r.cap = body_pair_use_ ? std::min(body_pair_use_, body_pair_cap_) : body_pair_cap_;
  r.none = (uint32_t)coupled_.size();
  r.key_bits = owner_bits_ + 1;
  r.key = pair_key_[0].ptr;
  r.value = pair_value_[0].ptr;
  r.key_sorted = pair_key_[1].ptr;
  r.value_sorted = pair_value_[1].ptr;
  r.force = pair_force_.ptr;
  r.torque = pair_torque_.ptr;
  r.temp = pair_temp_.ptr;
  r.temp_bytes = pair_temp_.cap;
  return r;
}

// fastest_near_water (src/fluid/couple.rs:321-366): the fastest point of
// each body and the box of its points on the device, the rest here.
int32_t World::fastest_near_water(float dt, float slow, float& most) {
  most = 0.0f;
  const uint32_t nb = (uint32_t)coupled_.size();
  if (nb == 0) return SFL_OK;
  SFL_HIP(wet_.reserve(nb, 0, stream_), "hipMalloc(body reach)");
  std::vector<BodyExtent> ext(nb);
  SFL_HIP(hipMemcpyAsync(ext.data(), extent_.ptr, nb * sizeof(BodyExtent), hipMemcpyDeviceToHost, stream_),
          "read body extents");
  SFL_HIP(hipStreamSynchronize(stream_), "body extents");
  const float water = vmax_, r = radius_;
  const bool unsorted = !grid_any_ && n_ > 0;
  std::vector<WetQuery> q;
  for (uint32_t b = 0; b < nb && !unsorted; ++b) {
    float speed;
    std::memcpy(&speed, &ext[b].speed_bits, 4);
    if (!(speed > slow) || !grid_any_) continue;
    if (grid_dropped_) {
      detail_ = "a fast body needs the water's last grid, which an allocation-free step could not keep "
                "(sfl_step or sfl_update_neighbors builds one)";
      return SFL_UNSUPPORTED;
    }
    float reach = r + (speed + water) * dt;
    WetQuery w{b, {}, {}};
    for (int a = 0; a < 3; ++a) {
      w.lo[a] = cell_of(from_ordered(ext[b].lo[a]) - reach, r);
      w.hi[a] = cell_of(from_ordered(ext[b].hi[a]) + reach, r);
    }
    q.push_back(w);
  }
  if (!q.empty()) {
    SFL_HIP(hipMemcpyAsync(wet_.ptr, q.data(), q.size() * sizeof(WetQuery), hipMemcpyHostToDevice, stream_),
            "body reach");
    launch_extent_wet(stream_, searched_ ? sstart_.ptr : cell_start_.ptr, searched_ ? cell_keys_.ptr : nullptr, ctl_,
                      wet_.ptr, (uint32_t)q.size(), extent_.ptr);
    SFL_HIP(hipMemcpyAsync(ext.data(), extent_.ptr, nb * sizeof(BodyExtent), hipMemcpyDeviceToHost, stream_),
            "read body reach");
    SFL_HIP(hipStreamSynchronize(stream_), "body reach");
  }
  for (uint32_t b = 0; b < nb; ++b) {
    float speed;
    std::memcpy(&speed, &ext[b].speed_bits, 4);
    if (speed <= most) continue;
    if (speed <= slow) {
      most = speed;
      continue;
    }
    if (unsorted || ext[b].wet) most = speed;
  }
  return SFL_OK;
}

// lifts (src/fluid/lift.rs:25-152) of every coupled body over the water as
// the step left it, and give_back (lift.rs:154-166) for the dynamic ones.
int32_t World::lifts(float dt, const sfl_body_pose* poses, size_t count, sfl_body_push* pushes) {
  const uint32_t nb = (uint32_t)coupled_.size();
  if (nb == 0 || n_ == 0) return SFL_OK;
  const sfl_vec3& g = settings_.gravity;
  const float gravity = length3(g.x, g.y, g.z);
  float up[3] = {0.0f, 1.0f, 0.0f};
  if (gravity > 0.0f) {
    up[0] = -g.x / gravity;
    up[1] = -g.y / gravity;
    up[2] = -g.z / gravity;
  }
  const float reach = kReadReach * radius_, column = radius_;
  uint32_t total_probes = 0;
  for (const Coupled& k : coupled_) total_probes += (uint32_t)(k.probes.size() / 3);
  SFL_HIP(lift_.reserve(nb, 0, stream_), "hipMalloc(lifts)");
  SFL_HIP(probe_arm_.reserve(total_probes, 0, stream_), "hipMalloc(probe arms)");
  SFL_HIP(probe_vel_.reserve(total_probes, 0, stream_), "hipMalloc(probe velocities)");
  // the water in each body's box: counted per (body, 1024-particle chunk),
  // chunks whose box misses the body's skipped, and scanned into offsets
  const uint64_t chunks = (n_ + 1023) / 1024, cells = (uint64_t)nb * chunks;
  SFL_HIP(lift_boxes_.reserve(2 * chunks, 0, stream_), "hipMalloc(lift chunks)");
  SFL_HIP(lift_counts_.reserve(cells + 1, 0, stream_), "hipMalloc(lift counts)");
  SFL_HIP(lift_offsets_.reserve(cells + 1, 0, stream_), "hipMalloc(lift offsets)");
  SFL_HIP(lift_tiles_.reserve(cells / 2048 + 2, 0, stream_), "hipMalloc(lift offsets)");
  launch_lift_box(stream_, probe_local_.ptr, bodies_d_.ptr, nb, most_probes_, probe_arm_.ptr, probe_vel_.ptr,
                  lift_.ptr, reach);
  launch_lift_count(stream_, pa_.ptr, (uint32_t)n_, lift_.ptr, nb, lift_boxes_.ptr, lift_counts_.ptr,
                    lift_offsets_.ptr, lift_tiles_.ptr, lift_tiles_.ptr + cells / 2048 + 1);
  std::vector<LiftState> ls(nb);
  auto pull = [&]() -> int32_t {
    SFL_HIP(hipMemcpyAsync(ls.data(), lift_.ptr, nb * sizeof(LiftState), hipMemcpyDeviceToHost, stream_),
            "read lifts");
    SFL_HIP(hipStreamSynchronize(stream_), "lifts");
    return SFL_OK;
  };
  // `ls` is next written by a pull, whose wait on the stream comes after
  // this copy in stream order
  auto push = [&]() -> int32_t {
    SFL_HIP(hipMemcpyAsync(lift_.ptr, ls.data(), nb * sizeof(LiftState), hipMemcpyHostToDevice, stream_),
            "write lifts");
    return SFL_OK;
  };
  int32_t st = pull();
  if (st != SFL_OK) return st;
  // A body's columns go to a table of its column box (one slot a column)
  // while the box holds at most 8 columns a water particle plus 65536 and
  // the step's tables stay under kMaxColumns; a sparser box sorts its
  // water's column keys instead (lift.rs:73-84 sorts them too).
  uint64_t near = 0, cols = 0, sort = 0, sorted_most = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    LiftState& l = ls[b];
    near += l.near;
    int64_t span[2] = {1, 1};
    if (l.near) {
      l.col_lo[0] = cell_of(l.low[0], column);
      l.col_lo[1] = cell_of(l.low[2], column);
      span[0] = (int64_t)cell_of(l.high[0], column) - l.col_lo[0] + 1;
      span[1] = (int64_t)cell_of(l.high[2], column) - l.col_lo[1] + 1;
    }
    const uint64_t here = (uint64_t)span[0] * (uint64_t)span[1];
    const bool table = lift_tables_ && span[0] <= (int64_t)kMaxColumns && span[1] <= (int64_t)kMaxColumns &&
                       here <= 8 * (uint64_t)l.near + 65536 && cols + here <= kMaxColumns;
    l.sorted = table ? 0u : 1u;
    l.tops = 0;
    uint64_t size = 1;
    if (table) {
      l.col_span[0] = (uint32_t)span[0];
      l.col_span[1] = (uint32_t)span[1];
      l.column_offset = (uint32_t)cols;
      cols += here;
      while (size < here) size <<= 1;
    } else {
      // at most one top a water particle
      l.col_span[0] = l.col_span[1] = 0;
      l.column_offset = 0;
      sorted_most = std::max<uint64_t>(sorted_most, l.near);
      while (size < l.near) size <<= 1;
    }
    l.sort_offset = (uint32_t)sort;
    l.sort_size = (uint32_t)size;
    sort += size;
    if (sort >= (uint64_t(1) << 32)) {
      detail_ = "the coupled bodies read more water columns than the lift's 32-bit offsets address";
      return SFL_UNSUPPORTED;
    }
  }
  SFL_HIP(lift_near_.reserve(std::max<uint64_t>(near, 1), 0, stream_), "hipMalloc(lift water)");
  SFL_HIP(lift_columns_.reserve(std::max<uint64_t>(cols, 1), 0, stream_), "hipMalloc(lift columns)");
  SFL_HIP(lift_sort_.reserve(sort, 0, stream_), "hipMalloc(lift levels)");
  if (cols)
    SFL_HIP(hipMemsetAsync(lift_columns_.ptr, 0, cols * sizeof(unsigned long long), stream_), "clear lift columns");
  LiftSort lsort{};
  if (sorted_most) {
    const uint64_t room = std::max<uint64_t>(near, 1);
    SFL_HIP(lift_ckeys_.reserve(room, 0, stream_), "hipMalloc(lift column keys)");
    SFL_HIP(lift_ckeys2_.reserve(room, 0, stream_), "hipMalloc(lift column keys)");
    SFL_HIP(lift_ukeys_.reserve(room, 0, stream_), "hipMalloc(lift column keys)");
    SFL_HIP(lift_cvals_.reserve(room, 0, stream_), "hipMalloc(lift column tops)");
    SFL_HIP(lift_cvals2_.reserve(room, 0, stream_), "hipMalloc(lift column tops)");
    size_t temp = 0;
    SFL_HIP(lift_sort_bytes((uint32_t)sorted_most, &temp), "lift column sort size");
    SFL_HIP(lift_sort_temp_.reserve(std::max<size_t>(temp, 1), 0, stream_), "hipMalloc(lift column sort)");
    lsort = LiftSort{lift_ckeys_.ptr, lift_cvals_.ptr, lift_ckeys2_.ptr, lift_cvals2_.ptr, lift_ukeys_.ptr,
                     lift_sort_temp_.ptr, lift_sort_temp_.cap};
  }
  st = push();
  if (st != SFL_OK) return st;
  launch_lift_near(stream_, pa_.ptr, (uint32_t)n_, lift_.ptr, nb, lift_boxes_.ptr, lift_offsets_.ptr, lift_near_.ptr,
                   lift_columns_.ptr, lsort.keys, lsort.values, column, up);
  for (uint32_t b = 0; b < nb && sorted_most; ++b)
    if (ls[b].sorted) SFL_HIP(launch_lift_tops(stream_, lsort, lift_.ptr, b, ls[b], lift_sort_.ptr), "lift column sort");
  launch_lift_level(stream_, pa_.ptr, va_.ptr, lift_.ptr, bodies_d_.ptr, nb, lift_near_.ptr, lift_columns_.ptr,
                    lift_sort_.ptr, probe_arm_.ptr, probe_vel_.ptr, settings_.spacing / 2.0f, up);
  st = pull();
  if (st != SFL_OK) return st;
  for (LiftState& l : ls) l.area = std::pow(l.volume, 2.0f / 3.0f) / l.wetted;
  st = push();
  if (st != SFL_OK) return st;
  launch_lift_force(stream_, lift_.ptr, bodies_d_.ptr, nb, probe_arm_.ptr, probe_vel_.ptr, settings_.rest_density,
                    gravity, up, dt);
  st = pull();
  if (st != SFL_OK) return st;
  for (uint32_t b = 0; b < nb; ++b) {
    const LiftState& l = ls[b];
    const Coupled& k = coupled_[b];
    if (!l.lifted || !pose_of(poses, count, k.index, k.generation)->dynamic) continue;
    sfl_body_push& out = pushes[b];
    out.lifted = 1;
    float change[3];
    for (int a = 0; a < 3; ++a) {
      out.lift_impulse[a] = l.impulse[a];
      out.lift_angular[a] = l.angular[a];
      change[a] = l.impulse[a] / (mass_ * (float)l.near);
    }
    launch_give_back(stream_, lift_near_.ptr + l.near_offset, l.near, change, va_.ptr);
  }
  // the impulses' read-back after this waits on the give-back launches
  vmax_valid_ = false;
  return SFL_OK;
}

// step_with_bodies (src/fluid/couple.rs:207-252)
int32_t World::step_with_bodies(float dt, const sfl_body_pose* poses, size_t count, sfl_report* report,
                                sfl_body_push* pushes, size_t cap, size_t* out_count) {
  DeviceGuard guard(device_);
  if (settings_.solver != SFL_WCSPH && !coupled_.empty()) return SFL_INVALID_SETTINGS;
  for (size_t a = 0; a < count; ++a) {
    const sfl_body_pose& p = poses[a];
    bool finite = true;
    for (int c = 0; c < 3; ++c)
      finite = finite && std::isfinite(p.position[c]) && std::isfinite((float)p.position[c]) &&
               std::isfinite(p.linear_velocity[c]) && std::isfinite(p.angular_velocity[c]);
    for (int c = 0; c < 4; ++c) finite = finite && std::isfinite(p.rotation[c]);
    if (!finite) {
      detail_ = "fluid body pose";
      return SFL_NON_FINITE;
    }
    for (size_t b = 0; b < a; ++b)
      if (poses[b].index == p.index && poses[b].generation == p.generation) return SFL_INVALID_HANDLE;
  }
  size_t kept = 0;
  for (const Coupled& k : coupled_) kept += pose_of(poses, count, k.index, k.generation) ? 1 : 0;
  if (out_count) *out_count = kept;
  if (cap < kept) return SFL_OUT_OF_RANGE;
  ++version_;
  launch_step_begin(stream_, ctl_);
  int32_t st = place_movers(poses, count, dt);
  if (st == SFL_OK) st = step(dt, report);
  movers_m_ = 0;
  if (st != SFL_OK) return st;
  for (size_t b = 0; b < coupled_.size(); ++b) {
    pushes[b] = sfl_body_push{};
    pushes[b].index = coupled_[b].index;
    pushes[b].generation = coupled_[b].generation;
  }
  if (dt > 0.0f && !coupled_.empty()) {
    phase_begin(SFL_PHASE_BODIES);
    st = lifts(dt, poses, count, pushes);
    phase_end(SFL_PHASE_BODIES);
    if (st != SFL_OK) return st;
    std::vector<BodyState> sums(coupled_.size());
    SFL_HIP(hipMemcpyAsync(sums.data(), bodies_d_.ptr, sums.size() * sizeof(BodyState), hipMemcpyDeviceToHost,
                           stream_),
            "read body impulses");
    SFL_HIP(hipStreamSynchronize(stream_), "read body impulses");
    for (size_t b = 0; b < coupled_.size(); ++b)
      for (int a = 0; a < 3; ++a) {
        pushes[b].impulse[a] = sums[b].impulse[a];
        pushes[b].angular[a] = sums[b].angular_impulse[a];
      }
  }
  return SFL_OK;
}

}  // namespace sfl
