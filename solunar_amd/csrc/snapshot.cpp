// Fluid snapshots, SLFL version 3 (src/fluid/save.rs:8-195), for the device
// world. The format is the reference's, byte for byte: little-endian
// integers, f32 as their bits, lengths as u32, an FNV-1a 64 checksum of
// everything before it. Loading treats the bytes as untrusted: every count
// is checked against its limit and the bytes left before anything is kept,
// and the checks come in the reference's order, so the same bad input gets
// the same refusal.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <tuple>
#include <vector>

#include "world.hpp"

namespace sfl {

namespace {

constexpr uint8_t kMagic[4] = {'S', 'L', 'F', 'L'};
constexpr uint32_t kVersion = 3;  // FLUID_SNAPSHOT_VERSION
constexpr size_t kChecksumBytes = 8;
constexpr size_t kMaxSnapshotBytes = size_t(1) << 30;  // MAX_FLUID_SNAPSHOT_BYTES
constexpr size_t kMaxParts = size_t(1) << 16;          // MAX_PARTS

// fnv1a (src/snapshot/bytes.rs:6-13)
uint64_t fnv1a(const uint8_t* data, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= data[i];
    h *= 0x00000100000001b3ull;
  }
  return h;
}

// Writer (src/snapshot/bytes.rs:15-70), into a buffer of the exact size
struct Out {
  uint8_t* at;
  uint8_t* end;  // never written past: a short buffer leaves `over` set
  bool over = false;
  bool room(size_t n) {
    if ((size_t)(end - at) < n) over = true;
    return !over;
  }
  void bytes(const void* p, size_t n) {
    if (!room(n)) return;
    std::memcpy(at, p, n);
    at += n;
  }
  void u8(uint8_t v) {
    if (room(1)) *at++ = v;
  }
  void u32(uint32_t v) {
    if (!room(4)) return;
    at[0] = (uint8_t)v, at[1] = (uint8_t)(v >> 8), at[2] = (uint8_t)(v >> 16), at[3] = (uint8_t)(v >> 24);
    at += 4;
  }
  void u64(uint64_t v) {
    u32((uint32_t)v);
    u32((uint32_t)(v >> 32));
  }
  void f32(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    u32(b);
  }
  void vec3(float x, float y, float z) {
    f32(x);
    f32(y);
    f32(z);
  }
  void vec3(const sfl_vec3& v) { vec3(v.x, v.y, v.z); }
  void len(size_t n) { u32((uint32_t)n); }
};

// Reader (src/snapshot/bytes.rs:72-150): false means Truncated unless `fail`
// says otherwise.
struct In {
  const uint8_t* data;
  size_t size;
  size_t at;
  SnapshotFailure& fail;
  bool take(void* dst, size_t n) {
    if (size - at < n) return truncated();
    std::memcpy(dst, data + at, n);
    at += n;
    return true;
  }
  bool truncated() {
    fail = {SFL_SNAPSHOT_TRUNCATED, 0, nullptr};
    return false;
  }
  bool invalid(const char* what) {
    fail = {SFL_SNAPSHOT_INVALID, 0, what};
    return false;
  }
  bool u8(uint8_t& v) { return take(&v, 1); }
  bool u32(uint32_t& v) {
    uint8_t b[4];
    if (!take(b, 4)) return false;
    v = (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24;
    return true;
  }
  bool u64(uint64_t& v) {
    uint32_t lo, hi;
    if (!u32(lo) || !u32(hi)) return false;
    v = (uint64_t)lo | (uint64_t)hi << 32;
    return true;
  }
  bool f32(float& v) {
    uint32_t b;
    if (!u32(b)) return false;
    std::memcpy(&v, &b, 4);
    return true;
  }
  bool vec3(sfl_vec3& v) { return f32(v.x) && f32(v.y) && f32(v.z); }
  // Reader::count: checked against `max` and the bytes left
  bool count(size_t record, size_t max, size_t& out) {
    uint32_t c;
    if (!u32(c)) return false;
    if (c > max) return invalid("count over limit");
    if ((uint64_t)c * record > size - at) return truncated();
    out = c;
    return true;
  }
};

bool finite3(const sfl_vec3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

struct Device {
  explicit Device(int d) { (void)hipSetDevice(d); }
};
bool inside_out(const sfl_vec3& lo, const sfl_vec3& hi) { return hi.x <= lo.x || hi.y <= lo.y || hi.z <= lo.z; }

}  // namespace

#define SFL_HIP(call, what)                              \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return device_error(what, e_); \
  } while (0)

// The size save_snapshot writes.
size_t World::snapshot_size() const {
  size_t flows = 0;
  for (const Flow& f : flows_) flows += f.drain ? 29 : 41;
  // a coupled body: handle, points, probe spacing, probes (couple.rs:50-55)
  size_t bodies = 0;
  for (const Coupled& k : coupled_) bodies += 8 + 4 + (k.points.size() / 3) * 12 + 4 + 4 + (k.probes.size() / 3) * 12;
  // a terrain region: owner, low, nodes, step, distances (terrain.rs:45-56)
  size_t ground = 0;
  for (const Terrain& t : terrains_) ground += 4 + 12 + 12 + 4 + 4 + t.distances.size() * 4;
  return 4 + 4 + 1 + 24 + 8 + 12 + 4 + 4 + n_ * 28 + 4 + 4 + wall_owner_.size() * 16 + 4 + 4 +
         containers_.size() * 28 + 4 + flows + 4 + 4 + bodies + 4 + ground + (has_domain_ ? 25 : 1) + kChecksumBytes;
}

// save_snapshot (save.rs:17-66) into `bytes`, snapshot_size() of them
int32_t World::save_snapshot(uint8_t* bytes) {
  Device g(device_);
  std::vector<float4> p(n_), v(n_);
  if (n_) {
    SFL_HIP(hipMemcpyAsync(p.data(), pa_.ptr, n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_), "read positions");
    SFL_HIP(hipMemcpyAsync(v.data(), va_.ptr, n_ * sizeof(float4), hipMemcpyDeviceToHost, stream_),
            "read velocities");
    SFL_HIP(hipStreamSynchronize(stream_), "snapshot");
  }
  const size_t size = snapshot_size();
  Out out{bytes, bytes + size};
  const sfl_settings& s = settings_;
  out.bytes(kMagic, 4);
  out.u32(kVersion);
  out.u8(s.solver == SFL_DFSPH ? 1 : 0);
  for (float f : {s.spacing, s.rest_density, s.support, s.sound_speed, s.viscosity, s.courant}) out.f32(f);
  out.u64(s.max_particles);
  out.vec3(s.gravity);
  out.u32(s.max_substeps);
  out.len(n_);
  for (size_t i = 0; i < n_; ++i) {
    uint32_t id;
    std::memcpy(&id, &p[i].w, 4);
    out.u32(id);
    out.vec3(p[i].x, p[i].y, p[i].z);
    out.vec3(v[i].x, v[i].y, v[i].z);
  }
  out.u32(next_id_);
  // Walls::save (boundary.rs:182-195)
  out.len(wall_owner_.size());
  for (size_t i = 0; i < wall_owner_.size(); ++i) {
    out.vec3(wall_pos_[3 * i], wall_pos_[3 * i + 1], wall_pos_[3 * i + 2]);
    out.u32(wall_owner_[i]);
  }
  out.u32(wall_next_);
  out.len(containers_.size());
  for (const Container& c : containers_) {
    out.u32(c.id);
    out.vec3(c.lo[0], c.lo[1], c.lo[2]);
    out.vec3(c.hi[0], c.hi[1], c.hi[2]);
  }
  // Flows::save (flow.rs:36-52)
  out.len(flows_.size());
  for (const Flow& f : flows_) {
    out.u32(f.id);
    out.vec3(f.low);
    out.vec3(f.high);
    if (f.drain) {
      out.u8(1);
    } else {
      out.u8(0);
      out.vec3(f.velocity);
    }
  }
  out.u32(flow_next_);
  // Coupled::save (couple.rs:50-55)
  out.len(coupled_.size());
  for (const Coupled& k : coupled_) {
    out.u32(k.index);
    out.u32(k.generation);
    out.len(k.points.size() / 3);
    for (size_t i = 0; i < k.points.size(); i += 3) out.vec3(k.points[i], k.points[i + 1], k.points[i + 2]);
    out.f32(k.probe_spacing);
    out.len(k.probes.size() / 3);
    for (size_t i = 0; i < k.probes.size(); i += 3) out.vec3(k.probes[i], k.probes[i + 1], k.probes[i + 2]);
  }
  // TerrainField::save (terrain.rs:45-56)
  out.len(terrains_.size());
  for (const Terrain& t : terrains_) {
    out.u32(t.owner);
    out.vec3(t.low[0], t.low[1], t.low[2]);
    for (uint32_t n : t.nodes) out.u32(n);
    out.f32(t.step);
    out.len(t.distances.size());
    for (float d : t.distances) out.f32(d);
  }
  if (has_domain_) {
    out.u8(1);
    out.vec3(domain_.lo[0], domain_.lo[1], domain_.lo[2]);
    out.vec3(domain_.hi[0], domain_.hi[1], domain_.hi[2]);
  } else {
    out.u8(0);
  }
  if (!out.over) out.u64(fnv1a(bytes, (size_t)(out.at - bytes)));
  if (out.over || (size_t)(out.at - bytes) != size) {
    detail_ = "snapshot size";
    return SFL_DEVICE_ERROR;
  }
  return SFL_OK;
}

// load_snapshot (save.rs:68-195)
int32_t World::load_snapshot(const uint8_t* data, size_t size, const sfl_settings& base, World** world_out,
                             SnapshotFailure& fail, std::string& why) {
  *world_out = nullptr;
  fail = {0, 0, nullptr};
  auto refuse = [&](int32_t kind) {
    fail = {kind, 0, nullptr};
    return SFL_SNAPSHOT_REFUSED;
  };
  if (size > kMaxSnapshotBytes) return refuse(SFL_SNAPSHOT_TOO_LARGE);
  if (size < 4 + 4 + kChecksumBytes) return refuse(SFL_SNAPSHOT_TRUNCATED);
  if (std::memcmp(data, kMagic, 4) != 0) return refuse(SFL_SNAPSHOT_BAD_MAGIC);
  const size_t body = size - kChecksumBytes;
  {
    In sum{data + body, kChecksumBytes, 0, fail};
    uint64_t stored = 0;
    if (!sum.u64(stored)) return SFL_SNAPSHOT_REFUSED;
    if (stored != fnv1a(data, body)) return refuse(SFL_SNAPSHOT_BAD_CHECKSUM);
  }
  In r{data, body, 4, fail};
  uint32_t version = 0;
  if (!r.u32(version)) return SFL_SNAPSHOT_REFUSED;
  if (version < 1 || version > kVersion) {
    fail = {SFL_SNAPSHOT_UNSUPPORTED_VERSION, version, nullptr};
    return SFL_SNAPSHOT_REFUSED;
  }
  uint8_t solver = 0;
  if (!r.u8(solver)) return SFL_SNAPSHOT_REFUSED;
  if (solver > 1) {
    r.invalid("fluid solver");
    return SFL_SNAPSHOT_REFUSED;
  }
  sfl_settings s = base;
  s.solver = solver;
  float f[6];
  for (float& x : f)
    if (!r.f32(x)) return SFL_SNAPSHOT_REFUSED;
  s.spacing = f[0];
  s.rest_density = f[1];
  s.support = f[2];
  s.sound_speed = f[3];
  s.viscosity = f[4];
  s.courant = f[5];
  uint64_t max_particles = 0;
  if (!r.u64(max_particles)) return SFL_SNAPSHOT_REFUSED;
  s.max_particles = max_particles;
  if (!r.vec3(s.gravity) || !r.u32(s.max_substeps)) return SFL_SNAPSHOT_REFUSED;
  // Self::new(settings) in the reference: its refusal is Invalid("fluid settings")
  if (!valid_settings(s)) {
    r.invalid("fluid settings");
    return SFL_SNAPSHOT_REFUSED;
  }
  const size_t max = (size_t)std::min<uint64_t>(max_particles, SIZE_MAX);
  size_t count = 0;
  if (!r.count(28, max, count)) return SFL_SNAPSHOT_REFUSED;
  std::vector<uint32_t> ids(count);
  std::vector<float> xyz(3 * count), vxyz(3 * count);
  for (size_t i = 0; i < count; ++i) {
    sfl_vec3 at, vel;
    if (!r.u32(ids[i]) || !r.vec3(at) || !r.vec3(vel)) return SFL_SNAPSHOT_REFUSED;
    if (!finite3(at) || !finite3(vel)) {
      r.invalid("fluid particle");
      return SFL_SNAPSHOT_REFUSED;
    }
    xyz[3 * i] = at.x, xyz[3 * i + 1] = at.y, xyz[3 * i + 2] = at.z;
    vxyz[3 * i] = vel.x, vxyz[3 * i + 1] = vel.y, vxyz[3 * i + 2] = vel.z;
  }
  uint32_t next_id = 0;
  if (!r.u32(next_id)) return SFL_SNAPSHOT_REFUSED;
  {
    std::vector<uint32_t> seen(ids);
    std::sort(seen.begin(), seen.end());
    bool bad = std::adjacent_find(seen.begin(), seen.end()) != seen.end() || (!seen.empty() && seen.back() >= next_id);
    if (bad) {
      r.invalid("fluid particle id");
      return SFL_SNAPSHOT_REFUSED;
    }
  }
  // Walls::load (boundary.rs:202-227)
  size_t walls = 0;
  if (!r.count(16, max, walls)) return SFL_SNAPSHOT_REFUSED;
  std::vector<float> wall_pos(3 * walls);
  std::vector<uint32_t> wall_owner(walls);
  for (size_t i = 0; i < walls; ++i) {
    sfl_vec3 at;
    if (!r.vec3(at)) return SFL_SNAPSHOT_REFUSED;
    if (!finite3(at)) {
      r.invalid("fluid wall");
      return SFL_SNAPSHOT_REFUSED;
    }
    wall_pos[3 * i] = at.x, wall_pos[3 * i + 1] = at.y, wall_pos[3 * i + 2] = at.z;
    if (!r.u32(wall_owner[i])) return SFL_SNAPSHOT_REFUSED;
  }
  uint32_t wall_next = 0;
  if (!r.u32(wall_next)) return SFL_SNAPSHOT_REFUSED;
  size_t ncont = 0;
  if (!r.count(28, max, ncont)) return SFL_SNAPSHOT_REFUSED;
  std::vector<Container> containers;
  for (size_t i = 0; i < ncont; ++i) {
    uint32_t id;
    sfl_vec3 lo, hi;
    if (!r.u32(id) || !r.vec3(lo) || !r.vec3(hi)) return SFL_SNAPSHOT_REFUSED;
    if (!finite3(lo) || !finite3(hi) || inside_out(lo, hi) || id >= wall_next) {
      r.invalid("fluid container");
      return SFL_SNAPSHOT_REFUSED;
    }
    containers.push_back(Container{id, {lo.x, lo.y, lo.z}, {hi.x, hi.y, hi.z}});
  }
  if (std::any_of(wall_owner.begin(), wall_owner.end(), [&](uint32_t o) { return o >= wall_next; })) {
    r.invalid("fluid wall owner");
    return SFL_SNAPSHOT_REFUSED;
  }
  // Flows::load (flow.rs:54-89)
  size_t nflows = 0;
  if (!r.count(29, kMaxParts, nflows)) return SFL_SNAPSHOT_REFUSED;
  std::vector<Flow> flows;
  for (size_t i = 0; i < nflows; ++i) {
    Flow fl{};
    uint8_t kind = 0;
    if (!r.u32(fl.id) || !r.vec3(fl.low) || !r.vec3(fl.high) || !r.u8(kind)) return SFL_SNAPSHOT_REFUSED;
    if (kind == 0) {
      if (!r.vec3(fl.velocity)) return SFL_SNAPSHOT_REFUSED;
    } else if (kind == 1) {
      fl.drain = true;
    } else {
      r.invalid("fluid flow kind");
      return SFL_SNAPSHOT_REFUSED;
    }
    const bool finite = fl.drain || finite3(fl.velocity);
    if (!(finite3(fl.low) && finite3(fl.high) && finite) || inside_out(fl.low, fl.high)) {
      r.invalid("fluid flow");
      return SFL_SNAPSHOT_REFUSED;
    }
    flows.push_back(fl);
  }
  uint32_t flow_next = 0;
  if (!r.u32(flow_next)) return SFL_SNAPSHOT_REFUSED;
  if (std::any_of(flows.begin(), flows.end(), [&](const Flow& fl) { return fl.id >= flow_next; })) {
    r.invalid("fluid flow id");
    return SFL_SNAPSHOT_REFUSED;
  }
  // Coupled::load (couple.rs:57-75) in save.rs:158-171's order
  size_t coupled = 0;
  if (!r.count(12, kMaxParts, coupled)) return SFL_SNAPSHOT_REFUSED;
  std::vector<Coupled> bodies;
  size_t room = max;
  auto points = [&](size_t most, std::vector<float>& pts) -> bool {
    size_t n = 0;
    if (!r.count(12, most, n)) return false;
    for (size_t k = 0; k < n; ++k) {
      sfl_vec3 p;
      if (!r.vec3(p)) return false;
      if (!finite3(p)) return r.invalid("fluid body point");
      pts.insert(pts.end(), {p.x, p.y, p.z});
    }
    if (pts.empty()) return r.invalid("fluid body");
    return true;
  };
  for (size_t b = 0; b < coupled; ++b) {
    Coupled k{};
    if (!r.u32(k.index) || !r.u32(k.generation)) return SFL_SNAPSHOT_REFUSED;
    if (!points(room, k.points)) return SFL_SNAPSHOT_REFUSED;
    if (version >= 3) {
      if (!r.f32(k.probe_spacing)) return SFL_SNAPSHOT_REFUSED;
      if (!std::isfinite(k.probe_spacing) || k.probe_spacing <= 0.0f) {
        r.invalid("fluid body probes");
        return SFL_SNAPSHOT_REFUSED;
      }
      if (!points(4096, k.probes)) return SFL_SNAPSHOT_REFUSED;  // MAX_PROBES, lift.rs:8
    } else {
      k.probe_spacing = s.spacing;
      k.probes = k.points;
    }
    room -= k.points.size() / 3;
    bodies.push_back(std::move(k));
  }
  for (size_t b = 1; b < bodies.size(); ++b)
    if (std::tie(bodies[b - 1].index, bodies[b - 1].generation) >= std::tie(bodies[b].index, bodies[b].generation)) {
      r.invalid("fluid body order");
      return SFL_SNAPSHOT_REFUSED;
    }
  // TerrainField::load (terrain.rs:58-95) in save.rs:172-179's order
  size_t terrains = 0;
  if (!r.count(32, kMaxParts, terrains)) return SFL_SNAPSHOT_REFUSED;
  std::vector<Terrain> ground;
  for (size_t t = 0; t < terrains; ++t) {
    Terrain f{};
    sfl_vec3 low;
    if (!r.u32(f.owner) || !r.vec3(low) || !r.u32(f.nodes[0]) || !r.u32(f.nodes[1]) || !r.u32(f.nodes[2]) ||
        !r.f32(f.step))
      return SFL_SNAPSHOT_REFUSED;
    size_t count = 0;
    if (!r.count(4, size_t(1) << 24, count)) return SFL_SNAPSHOT_REFUSED;  // MAX_NODES, terrain.rs:9
    const unsigned __int128 total = (unsigned __int128)f.nodes[0] * f.nodes[1] * f.nodes[2];
    if (!finite3(low) || !std::isfinite(f.step) || f.step <= 0.0f || f.nodes[0] < 2 || f.nodes[1] < 2 ||
        f.nodes[2] < 2 || total != count) {
      r.invalid("fluid terrain");
      return SFL_SNAPSHOT_REFUSED;
    }
    f.low[0] = low.x, f.low[1] = low.y, f.low[2] = low.z;
    f.distances.resize(count);
    for (size_t k = 0; k < count; ++k) {
      if (!r.f32(f.distances[k])) return SFL_SNAPSHOT_REFUSED;
      if (!std::isfinite(f.distances[k])) {
        r.invalid("fluid terrain distance");
        return SFL_SNAPSHOT_REFUSED;
      }
    }
    if (f.owner >= wall_next) {
      r.invalid("fluid terrain owner");
      return SFL_SNAPSHOT_REFUSED;
    }
    ground.push_back(std::move(f));
  }
  bool has_domain = false;
  sfl_vec3 dlo{0, 0, 0}, dhi{0, 0, 0};
  if (version >= 2) {
    uint8_t flag = 0;
    if (!r.u8(flag)) return SFL_SNAPSHOT_REFUSED;
    if (flag == 1) {
      if (!r.vec3(dlo) || !r.vec3(dhi)) return SFL_SNAPSHOT_REFUSED;
      if (!finite3(dlo) || !finite3(dhi) || inside_out(dlo, dhi)) {
        r.invalid("fluid domain");
        return SFL_SNAPSHOT_REFUSED;
      }
      has_domain = true;
    } else if (flag != 0) {
      r.invalid("fluid domain flag");
      return SFL_SNAPSHOT_REFUSED;
    }
  }
  if (r.at != body) {
    r.invalid("trailing bytes");
    return SFL_SNAPSHOT_REFUSED;
  }

  // accepted: a world with this state
  World* w = nullptr;
  int32_t st = create(s, &w, why);
  if (st != SFL_OK) return st;
  st = w->restore(count, ids.data(), xyz.data(), vxyz.data(), next_id);
  if (st == SFL_OK) {
    w->wall_pos_ = std::move(wall_pos);
    w->wall_owner_ = std::move(wall_owner);
    w->wall_next_ = wall_next;
    w->containers_ = std::move(containers);
    st = w->settle_walls();
  }
  if (st != SFL_OK) {
    why = w->detail_;
    delete w;
    return st;
  }
  w->flows_ = std::move(flows);
  w->flow_next_ = flow_next;
  w->coupled_ = std::move(bodies);
  w->coupled_dirty_ = true;
  w->terrains_ = std::move(ground);
  w->terrain_dirty_ = true;
  w->has_domain_ = has_domain;
  w->domain_ = FlowBox{{dlo.x, dlo.y, dlo.z}, {dhi.x, dhi.y, dhi.z}};
  *world_out = w;
  return SFL_OK;
}

}  // namespace sfl
