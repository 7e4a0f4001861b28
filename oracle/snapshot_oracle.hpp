// TEST INFRASTRUCTURE: the CPU oracle's fluid snapshot, SLFL version 3
// (src/fluid/save.rs:8-195), restated from the reference: the Writer and
// Reader of src/snapshot/bytes.rs:6-150, Walls::save/load
// (src/fluid/boundary.rs:182-227), Flows::save/load (src/fluid/flow.rs:36-89).
// Pinned by the reference's own snapshot tests (save.rs:237-314) and by the
// device writing the same bytes.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "fluid_oracle.hpp"

namespace oracle {

enum SnapshotKind : int32_t {
  kSnapOk = 0,
  kSnapTooLarge = 1,
  kSnapTruncated = 2,
  kSnapBadMagic = 3,
  kSnapUnsupportedVersion = 4,
  kSnapBadChecksum = 5,
  kSnapInvalid = 6,
  kSnapUnsupported = 101,
};

struct SnapshotResult {
  int32_t kind = kSnapOk;
  uint32_t version = 0;
  const char* what = nullptr;
};

// fnv1a (src/snapshot/bytes.rs:6-13)
inline uint64_t snapshot_fnv1a(const uint8_t* d, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ d[i]) * 0x00000100000001b3ull;
  return h;
}

struct SnapWriter {
  std::vector<uint8_t> data;
  void u8(uint8_t v) { data.push_back(v); }
  void u32(uint32_t v) {
    for (int k = 0; k < 4; ++k) data.push_back((uint8_t)(v >> (8 * k)));
  }
  void u64(uint64_t v) {
    for (int k = 0; k < 8; ++k) data.push_back((uint8_t)(v >> (8 * k)));
  }
  void f32(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    u32(b);
  }
  void vec3(Vec3<float> v) {
    f32(v.x);
    f32(v.y);
    f32(v.z);
  }
};

struct SnapReader {
  const uint8_t* data;
  size_t len;
  size_t at;
  SnapshotResult& res;
  bool fail(int32_t kind, const char* what = nullptr) {
    res = {kind, 0, what};
    return false;
  }
  bool take(uint8_t* out, size_t n) {
    if (len - at < n) return fail(kSnapTruncated);
    std::memcpy(out, data + at, n);
    at += n;
    return true;
  }
  bool u8(uint8_t& v) { return take(&v, 1); }
  bool u32(uint32_t& v) {
    uint8_t b[4];
    if (!take(b, 4)) return false;
    v = 0;
    for (int k = 3; k >= 0; --k) v = v << 8 | b[k];
    return true;
  }
  bool u64(uint64_t& v) {
    uint8_t b[8];
    if (!take(b, 8)) return false;
    v = 0;
    for (int k = 7; k >= 0; --k) v = v << 8 | b[k];
    return true;
  }
  bool f32(float& v) {
    uint32_t b;
    if (!u32(b)) return false;
    std::memcpy(&v, &b, 4);
    return true;
  }
  bool vec3(Vec3<float>& v) { return f32(v.x) && f32(v.y) && f32(v.z); }
  // Reader::count (bytes.rs:129-149)
  bool count(size_t record, uint64_t max, size_t& out) {
    uint32_t c;
    if (!u32(c)) return false;
    if (c > max) return fail(kSnapInvalid, "count over limit");
    if ((uint64_t)c * record > len - at) return fail(kSnapTruncated);
    out = c;
    return true;
  }
};

// What a snapshot holds.
struct SnapshotState {
  Settings settings;
  std::vector<uint32_t> ids;
  std::vector<Vec3<float>> positions, velocities;
  uint32_t next_id = 0;
  std::vector<Vec3<float>> wall_positions;
  std::vector<uint32_t> wall_owners;
  uint32_t wall_next = 0;
  std::vector<std::tuple<uint32_t, Vec3<float>, Vec3<float>>> containers;
  struct FlowRec {
    uint32_t id;
    Vec3<float> low, high;
    bool drain;
    Vec3<float> velocity;
  };
  std::vector<FlowRec> flows;
  uint32_t flow_next = 0;
  bool has_domain = false;
  Vec3<float> domain_low{0, 0, 0}, domain_high{0, 0, 0};
  struct CoupledRec {
    uint32_t index, generation;
    std::vector<Vec3<float>> points, probes;
    float probe_spacing;
  };
  std::vector<CoupledRec> coupled;
  std::vector<TerrainField> terrains;
};

inline bool snap_box_ok(Vec3<float> lo, Vec3<float> hi) {
  return lo.finite() && hi.finite() && !(hi.x <= lo.x || hi.y <= lo.y || hi.z <= lo.z);
}

// save_snapshot (save.rs:17-66)
template <class R>
std::vector<uint8_t> save_snapshot(const World<R>& w) {
  SnapWriter out;
  const Settings& s = w.settings;
  for (char c : {'S', 'L', 'F', 'L'}) out.u8((uint8_t)c);
  out.u32(3);
  out.u8(s.solver == kDfsph ? 1 : 0);
  for (float v : {s.spacing, s.rest_density, s.support, s.sound_speed, s.viscosity, s.courant}) out.f32(v);
  out.u64(s.max_particles);
  out.vec3(s.gravity);
  out.u32(s.max_substeps);
  auto f3 = [](const Vec3<R>& v) { return Vec3<float>{(float)v.x, (float)v.y, (float)v.z}; };
  out.u32((uint32_t)w.ids.size());
  for (size_t i = 0; i < w.ids.size(); ++i) {
    out.u32(w.ids[i]);
    out.vec3(f3(w.positions[i]));
    out.vec3(f3(w.velocities[i]));
  }
  out.u32(w.next_id);
  out.u32((uint32_t)w.walls.positions.size());
  for (size_t i = 0; i < w.walls.positions.size(); ++i) {
    out.vec3(f3(w.walls.positions[i]));
    out.u32(w.walls.owners[i]);
  }
  out.u32(w.walls.next);
  out.u32((uint32_t)w.walls.containers.size());
  for (const auto& c : w.walls.containers) {
    out.u32(std::get<0>(c));
    out.vec3(f3(std::get<1>(c)));
    out.vec3(f3(std::get<2>(c)));
  }
  out.u32((uint32_t)w.flow_list.size());
  for (const auto& f : w.flow_list) {
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
  out.u32(w.flow_next);
  // Coupled::save (src/fluid/couple.rs:50-55)
  out.u32((uint32_t)w.coupled.size());
  for (const auto& k : w.coupled) {
    out.u32(k.index);
    out.u32(k.generation);
    out.u32((uint32_t)k.points.size());
    for (const auto& p : k.points) out.vec3(p);
    out.f32(k.probe_spacing);
    out.u32((uint32_t)k.probes.size());
    for (const auto& p : k.probes) out.vec3(p);
  }
  // TerrainField::save (terrain.rs:45-56)
  out.u32((uint32_t)w.terrains.size());
  for (const auto& t : w.terrains) {
    out.u32(t.owner);
    out.vec3(t.low);
    for (size_t n : t.nodes) out.u32((uint32_t)n);
    out.f32(t.step);
    out.u32((uint32_t)t.distances.size());
    for (float d : t.distances) out.f32(d);
  }
  if (w.has_domain) {
    out.u8(1);
    out.vec3(w.domain_low);
    out.vec3(w.domain_high);
  } else {
    out.u8(0);
  }
  const uint64_t sum = snapshot_fnv1a(out.data.data(), out.data.size());
  out.u64(sum);
  return out.data;
}

// load_snapshot (save.rs:68-195) up to building the world: the checks in
// the reference's order. `base` gives the settings the format does not hold.
inline SnapshotResult parse_snapshot(const uint8_t* data, size_t len, const Settings& base, SnapshotState& st) {
  SnapshotResult res;
  if (len > (size_t(1) << 30)) return {kSnapTooLarge, 0, nullptr};
  if (len < 4 + 4 + 8) return {kSnapTruncated, 0, nullptr};
  if (std::memcmp(data, "SLFL", 4) != 0) return {kSnapBadMagic, 0, nullptr};
  const size_t body = len - 8;
  {
    SnapReader sum{data + body, 8, 0, res};
    uint64_t stored = 0;
    if (!sum.u64(stored)) return res;
    if (stored != snapshot_fnv1a(data, body)) return {kSnapBadChecksum, 0, nullptr};
  }
  SnapReader r{data, body, 4, res};
  uint32_t version = 0;
  if (!r.u32(version)) return res;
  if (version < 1 || version > 3) return {kSnapUnsupportedVersion, version, nullptr};
  uint8_t solver = 0;
  if (!r.u8(solver)) return res;
  if (solver > 1) return {kSnapInvalid, 0, "fluid solver"};
  Settings s = base;
  s.solver = solver == 1 ? kDfsph : kWcsph;
  float f[6];
  for (float& v : f)
    if (!r.f32(v)) return res;
  s.spacing = f[0], s.rest_density = f[1], s.support = f[2], s.sound_speed = f[3], s.viscosity = f[4],
  s.courant = f[5];
  if (!r.u64(s.max_particles) || !r.vec3(s.gravity) || !r.u32(s.max_substeps)) return res;
  if (!s.valid()) return {kSnapInvalid, 0, "fluid settings"};
  st.settings = s;
  const uint64_t max = s.max_particles;
  size_t count = 0;
  if (!r.count(28, max, count)) return res;
  for (size_t i = 0; i < count; ++i) {
    uint32_t id;
    Vec3<float> at, vel;
    if (!r.u32(id) || !r.vec3(at) || !r.vec3(vel)) return res;
    if (!at.finite() || !vel.finite()) return {kSnapInvalid, 0, "fluid particle"};
    st.ids.push_back(id);
    st.positions.push_back(at);
    st.velocities.push_back(vel);
  }
  if (!r.u32(st.next_id)) return res;
  std::vector<uint32_t> seen = st.ids;
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end() || (!seen.empty() && seen.back() >= st.next_id))
    return {kSnapInvalid, 0, "fluid particle id"};
  size_t walls = 0;
  if (!r.count(16, max, walls)) return res;
  for (size_t i = 0; i < walls; ++i) {
    Vec3<float> at;
    uint32_t owner;
    if (!r.vec3(at)) return res;
    if (!at.finite()) return {kSnapInvalid, 0, "fluid wall"};
    if (!r.u32(owner)) return res;
    st.wall_positions.push_back(at);
    st.wall_owners.push_back(owner);
  }
  if (!r.u32(st.wall_next)) return res;
  size_t containers = 0;
  if (!r.count(28, max, containers)) return res;
  for (size_t i = 0; i < containers; ++i) {
    uint32_t id;
    Vec3<float> lo, hi;
    if (!r.u32(id) || !r.vec3(lo) || !r.vec3(hi)) return res;
    if (!snap_box_ok(lo, hi) || id >= st.wall_next) return {kSnapInvalid, 0, "fluid container"};
    st.containers.push_back({id, lo, hi});
  }
  for (uint32_t o : st.wall_owners)
    if (o >= st.wall_next) return {kSnapInvalid, 0, "fluid wall owner"};
  size_t flows = 0;
  if (!r.count(29, uint64_t(1) << 16, flows)) return res;
  for (size_t i = 0; i < flows; ++i) {
    SnapshotState::FlowRec fl{};
    uint8_t kind = 0;
    if (!r.u32(fl.id) || !r.vec3(fl.low) || !r.vec3(fl.high) || !r.u8(kind)) return res;
    if (kind == 0) {
      if (!r.vec3(fl.velocity)) return res;
    } else if (kind == 1) {
      fl.drain = true;
    } else {
      return {kSnapInvalid, 0, "fluid flow kind"};
    }
    const bool finite = fl.drain || fl.velocity.finite();
    if (!(fl.low.finite() && fl.high.finite() && finite) || !snap_box_ok(fl.low, fl.high))
      return {kSnapInvalid, 0, "fluid flow"};
    st.flows.push_back(fl);
  }
  if (!r.u32(st.flow_next)) return res;
  for (const auto& fl : st.flows)
    if (fl.id >= st.flow_next) return {kSnapInvalid, 0, "fluid flow id"};
  // Coupled::load (src/fluid/couple.rs:57-75), save.rs:158-171
  size_t coupled = 0;
  if (!r.count(12, uint64_t(1) << 16, coupled)) return res;
  uint64_t room = max;
  auto points = [&](uint64_t most, std::vector<Vec3<float>>& out) -> int {
    size_t n = 0;
    if (!r.count(12, most, n)) return 1;
    for (size_t k = 0; k < n; ++k) {
      Vec3<float> p;
      if (!r.vec3(p)) return 1;
      if (!p.finite()) return 2;
      out.push_back(p);
    }
    return out.empty() ? 3 : 0;
  };
  for (size_t b = 0; b < coupled; ++b) {
    SnapshotState::CoupledRec k{};
    if (!r.u32(k.index) || !r.u32(k.generation)) return res;
    int e = points(room, k.points);
    if (e == 1) return res;
    if (e == 2) return {kSnapInvalid, 0, "fluid body point"};
    if (e == 3) return {kSnapInvalid, 0, "fluid body"};
    if (version >= 3) {
      if (!r.f32(k.probe_spacing)) return res;
      if (!std::isfinite(k.probe_spacing) || k.probe_spacing <= 0.0f) return {kSnapInvalid, 0, "fluid body probes"};
      e = points(kMaxProbes, k.probes);
      if (e == 1) return res;
      if (e == 2) return {kSnapInvalid, 0, "fluid body point"};
      if (e == 3) return {kSnapInvalid, 0, "fluid body"};
    } else {
      k.probe_spacing = s.spacing;
      k.probes = k.points;
    }
    room -= k.points.size();
    st.coupled.push_back(std::move(k));
  }
  for (size_t b = 1; b < st.coupled.size(); ++b)
    if (std::tie(st.coupled[b - 1].index, st.coupled[b - 1].generation) >=
        std::tie(st.coupled[b].index, st.coupled[b].generation))
      return {kSnapInvalid, 0, "fluid body order"};
  // TerrainField::load (terrain.rs:58-95), save.rs:172-179
  size_t terrains = 0;
  if (!r.count(32, uint64_t(1) << 16, terrains)) return res;
  for (size_t t = 0; t < terrains; ++t) {
    TerrainField f{};
    uint32_t n[3];
    if (!r.u32(f.owner) || !r.vec3(f.low) || !r.u32(n[0]) || !r.u32(n[1]) || !r.u32(n[2]) || !r.f32(f.step))
      return res;
    size_t count = 0;
    if (!r.count(4, kMaxTerrainNodes, count)) return res;
    unsigned __int128 total = (unsigned __int128)n[0] * n[1] * n[2];
    if (!f.low.finite() || !std::isfinite(f.step) || f.step <= 0.0f || n[0] < 2 || n[1] < 2 || n[2] < 2 ||
        total != count)
      return {kSnapInvalid, 0, "fluid terrain"};
    for (int a = 0; a < 3; ++a) f.nodes[a] = n[a];
    f.distances.resize(count);
    for (size_t k = 0; k < count; ++k) {
      if (!r.f32(f.distances[k])) return res;
      if (!std::isfinite(f.distances[k])) return {kSnapInvalid, 0, "fluid terrain distance"};
    }
    if (f.owner >= st.wall_next) return {kSnapInvalid, 0, "fluid terrain owner"};
    st.terrains.push_back(std::move(f));
  }
  if (version >= 2) {
    uint8_t flag = 0;
    if (!r.u8(flag)) return res;
    if (flag == 1) {
      if (!r.vec3(st.domain_low) || !r.vec3(st.domain_high)) return res;
      if (!snap_box_ok(st.domain_low, st.domain_high)) return {kSnapInvalid, 0, "fluid domain"};
      st.has_domain = true;
    } else if (flag != 0) {
      return {kSnapInvalid, 0, "fluid domain flag"};
    }
  }
  if (r.at != body) return {kSnapInvalid, 0, "trailing bytes"};
  return {kSnapOk, 0, nullptr};
}

// The rest of load_snapshot: a world (built from st.settings) takes the
// state; densities start at the rest density (save.rs:143).
template <class R>
void apply_snapshot(World<R>& w, const SnapshotState& st) {
  using V = Vec3<R>;
  auto cv = [](Vec3<float> v) { return V{R(v.x), R(v.y), R(v.z)}; };
  w.ids = st.ids;
  w.positions.clear();
  w.velocities.clear();
  for (size_t i = 0; i < st.ids.size(); ++i) {
    w.positions.push_back(cv(st.positions[i]));
    w.velocities.push_back(cv(st.velocities[i]));
  }
  w.densities.assign(st.ids.size(), R(st.settings.rest_density));
  w.next_id = st.next_id;
  w.walls.positions.clear();
  for (const auto& p : st.wall_positions) w.walls.positions.push_back(cv(p));
  w.walls.owners = st.wall_owners;
  w.walls.next = st.wall_next;
  w.walls.containers.clear();
  for (const auto& c : st.containers) w.walls.containers.push_back({std::get<0>(c), cv(std::get<1>(c)), cv(std::get<2>(c))});
  w.walls.bounds = w.union_of(w.walls.containers);
  w.flow_list.clear();
  for (const auto& f : st.flows) w.flow_list.push_back({f.id, f.low, f.high, f.drain, f.velocity});
  w.flow_next = st.flow_next;
  w.has_domain = st.has_domain;
  w.domain_low = st.domain_low;
  w.domain_high = st.domain_high;
  w.coupled.clear();
  for (const auto& k : st.coupled) {
    typename World<R>::Coupled c{k.index, k.generation, k.points, k.probes, k.probe_spacing};
    w.coupled.push_back(std::move(c));
  }
  w.terrains = st.terrains;
  w.neighbors_built = false;
  w.settle_walls();
}

}  // namespace oracle
