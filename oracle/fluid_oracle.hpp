// TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// CPU restatement of solunar's SPH particle-fluid step, the oracle that the
// HIP path is checked against. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product library never links
// or calls it.
//
// It follows, in operation order, the reference files
//   src/fluid/mod.rs       (settings, kernel, grid, neighbour lists, density,
//                           state hash, rest mass, chunked jobs)
//   src/fluid/step.rs      (WCSPH step, viscosity, wall viscosity, advance)
//   src/fluid/dfsph.rs     (DFSPH step, factors, density rate, Jacobi solve)
//   src/fluid/boundary.rs  (walls, containers, settle order)
//   src/snapshot/bytes.rs:6-13 (FNV-1a)
//   src/dispatch.rs, src/error.rs
// with each function citing the lines it restates. `R` is float for the
// reference arithmetic and double for the FP64 yardstick; with R = float and
// -ffp-contract=off every operation is the IEEE single operation the Rust
// code performs, under these assumptions about the `algebrix` crate that the
// reference cannot pin (its source is not available):
//   * Vec3 / f32 divides each component; dot and length_squared are
//     (x*x' + y*y') + z*z'; length is the sqrt of that;
//   * normalize_or(v, f) is v * (1/|v|) when 1/|v| is finite and > 0, else f;
//   * clamp is max then min per component (NaN-ignoring, as f32::max/min);
//   * powi(n) is LLVM's square-and-multiply: x^3 = x*x^2, x^4 = (x^2)^2,
//     x^7 = (x*x^2)*x^4;
//   * f32 `sum` is a left fold.
//
// Pinning: the tests check this restatement against the reference's own
// known answers and invariants (kernel integral and slope, dense table versus
// search, brute-force neighbours, rest lattice density, thread-count
// invariance, sort invariance of the hash, refused input, Tait hydrostatic
// squeeze, ideal dam-break front, Martin-Moyce front, gate), plus FNV-1a
// known answers. The DFSPH restatement has no known answer in the reference
// (it is marked "Not ready" there); it is pinned from outside: the density
// rate is zero under rigid motion and follows the continuity equation under
// a dilation, push_apart keeps momentum and angular momentum, factors, rates
// and velocity changes equal hand-worked FP64 values for two and three
// particles and a particle by one wall point, the divergence term is zero
// below 20 neighbours and for a negative rate, and the solves end under
// their tolerance; planted mistakes 4-8 each fail one of these checks.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <optional>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

namespace oracle {

// Planted mistakes, for tests that show the oracle's checks are not blind.
// 0 = none; 1 = kernel value without its (1 + 4q) factor; 2 = pressure force
// with the sign of the pair gradient flipped; 3 = neighbour cells visited dx
// outer instead of dz outer. DFSPH: 4 = push_apart with the particle's own
// stiffness only (kappa_i for kappa_i + kappa_j); 5 = factors without the
// sum of squares; 6 = divergence rate not clamped at zero; 7 = no free-surface
// rule (crowd < 20); 8 = density rate without the walls' term.
inline int& planted_mistake() {
  static int m = 0;
  return m;
}

// src/error.rs:5-22, offset by one (0 is success)
enum Status : int32_t {
  kOk = 0,
  kInvalidHandle = 1,
  kNonFinite = 4,
  kInvalidShape = 5,
  kInvalidSettings = 7,
  kOutOfRange = 10,
  kDispatchIncomplete = 11,
  kFluidFull = 16,
  kUnsupported = 101,
};

template <class R>
struct Vec3 {
  R x, y, z;
  bool finite() const { return std::isfinite(x) && std::isfinite(y) && std::isfinite(z); }
  bool operator==(const Vec3& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const Vec3& o) const { return !(*this == o); }
};
template <class R> inline Vec3<R> operator+(Vec3<R> a, Vec3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline Vec3<R> operator-(Vec3<R> a, Vec3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline Vec3<R> operator-(Vec3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline Vec3<R> operator*(Vec3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline Vec3<R> operator/(Vec3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(Vec3<R> a, Vec3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R length_squared(Vec3<R> a) { return dot(a, a); }
template <class R> inline R length(Vec3<R> a) { return std::sqrt(length_squared(a)); }

// f32::max / f32::min: a NaN operand yields the other one.
template <class R> inline R rmax(R a, R b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
template <class R> inline R rmin(R a, R b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }
template <class R> inline Vec3<R> vmax(Vec3<R> a, Vec3<R> b) { return {rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)}; }
template <class R> inline Vec3<R> vmin(Vec3<R> a, Vec3<R> b) { return {rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)}; }
template <class R> inline Vec3<R> clamp(Vec3<R> v, Vec3<R> lo, Vec3<R> hi) { return vmin(vmax(v, lo), hi); }
template <class R> inline Vec3<R> normalize_or(Vec3<R> v, Vec3<R> fallback) {
  R rcp = R(1) / length(v);
  if (std::isfinite(rcp) && rcp > R(0)) return v * rcp;
  return fallback;
}

// `x as i32` in Rust: saturating, NaN -> 0.
inline int32_t sat_i32(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483647.0) return std::numeric_limits<int32_t>::max();
  if (v <= -2147483648.0) return std::numeric_limits<int32_t>::min();
  return static_cast<int32_t>(v);
}

enum Solver : int32_t { kWcsph = 0, kDfsph = 1 };

// FluidSettings (src/fluid/mod.rs:55-95) and the DFSPH constants of
// src/fluid/dfsph.rs:8-16 made settable. Always single precision, as in the
// reference; the FP64 world converts them exactly.
struct Settings {
  int32_t solver = kWcsph;
  float spacing = 0.05f;
  float rest_density = 1000.0f;
  float support = 2.0f;
  uint64_t max_particles = uint64_t(1) << 22;
  float sound_speed = 20.0f;
  float viscosity = 0.15f;
  Vec3<float> gravity{0.0f, -9.81f, 0.0f};
  float courant = 0.4f;
  uint32_t max_substeps = 2000;
  float divergence_tolerance = 1e-3f;
  float density_tolerance = 5e-4f;
  uint32_t min_iterations = 2;
  uint32_t max_iterations = 100;

  // FluidSettings::validate (src/fluid/mod.rs:98-117)
  bool valid() const {
    auto positive = [](float v) { return std::isfinite(v) && v > 0.0f; };
    bool support_ok = support >= 1.5f && support <= 4.0f;  // SUPPORT_RANGE, mod.rs:37
    return positive(spacing) && positive(rest_density) && support_ok && max_particles > 0 &&
           positive(sound_speed) && std::isfinite(viscosity) && viscosity >= 0.0f &&
           gravity.finite() && courant > 0.0f && courant <= 1.0f && max_substeps > 0 &&
           (solver == kWcsph || solver == kDfsph) && std::isfinite(divergence_tolerance) &&
           std::isfinite(density_tolerance) && max_iterations >= min_iterations;
  }
  float radius() const { return spacing * support; }  // mod.rs:120-122
};

// Wendland C2 in 3D (src/fluid/mod.rs:126-165).
template <class R>
struct Kernel {
  R radius, scale;
  explicit Kernel(R r) : radius(r) {
    R r3 = r * (r * r);  // powi(3)
    R two_pi = R(2) * R(kPi);
    scale = R(21) / (two_pi * r3);
  }
  static constexpr double kPi = std::is_same<R, float>::value ? double(3.14159265358979323846f) : 3.14159265358979323846;
  R value(R distance) const {
    R q = distance / radius;
    if (q >= R(1)) return R(0);
    R t = R(1) - q;
    R t2 = t * t;
    R t4 = t2 * t2;  // powi(4)
    if (planted_mistake() == 1) return scale * t4;
    return scale * t4 * (R(1) + R(4) * q);
  }
  R slope(R distance) const {
    R q = distance / radius;
    if (q >= R(1)) return R(0);
    R t = R(1) - q;
    R t3 = t * (t * t);  // powi(3)
    return scale * R(-20) * q * t3 / radius;
  }
  Vec3<R> gradient(Vec3<R> offset) const {
    R distance = length(offset);
    if (distance <= R(0)) return {R(0), R(0), R(0)};
    return offset * (slope(distance) / distance);
  }
};

struct CellKey {
  int32_t x, y, z;
  bool operator==(const CellKey& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const CellKey& o) const { return !(*this == o); }
  bool operator<(const CellKey& o) const { return std::tie(x, y, z) < std::tie(o.x, o.y, o.z); }
};

struct Range {
  uint32_t begin, end;
  bool operator==(const Range& o) const { return begin == o.begin && end == o.end; }
};

// Grid (src/fluid/mod.rs:169-262): occupied cells in key order with their
// ranges, plus a dense slot table when the box is small enough.
struct Grid {
  std::vector<std::pair<CellKey, Range>> cells;
  std::vector<uint32_t> dense;
  CellKey low{0, 0, 0};
  int64_t span[3] = {0, 0, 0};

  // Grid::build (mod.rs:182-197); keys must be sorted.
  template <class It>
  static Grid build(It first, It last) {
    Grid g;
    uint32_t index = 0;
    for (It it = first; it != last; ++it, ++index) {
      CellKey cell = *it;
      if (!g.cells.empty() && g.cells.back().first == cell) {
        g.cells.back().second.end = index + 1;
      } else {
        g.cells.push_back({cell, {index, index + 1}});
      }
    }
    g.index();
    return g;
  }

  // Grid::index (mod.rs:199-227)
  void index() {
    if (cells.empty()) return;
    CellKey lo = cells.front().first, hi = lo;
    for (auto& c : cells) {
      lo = {std::min(lo.x, c.first.x), std::min(lo.y, c.first.y), std::min(lo.z, c.first.z)};
      hi = {std::max(hi.x, c.first.x), std::max(hi.y, c.first.y), std::max(hi.z, c.first.z)};
    }
    auto side = [](int32_t a, int32_t b) { return int64_t(b) - int64_t(a) + 1; };
    int64_t s[3] = {side(lo.x, hi.x), side(lo.y, hi.y), side(lo.z, hi.z)};
    // saturating products
    auto sat_mul = [](int64_t a, int64_t b) -> int64_t {
      __int128 p = (__int128)a * b;
      if (p > std::numeric_limits<int64_t>::max()) return std::numeric_limits<int64_t>::max();
      return (int64_t)p;
    };
    int64_t volume = sat_mul(sat_mul(s[0], s[1]), s[2]);
    if (volume > 64 * (int64_t)cells.size() + 4096) return;  // DENSE_PER_CELL, DENSE_SLACK
    dense.assign((size_t)volume, UINT32_MAX);
    for (size_t slot = 0; slot < cells.size(); ++slot) {
      const CellKey& c = cells[slot].first;
      int64_t at = (int64_t(c.x - lo.x) * s[1] + int64_t(c.y - lo.y)) * s[2] + int64_t(c.z - lo.z);
      dense[(size_t)at] = (uint32_t)slot;
    }
    low = lo;
    span[0] = s[0];
    span[1] = s[1];
    span[2] = s[2];
  }

  // Grid::range (mod.rs:229-247)
  std::optional<Range> range(CellKey key) const {
    size_t slot;
    if (dense.empty()) {
      auto it = std::lower_bound(cells.begin(), cells.end(), key,
                                 [](const std::pair<CellKey, Range>& c, const CellKey& k) { return c.first < k; });
      if (it == cells.end() || it->first != key) return std::nullopt;
      slot = (size_t)(it - cells.begin());
    } else {
      auto offset = [](int32_t v, int32_t lo, int64_t sp, int64_t& out) {
        int64_t at = int64_t(v) - int64_t(lo);
        if (at < 0 || at >= sp) return false;
        out = at;
        return true;
      };
      int64_t x, y, z;
      if (!offset(key.x, low.x, span[0], x) || !offset(key.y, low.y, span[1], y) || !offset(key.z, low.z, span[2], z))
        return std::nullopt;
      uint32_t s = dense[(size_t)((x * span[1] + y) * span[2] + z)];
      if (s == UINT32_MAX) return std::nullopt;
      slot = s;
    }
    return cells[slot].second;
  }

  // Grid::around (mod.rs:250-261): dz outer, dy, dx inner.
  void around(CellKey c, std::vector<Range>& out) const {
    out.clear();
    if (planted_mistake() == 3) {
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            auto r = range({c.x + dx, c.y + dy, c.z + dz});
            if (r) out.push_back(*r);
          }
      return;
    }
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          auto r = range({c.x + dx, c.y + dy, c.z + dz});
          if (r) out.push_back(*r);
        }
  }
};

// Lists (src/fluid/mod.rs:279-308)
struct Lists {
  std::vector<uint32_t> start;
  std::vector<uint32_t> items;
  size_t count(size_t index) const {
    if (index + 1 >= start.size()) return 0;
    return start[index + 1] - start[index];
  }
};

// in_chunks / in_chunks_of (src/fluid/mod.rs:711-740) on `threads` workers
// with the CHUNK of mod.rs:36. Work writes per-index results, so the merge in
// chunk order is the identity. `skip` drops jobs, like dispatch::testing::
// SkipsJobs (src/dispatch.rs:50-65), to exercise DispatchIncomplete.
struct Dispatch {
  static constexpr size_t kChunk = 512;
  int threads = 1;
  std::function<bool(size_t)> skip;

  // Returns false when a job was skipped (DispatchIncomplete).
  template <class F>
  bool run(size_t count, F&& work) const {
    size_t jobs = (count + kChunk - 1) / kChunk;
    std::vector<uint8_t> done(jobs, 0);
    auto job = [&](size_t j) {
      size_t b = j * kChunk, e = std::min(count, b + kChunk);
      work(b, e);
      done[j] = 1;
    };
    if (threads <= 1 || jobs <= 1) {
      for (size_t j = 0; j < jobs; ++j)
        if (!(skip && skip(j))) job(j);
    } else {
      // run backwards like dispatch::testing::Threads (src/dispatch.rs:28-48)
      std::atomic<size_t> next{0};
      std::vector<std::thread> pool;
      int t = std::min<int>(threads, (int)jobs);
      for (int w = 0; w < t; ++w)
        pool.emplace_back([&]() {
          for (;;) {
            size_t k = next.fetch_add(1);
            if (k >= jobs) break;
            size_t j = jobs - 1 - k;
            if (!(skip && skip(j))) job(j);
          }
        });
      for (auto& th : pool) th.join();
    }
    for (uint8_t d : done)
      if (!d) return false;
    return true;
  }
};

// ---- coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) ----

// Shapes a body couples with (src/shape.rs:8-49): a cuboid (a, b, c its half
// extents), a sphere (a its radius), a capsule along local y (a its radius,
// b the half height of its cylinder).
enum ShapeKind : int32_t { kCuboid = 0, kSphere = 1, kCapsule = 2 };
struct Shape {
  int32_t kind;
  float a, b, c;
};

// Where the rigid world has a coupled body this step (SubstepPose,
// src/substep.rs) and whether it takes pushes (a dynamic body;
// PhysicsWorld::apply_force fails on the others).
struct Pose {
  uint32_t index, generation;
  double position[3];
  float rotation[4];  // x, y, z, w
  float linear[3], angular[3];
  int32_t dynamic;
};

// What the water gave one coupled body over a step: the lift of lift.rs
// (when the body is dynamic and so took it) and the damping of mover_push,
// both as impulses (src/fluid/couple.rs:236-251).
struct BodyPush {
  uint32_t index, generation;
  int32_t lifted;
  float lift_impulse[3], lift_angular[3];
  float impulse[3], angular[3];
};

// Most probes one body takes (src/fluid/lift.rs:8).
constexpr size_t kMaxProbes = 4096;

// Vec3::cross as glam writes it: (y z' - y' z, z x' - z' x, x y' - x' y).
template <class R>
inline Vec3<R> cross(Vec3<R> a, Vec3<R> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

// Quat::rotate_vec3 of `algebrix`, whose source is not available: taken to
// be glam's v (w^2 - |u|^2) + u (2 (v.u)) + (u x v)(2 w), u = (x, y, z).
template <class R>
inline Vec3<R> rotate(const float q[4], Vec3<R> v) {
  Vec3<R> u{R(q[0]), R(q[1]), R(q[2])};
  R w = R(q[3]);
  R b2 = dot(u, u);
  return v * (w * w - b2) + u * (dot(v, u) * R(2)) + cross(u, v) * (w * R(2));
}

// f32::total_cmp / f64::total_cmp as an integer key.
inline int32_t total_key(float v) {
  int32_t b;
  std::memcpy(&b, &v, 4);
  return b ^ (int32_t)((uint32_t)(b >> 31) >> 1);
}
inline int64_t total_key(double v) {
  int64_t b;
  std::memcpy(&b, &v, 8);
  return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);
}

// f32::clamp: NaN passes through.
template <class R>
inline R clamp_rust(R v, R lo, R hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

// shape_points (src/fluid/couple.rs:443-502): lattice points `spacing`
// apart inside the shape, no deeper than `reach`, centred on it, loops y, z,
// x. A capsule with a negative or non-finite half height is refused as
// Shape::validate refuses it (src/shape.rs:59-62).
inline int32_t shape_points(const Shape& s, float spacing, float reach, size_t limit, std::vector<Vec3<float>>& out) {
  using F = Vec3<float>;
  out.clear();
  F half;
  switch (s.kind) {
    case kCuboid:
      half = {s.a, s.b, s.c};
      break;
    case kSphere:
      half = {s.a, s.a, s.a};
      break;
    case kCapsule:
      if (!std::isfinite(s.b) || s.b < 0.0f) return kInvalidShape;
      half = {s.a, s.b + s.a, s.a};
      break;
    default:
      return kInvalidShape;
  }
  if (!half.finite() || rmin(half.x, rmin(half.y, half.z)) <= 0.0f) return kInvalidShape;
  auto count = [&](float extent) -> int64_t {
    float f = std::floor(2.0f * extent / spacing);
    int64_t c = std::isnan(f) ? 0 : (f >= 9.2e18f ? INT64_MAX : (f <= -9.2e18f ? INT64_MIN : (int64_t)f));
    return std::min<int64_t>(std::max<int64_t>(c, 0), 1 << 12);
  };
  int64_t cx = count(half.x), cy = count(half.y), cz = count(half.z);
  if (cx * cy * cz > (int64_t)limit) return kFluidFull;
  auto start = [&](int64_t cells) { return -((float)cells - 1.0f) * spacing / 2.0f; };
  F origin{start(cx), start(cy), start(cz)};
  for (int64_t y = 0; y < cy; ++y)
    for (int64_t z = 0; z < cz; ++z)
      for (int64_t x = 0; x < cx; ++x) {
        F at = origin + F{(float)x, (float)y, (float)z} * spacing;
        float inside;
        if (s.kind == kCuboid) {
          F room{half.x - std::fabs(at.x), half.y - std::fabs(at.y), half.z - std::fabs(at.z)};
          inside = rmin(rmin(room.x, room.y), room.z);
        } else if (s.kind == kSphere) {
          inside = s.a - length(at);
        } else {
          float yy = clamp_rust(at.y, -s.b, s.b);
          inside = s.a - length(at - F{0.0f, yy, 0.0f});
        }
        if (inside >= 0.0f && inside <= reach) out.push_back(at);
      }
  return kOk;
}

// ---- terrain ground (src/fluid/terrain.rs) ----

// Most lattice nodes one terrain region may sample (terrain.rs:9), and the
// samples of the ground's damping table (terrain.rs:11).
constexpr size_t kMaxTerrainNodes = size_t(1) << 24;
constexpr int kMapSamples = 256;

// TerrainField (terrain.rs:35-42): the signed distance at nodes `step`
// apart from `low`, node (x, y, z) at (y * nz + z) * nx + x.
struct TerrainField {
  uint32_t owner;
  Vec3<float> low;
  size_t nodes[3];
  float step;
  std::vector<float> distances;

  // read (terrain.rs:101-136): trilinear value and its change per axis
  template <class R>
  bool read(Vec3<R> point, R& value, Vec3<R>& change) const {
    R st = R(step);
    Vec3<R> local = Vec3<R>{point.x - R(low.x), point.y - R(low.y), point.z - R(low.z)} / st;
    auto inside = [](R v, size_t n) { return v >= R(0) && v <= R(n - 1); };
    if (!(inside(local.x, nodes[0]) && inside(local.y, nodes[1]) && inside(local.z, nodes[2]))) return false;
    auto base = [](R v, size_t n) { return std::min<size_t>((size_t)std::floor(v), n - 2); };
    size_t x = base(local.x, nodes[0]), y = base(local.y, nodes[1]), z = base(local.z, nodes[2]);
    R fx = local.x - R(x), fy = local.y - R(y), fz = local.z - R(z);
    auto corner = [&](size_t dx, size_t dy, size_t dz) {
      return R(distances[((y + dy) * nodes[2] + (z + dz)) * nodes[0] + (x + dx)]);
    };
    auto lerp = [](R a, R b, R t) { return a + (b - a) * t; };
    auto face = [&](size_t dy, size_t dz, R& v, R& g) {
      R a = corner(0, dy, dz), b = corner(1, dy, dz);
      v = lerp(a, b, fx);
      g = b - a;
    };
    R v00, g00, v10, g10, v01, g01, v11, g11;
    face(0, 0, v00, g00);
    face(1, 0, v10, g10);
    face(0, 1, v01, g01);
    face(1, 1, v11, g11);
    R v0 = lerp(v00, v10, fy), v1 = lerp(v01, v11, fy);
    value = lerp(v0, v1, fz);
    change = Vec3<R>{lerp(lerp(g00, g10, fy), lerp(g01, g11, fy), fz), lerp(v10 - v00, v11 - v01, fz), v1 - v0} / st;
    return true;
  }

  // sample (terrain.rs:138-142): distance and the way out of the ground
  template <class R>
  bool sample(Vec3<R> point, R& distance, Vec3<R>& normal) const {
    Vec3<R> change;
    if (!read(point, distance, change)) return false;
    normal = normalize_or(change, Vec3<R>{R(0), R(1), R(0)});
    return true;
  }
};

// GroundSites (terrain.rs:173-247): the water's lattice around a particle
// within the kernel with the density each site adds, and the damping weight
// into flat ground by distance.
template <class R>
struct GroundSites {
  R spacing, reach, low, step;
  struct Site {
    Vec3<R> offset;
    R weight, far;
  };
  std::vector<Site> offsets;
  std::vector<R> damping;

  template <class K>
  GroundSites(const K& kernel, R mass, R spacing_, R radius) : spacing(spacing_) {
    int r = (int)std::ceil(radius / spacing);
    for (int z = -r; z <= r; ++z)
      for (int y = -r; y <= r; ++y)
        for (int x = -r; x <= r; ++x) {
          Vec3<R> offset = Vec3<R>{R(x), R(y), R(z)} * spacing;
          R far = length(offset);
          if (far < radius) offsets.push_back({offset, mass * kernel.value(far), far});
        }
    low = -spacing / R(2);
    step = (radius - low) / R(kMapSamples - 1);
    R soften = R(0.01f) * radius * radius;
    int across = r + 1;
    for (int sample = 0; sample < kMapSamples; ++sample) {
      R distance = low + step * R(sample);
      R weight = R(0);
      for (int layer = 0;; ++layer) {
        R depth = distance + spacing * (R(layer) + R(0.5f));
        if (depth >= radius) break;
        for (int a = -across; a <= across; ++a)
          for (int b = -across; b <= across; ++b) {
            R flat = R(a * a + b * b) * spacing * spacing;
            R squared = flat + depth * depth;
            R far = std::sqrt(squared);
            if (far >= radius || far <= R(0)) continue;
            weight += mass * depth * depth * -kernel.slope(far) / (far * (squared + soften));
          }
      }
      damping.push_back(weight);
    }
    reach = radius + spacing / R(2);
  }

  // damping (terrain.rs:236-245)
  R damping_at(R distance) const {
    R at = (rmax(distance, low) - low) / step;
    R f = std::floor(at);
    size_t below = std::isnan(f) ? 0 : (f >= R(1e18) ? SIZE_MAX : (size_t)f);
    if (below >= size_t(kMapSamples) - 1) return R(0);
    R part = at - R(below);
    return damping[below] + (damping[below + 1] - damping[below]) * part;
  }
};

// solid (terrain.rs:249-252)
template <class R>
inline R solid(R distance, R size) {
  R v = R(0.5f) - distance / size;
  return clamp_rust(v, R(0), R(1));
}

// FluidReport (src/fluid/step.rs:28-49), iterations split by solve.
struct Report {
  uint32_t substeps = 0;
  uint32_t capped = 0;
  float max_compression = 0;
  float max_speed = 0;
  uint32_t held_back = 0;
  uint32_t iterations = 0;
  uint32_t emitted = 0, drained = 0, lost = 0;
  float sound_speed = 0;
  uint32_t iterations_divergence = 0;
  uint32_t iterations_density = 0;
};

// How the DFSPH solve averages its per-particle errors.
//  kReferenceF32: the reference's f32 left fold in particle order
//                 (src/fluid/dfsph.rs:263-267).
//  kFixedTreeF64: the device path's FP64 sum of fixed shape (tile_tree_sum).
//                 Count of errors > 0 is exact.
enum ResidualMode : int32_t { kReferenceF32 = 0, kFixedTreeF64 = 1 };

// An emitter's lattice slots, at most (the reference sizes its slot flags
// by the box, up to 4096 per axis; past this a box is refused).
constexpr int64_t kEmitterSlotsMax = int64_t(1) << 28;

// The device's FP64 stop sum (kernels_dfsph.hip, fold_tiles and the tile
// sums of its rate pass), over the particles' cells in particle order (sorted
// by cell, then id). The cell box is the fluid's: span s = max - min + 1 per
// axis. Solve tiles are 4 x 2 cell columns by tz cells, ntx = ceil(sx/4),
// nty = ceil(sy/2), ntz = ceil(sz/16), tz = ceil(sz/ntz), numbered
// (tx * nty + ty) * ntz + tz_index. A tile's particles in particle order are
// q = 0, 1, ...: accumulator t of 1024 adds e[q] for q = t, t + 1024, ... in
// order from 0.0; the 16 groups of 64 fold by halving (s[l] += s[l + h],
// h = 32..1), their 16 results likewise (h = 8..1): the tile's sum. Lane l of
// 64 adds tile sums l, l + 64, ... in order from 0.0; the lanes fold by
// halving (h = 32..1).
template <class E>
inline double tile_tree_sum(const std::vector<E>& e, const std::vector<CellKey>& cells) {
  const size_t n = e.size();
  if (n == 0) return 0.0;
  int64_t lo[3] = {cells[0].x, cells[0].y, cells[0].z}, hi[3] = {lo[0], lo[1], lo[2]};
  for (const CellKey& c : cells) {
    const int64_t v[3] = {c.x, c.y, c.z};
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], v[a]);
      hi[a] = std::max(hi[a], v[a]);
    }
  }
  const uint64_t sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
  const uint64_t ntx = (sx + 3) / 4, nty = (sy + 1) / 2, ntz = (sz + 15) / 16, tz = (sz + ntz - 1) / ntz;
  const uint64_t tiles = ntx * nty * ntz;
  // each tile's particles, in particle order (a stable counting sort)
  std::vector<uint64_t> of(n), start(tiles + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    const uint64_t x = cells[i].x - lo[0], y = cells[i].y - lo[1], z = cells[i].z - lo[2];
    of[i] = ((x / 4) * nty + y / 2) * ntz + z / tz;
    ++start[of[i] + 1];
  }
  for (uint64_t t = 0; t < tiles; ++t) start[t + 1] += start[t];
  std::vector<uint64_t> at(start.begin(), start.end() - 1);
  std::vector<size_t> members(n);
  for (size_t i = 0; i < n; ++i) members[at[of[i]]++] = i;
  std::vector<double> tsum(tiles);
  std::vector<double> acc(1024);
  for (uint64_t t = 0; t < tiles; ++t) {
    std::fill(acc.begin(), acc.end(), 0.0);
    for (uint64_t q = 0; q < start[t + 1] - start[t]; ++q) acc[q % 1024] += (double)e[members[start[t] + q]];
    double ws[16];
    for (int w = 0; w < 16; ++w) {
      double* s = &acc[64 * w];
      for (int h = 32; h >= 1; h >>= 1)
        for (int l = 0; l < h; ++l) s[l] = s[l] + s[l + h];
      ws[w] = s[0];
    }
    for (int h = 8; h >= 1; h >>= 1)
      for (int w = 0; w < h; ++w) ws[w] = ws[w] + ws[w + h];
    tsum[t] = ws[0];
  }
  double lane[64];
  for (int l = 0; l < 64; ++l) {
    double s = 0.0;
    for (uint64_t k = (uint64_t)l; k < tiles; k += 64) s += tsum[k];
    lane[l] = s;
  }
  for (int h = 32; h >= 1; h >>= 1)
    for (int l = 0; l < h; ++l) lane[l] = lane[l] + lane[l + h];
  return lane[0];
}

// One Jacobi iteration's stop inputs, for logging the two averages side by
// side (the f32 fold the reference uses and the FP64 sum the device uses).
struct SolveLog {
  int32_t solve;  // 0 divergence, 1 density
  uint32_t iteration;
  uint32_t loaded;
  float average_f32;
  double average_f64;
};

template <class R>
class World {
 public:
  using V = Vec3<R>;
  using Real = R;
  static V cvt(Vec3<float> v) { return {R(v.x), R(v.y), R(v.z)}; }

  Settings settings;
  Kernel<R> kernel;
  R mass;
  std::vector<uint32_t> ids;
  std::vector<V> positions, velocities;
  std::vector<R> densities;
  std::vector<V> forces;  // last pushes, for parity readback
  uint32_t next_id = 0;
  Grid grid;
  Lists neighbors, wall_neighbors;
  bool neighbors_built = false;

  // Walls (src/fluid/boundary.rs:23-33)
  struct Walls {
    std::vector<V> positions;
    std::vector<uint32_t> owners;
    Grid grid;
    uint32_t next = 0;
    std::optional<std::pair<V, V>> bounds;
    std::vector<std::tuple<uint32_t, V, V>> containers;
  } walls;

  // Coupled (src/fluid/couple.rs:15-32): wall points and probes in body
  // axes, where the body was and how fast it went this step, what the
  // damping did to it, how far its centre has gone.
  struct Coupled {
    uint32_t index, generation;
    std::vector<Vec3<float>> points, probes;
    float probe_spacing;
    V center{}, velocity{}, impulse{}, angular{}, moved{};
  };
  std::vector<Coupled> coupled;
  // Movers (src/fluid/couple.rs:34-47): every coupled body's points this
  // step, sorted by (cell, owner, place in the previous order).
  struct Movers {
    std::vector<V> start, positions, velocities;
    std::vector<uint32_t> owners;
    Grid grid;
  } movers;
  Lists mover_neighbors;
  // terrain ground (src/fluid/mod.rs:320-335)
  std::vector<TerrainField> terrains;
  GroundSites<R> ground_sites;
  std::vector<V> ground_slopes;

  Dispatch dispatch;
  // called before (false) and after (true) every WCSPH substep with its
  // length; the adjoint tape (adjoint_oracle.hpp) records through it
  std::function<void(bool, R)> substep_hook;
  ResidualMode residual_mode = kReferenceF32;
  std::vector<SolveLog> solve_log;
  bool log_solves = false;
  std::string detail;

  explicit World(const Settings& s)
      : settings(s),
        kernel(R(s.radius())),
        mass(rest_mass()),
        ground_sites(kernel, mass, R(s.spacing), R(s.radius())) {}

  R radius() const { return R(settings.radius()); }
  size_t len() const { return ids.size(); }

  // rest_mass (src/fluid/mod.rs:694-709)
  R rest_mass() const {
    Kernel<R> k(radius());
    int reach = (int)std::ceil(settings.support);
    R sum = R(0);
    R spacing = R(settings.spacing);
    for (int z = -reach; z <= reach; ++z)
      for (int y = -reach; y <= reach; ++y)
        for (int x = -reach; x <= reach; ++x) {
          V offset = V{R(x), R(y), R(z)} * spacing;
          sum += k.value(length(offset));
        }
    return R(settings.rest_density) / sum;
  }

  // add_particle (src/fluid/mod.rs:412-426)
  int32_t add_particle(V position, V velocity, uint32_t* id_out) {
    if (!position.finite() || !velocity.finite()) {
      detail = "fluid particle";
      return kNonFinite;
    }
    if (len() >= settings.max_particles || next_id == UINT32_MAX) return kFluidFull;
    uint32_t id = next_id++;
    ids.push_back(id);
    positions.push_back(position);
    velocities.push_back(velocity);
    densities.push_back(R(0));
    neighbors_built = false;
    if (id_out) *id_out = id;
    return kOk;
  }

  // ---- flows and the domain (src/fluid/flow.rs, src/fluid/mod.rs:428-486) ----

  struct Flow {
    uint32_t id;
    Vec3<float> low, high;
    bool drain;
    Vec3<float> velocity;  // an emitter's
  };
  std::vector<Flow> flow_list;
  uint32_t flow_next = 0;
  bool has_domain = false;
  Vec3<float> domain_low{0, 0, 0}, domain_high{0, 0, 0};

  // set_domain (src/fluid/mod.rs:433-444); null bounds clear it
  int32_t set_domain(const Vec3<float>* low, const Vec3<float>* high) {
    if (low && high) {
      if (!low->finite() || !high->finite()) {
        detail = "fluid domain";
        return kNonFinite;
      }
      if (high->x <= low->x || high->y <= low->y || high->z <= low->z) return kOutOfRange;
      has_domain = true;
      domain_low = *low;
      domain_high = *high;
    } else {
      has_domain = false;
    }
    return kOk;
  }

  // add_emitter / add_drain / add_flow (src/fluid/flow.rs:92-139)
  int32_t add_flow(Vec3<float> low, Vec3<float> high, bool drain, Vec3<float> velocity, uint32_t* id_out) {
    if (!drain && !velocity.finite()) {
      detail = "emitter velocity";
      return kNonFinite;
    }
    if (!low.finite() || !high.finite()) {
      detail = "flow box";
      return kNonFinite;
    }
    if (high.x <= low.x || high.y <= low.y || high.z <= low.z) return kOutOfRange;
    if (flow_next == UINT32_MAX) return kFluidFull;
    uint32_t id = flow_next++;
    flow_list.push_back({id, low, high, drain, velocity});
    if (id_out) *id_out = id;
    return kOk;
  }

  // remove_flow (src/fluid/flow.rs:113-118)
  bool remove_flow(uint32_t id) {
    size_t before = flow_list.size();
    flow_list.erase(std::remove_if(flow_list.begin(), flow_list.end(), [&](const Flow& f) { return f.id == id; }),
                    flow_list.end());
    return flow_list.size() != before;
  }

  // keep_only (src/fluid/mod.rs:468-489): the kept particles stay in order
  void keep_only(const std::vector<char>& keep) {
    size_t w = 0;
    for (size_t i = 0; i < len(); ++i) {
      if (!keep[i]) continue;
      ids[w] = ids[i];
      positions[w] = positions[i];
      velocities[w] = velocities[i];
      densities[w] = densities[i];
      ++w;
    }
    ids.resize(w);
    positions.resize(w);
    velocities.resize(w);
    densities.resize(w);
    neighbors_built = false;
  }

  // run_flows (src/fluid/flow.rs:144-223): drains, then emitters, each in
  // the order they were added. A full world stops an emitter with an error,
  // the particles it added before staying.
  int32_t run_flows(uint32_t& emitted, uint32_t& drained) {
    emitted = drained = 0;
    if (flow_list.empty()) return kOk;
    const std::vector<Flow> list = flow_list;
    for (const Flow& f : list) {
      if (!f.drain) continue;
      std::vector<char> keep(len());
      size_t gone = 0;
      for (size_t i = 0; i < len(); ++i) {
        const V& at = positions[i];
        bool inside = at.x >= R(f.low.x) && at.y >= R(f.low.y) && at.z >= R(f.low.z) && at.x <= R(f.high.x) &&
                      at.y <= R(f.high.y) && at.z <= R(f.high.z);
        keep[i] = !inside;
        gone += inside;
      }
      if (gone > 0) {
        drained += (uint32_t)gone;
        keep_only(keep);
      }
    }
    const R spacing = R(settings.spacing);
    for (const Flow& f : list) {
      if (f.drain) continue;
      const V low = cvt(f.low), size = cvt(f.high) - cvt(f.low);
      auto count = [&](R extent) -> int64_t {
        R c = std::floor(extent / spacing);
        int64_t v = std::isnan(c) ? 0 : (c >= R(9.2e18) ? INT64_MAX : (c <= R(-9.2e18) ? INT64_MIN : (int64_t)c));
        return std::min<int64_t>(std::max<int64_t>(v, 0), 1 << 12);
      };
      const int64_t cx = count(size.x), cy = count(size.y), cz = count(size.z);
      const int64_t total = cx * cy * cz;
      if (total == 0) continue;
      if (total > (int64_t)kEmitterSlotsMax) {
        detail = "emitter box";
        return kOutOfRange;
      }
      // CLEAR_SHARE = 0.9 (flow.rs:8); powi(2) as one product
      const R near = R(0.9f) * spacing;
      const R clear = near * near;
      std::vector<char> taken((size_t)total, 0);
      for (const V& at : positions) {
        const V local = (at - low) / spacing - V{R(0.5), R(0.5), R(0.5)};
        auto nearest = [](R value) -> int64_t {
          R r = std::round(value);
          return std::isnan(r) ? 0 : (r >= R(9.2e18) ? INT64_MAX : (r <= R(-9.2e18) ? INT64_MIN : (int64_t)r));
        };
        const int64_t px = nearest(local.x), py = nearest(local.y), pz = nearest(local.z);
        // a slot index this far out has no slot within one of it (and the
        // reference's +-1 around a saturated value marks nothing either)
        if (px < -1 || px > cx || py < -1 || py > cy || pz < -1 || pz > cz) continue;
        for (int64_t y = py - 1; y <= py + 1; ++y)
          for (int64_t z = pz - 1; z <= pz + 1; ++z)
            for (int64_t x = px - 1; x <= px + 1; ++x) {
              if (!(x >= 0 && x < cx && y >= 0 && y < cy && z >= 0 && z < cz)) continue;
              const V center = low + V{R(x) + R(0.5), R(y) + R(0.5), R(z) + R(0.5)} * spacing;
              if (length_squared(at - center) < clear) taken[(size_t)((y * cz + z) * cx + x)] = 1;
            }
      }
      for (int64_t y = 0; y < cy; ++y)
        for (int64_t z = 0; z < cz; ++z)
          for (int64_t x = 0; x < cx; ++x) {
            if (taken[(size_t)((y * cz + z) * cx + x)]) continue;
            const V at = low + V{R(x) + R(0.5), R(y) + R(0.5), R(z) + R(0.5)} * spacing;
            int32_t st = add_particle(at, cvt(f.velocity), nullptr);
            if (st != kOk) return st;
            ++emitted;
          }
    }
    return kOk;
  }

  // keep_in_domain (src/fluid/mod.rs:451-465): water whose clamp to the
  // domain moves it is taken out
  uint32_t keep_in_domain() {
    if (!has_domain) return 0;
    const V lo = cvt(domain_low), hi = cvt(domain_high);
    std::vector<char> keep(len());
    uint32_t lost = 0;
    for (size_t i = 0; i < len(); ++i) {
      keep[i] = clamp(positions[i], lo, hi) == positions[i];
      lost += !keep[i];
    }
    if (lost > 0) keep_only(keep);
    return lost;
  }

  // fill_box (src/fluid/mod.rs:493-524); jitter > 0 adds the seeded
  // xorshift32 offsets described in include/solunar_fluid.h
  // (sfl_fill_box_jittered). Arithmetic stays in f32: the lattice is input.
  int32_t fill_box(Vec3<float> low, Vec3<float> high, float jitter, uint32_t seed, size_t* count_out) {
    if (!low.finite() || !high.finite()) {
      detail = "fluid box";
      return kNonFinite;
    }
    if (high.x < low.x || high.y < low.y || high.z < low.z) return kOutOfRange;
    float spacing = settings.spacing;
    auto count = [&](float extent) -> uint64_t {
      float c = std::floor(extent / spacing);
      c = std::min(c, 4294967295.0f);
      return (uint64_t)sat_u64(c);
    };
    Vec3<float> size = {high.x - low.x, high.y - low.y, high.z - low.z};
    uint64_t nx = count(size.x), ny = count(size.y), nz = count(size.z);
    unsigned __int128 t = (unsigned __int128)nx * ny * nz;
    uint64_t total = t > UINT64_MAX ? UINT64_MAX : (uint64_t)t;
    uint64_t room = settings.max_particles > len() ? settings.max_particles - len() : 0;
    uint64_t ids_left = (uint64_t)(UINT32_MAX - next_id);
    if (total > room || total > ids_left) return kFluidFull;
    float half = spacing * 0.5f;
    Vec3<float> start = {low.x + half, low.y + half, low.z + half};
    uint32_t state = seed ? seed : 0x9E3779B9u;
    float amp = 2.0f * jitter * spacing;
    auto draw = [&]() {
      state ^= state << 13;
      state ^= state >> 17;
      state ^= state << 5;
      return ((float)state / 4294967295.0f - 0.5f) * amp;
    };
    for (uint64_t y = 0; y < ny; ++y)
      for (uint64_t z = 0; z < nz; ++z)
        for (uint64_t x = 0; x < nx; ++x) {
          Vec3<float> step = {(float)x * spacing, (float)y * spacing, (float)z * spacing};
          Vec3<float> at = {start.x + step.x, start.y + step.y, start.z + step.z};
          if (jitter != 0.0f) {
            float jx = draw(), jy = draw(), jz = draw();
            at = {at.x + jx, at.y + jy, at.z + jz};
          }
          int32_t st = add_particle(cvt(at), V{R(0), R(0), R(0)}, nullptr);
          if (st != kOk) return st;
        }
    if (count_out) *count_out = (size_t)total;
    return kOk;
  }

  static double sat_u64(float c) { return c < 0 ? 0.0 : (double)c; }

  // cell_of (src/fluid/mod.rs:526-533): per-component division, floor,
  // saturating cast.
  CellKey cell_of(V p) const {
    R r = radius();
    R x = p.x / r, y = p.y / r, z = p.z / r;
    return {sat_i32((double)std::floor(x)), sat_i32((double)std::floor(y)), sat_i32((double)std::floor(z))};
  }

  // `near` (src/fluid/mod.rs:264-277): candidates kept when
  // |x_j - x_i|^2 < reach, appended in range order.
  static void near(const std::vector<V>& pos, const std::vector<Range>& ranges, V at, R reach,
                   std::vector<uint32_t>& out) {
    for (const Range& r : ranges)
      for (uint32_t other = r.begin; other < r.end; ++other)
        if (length_squared(pos[other] - at) < reach) out.push_back(other);
  }

  // update_neighbors (src/fluid/mod.rs:540-608)
  int32_t update_neighbors() {
    size_t n = len();
    struct Key {
      CellKey cell;
      uint32_t id, index;
    };
    std::vector<Key> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = {cell_of(positions[i]), ids[i], (uint32_t)i};
    std::sort(order.begin(), order.end(), [](const Key& a, const Key& b) {
      return std::tie(a.cell.x, a.cell.y, a.cell.z, a.id, a.index) <
             std::tie(b.cell.x, b.cell.y, b.cell.z, b.id, b.index);
    });
    std::vector<uint32_t> nid(n);
    std::vector<V> np(n), nv(n);
    std::vector<R> nd(n);
    for (size_t k = 0; k < n; ++k) {
      uint32_t j = order[k].index;
      nid[k] = ids[j];
      np[k] = positions[j];
      nv[k] = velocities[j];
      nd[k] = densities[j];
    }
    ids.swap(nid);
    positions.swap(np);
    velocities.swap(nv);
    densities.swap(nd);
    std::vector<CellKey> keys(n);
    for (size_t k = 0; k < n; ++k) keys[k] = order[k].cell;
    grid = Grid::build(keys.begin(), keys.end());

    R r = radius();
    R reach = r * r;  // powi(2)
    std::vector<uint32_t> fcount(n), wcount(n), mcount(n);
    std::vector<std::vector<uint32_t>> fparts((n + Dispatch::kChunk - 1) / Dispatch::kChunk),
        wparts(fparts.size()), mparts(fparts.size());
    bool ok = dispatch.run(n, [&](size_t b, size_t e) {
      std::vector<Range> af, aw, am;
      std::optional<CellKey> last;
      auto& fl = fparts[b / Dispatch::kChunk];
      auto& wl = wparts[b / Dispatch::kChunk];
      auto& ml = mparts[b / Dispatch::kChunk];
      for (size_t i = b; i < e; ++i) {
        V at = positions[i];
        CellKey cell = cell_of(at);
        if (!last || *last != cell) {
          last = cell;
          grid.around(cell, af);
          walls.grid.around(cell, aw);
          movers.grid.around(cell, am);
        }
        size_t f0 = fl.size(), w0 = wl.size(), m0 = ml.size();
        near(positions, af, at, reach, fl);
        near(walls.positions, aw, at, reach, wl);
        near(movers.positions, am, at, reach, ml);
        fcount[i] = (uint32_t)(fl.size() - f0);
        wcount[i] = (uint32_t)(wl.size() - w0);
        mcount[i] = (uint32_t)(ml.size() - m0);
      }
    });
    if (!ok) return kDispatchIncomplete;
    auto join = [&](Lists& lists, const std::vector<uint32_t>& counts, std::vector<std::vector<uint32_t>>& parts) {
      lists.start.assign(1, 0);
      lists.items.clear();
      for (size_t i = 0; i < n; ++i) lists.start.push_back(lists.start.back() + counts[i]);
      for (auto& p : parts) lists.items.insert(lists.items.end(), p.begin(), p.end());
    };
    join(neighbors, fcount, fparts);
    join(wall_neighbors, wcount, wparts);
    join(mover_neighbors, mcount, mparts);
    neighbors_built = true;
    return kOk;
  }

  // compute_densities (src/fluid/mod.rs:621-675), no terrains: the ground
  // sum is empty, and adding it (+0) to a non-negative sum is exact.
  int32_t compute_densities() {
    if (!neighbors_built || neighbors.start.size() != len() + 1) return kOutOfRange;
    std::vector<R> out(len());
    std::vector<V> slopes(len());
    bool ok = dispatch.run(len(), [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        V at = positions[i];
        R sum = R(0);
        for (uint32_t k = neighbors.start[i]; k < neighbors.start[i + 1]; ++k)
          sum += kernel.value(length(positions[neighbors.items[k]] - at));
        R wsum = R(0);
        for (uint32_t k = wall_neighbors.start[i]; k < wall_neighbors.start[i + 1]; ++k)
          wsum += kernel.value(length(walls.positions[wall_neighbors.items[k]] - at));
        R msum = R(0);
        for (uint32_t k = mover_neighbors.start[i]; k < mover_neighbors.start[i + 1]; ++k)
          msum += kernel.value(length(movers.positions[mover_neighbors.items[k]] - at));
        R ground = R(0);
        V slope{R(0), R(0), R(0)};
        for (const TerrainField& field : terrains) {
          R gd;
          V gs;
          if (ground_at(field, at, gd, gs)) {
            ground = ground + gd;
            slope = slope + gs;
          }
        }
        out[i] = mass * (sum + wsum + msum) + ground;
        slopes[i] = slope;
      }
    });
    if (!ok) return kDispatchIncomplete;
    densities.swap(out);
    if (terrains.empty())
      ground_slopes.clear();
    else
      ground_slopes.swap(slopes);
    return kOk;
  }

  // TerrainField::ground (terrain.rs:144-170): the density the ground's
  // sites add at `point` and how it changes; false outside the region
  bool ground_at(const TerrainField& field, V point, R& density, V& gradient) const {
    R distance;
    V change;
    if (!field.read(point, distance, change)) return false;
    density = R(0);
    gradient = V{R(0), R(0), R(0)};
    const GroundSites<R>& g = ground_sites;
    if (distance < g.reach) {
      R open = distance - g.spacing / R(2);
      for (const auto& site : g.offsets) {
        if (site.far <= open) continue;
        R there;
        V toward;
        if (!field.read(point + site.offset, there, toward)) continue;
        R share = solid(there, g.spacing);
        density += site.weight * share;
        if (share > R(0) && share < R(1)) gradient = gradient - toward * (site.weight / g.spacing);
      }
    }
    return true;
  }

  // The lattice add_terrain samples over a box (terrain.rs:280-286):
  // ceil(extent / step) + 1 nodes per axis, step = spacing / 2.
  bool terrain_nodes(Vec3<float> low, Vec3<float> high, size_t nodes[3]) const {
    float step = settings.spacing / 2.0f;
    Vec3<float> size = high - low;
    float e[3] = {size.x, size.y, size.z};
    for (int a = 0; a < 3; ++a) {
      float c = std::fmin(std::ceil(e[a] / step), 1e7f);  // `as usize`: NaN and below 0 give 0
      nodes[a] = (std::isnan(c) || c <= 0.0f ? 0 : (size_t)c) + 1;
    }
    return true;
  }

  // add_terrain (terrain.rs:262-305) over distances the caller sampled at
  // low + (x, y, z) step, loops y, z, x (its signed-distance volume is not
  // part of this path)
  int32_t add_terrain(Vec3<float> low, Vec3<float> high, const std::vector<float>& distances, uint32_t* id_out) {
    if (!low.finite() || !high.finite()) {
      detail = "terrain box";
      return kNonFinite;
    }
    if (high.x <= low.x || high.y <= low.y || high.z <= low.z) return kOutOfRange;
    TerrainField f{};
    terrain_nodes(low, high, f.nodes);
    uint64_t total = 1;
    for (size_t n : f.nodes) total = total > UINT64_MAX / n ? UINT64_MAX : total * n;
    if (total > kMaxTerrainNodes) return kFluidFull;
    if (distances.size() != total) return kOutOfRange;
    for (float d : distances)
      if (!std::isfinite(d)) {
        detail = "terrain distance";
        return kNonFinite;
      }
    f.low = low;
    f.step = settings.spacing / 2.0f;
    f.distances = distances;
    f.owner = add_wall_points({});
    terrains.push_back(std::move(f));
    if (id_out) *id_out = terrains.back().owner;
    return kOk;
  }

  // state_hash (src/fluid/mod.rs:678-691) over the f32 values.
  uint64_t state_hash() const {
    std::vector<size_t> order(len());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ids[a] < ids[b]; });
    uint64_t h = 0xcbf29ce484222325ull;
    auto feed = [&](uint32_t w) {
      for (int b = 0; b < 4; ++b) {
        h ^= (uint64_t)((w >> (8 * b)) & 0xffu);
        h *= 0x00000100000001b3ull;
      }
    };
    for (size_t i : order) {
      feed(ids[i]);
      float vals[6] = {(float)positions[i].x, (float)positions[i].y, (float)positions[i].z,
                       (float)velocities[i].x, (float)velocities[i].y, (float)velocities[i].z};
      for (float f : vals) {
        uint32_t bits;
        std::memcpy(&bits, &f, 4);
        feed(bits);
      }
    }
    return h;
  }

  // fastest (src/fluid/step.rs:92-96)
  R fastest() const {
    R most = R(0);
    for (const V& v : velocities) most = rmax(most, length(v));
    return most;
  }

  // ---- coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) ----

  size_t coupled_points() const {
    size_t n = 0;
    for (const Coupled& k : coupled) n += k.points.size();
    return n;
  }

  // couple (src/fluid/couple.rs:165-198): wall points as deep as the kernel
  // reaches, probes through the whole volume, coarser (x1.25 a time) until
  // at most kMaxProbes; the bodies kept in handle order.
  int32_t couple(uint32_t index, uint32_t generation, const Shape& shape) {
    std::vector<Vec3<float>> points;
    int32_t st = shape_points(shape, settings.spacing, settings.radius(), (size_t)settings.max_particles, points);
    if (st != kOk) return st;
    if (points.empty()) return kInvalidShape;
    size_t used = coupled_points();
    size_t room = settings.max_particles > used ? (size_t)settings.max_particles - used : 0;
    if (points.size() > room) return kFluidFull;
    float probe_spacing = settings.spacing;
    std::vector<Vec3<float>> probes;
    for (;;) {
      st = shape_points(shape, probe_spacing, std::numeric_limits<float>::infinity(), kMaxProbes, probes);
      if (st == kFluidFull) {
        probe_spacing *= 1.25f;
        continue;
      }
      if (st != kOk) return st;
      if (probes.empty()) {
        probe_spacing = settings.spacing;
        probes = points;
      }
      break;
    }
    uncouple(index, generation);
    Coupled k{index, generation, std::move(points), std::move(probes), probe_spacing};
    coupled.push_back(std::move(k));
    std::stable_sort(coupled.begin(), coupled.end(), [](const Coupled& a, const Coupled& b) {
      return std::tie(a.index, a.generation) < std::tie(b.index, b.generation);
    });
    return kOk;
  }

  // uncouple (src/fluid/couple.rs:200-205)
  bool uncouple(uint32_t index, uint32_t generation) {
    size_t before = coupled.size();
    coupled.erase(std::remove_if(coupled.begin(), coupled.end(),
                                 [&](const Coupled& k) { return k.index == index && k.generation == generation; }),
                  coupled.end());
    return coupled.size() != before;
  }

  static const Pose* pose_of(const std::vector<Pose>& poses, uint32_t index, uint32_t generation) {
    for (const Pose& p : poses)
      if (p.index == index && p.generation == generation) return &p;
    return nullptr;
  }
  static V center_of(const Pose& p) {  // position.to_dvec3().as_vec3()
    return V{R((float)p.position[0]), R((float)p.position[1]), R((float)p.position[2])};
  }
  static V v3(const float f[3]) { return V{R(f[0]), R(f[1]), R(f[2])}; }

  // place_movers_at (src/fluid/couple.rs:258-291): bodies without a pose
  // are uncoupled; every point placed where its body is, moving with it.
  void place_movers(const std::vector<Pose>& poses) {
    coupled.erase(std::remove_if(coupled.begin(), coupled.end(),
                                 [&](const Coupled& k) { return !pose_of(poses, k.index, k.generation); }),
                  coupled.end());
    movers = Movers{};
    for (size_t owner = 0; owner < coupled.size(); ++owner) {
      Coupled& k = coupled[owner];
      const Pose& p = *pose_of(poses, k.index, k.generation);
      V center = center_of(p);
      k.center = center;
      k.velocity = v3(p.linear);
      k.impulse = V{R(0), R(0), R(0)};
      k.angular = V{R(0), R(0), R(0)};
      for (const Vec3<float>& local : k.points) {
        V arm = rotate(p.rotation, cvt(local));
        V at = center + arm;
        V velocity = v3(p.linear) + cross(v3(p.angular), arm);
        movers.start.push_back(at);
        movers.positions.push_back(at);
        movers.velocities.push_back(velocity);
        movers.owners.push_back((uint32_t)owner);
      }
    }
    carry_movers(R(0));
  }

  // carry_movers (src/fluid/couple.rs:293-319)
  void carry_movers(R elapsed) {
    size_t m = movers.start.size();
    struct Key {
      CellKey cell;
      uint32_t owner, index;
    };
    std::vector<Key> order(m);
    for (size_t i = 0; i < m; ++i) {
      V at = movers.start[i] + movers.velocities[i] * elapsed;
      movers.positions[i] = at;
      order[i] = {cell_of(at), movers.owners[i], (uint32_t)i};
    }
    std::sort(order.begin(), order.end(), [](const Key& a, const Key& b) {
      return std::tie(a.cell.x, a.cell.y, a.cell.z, a.owner, a.index) <
             std::tie(b.cell.x, b.cell.y, b.cell.z, b.owner, b.index);
    });
    Movers next;
    std::vector<CellKey> keys(m);
    for (size_t k = 0; k < m; ++k) {
      uint32_t i = order[k].index;
      next.start.push_back(movers.start[i]);
      next.positions.push_back(movers.positions[i]);
      next.velocities.push_back(movers.velocities[i]);
      next.owners.push_back(order[k].owner);
      keys[k] = order[k].cell;
    }
    next.grid = Grid::build(keys.begin(), keys.end());
    movers = std::move(next);
    for (Coupled& k : coupled) k.moved = k.velocity * elapsed;
  }

  // fastest_near_water (src/fluid/couple.rs:321-366)
  R fastest_near_water(R dt, R slow) const {
    size_t count = coupled.size();
    const R big = R(std::numeric_limits<float>::max());
    std::vector<R> fast(count, R(0));
    std::vector<std::pair<V, V>> boxes(count, {V{big, big, big}, V{-big, -big, -big}});
    for (size_t i = 0; i < movers.positions.size(); ++i) {
      uint32_t o = movers.owners[i];
      fast[o] = rmax(fast[o], length(movers.velocities[i]));
      boxes[o] = {vmin(boxes[o].first, movers.positions[i]), vmax(boxes[o].second, movers.positions[i])};
    }
    R water = fastest();
    R r = radius();
    R most = R(0);
    for (size_t b = 0; b < count; ++b) {
      R speed = fast[b];
      if (speed <= most) continue;
      if (speed <= slow) {
        most = speed;
        continue;
      }
      R reach = r + (speed + water) * dt;
      V splat{reach, reach, reach};
      CellKey lo = cell_of(boxes[b].first - splat), hi = cell_of(boxes[b].second + splat);
      bool unsorted = grid.cells.empty() && !positions.empty();
      bool wet = unsorted;
      for (size_t c = 0; !wet && c < grid.cells.size(); ++c) {
        const CellKey& k = grid.cells[c].first;
        wet = k.x >= lo.x && k.x <= hi.x && k.y >= lo.y && k.y <= hi.y && k.z >= lo.z && k.z <= hi.z;
      }
      if (wet) most = speed;
    }
    return most;
  }

  // mover_push (src/fluid/couple.rs:368-408)
  void mover_push(size_t index, R mirrored, R damping, V& push, std::vector<std::pair<V, V>>& bodies) const {
    uint32_t b = mover_neighbors.start[index], e = mover_neighbors.start[index + 1];
    if (b == e) return;
    V at = positions[index], velocity = velocities[index];
    R density = densities[index];
    R r = radius();
    R soften = R(0.01f) * r * r;
    R settles = length(cvt(settings.gravity)) * R(settings.spacing);
    for (uint32_t k = b; k < e; ++k) {
      uint32_t point = mover_neighbors.items[k];
      V offset = at - movers.positions[point];
      V relative = velocity - movers.velocities[point];
      R toward = dot(relative, offset);
      R damp = toward < R(0) ? -damping * toward / (length_squared(offset) + soften) / density : R(0);
      V gradient = kernel.gradient(offset) * mass;
      push = push - gradient * (mirrored + damp);
      if (toward * toward < settles * length_squared(offset)) continue;
      V force = gradient * (damp * mass);
      uint32_t owner = movers.owners[point];
      const Coupled& known = coupled[owner];
      V arm = movers.positions[point] - (known.center + known.moved);
      bodies[owner].first = bodies[owner].first + force;
      bodies[owner].second = bodies[owner].second + cross(arm, force);
    }
  }

  // Lift (src/fluid/lift.rs:14-22)
  struct Lift {
    size_t body;
    V impulse, angular;
    std::vector<uint32_t> near;
  };

  // lifts (src/fluid/lift.rs:25-152): buoyancy from the probes' depth under
  // the water's level around the body, drag from the flow past them, and
  // damping on the waterplane; DRAG = 1, READ_REACH = 3 (lift.rs:9-12).
  std::vector<Lift> lifts(R dt, const std::vector<Pose>& poses) const {
    std::vector<Lift> out;
    V g = cvt(settings.gravity);
    R gravity = length(g);
    V up = gravity > R(0) ? (-g) / gravity : V{R(0), R(1), R(0)};
    R reach = R(3) * radius();
    R column = radius();
    const R big = R(std::numeric_limits<float>::max());
    for (size_t b = 0; b < coupled.size(); ++b) {
      const Coupled& known = coupled[b];
      const Pose* pose = pose_of(poses, known.index, known.generation);
      if (!pose) continue;
      V center = center_of(*pose);
      std::vector<std::pair<V, V>> probes;
      for (const Vec3<float>& local : known.probes) {
        V arm = rotate(pose->rotation, cvt(local));
        probes.push_back({arm, v3(pose->linear) + cross(v3(pose->angular), arm)});
      }
      V low{big, big, big}, high{-big, -big, -big};
      for (auto& pr : probes) {
        low = vmin(low, center + pr.first);
        high = vmax(high, center + pr.first);
      }
      V splat{reach, reach, reach};
      low = low - splat;
      high = high + splat;
      std::vector<uint32_t> near;
      for (size_t i = 0; i < len(); ++i)
        if (clamp(positions[i], low, high) == positions[i]) near.push_back((uint32_t)i);
      if (near.empty()) continue;
      // the level: mean of the middle half of the columns' tops
      struct Top {
        int32_t kx, kz;
        R height;
      };
      std::vector<Top> tops;
      for (uint32_t i : near) {
        V at = positions[i];
        tops.push_back({sat_i32((double)std::floor(at.x / column)), sat_i32((double)std::floor(at.z / column)),
                        dot(at, up)});
      }
      std::sort(tops.begin(), tops.end(), [](const Top& a, const Top& b) {
        if (a.kx != b.kx) return a.kx < b.kx;
        if (a.kz != b.kz) return a.kz < b.kz;
        return total_key(a.height) > total_key(b.height);
      });
      std::vector<R> heights;
      for (size_t t = 0; t < tops.size(); ++t)
        if (t == 0 || tops[t].kx != tops[t - 1].kx || tops[t].kz != tops[t - 1].kz) heights.push_back(tops[t].height);
      std::sort(heights.begin(), heights.end(), [](R a, R b) { return total_key(a) < total_key(b); });
      size_t quarter = heights.size() / 4;
      R sum = R(0);
      for (size_t t = quarter; t < heights.size() - quarter; ++t) sum += heights[t];
      R level = sum / R(heights.size() - 2 * quarter) + R(settings.spacing / 2.0f);
      V flow{R(0), R(0), R(0)};
      uint32_t wet = 0;
      for (uint32_t i : near)
        if (dot(positions[i], up) <= level) {
          flow = flow + velocities[i];
          wet += 1;
        }
      flow = flow / R(std::max<uint32_t>(wet, 1));
      R spacing = R(known.probe_spacing);
      struct Under {
        V arm, velocity;
        R share;
      };
      std::vector<Under> under;
      for (auto& pr : probes) {
        R depth = level - dot(center + pr.first, up);
        R share = clamp_rust(depth / spacing + R(0.5f), R(0), R(1));
        if (share > R(0)) under.push_back({pr.first, pr.second, share});
      }
      R wetted = R(0);
      for (auto& u : under) wetted += u.share;
      if (wetted <= R(0)) continue;
      R volume = wetted * (spacing * (spacing * spacing));  // powi(3)
      R density = R(settings.rest_density);
      V middle{R(0), R(0), R(0)};
      for (auto& u : under) middle = middle + u.arm * u.share;
      middle = middle / wetted;
      V force{R(0), R(0), R(0)}, torque{R(0), R(0), R(0)};
      V buoyancy = up * (density * gravity * volume);
      force = force + buoyancy;
      torque = torque + cross(middle, buoyancy);
      R area = R(std::pow(volume, R(2.0f / 3.0f))) / wetted;
      for (auto& u : under) {
        V past = flow - u.velocity;
        V drag = past * (R(0.5f) * R(1) * density * area * u.share * length(past));
        force = force + drag;
        torque = torque + cross(u.arm, drag);
      }
      size_t line = 0;
      for (auto& u : under) line += u.share < R(1) ? 1 : 0;
      R plane = R(line) * spacing * spacing;
      R waves = density * plane * std::sqrt(gravity * std::sqrt(plane));
      R share = waves / R(std::max<size_t>(line, 1));
      for (auto& u : under) {
        if (!(u.share < R(1))) continue;
        V calm = up * (share * dot(flow - u.velocity, up));
        force = force + calm;
        torque = torque + cross(u.arm, calm);
      }
      out.push_back({b, force * dt, torque * dt, std::move(near)});
    }
    return out;
  }

  // give_back (src/fluid/lift.rs:154-166)
  void give_back(const Lift& lift) {
    V change = lift.impulse / (mass * R(lift.near.size()));
    for (uint32_t i : lift.near) velocities[i] = velocities[i] - change;
  }

  // step_with_bodies (src/fluid/couple.rs:207-252): the bodies stand where
  // `poses` has them; each coupled body's lift (taken when it is dynamic,
  // and given back by the water) and damping impulse go to `pushes`, in
  // handle order. Poses must be finite and name each body once.
  int32_t step_with_bodies(float dt, const std::vector<Pose>& poses, Report& report, std::vector<BodyPush>& pushes) {
    pushes.clear();
    report = Report{};
    if (settings.solver != kWcsph && !coupled.empty()) return kInvalidSettings;
    for (size_t a = 0; a < poses.size(); ++a) {
      const Pose& p = poses[a];
      bool finite = true;
      for (int c = 0; c < 3; ++c)
        finite = finite && std::isfinite(p.position[c]) && std::isfinite((float)p.position[c]) &&
                 std::isfinite(p.linear[c]) && std::isfinite(p.angular[c]);
      for (int c = 0; c < 4; ++c) finite = finite && std::isfinite(p.rotation[c]);
      if (!finite) {
        detail = "fluid body pose";
        return kNonFinite;
      }
      for (size_t b = 0; b < a; ++b)
        if (poses[b].index == p.index && poses[b].generation == p.generation) return kInvalidHandle;
    }
    place_movers(poses);
    int32_t st = step(dt, report);
    movers = Movers{};
    if (st != kOk) return st;
    for (const Coupled& k : coupled) {
      BodyPush bp{};
      bp.index = k.index;
      bp.generation = k.generation;
      pushes.push_back(bp);
    }
    if (dt > 0.0f) {
      for (const Lift& lift : lifts(R(dt), poses)) {
        const Coupled& k = coupled[lift.body];
        if (!pose_of(poses, k.index, k.generation)->dynamic) continue;
        BodyPush& bp = pushes[lift.body];
        bp.lifted = 1;
        float li[3] = {(float)lift.impulse.x, (float)lift.impulse.y, (float)lift.impulse.z};
        float la[3] = {(float)lift.angular.x, (float)lift.angular.y, (float)lift.angular.z};
        std::memcpy(bp.lift_impulse, li, sizeof li);
        std::memcpy(bp.lift_angular, la, sizeof la);
        give_back(lift);
      }
      for (size_t b = 0; b < coupled.size(); ++b) {
        Coupled& k = coupled[b];
        float im[3] = {(float)k.impulse.x, (float)k.impulse.y, (float)k.impulse.z};
        float an[3] = {(float)k.angular.x, (float)k.angular.y, (float)k.angular.z};
        std::memcpy(pushes[b].impulse, im, sizeof im);
        std::memcpy(pushes[b].angular, an, sizeof an);
        k.impulse = V{R(0), R(0), R(0)};
        k.angular = V{R(0), R(0), R(0)};
      }
    }
    return kOk;
  }

  // step (src/fluid/step.rs:58-90)
  int32_t step(float dt, Report& report) {
    report = Report{};
    if (!std::isfinite(dt) || dt < 0.0f) {
      detail = "fluid step";
      return kNonFinite;
    }
    if (dt == 0.0f) return kOk;
    if (len() > 0) {
      int32_t st = settings.solver == kWcsph ? wcsph(dt, report) : dfsph(dt, report);
      if (st != kOk) return st;
    }
    uint32_t emitted = 0, drained = 0;
    int32_t flowed = run_flows(emitted, drained);
    if (flowed != kOk) return flowed;
    report.emitted = emitted;
    report.drained = drained;
    report.lost = keep_in_domain();
    report.max_speed = (float)fastest();
    R rest = R(settings.rest_density);
    R most = R(0);
    for (R d : densities) most = rmax(most, d / rest - R(1));
    report.max_compression = (float)most;
    return kOk;
  }

  // wcsph (src/fluid/step.rs:102-116): the sound speed of the step is
  // raised to IMPACT_SOUND (4, step.rs:20) times the fastest coupled body
  // near water.
  int32_t wcsph(float dt_in, Report& report) {
    float configured = settings.sound_speed;
    R mover = fastest_near_water(R(dt_in), R(configured) / R(4));
    settings.sound_speed = (float)rmax(R(configured), R(4) * mover);
    report.sound_speed = settings.sound_speed;
    int32_t st = wcsph_steps(dt_in, mover, report);
    settings.sound_speed = configured;
    return st;
  }

  // wcsph_steps (src/fluid/step.rs:118-142)
  int32_t wcsph_steps(float dt_in, R mover, Report& report) {
    R dt = R(dt_in);
    R fast = rmax(fastest(), mover);
    R largest = R(settings.courant) * radius() / (R(settings.sound_speed) + fast);
    R wanted = rmax(std::ceil(dt / largest), R(1));
    R cap = R(settings.max_substeps);
    uint32_t substeps = (uint32_t)rmin(wanted, cap);
    report.capped = wanted > cap;
    R small = dt / R(substeps);
    for (uint32_t s = 0; s < substeps; ++s) {
      uint32_t held = 0;
      if (!coupled.empty()) carry_movers(small * R(s));
      if (substep_hook) substep_hook(false, small);
      int32_t st = wcsph_substep(small, held);
      if (st != kOk) return st;
      if (substep_hook) substep_hook(true, small);
      report.held_back += held;
    }
    report.substeps = substeps;
    return kOk;
  }

  // viscosity (src/fluid/step.rs:254-259)
  R viscosity(R dt) const {
    R r = radius();
    return rmin(R(settings.viscosity) * R(settings.sound_speed) * r, R(0.05f) * r * r / dt);
  }

  // wall_viscous (src/fluid/step.rs:262-301), no terrains.
  V wall_viscous(size_t index, R dt) const {
    R r = radius();
    R soften = R(0.01f) * r * r;
    R stable = r * r / dt;
    V at = positions[index], velocity = velocities[index];
    R density = densities[index];
    V push{R(0), R(0), R(0)};
    V away{R(0), R(0), R(0)};
    for (uint32_t k = wall_neighbors.start[index]; k < wall_neighbors.start[index + 1]; ++k)
      away = away + (-kernel.gradient(at - walls.positions[wall_neighbors.items[k]]));
    V normal = normalize_or(away, V{R(0), R(0), R(0)});
    V into = normal * rmin(dot(velocity, normal), R(0));
    R wall_viscosity = rmin(R(1) * R(settings.sound_speed) * r, R(0.25f) * stable);
    for (const TerrainField& field : terrains) {
      R distance;
      V normal;
      if (!field.sample(at, distance, normal)) continue;
      R into_ground = dot(velocity, normal);
      if (into_ground < R(0)) {
        R weight = ground_sites.damping_at(distance);
        push = push - normal * (wall_viscosity * into_ground * weight / density);
      }
    }
    for (uint32_t k = wall_neighbors.start[index]; k < wall_neighbors.start[index + 1]; ++k) {
      V offset = at - walls.positions[wall_neighbors.items[k]];
      R toward = dot(into, offset);
      if (toward < R(0)) {
        R squeeze = toward / (length_squared(offset) + soften);
        R damp = -wall_viscosity * squeeze / density;
        push = push - kernel.gradient(offset) * (mass * damp);
      }
    }
    return push;
  }

  // viscous (src/fluid/step.rs:229-251)
  V viscous(size_t index, R dt) const {
    R r = radius();
    R soften = R(0.01f) * r * r;
    R nu = viscosity(dt);
    V at = positions[index], velocity = velocities[index];
    R density = densities[index];
    V push{R(0), R(0), R(0)};
    for (uint32_t k = neighbors.start[index]; k < neighbors.start[index + 1]; ++k) {
      uint32_t other = neighbors.items[k];
      if (other == index) continue;
      V offset = at - positions[other];
      R toward = dot(velocity - velocities[other], offset);
      if (toward < R(0)) {
        R squeeze = toward / (length_squared(offset) + soften);
        R damp = -nu * squeeze * R(2) / (density + densities[other]);
        push = push - kernel.gradient(offset) * (mass * damp);
      }
    }
    return push + wall_viscous(index, dt);
  }

  // wcsph_substep (src/fluid/step.rs:144-224)
  int32_t wcsph_substep(R dt, uint32_t& held) {
    int32_t st = update_neighbors();
    if (st != kOk) return st;
    st = compute_densities();
    if (st != kOk) return st;
    size_t n = len();
    R rest = R(settings.rest_density);
    R c = R(settings.sound_speed);
    R stiffness = rest * (c * c) / R(7);  // TAIT, step.rs:9
    std::vector<R> pressures(n), scaled(n);
    for (size_t i = 0; i < n; ++i) {
      R ratio = rmin(densities[i] / rest, R(1.1f));  // PRESSURE_SQUEEZE_CAP
      R r2 = ratio * ratio;
      R r4 = r2 * r2;
      R r7 = ratio * r2 * r4;  // powi(7)
      pressures[i] = rmax(stiffness * (r7 - R(1)), R(0));
    }
    R r = radius();
    R soften = R(0.01f) * r * r;
    R nu = viscosity(dt);
    for (size_t i = 0; i < n; ++i) scaled[i] = pressures[i] / (densities[i] * densities[i]);
    V gravity = cvt(settings.gravity);
    // step.rs:163-164
    R damping = rmin(R(1) * c * r, R(0.25f) * r * r / dt);
    std::vector<V> pushes(n);
    size_t bodies = coupled.size();
    std::vector<std::vector<std::pair<V, V>>> felt((n + Dispatch::kChunk - 1) / Dispatch::kChunk);
    bool ok = dispatch.run(n, [&](size_t b, size_t e) {
      auto& mine = felt[b / Dispatch::kChunk];
      mine.assign(bodies, {V{R(0), R(0), R(0)}, V{R(0), R(0), R(0)}});
      for (size_t i = b; i < e; ++i) {
        V at = positions[i], velocity = velocities[i];
        R density = densities[i];
        R own = scaled[i];
        V push{R(0), R(0), R(0)};
        for (uint32_t k = neighbors.start[i]; k < neighbors.start[i + 1]; ++k) {
          uint32_t other = neighbors.items[k];
          V offset = at - positions[other];
          R squared = length_squared(offset);
          if (squared <= R(0)) continue;
          R distance = std::sqrt(squared);
          V gradient = offset * (kernel.slope(distance) / distance);
          if (planted_mistake() == 2) gradient = -gradient;
          R toward = rmin(dot(velocity - velocities[other], offset), R(0));
          R damp = -nu * toward / (squared + soften) * R(2) / (density + densities[other]);
          push = push - gradient * (mass * (own + scaled[other] + damp));
        }
        R mirrored = own + pressures[i] / (rest * rest);
        for (uint32_t k = wall_neighbors.start[i]; k < wall_neighbors.start[i + 1]; ++k) {
          V offset = at - walls.positions[wall_neighbors.items[k]];
          push = push - kernel.gradient(offset) * (mass * mirrored);
        }
        if (i < ground_slopes.size()) push = push - ground_slopes[i] * mirrored;
        mover_push(i, mirrored, damping, push, mine);
        pushes[i] = push + wall_viscous(i, dt) + gravity;
      }
    });
    if (!ok) return kDispatchIncomplete;
    for (size_t i = 0; i < n; ++i) velocities[i] = velocities[i] + pushes[i] * dt;
    // each chunk's damping on the bodies, in chunk order (step.rs:213-221)
    for (auto& part : felt)
      for (size_t k = 0; k < bodies; ++k) {
        coupled[k].impulse = coupled[k].impulse + part[k].first * dt;
        coupled[k].angular = coupled[k].angular + part[k].second * dt;
      }
    forces = pushes;
    return advance(dt, held);
  }

  // advance (src/fluid/step.rs:305-339)
  int32_t advance(R dt, uint32_t& held) {
    held = 0;
    R half = R(settings.spacing / 2.0f);
    V rest{half, half, half};
    for (size_t i = 0; i < len(); ++i) {
      V& at = positions[i];
      V& velocity = velocities[i];
      V next = at + velocity * dt;
      if (walls.bounds) {
        V low = walls.bounds->first, high = walls.bounds->second;
        V inside = clamp(next, low + rest, high - rest);
        if (inside != next) {
          if (clamp(next, low, high) != next) held += 1;
          if (inside.x != next.x) velocity.x = R(0);
          if (inside.y != next.y) velocity.y = R(0);
          if (inside.z != next.z) velocity.z = R(0);
          next = inside;
        }
      }
      if (terrain_contact(half, next, velocity)) held += 1;
      if (!next.finite() || !velocity.finite()) {
        detail = "fluid particle";
        return kNonFinite;
      }
      at = next;
    }
    return kOk;
  }

  // terrain_contact (terrain.rs:315-343): water closer than `rest` to a
  // terrain's surface is put back there and stops moving into it; true when
  // it got under the surface
  bool terrain_contact(R rest, V& at, V& velocity) const {
    bool under = false;
    for (const TerrainField& field : terrains) {
      R distance;
      V normal;
      if (!field.sample(at, distance, normal)) continue;
      if (distance < rest) {
        under = under || distance < R(0);
        at = at + normal * (rest - distance);
        R into = dot(velocity, normal);
        if (into < R(0)) velocity = velocity - normal * into;
      }
    }
    return under;
  }

  // dfsph (src/fluid/dfsph.rs:46-77)
  int32_t dfsph(float dt_in, Report& report) {
    // ground pushes through the weakly compressible solver's pressure only
    if (!terrains.empty()) return kInvalidSettings;
    R spacing = R(settings.spacing);
    R courant = R(settings.courant);
    R pull = length(cvt(settings.gravity));
    R left = R(dt_in);
    while (left > R(0)) {
      R small = courant * spacing / rmax(fastest(), R(1e-6f));
      if (pull > R(0)) small = rmin(small, std::sqrt(R(2) * courant * spacing / pull));
      if (report.substeps + 1 >= settings.max_substeps && small < left) {
        report.capped = 1;
        small = left;
      }
      small = rmin(small, left);
      uint32_t idiv = 0, iden = 0, held = 0;
      int32_t st = dfsph_substep(small, idiv, iden, held);
      report.held_back += held;
      report.iterations += idiv + iden;
      report.iterations_divergence += idiv;
      report.iterations_density += iden;
      if (st != kOk) return st;
      report.substeps += 1;
      left -= small;
    }
    return kOk;
  }

  // gradients (src/fluid/dfsph.rs:108-136): every listed pair's kernel
  // gradient, fluid then walls, in list order.
  std::vector<V> gfluid, gwalls;
  void gradients() {
    gfluid.resize(neighbors.items.size());
    gwalls.resize(wall_neighbors.items.size());
    dispatch.run(len(), [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        V at = positions[i];
        for (uint32_t k = neighbors.start[i]; k < neighbors.start[i + 1]; ++k)
          gfluid[k] = kernel.gradient(at - positions[neighbors.items[k]]);
        for (uint32_t k = wall_neighbors.start[i]; k < wall_neighbors.start[i + 1]; ++k)
          gwalls[k] = kernel.gradient(at - walls.positions[wall_neighbors.items[k]]);
      }
    });
  }

  // factors (src/fluid/dfsph.rs:138-163)
  std::vector<R> factors() {
    std::vector<R> out(len());
    dispatch.run(len(), [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        V sum{R(0), R(0), R(0)};
        R squares = R(0);
        for (uint32_t k = neighbors.start[i]; k < neighbors.start[i + 1]; ++k) {
          V part = gfluid[k] * mass;
          sum = sum + part;
          if (planted_mistake() != 5) squares += length_squared(part);
        }
        for (uint32_t k = wall_neighbors.start[i]; k < wall_neighbors.start[i + 1]; ++k)
          sum = sum + gwalls[k] * mass;
        R total = length_squared(sum) + squares;
        out[i] = total > R(0) ? R(1) / total : R(0);
      }
    });
    return out;
  }

  // density_rate (src/fluid/dfsph.rs:166-181)
  R density_rate(size_t i) const {
    V velocity = velocities[i];
    R rate = R(0);
    for (uint32_t k = neighbors.start[i]; k < neighbors.start[i + 1]; ++k)
      rate += dot(velocity - velocities[neighbors.items[k]], gfluid[k]);
    if (planted_mistake() != 8)
      for (uint32_t k = wall_neighbors.start[i]; k < wall_neighbors.start[i + 1]; ++k)
        rate += dot(velocity, gwalls[k]);
    return rate * mass;
  }

  // push_apart (src/fluid/dfsph.rs:184-216): the velocity changes, then
  // their addition to the velocities.
  std::vector<V> push_changes(const std::vector<R>& stiffness, R dt) const {
    std::vector<V> changes(len());
    dispatch.run(len(), [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) {
        R own = stiffness[i];
        V change{R(0), R(0), R(0)};
        for (uint32_t k = neighbors.start[i]; k < neighbors.start[i + 1]; ++k) {
          R other = planted_mistake() == 4 ? R(0) : stiffness[neighbors.items[k]];
          change = change - gfluid[k] * (mass * (own + other));
        }
        for (uint32_t k = wall_neighbors.start[i]; k < wall_neighbors.start[i + 1]; ++k)
          change = change - gwalls[k] * (mass * own);
        changes[i] = change * dt;
      }
    });
    return changes;
  }
  void push_apart(const std::vector<R>& stiffness, R dt) {
    std::vector<V> changes = push_changes(stiffness, dt);
    for (size_t i = 0; i < len(); ++i) velocities[i] = velocities[i] + changes[i];
  }

  // One particle's term of a solve iteration (src/fluid/dfsph.rs:236-256):
  // its stiffness (target times factor) and its error. `which` 0 =
  // divergence, 1 = density.
  void solve_term(int which, R dt, R alpha, size_t i, R& stiffness, R& error) const {
    R rest = R(settings.rest_density);
    R rate = density_rate(i);
    R target;
    if (which == 1) {
      R over = rmax(densities[i] + dt * rate - rest, R(0));
      target = over / (dt * dt);
      error = over / rest;
    } else {
      size_t crowd = neighbors.count(i) + wall_neighbors.count(i);
      R r = planted_mistake() == 6 ? rate : rmax(rate, R(0));
      if (crowd < 20 && planted_mistake() != 7) r = R(0);  // SURFACE_NEIGHBORS
      target = r / dt;
      error = dt * r / rest;
    }
    stiffness = target * alpha;
  }

  // Test hook: the DFSPH quantities of one substep's first solve iteration
  // at the current state, with the lists already built: densities, factors,
  // density rates, both solves' stiffness and error, and the velocity
  // changes push_apart makes for the caller's stiffness (velocities are left
  // as they are).
  struct Probe {
    std::vector<R> alpha, rate, div_stiffness, div_error, den_stiffness, den_error;
    std::vector<V> change;
  };
  int32_t dfsph_probe(R dt, const std::vector<R>* stiffness, Probe& out) {
    if (!neighbors_built) return kOutOfRange;
    int32_t st = compute_densities();
    if (st != kOk) return st;
    gradients();
    out.alpha = factors();
    size_t n = len();
    out.rate.resize(n);
    out.div_stiffness.resize(n);
    out.div_error.resize(n);
    out.den_stiffness.resize(n);
    out.den_error.resize(n);
    for (size_t i = 0; i < n; ++i) {
      out.rate[i] = density_rate(i);
      solve_term(0, dt, out.alpha[i], i, out.div_stiffness[i], out.div_error[i]);
      solve_term(1, dt, out.alpha[i], i, out.den_stiffness[i], out.den_error[i]);
    }
    if (stiffness) out.change = push_changes(*stiffness, dt);
    return kOk;
  }

  // solve (src/fluid/dfsph.rs:222-275); `which` 0 = divergence, 1 = density.
  uint32_t solve(int which, R dt, const std::vector<R>& alpha) {
    float tolerance = which == 0 ? settings.divergence_tolerance : settings.density_tolerance;
    uint32_t iterations = 0;
    size_t n = len();
    std::vector<R> stiffness(n);
    std::vector<R> errors(n);
    // the cells of the device's solve tiles (positions hold still in a solve)
    std::vector<CellKey> cells;
    if (residual_mode == kFixedTreeF64 || log_solves) {
      cells.resize(n);
      for (size_t i = 0; i < n; ++i) cells[i] = cell_of(positions[i]);
    }
    for (;;) {
      dispatch.run(n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) solve_term(which, dt, alpha[i], i, stiffness[i], errors[i]);
      });
      uint32_t loaded = 0;
      for (R e : errors) loaded += e > R(0);
      uint32_t denom = std::max<uint32_t>(loaded, 1);
      R sum32 = R(0);
      bool need32 = residual_mode == kReferenceF32 || log_solves;
      if (need32)
        for (R e : errors) sum32 += e;
      R avg32 = sum32 / R(denom);
      double avg64 = 0.0;
      bool need64 = residual_mode == kFixedTreeF64 || log_solves;
      if (need64) avg64 = tile_tree_sum(errors, cells) / (double)denom;
      if (log_solves) solve_log.push_back({which, iterations, loaded, (float)avg32, avg64});
      bool below = residual_mode == kReferenceF32 ? (avg32 <= R(tolerance)) : (avg64 <= (double)tolerance);
      bool settled = iterations >= settings.min_iterations && below;
      if (settled || iterations >= settings.max_iterations) return iterations;
      push_apart(stiffness, dt);
      iterations += 1;
    }
  }

  // dfsph_substep (src/fluid/dfsph.rs:79-103)
  int32_t dfsph_substep(R dt, uint32_t& idiv, uint32_t& iden, uint32_t& held) {
    int32_t st = update_neighbors();
    if (st != kOk) return st;
    st = compute_densities();
    if (st != kOk) return st;
    gradients();
    std::vector<R> alpha = factors();
    idiv = solve(0, dt, alpha);
    V gravity = cvt(settings.gravity);
    std::vector<V> pushes(len());
    dispatch.run(len(), [&](size_t b, size_t e) {
      for (size_t i = b; i < e; ++i) pushes[i] = viscous(i, dt) + gravity;
    });
    for (size_t i = 0; i < len(); ++i) velocities[i] = velocities[i] + pushes[i] * dt;
    forces = pushes;
    iden = solve(1, dt, alpha);
    return advance(dt, held);
  }

  // ---- walls (src/fluid/boundary.rs) ----

  static int32_t check_box(Vec3<float> low, Vec3<float> high, std::string& detail) {
    if (!low.finite() || !high.finite()) {
      detail = "fluid box";
      return kNonFinite;
    }
    if (high.x <= low.x || high.y <= low.y || high.z <= low.z) return kOutOfRange;
    return kOk;
  }

  // cells_in (boundary.rs:122-126)
  void cells_in(Vec3<float> extent, int64_t out[3]) const {
    auto count = [&](float length) -> int64_t {
      float c = std::round(length / settings.spacing);
      c = std::min(c, 1e9f);
      return (int64_t)c;
    };
    out[0] = count(extent.x);
    out[1] = count(extent.y);
    out[2] = count(extent.z);
  }

  // check_room (boundary.rs:128-137)
  int32_t check_room(int64_t count) const {
    uint64_t have = walls.positions.size();
    uint64_t room = settings.max_particles > have ? settings.max_particles - have : 0;
    if (count < 0 || (uint64_t)count > room || walls.next == UINT32_MAX) return kFluidFull;
    return kOk;
  }

  // add_box_container (boundary.rs:41-68)
  int32_t add_box_container(Vec3<float> low, Vec3<float> high, uint32_t* out) {
    int32_t st = check_box(low, high, detail);
    if (st != kOk) return st;
    int64_t c[3];
    cells_in({high.x - low.x, high.y - low.y, high.z - low.z}, c);
    int64_t layers = (int64_t)std::ceil(settings.support);
    int64_t shell = (c[0] + 2 * layers) * (c[1] + 2 * layers) * (c[2] + 2 * layers) - c[0] * c[1] * c[2];
    st = check_room(shell);
    if (st != kOk) return st;
    float spacing = settings.spacing;
    std::vector<Vec3<float>> points;
    points.reserve((size_t)shell);
    for (int64_t y = -layers; y < c[1] + layers; ++y)
      for (int64_t z = -layers; z < c[2] + layers; ++z)
        for (int64_t x = -layers; x < c[0] + layers; ++x) {
          bool inside = x >= 0 && x < c[0] && y >= 0 && y < c[1] && z >= 0 && z < c[2];
          if (!inside) {
            Vec3<float> at = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
            points.push_back({low.x + at.x * spacing, low.y + at.y * spacing, low.z + at.z * spacing});
          }
        }
    uint32_t id = add_wall_points(points);
    walls.containers.push_back({id, cvt(low), cvt(high)});
    walls.bounds = union_of(walls.containers);
    if (out) *out = id;
    return kOk;
  }

  // add_wall (boundary.rs:74-91)
  int32_t add_wall(Vec3<float> low, Vec3<float> high, uint32_t* out) {
    int32_t st = check_box(low, high, detail);
    if (st != kOk) return st;
    int64_t c[3];
    cells_in({high.x - low.x, high.y - low.y, high.z - low.z}, c);
    for (auto& v : c) v = std::max<int64_t>(v, 1);
    int64_t count = c[0] * c[1] * c[2];
    st = check_room(count);
    if (st != kOk) return st;
    float spacing = settings.spacing;
    std::vector<Vec3<float>> points;
    points.reserve((size_t)count);
    for (int64_t y = 0; y < c[1]; ++y)
      for (int64_t z = 0; z < c[2]; ++z)
        for (int64_t x = 0; x < c[0]; ++x) {
          Vec3<float> at = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
          points.push_back({low.x + at.x * spacing, low.y + at.y * spacing, low.z + at.z * spacing});
        }
    uint32_t id = add_wall_points(points);
    if (out) *out = id;
    return kOk;
  }

  // add_wall_points for caller points (the device path's boundary points).
  int32_t add_boundary_points(const std::vector<Vec3<float>>& points, uint32_t* out) {
    for (auto& p : points)
      if (!p.finite()) {
        detail = "fluid wall";
        return kNonFinite;
      }
    int32_t st = check_room((int64_t)points.size());
    if (st != kOk) return st;
    uint32_t id = add_wall_points(points);
    if (out) *out = id;
    return kOk;
  }

  // add_wall_points (boundary.rs:139-151)
  uint32_t add_wall_points(const std::vector<Vec3<float>>& points) {
    uint32_t id = walls.next++;
    for (auto& p : points) {
      walls.owners.push_back(id);
      walls.positions.push_back(cvt(p));
    }
    settle_walls();
    return id;
  }

  // remove_boundary (boundary.rs:94-115)
  bool remove_boundary(uint32_t id) {
    size_t before = walls.positions.size();
    std::vector<V> kp;
    std::vector<uint32_t> ko;
    for (size_t i = 0; i < walls.positions.size(); ++i)
      if (walls.owners[i] != id) {
        kp.push_back(walls.positions[i]);
        ko.push_back(walls.owners[i]);
      }
    walls.positions.swap(kp);
    walls.owners.swap(ko);
    bool had = false;
    std::vector<std::tuple<uint32_t, V, V>> kc;
    for (auto& c : walls.containers) {
      if (std::get<0>(c) == id)
        had = true;
      else
        kc.push_back(c);
    }
    walls.containers.swap(kc);
    walls.bounds = union_of(walls.containers);
    size_t fields = terrains.size();
    terrains.erase(std::remove_if(terrains.begin(), terrains.end(),
                                  [&](const TerrainField& f) { return f.owner == id; }),
                   terrains.end());
    bool gone = walls.positions.size() != before || had || terrains.size() != fields;
    if (gone) settle_walls();
    return gone;
  }

  // settle_walls (boundary.rs:155-178): sort by (cell, owner, x/y/z bits,
  // index).
  void settle_walls() {
    size_t n = walls.positions.size();
    struct Key {
      CellKey cell;
      uint32_t owner, bx, by, bz;
      size_t index;
    };
    auto bits = [](R v) {
      float f = (float)v;
      uint32_t b;
      std::memcpy(&b, &f, 4);
      return b;
    };
    std::vector<Key> order(n);
    for (size_t i = 0; i < n; ++i) {
      V p = walls.positions[i];
      order[i] = {cell_of(p), walls.owners[i], bits(p.x), bits(p.y), bits(p.z), i};
    }
    std::sort(order.begin(), order.end(), [](const Key& a, const Key& b) {
      return std::tie(a.cell.x, a.cell.y, a.cell.z, a.owner, a.bx, a.by, a.bz, a.index) <
             std::tie(b.cell.x, b.cell.y, b.cell.z, b.owner, b.bx, b.by, b.bz, b.index);
    });
    std::vector<V> np(n);
    std::vector<uint32_t> no(n);
    std::vector<CellKey> keys(n);
    for (size_t k = 0; k < n; ++k) {
      np[k] = walls.positions[order[k].index];
      no[k] = order[k].owner;
      keys[k] = order[k].cell;
    }
    walls.positions.swap(np);
    walls.owners.swap(no);
    walls.grid = Grid::build(keys.begin(), keys.end());
  }

  // union (boundary.rs:240-247)
  static std::optional<std::pair<V, V>> union_of(const std::vector<std::tuple<uint32_t, V, V>>& cs) {
    std::optional<std::pair<V, V>> sum;
    for (auto& c : cs) {
      V lo = std::get<1>(c), hi = std::get<2>(c);
      if (!sum)
        sum = std::make_pair(lo, hi);
      else
        sum = std::make_pair(vmin(sum->first, lo), vmax(sum->second, hi));
    }
    return sum;
  }
};

}  // namespace oracle
