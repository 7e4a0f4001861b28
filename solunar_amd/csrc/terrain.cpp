// Host side of terrain ground (src/fluid/terrain.rs): the regions a world
// samples its ground from, checked as add_terrain checks them, and the
// world's ground sites and damping table, built in the reference's f32
// arithmetic (compiled with -ffp-contract=off). The caller samples its
// signed-distance field at the nodes terrain_nodes names: the field itself
// (the reference's crate::sdf) is not part of this path.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "world.hpp"

namespace sfl {

namespace {

struct DeviceGuard {
  explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

constexpr uint64_t kMaxNodes = uint64_t(1) << 24;  // MAX_NODES, terrain.rs:9

float value_host(float distance, float radius, float scale) {
  float q = distance / radius;
  if (q >= 1.0f) return 0.0f;
  float t = 1.0f - q;
  float t2 = t * t;
  float t4 = t2 * t2;
  return scale * t4 * (1.0f + 4.0f * q);
}

float slope_host(float distance, float radius, float scale) {
  float q = distance / radius;
  if (q >= 1.0f) return 0.0f;
  float t = 1.0f - q;
  float t3 = t * (t * t);
  return scale * -20.0f * q * t3 / radius;
}

bool finite3(const sfl_vec3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

}  // namespace

#define SFL_HIP(call, what)                              \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return device_error(what, e_); \
  } while (0)

// The lattice add_terrain samples (terrain.rs:280-286): ceil(extent / step)
// + 1 nodes per axis, step = spacing / 2, `as usize` taking NaN and
// negatives to 0.
void World::terrain_nodes(sfl_vec3 low, sfl_vec3 high, uint64_t nodes[3]) const {
  const float step = settings_.spacing / 2.0f;
  const float extent[3] = {high.x - low.x, high.y - low.y, high.z - low.z};
  for (int a = 0; a < 3; ++a) {
    float c = std::fmin(std::ceil(extent[a] / step), 1e7f);
    nodes[a] = (std::isnan(c) || c <= 0.0f ? 0 : (uint64_t)c) + 1;
  }
}

// GroundSites::new (terrain.rs:186-233)
void World::build_ground_sites() {
  const float spacing = settings_.spacing, radius = radius_, mass = mass_;
  const int reach = (int)std::ceil(radius / spacing);
  sites_h_.clear();
  far_h_.clear();
  for (int z = -reach; z <= reach; ++z)
    for (int y = -reach; y <= reach; ++y)
      for (int x = -reach; x <= reach; ++x) {
        float ox = (float)x * spacing, oy = (float)y * spacing, oz = (float)z * spacing;
        float far = std::sqrt(ox * ox + oy * oy + oz * oz);
        if (far < radius) {
          sites_h_.push_back(make_float4(ox, oy, oz, mass * value_host(far, radius, scale_)));
          far_h_.push_back(far);
        }
      }
  ground_low_ = -spacing / 2.0f;
  ground_step_ = (radius - ground_low_) / (float)(kMapSamples - 1);
  const float soften = 0.01f * radius * radius;
  const int across = reach + 1;
  damping_h_.clear();
  for (int sample = 0; sample < kMapSamples; ++sample) {
    float distance = ground_low_ + ground_step_ * (float)sample;
    float weight = 0.0f;
    for (int layer = 0;; ++layer) {
      float depth = distance + spacing * ((float)layer + 0.5f);
      if (depth >= radius) break;
      for (int a = -across; a <= across; ++a)
        for (int b = -across; b <= across; ++b) {
          float flat = (float)(a * a + b * b) * spacing * spacing;
          float squared = flat + depth * depth;
          float far = std::sqrt(squared);
          if (far >= radius || far <= 0.0f) continue;
          weight += mass * depth * depth * -slope_host(far, radius, scale_) / (far * (squared + soften));
        }
    }
    damping_h_.push_back(weight);
  }
  ground_reach_ = radius + spacing / 2.0f;
}

// add_terrain (terrain.rs:262-305) over the caller's samples of its field
// at low + (x, y, z) step, loops y, z, x
int32_t World::add_terrain(sfl_vec3 low, sfl_vec3 high, const float* distances, size_t count, uint32_t* id) {
  ++version_;
  DeviceGuard g(device_);
  if (!finite3(low) || !finite3(high)) {
    detail_ = "terrain box";
    return SFL_NON_FINITE;
  }
  if (high.x <= low.x || high.y <= low.y || high.z <= low.z) return SFL_OUT_OF_RANGE;
  uint64_t nodes[3];
  terrain_nodes(low, high, nodes);
  uint64_t total = 1;
  for (uint64_t n : nodes) total = total > UINT64_MAX / n ? UINT64_MAX : total * n;
  if (total > kMaxNodes) return SFL_FLUID_FULL;
  if (count != total || (count && !distances)) return SFL_OUT_OF_RANGE;
  for (size_t k = 0; k < count; ++k)
    if (!std::isfinite(distances[k])) {
      detail_ = "terrain distance";
      return SFL_NON_FINITE;
    }
  uint32_t owner = 0;
  int32_t st = add_wall_points({}, &owner);
  if (st != SFL_OK) return st;
  Terrain t;
  t.owner = owner;
  t.low[0] = low.x, t.low[1] = low.y, t.low[2] = low.z;
  for (int a = 0; a < 3; ++a) t.nodes[a] = (uint32_t)nodes[a];
  t.step = settings_.spacing / 2.0f;
  t.distances.assign(distances, distances + count);
  terrains_.push_back(std::move(t));
  terrain_dirty_ = true;
  if (id) *id = owner;
  return SFL_OK;
}

int32_t World::upload_terrains() {
  if (!terrain_dirty_) return SFL_OK;
  terrain_dirty_ = false;
  if (terrains_.empty()) return SFL_OK;
  if (sites_h_.empty()) {
    build_ground_sites();
    SFL_HIP(sites_d_.reserve(sites_h_.size(), 0, stream_), "hipMalloc(ground sites)");
    SFL_HIP(far_d_.reserve(far_h_.size(), 0, stream_), "hipMalloc(ground sites)");
    SFL_HIP(damping_d_.reserve(damping_h_.size(), 0, stream_), "hipMalloc(ground damping)");
    SFL_HIP(hipMemcpyAsync(sites_d_.ptr, sites_h_.data(), sites_h_.size() * sizeof(float4), hipMemcpyHostToDevice,
                           stream_),
            "upload ground sites");
    SFL_HIP(hipMemcpyAsync(far_d_.ptr, far_h_.data(), far_h_.size() * sizeof(float), hipMemcpyHostToDevice, stream_),
            "upload ground sites");
    SFL_HIP(hipMemcpyAsync(damping_d_.ptr, damping_h_.data(), damping_h_.size() * sizeof(float),
                           hipMemcpyHostToDevice, stream_),
            "upload ground damping");
  }
  std::vector<TerrainDesc> desc;
  size_t total = 0;
  for (const Terrain& t : terrains_) {
    TerrainDesc d{};
    for (int a = 0; a < 3; ++a) {
      d.low[a] = t.low[a];
      d.nodes[a] = t.nodes[a];
    }
    d.step = t.step;
    d.first = (uint32_t)total;
    total += t.distances.size();
    desc.push_back(d);
  }
  SFL_HIP(terrain_desc_.reserve(desc.size(), 0, stream_), "hipMalloc(terrain)");
  SFL_HIP(terrain_dist_.reserve(total, 0, stream_), "hipMalloc(terrain distances)");
  SFL_HIP(hipMemcpyAsync(terrain_desc_.ptr, desc.data(), desc.size() * sizeof(TerrainDesc), hipMemcpyHostToDevice,
                         stream_),
          "upload terrain");
  size_t at = 0;
  for (const Terrain& t : terrains_) {
    SFL_HIP(hipMemcpyAsync(terrain_dist_.ptr + at, t.distances.data(), t.distances.size() * sizeof(float),
                           hipMemcpyHostToDevice, stream_),
            "upload terrain distances");
    at += t.distances.size();
  }
  SFL_HIP(hipStreamSynchronize(stream_), "upload terrain");
  return SFL_OK;
}

TerrainView World::terrain_view() const {
  TerrainView v{};
  if (terrains_.empty()) return v;
  v.fields = terrain_desc_.ptr;
  v.distances = terrain_dist_.ptr;
  v.count = (uint32_t)terrains_.size();
  v.nsites = (uint32_t)sites_h_.size();
  v.sites = sites_d_.ptr;
  v.far = far_d_.ptr;
  v.damping = damping_d_.ptr;
  v.low = ground_low_;
  v.step = ground_step_;
  v.reach = ground_reach_;
  v.spacing = settings_.spacing;
  v.slopes = slopes_.ptr;
  return v;
}

}  // namespace sfl
