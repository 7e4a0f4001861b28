// Host side of the device fluid world (see world.hpp). All scalar arithmetic
// that feeds the kernels (kernel scale, particle mass, substep counts and
// lengths, viscosities) is the reference's f32 arithmetic in its order; this
// file is compiled with -ffp-contract=off like the kernels.
#include "world.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <tuple>

#include "launch.hpp"

namespace sfl {

namespace {

constexpr float kPi = 3.14159265358979323846f;  // std::f32::consts::PI
constexpr uint64_t kMaxCells = uint64_t(1) << 31;
// A dense table is always allowed this many cells; a searched table's cell
// key (the slot in the box) stays below 2^63.
constexpr uint64_t kDenseFloor = uint64_t(1) << 26;
constexpr uint64_t kMaxKeyedCells = uint64_t(1) << 63;

// Kernel::value (src/fluid/mod.rs:140-146) on the host.
float kernel_value_host(float distance, float radius, float scale) {
  float q = distance / radius;
  if (q >= 1.0f) return 0.0f;
  float t = 1.0f - q;
  float t2 = t * t;
  float t4 = t2 * t2;
  return scale * t4 * (1.0f + 4.0f * q);
}

float length3(float x, float y, float z) { return std::sqrt(x * x + y * y + z * z); }

float fmax_rust(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a > b ? a : b)); }
float fmin_rust(float a, float b) { return std::isnan(a) ? b : (std::isnan(b) ? a : (a < b ? a : b)); }

int32_t sat_i32(float v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483648.0f) return std::numeric_limits<int32_t>::max();
  if (v <= -2147483648.0f) return std::numeric_limits<int32_t>::min();
  return (int32_t)v;
}

int32_t cell_coord_host(float v, float radius) { return sat_i32(std::floor(v / radius)); }

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
bool finite_vec(sfl_vec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// FluidSettings::validate (src/fluid/mod.rs:98-117) plus the extensions.
bool settings_valid(const sfl_settings& s) {
  auto positive = [](float v) { return std::isfinite(v) && v > 0.0f; };
  bool support = s.support >= 1.5f && s.support <= 4.0f;
  return (s.solver == SFL_WCSPH || s.solver == SFL_DFSPH) && positive(s.spacing) && positive(s.rest_density) &&
         support && s.max_particles > 0 && positive(s.sound_speed) && std::isfinite(s.viscosity) &&
         s.viscosity >= 0.0f && finite_vec(s.gravity) && s.courant > 0.0f && s.courant <= 1.0f &&
         s.max_substeps > 0 && std::isfinite(s.divergence_tolerance) && std::isfinite(s.density_tolerance) &&
         s.max_iterations >= s.min_iterations;
}

// check_box (src/fluid/boundary.rs:230-238)
int32_t check_box(sfl_vec3 low, sfl_vec3 high, std::string& detail) {
  if (!finite_vec(low) || !finite_vec(high)) {
    detail = "fluid box";
    return SFL_NON_FINITE;
  }
  if (high.x <= low.x || high.y <= low.y || high.z <= low.z) return SFL_OUT_OF_RANGE;
  return SFL_OK;
}

struct DeviceGuard {
  explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

}  // namespace

int32_t World::device_error(const char* what, hipError_t e) {
  detail_ = std::string(what) + ": " + hipGetErrorString(e);
  return SFL_DEVICE_ERROR;
}

#define SFL_HIP(call, what)                          \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return device_error(what, e_); \
  } while (0)

bool World::valid_settings(const sfl_settings& s) { return settings_valid(s); }

// The particles of a snapshot, in its order, with its ids; densities at the
// rest density (save.rs:143).
int32_t World::restore(size_t count, const uint32_t* ids, const float* xyz, const float* vxyz, uint32_t next_id) {
  DeviceGuard g(device_);
  int32_t st = ensure_particles(count);
  if (st != SFL_OK) return st;
  std::vector<float4> hp(count), hv(count);
  float vmax = 0.0f;
  for (size_t i = 0; i < count; ++i) {
    float idf;
    std::memcpy(&idf, &ids[i], 4);
    hp[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], idf);
    hv[i] = make_float4(vxyz[3 * i], vxyz[3 * i + 1], vxyz[3 * i + 2], settings_.rest_density);
    vmax = fmax_rust(vmax, length3(vxyz[3 * i], vxyz[3 * i + 1], vxyz[3 * i + 2]));
  }
  if (count) {
    SFL_HIP(hipMemcpyAsync(pa_.ptr, hp.data(), count * sizeof(float4), hipMemcpyHostToDevice, stream_),
            "upload positions");
    SFL_HIP(hipMemcpyAsync(va_.ptr, hv.data(), count * sizeof(float4), hipMemcpyHostToDevice, stream_),
            "upload velocities");
    SFL_HIP(hipStreamSynchronize(stream_), "upload");
  }
  n_ = count;
  next_id_ = next_id;
  vmax_ = vmax;
  vmax_valid_ = true;
  invalidate_lists();
  forces_valid_ = false;
  ++version_;
  return SFL_OK;
}


int32_t World::create(const sfl_settings& s, World** out, std::string& why) {
  if (!settings_valid(s)) return SFL_INVALID_SETTINGS;
  World* w = new World();
  w->settings_ = s;
  // radius (mod.rs:120-122), Kernel::new (mod.rs:133-138)
  w->radius_ = s.spacing * s.support;
  w->reach_ = w->radius_ * w->radius_;
  float r = w->radius_;
  float r3 = r * (r * r);
  w->scale_ = 21.0f / (2.0f * kPi * r3);
  // rest_mass (mod.rs:696-709)
  int reach = (int)std::ceil(s.support);
  float sum = 0.0f;
  for (int z = -reach; z <= reach; ++z)
    for (int y = -reach; y <= reach; ++y)
      for (int x = -reach; x <= reach; ++x) {
        float ox = (float)x * s.spacing, oy = (float)y * s.spacing, oz = (float)z * s.spacing;
        sum += kernel_value_host(length3(ox, oy, oz), r, w->scale_);
      }
  w->mass_ = s.rest_density / sum;
  int32_t st = w->init_device();
  if (st != SFL_OK) {
    why = w->detail_;
    delete w;
    return st;
  }
  *out = w;
  return SFL_OK;
}

int32_t World::init_device() {
  int count = 0;
  SFL_HIP(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (settings_.device < 0 || settings_.device >= count) {
    detail_ = "no such HIP device";
    return SFL_DEVICE_ERROR;
  }
  device_ = settings_.device;
  SFL_HIP(hipSetDevice(device_), "hipSetDevice");
  SFL_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  SFL_HIP(hipMalloc(reinterpret_cast<void**>(&ctl_), sizeof(Control)), "hipMalloc(control)");
  SFL_HIP(hipMemset(ctl_, 0, sizeof(Control)), "hipMemset(control)");
  SFL_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctl_host_), sizeof(Control), hipHostMallocDefault),
          "hipHostMalloc(control)");
  std::memset(ctl_host_, 0, sizeof(Control));
  SFL_HIP(scan_total_.reserve(1, 0, stream_), "hipMalloc(scan total)");
  // SOLUNAR_SOLVE_BLOCKS_PER_CU (1..3) narrows the persistent solve's grid,
  // a tuning knob; the default is the widest grid that stays resident
  const char* knob = std::getenv("SOLUNAR_SOLVE_BLOCKS_PER_CU");
  // SOLUNAR_SOLVE_TILES=0 makes the list solve gather from global memory
  // instead of staging each tile's neighbourhood in LDS
  const char* tiles = std::getenv("SOLUNAR_SOLVE_TILES");
  solve_grid_ = solve_grid_limit(knob ? std::atoi(knob) : 0, !(tiles && std::atoi(tiles) == 0), &solve_lds_);
  // SOLUNAR_SOLVE_TILE_PARTICLES caps the particles one block stages
  // (rounded down to 64; tiles with more gather from global memory)
  const char* tile_cap = std::getenv("SOLUNAR_SOLVE_TILE_PARTICLES");
  if (tile_cap) {
    long c = std::max(std::atol(tile_cap), 0L) / 64 * 64;
    solve_lds_ = (uint32_t)std::min<uint64_t>(solve_lds_, (uint64_t)c * kTileBytes);
  }
  // SOLUNAR_NEIGHBOUR_FORMAT (0: 16-bit masks, 1: 32-bit masks, 2: lists,
  // 3: compact lists, the default where the solve has its LDS tiles) picks
  // the DFSPH neighbour format a world starts with; each widens on demand
  const char* fmt = std::getenv("SOLUNAR_NEIGHBOUR_FORMAT");
  format_ = solve_lds_ ? kCompact : kLists;
  if (fmt) format_ = std::min(std::max(std::atoi(fmt), (int)kMasks16), (int)kCompact);
  if (format_ == kCompact && !solve_lds_) format_ = kLists;  // compact lists need the tiled solve
  // SOLUNAR_SOLVE_FLOW=0 makes the tiled solve's passes meet at grid
  // barriers instead of waiting tile by tile on their neighbours
  const char* flow = std::getenv("SOLUNAR_SOLVE_FLOW");
  flow_ = !(flow && std::atoi(flow) == 0);
  // SOLUNAR_BODY_PAIR_SLACK (default 65536) is the smallest room for
  // body-point pair records and the slack a step adds to the last step's
  // count; a small value makes coupled steps outgrow their room (tests)
  const char* slack = std::getenv("SOLUNAR_BODY_PAIR_SLACK");
  if (slack) pair_slack_ = (uint32_t)std::min<long>(std::max(std::atol(slack), 0L), 1L << 30);
  // SOLUNAR_LIFT_COLUMN_TABLES=0 reads every body's columns by sorting, the
  // path a body too spread out for a column table takes (tests)
  const char* tables = std::getenv("SOLUNAR_LIFT_COLUMN_TABLES");
  if (tables && std::atoi(tables) == 0) lift_tables_ = false;
  if (solve_grid_ <= 0) {
    detail_ = "occupancy query for the persistent solve failed";
    return SFL_DEVICE_ERROR;
  }
  return SFL_OK;
}

World::~World() {
  DeviceGuard g(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : free_events_) (void)hipEventDestroy(e);
  if (ctl_) (void)hipFree(ctl_);
  if (ctl_host_) (void)hipHostFree(ctl_host_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// ---- particles ---------------------------------------------------------------

int32_t World::ensure_particles(size_t n) {
  SFL_HIP(pa_.reserve(n, n_, stream_), "hipMalloc(positions)");
  SFL_HIP(va_.reserve(n, n_, stream_), "hipMalloc(velocities)");
  SFL_HIP(pb_.reserve(n, 0, stream_), "hipMalloc(sorted positions)");
  SFL_HIP(vb_.reserve(n, 0, stream_), "hipMalloc(sorted velocities)");
  SFL_HIP(aux_.reserve(n, 0, stream_), "hipMalloc(aux)");
  SFL_HIP(kappa_.reserve(n, 0, stream_), "hipMalloc(kappa)");
  SFL_HIP(errors_.reserve(n, 0, stream_), "hipMalloc(errors)");
  SFL_HIP(slot_rank_.reserve(n, 0, stream_), "hipMalloc(slot ranks)");
  SFL_HIP(scatter_.reserve(n, 0, stream_), "hipMalloc(scatter)");
  size_t tiles = (n + 1 + 2047) / 2048;
  SFL_HIP(tile_sums_.reserve(std::max<size_t>(tiles, (cell_cap_ + 2047) / 2048 + 1), tile_sums_.cap ? 0 : 0, stream_),
          "hipMalloc(tile sums)");
  if (keep_forces_) SFL_HIP(forces_.reserve(3 * n, 0, stream_), "hipMalloc(forces)");
  return SFL_OK;
}

// add_particle (src/fluid/mod.rs:412-426), in order, stopping at the first
// refusal.
int32_t World::add_particles(size_t count, const float* xyz, const float* vxyz, uint32_t* ids) {
  ++version_;
  DeviceGuard g(device_);
  size_t ok = 0;
  int32_t status = SFL_OK;
  for (; ok < count; ++ok) {
    const float* p = xyz + 3 * ok;
    static const float zero[3] = {0.0f, 0.0f, 0.0f};
    const float* v = vxyz ? vxyz + 3 * ok : zero;
    if (!finite3(p) || !finite3(v)) {
      detail_ = "fluid particle";
      status = SFL_NON_FINITE;
      break;
    }
    if (n_ + ok >= settings_.max_particles || next_id_ + ok >= 0xffffffffull) {
      status = SFL_FLUID_FULL;
      break;
    }
  }
  if (ok == 0) return status;
  int32_t st = ensure_particles(n_ + ok);
  if (st != SFL_OK) return st;
  std::vector<float4> hp(ok), hv(ok);
  for (size_t i = 0; i < ok; ++i) {
    uint32_t id = next_id_ + (uint32_t)i;
    float idf;
    std::memcpy(&idf, &id, 4);
    hp[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], idf);
    float vx = vxyz ? vxyz[3 * i] : 0.0f, vy = vxyz ? vxyz[3 * i + 1] : 0.0f, vz = vxyz ? vxyz[3 * i + 2] : 0.0f;
    hv[i] = make_float4(vx, vy, vz, 0.0f);
    if (ids) ids[i] = id;
    if (vmax_valid_) vmax_ = fmax_rust(vmax_, length3(vx, vy, vz));
  }
  SFL_HIP(hipMemcpyAsync(pa_.ptr + n_, hp.data(), ok * sizeof(float4), hipMemcpyHostToDevice, stream_),
          "upload positions");
  SFL_HIP(hipMemcpyAsync(va_.ptr + n_, hv.data(), ok * sizeof(float4), hipMemcpyHostToDevice, stream_),
          "upload velocities");
  SFL_HIP(hipStreamSynchronize(stream_), "upload");
  n_ += ok;
  next_id_ += (uint32_t)ok;
  invalidate_lists();
  forces_valid_ = false;
  return status;
}

// fill_box (src/fluid/mod.rs:493-524), optionally jittered (see
// sfl_fill_box_jittered).
int32_t World::fill_box(sfl_vec3 low, sfl_vec3 high, float jitter, uint32_t seed, size_t* count_out) {
  ++version_;
  if (!finite_vec(low) || !finite_vec(high) || !std::isfinite(jitter)) {
    detail_ = "fluid box";
    return SFL_NON_FINITE;
  }
  if (high.x < low.x || high.y < low.y || high.z < low.z) return SFL_OUT_OF_RANGE;
  float spacing = settings_.spacing;
  auto count = [&](float extent) -> uint64_t {
    float c = std::min(std::floor(extent / spacing), 4294967295.0f);
    return c <= 0.0f ? 0 : (uint64_t)c;
  };
  uint64_t nx = count(high.x - low.x), ny = count(high.y - low.y), nz = count(high.z - low.z);
  unsigned __int128 t = (unsigned __int128)nx * ny * nz;
  uint64_t total = t > UINT64_MAX ? UINT64_MAX : (uint64_t)t;
  uint64_t room = settings_.max_particles > n_ ? settings_.max_particles - n_ : 0;
  uint64_t ids_left = 0xffffffffull - next_id_;
  if (total > room || total > ids_left) return SFL_FLUID_FULL;
  float half = spacing * 0.5f;
  float sx = low.x + half, sy = low.y + half, sz = low.z + half;
  uint32_t state = seed ? seed : 0x9E3779B9u;
  float amp = 2.0f * jitter * spacing;
  auto draw = [&]() {
    state ^= state << 13;
    state ^= state >> 17;
    state ^= state << 5;
    return ((float)state / 4294967295.0f - 0.5f) * amp;
  };
  std::vector<float> pts;
  pts.reserve((size_t)total * 3);
  for (uint64_t y = 0; y < ny; ++y)
    for (uint64_t z = 0; z < nz; ++z)
      for (uint64_t x = 0; x < nx; ++x) {
        float px = sx + (float)x * spacing, py = sy + (float)y * spacing, pz = sz + (float)z * spacing;
        if (jitter != 0.0f) {
          float jx = draw(), jy = draw(), jz = draw();
          px = px + jx;
          py = py + jy;
          pz = pz + jz;
        }
        pts.push_back(px);
        pts.push_back(py);
        pts.push_back(pz);
      }
  int32_t st = total ? add_particles((size_t)total, pts.data(), nullptr, nullptr) : SFL_OK;
  if (st == SFL_OK && count_out) *count_out = (size_t)total;
  return st;
}

// set_domain (src/fluid/mod.rs:433-444): validated, but a box is refused as
// unsupported on the device path.
int32_t World::set_domain(const sfl_vec3* low, const sfl_vec3* high) {
  ++version_;
  if (!low && !high) {
    has_domain_ = false;
    return SFL_OK;
  }
  if (!low || !high) return SFL_OUT_OF_RANGE;
  if (!finite_vec(*low) || !finite_vec(*high)) {
    detail_ = "fluid domain";
    return SFL_NON_FINITE;
  }
  if (high->x <= low->x || high->y <= low->y || high->z <= low->z) return SFL_OUT_OF_RANGE;
  has_domain_ = true;
  domain_ = FlowBox{{low->x, low->y, low->z}, {high->x, high->y, high->z}};
  return SFL_OK;
}

bool World::domain(sfl_vec3* low, sfl_vec3* high) const {
  if (!has_domain_) return false;
  if (low) *low = sfl_vec3{domain_.lo[0], domain_.lo[1], domain_.lo[2]};
  if (high) *high = sfl_vec3{domain_.hi[0], domain_.hi[1], domain_.hi[2]};
  return true;
}

sfl_vec3 World::kernel_gradient_at(sfl_vec3 o) const {
  sfl_vec3 g{0.0f, 0.0f, 0.0f};
  kernel_gradient(o.x, o.y, o.z, dot3(o.x, o.y, o.z, o.x, o.y, o.z), radius_, scale_, g.x, g.y, g.z);
  return g;
}

// ---- flows (src/fluid/flow.rs) ----------------------------------------------

// add_emitter / add_drain / add_flow (flow.rs:92-139)
int32_t World::add_flow(sfl_vec3 low, sfl_vec3 high, bool drain, sfl_vec3 velocity, uint32_t* id_out) {
  if (!drain && !finite_vec(velocity)) {
    detail_ = "emitter velocity";
    return SFL_NON_FINITE;
  }
  if (!finite_vec(low) || !finite_vec(high)) {
    detail_ = "flow box";
    return SFL_NON_FINITE;
  }
  if (high.x <= low.x || high.y <= low.y || high.z <= low.z) return SFL_OUT_OF_RANGE;
  if (flow_next_ == 0xffffffffu) return SFL_FLUID_FULL;
  const uint32_t id = flow_next_++;
  flows_.push_back(Flow{id, low, high, drain, velocity});
  if (id_out) *id_out = id;
  return SFL_OK;
}

int32_t World::add_emitter(sfl_vec3 low, sfl_vec3 high, sfl_vec3 velocity, uint32_t* id) {
  return add_flow(low, high, false, velocity, id);
}

int32_t World::add_drain(sfl_vec3 low, sfl_vec3 high, uint32_t* id) {
  return add_flow(low, high, true, sfl_vec3{0.0f, 0.0f, 0.0f}, id);
}

// remove_flow (flow.rs:113-118)
int32_t World::remove_flow(uint32_t id, int32_t* removed) {
  const size_t before = flows_.size();
  flows_.erase(std::remove_if(flows_.begin(), flows_.end(), [&](const Flow& f) { return f.id == id; }),
               flows_.end());
  if (removed) *removed = flows_.size() != before ? 1 : 0;
  return SFL_OK;
}

int32_t World::read_count(const uint32_t* dev, uint32_t* out) {
  SFL_HIP(hipMemcpyAsync(out, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "read count");
  SFL_HIP(hipStreamSynchronize(stream_), "flows");
  return SFL_OK;
}

// keep_only (src/fluid/mod.rs:468-489) of the particles outside a drain box
// or inside the domain: flags, their exclusive scan, the kept particles
// gathered in order to B and copied back to A.
int32_t World::keep_only(const FlowBox& box, bool domain, uint32_t& gone) {
  gone = 0;
  if (n_ == 0) return SFL_OK;
  const uint32_t n = (uint32_t)n_;
  SFL_HIP(flow_flags_.reserve(n_ + 1, 0, stream_), "hipMalloc(flow flags)");
  SFL_HIP(flow_start_.reserve(n_ + 1, 0, stream_), "hipMalloc(flow starts)");
  SFL_HIP(flow_tiles_.reserve((n_ + 2047) / 2048 + 1, 0, stream_), "hipMalloc(flow tiles)");
  SFL_HIP(flow_total_.reserve(1, 0, stream_), "hipMalloc(flow total)");
  launch_flow_keep(stream_, pa_.ptr, n, box, domain ? 1 : 0, flow_flags_.ptr);
  launch_scan(stream_, flow_flags_.ptr, flow_start_.ptr, flow_tiles_.ptr, flow_total_.ptr, nullptr, n, 0, nullptr);
  uint32_t kept = 0;
  int32_t st = read_count(flow_total_.ptr, &kept);
  if (st != SFL_OK) return st;
  gone = n - kept;
  if (gone == 0) return SFL_OK;
  launch_flow_gather(stream_, pa_.ptr, va_.ptr, n, flow_flags_.ptr, flow_start_.ptr, pb_.ptr, vb_.ptr);
  SFL_HIP(hipMemcpyAsync(pa_.ptr, pb_.ptr, (size_t)kept * sizeof(float4), hipMemcpyDeviceToDevice, stream_),
          "keep positions");
  SFL_HIP(hipMemcpyAsync(va_.ptr, vb_.ptr, (size_t)kept * sizeof(float4), hipMemcpyDeviceToDevice, stream_),
          "keep velocities");
  SFL_HIP(hipGetLastError(), "flow kernels");
  n_ = kept;
  invalidate_lists();
  forces_valid_ = false;
  return SFL_OK;
}

// One emitter (flow.rs:170-220): the lattice slots of its box with no water
// nearer than 0.9 spacing, in y, z, x order, become particles moving at its
// velocity. A full world stops it after the particles that fit.
int32_t World::emit(const sfl_vec3& low, const sfl_vec3& high, const sfl_vec3& velocity, uint32_t& added) {
  added = 0;
  const float spacing = settings_.spacing;
  auto count = [&](float extent) -> int64_t {
    const float c = std::floor(extent / spacing);
    int64_t v = std::isnan(c) ? 0 : (c >= 9.2e18f ? INT64_MAX : (c <= -9.2e18f ? INT64_MIN : (int64_t)c));
    return std::min<int64_t>(std::max<int64_t>(v, 0), 1 << 12);
  };
  const int64_t cx = count(high.x - low.x), cy = count(high.y - low.y), cz = count(high.z - low.z);
  const int64_t total = cx * cy * cz;
  if (total == 0) return SFL_OK;
  if (total > kEmitterSlotsMax) {
    detail_ = "emitter box";
    return SFL_OUT_OF_RANGE;
  }
  const float near = 0.9f * spacing;  // CLEAR_SHARE (flow.rs:8), squared as one product
  EmitBox e{{low.x, low.y, low.z}, spacing, {cx, cy, cz}, near * near, {velocity.x, velocity.y, velocity.z}};
  const size_t slots = (size_t)total;
  SFL_HIP(flow_flags_.reserve(slots + 1, 0, stream_), "hipMalloc(emitter slots)");
  SFL_HIP(flow_start_.reserve(slots + 1, 0, stream_), "hipMalloc(emitter ranks)");
  SFL_HIP(flow_tiles_.reserve((slots + 2047) / 2048 + 1, 0, stream_), "hipMalloc(flow tiles)");
  SFL_HIP(flow_total_.reserve(1, 0, stream_), "hipMalloc(flow total)");
  launch_emit_mark(stream_, pa_.ptr, (uint32_t)n_, e, flow_flags_.ptr, slots);
  launch_scan(stream_, flow_flags_.ptr, flow_start_.ptr, flow_tiles_.ptr, flow_total_.ptr, nullptr, slots, 0,
              nullptr);
  uint32_t open = 0;
  int32_t st = read_count(flow_total_.ptr, &open);
  if (st != SFL_OK) return st;
  // add_particle's room (src/fluid/mod.rs:416): below max_particles, ids below u32::MAX
  const uint64_t room = n_ < settings_.max_particles ? settings_.max_particles - n_ : 0;
  const uint64_t ids = 0xffffffffull - next_id_;
  const uint32_t limit = (uint32_t)std::min<uint64_t>({room, ids, (uint64_t)open});
  if (limit > 0) {
    st = ensure_particles(n_ + limit);
    if (st != SFL_OK) return st;
    launch_emit_append(stream_, flow_flags_.ptr, flow_start_.ptr, slots, e, (uint32_t)n_, limit, next_id_, pa_.ptr,
                       va_.ptr);
    SFL_HIP(hipGetLastError(), "emitter kernels");
    n_ += limit;
    next_id_ += limit;
    added = limit;
    invalidate_lists();
    forces_valid_ = false;
  }
  return open > limit ? SFL_FLUID_FULL : SFL_OK;
}

// run_flows (flow.rs:144-223): drains, then emitters, each in the order
// they were added.
int32_t World::run_flows(uint32_t& emitted, uint32_t& drained, bool& changed) {
  emitted = drained = 0;
  if (flows_.empty()) return SFL_OK;
  const std::vector<Flow> list = flows_;
  for (const Flow& f : list) {
    if (!f.drain) continue;
    uint32_t gone = 0;
    int32_t st = keep_only(FlowBox{{f.low.x, f.low.y, f.low.z}, {f.high.x, f.high.y, f.high.z}}, false, gone);
    if (st != SFL_OK) return st;
    drained += gone;
    changed = changed || gone > 0;
  }
  for (const Flow& f : list) {
    if (f.drain) continue;
    uint32_t added = 0;
    int32_t st = emit(f.low, f.high, f.velocity, added);
    emitted += added;
    changed = changed || added > 0;
    if (st != SFL_OK) return st;
  }
  return SFL_OK;
}

// ---- walls (src/fluid/boundary.rs) ------------------------------------------

int32_t World::add_box_container(sfl_vec3 low, sfl_vec3 high, uint32_t* id_out) {
  ++version_;
  DeviceGuard g(device_);
  int32_t st = check_box(low, high, detail_);
  if (st != SFL_OK) return st;
  // cells_in (boundary.rs:122-126)
  auto cells = [&](float length) -> int64_t {
    float c = std::min(std::round(length / settings_.spacing), 1e9f);
    return (int64_t)c;
  };
  int64_t c[3] = {cells(high.x - low.x), cells(high.y - low.y), cells(high.z - low.z)};
  int64_t layers = (int64_t)std::ceil(settings_.support);
  int64_t shell = (c[0] + 2 * layers) * (c[1] + 2 * layers) * (c[2] + 2 * layers) - c[0] * c[1] * c[2];
  uint64_t have = wall_owner_.size();
  uint64_t room = settings_.max_particles > have ? settings_.max_particles - have : 0;
  if (shell < 0 || (uint64_t)shell > room || wall_next_ == 0xffffffffu) return SFL_FLUID_FULL;
  float spacing = settings_.spacing;
  std::vector<float> pts;
  pts.reserve((size_t)shell * 3);
  for (int64_t y = -layers; y < c[1] + layers; ++y)
    for (int64_t z = -layers; z < c[2] + layers; ++z)
      for (int64_t x = -layers; x < c[0] + layers; ++x) {
        bool inside = x >= 0 && x < c[0] && y >= 0 && y < c[1] && z >= 0 && z < c[2];
        if (inside) continue;
        float ax = (float)x + 0.5f, ay = (float)y + 0.5f, az = (float)z + 0.5f;
        pts.push_back(low.x + ax * spacing);
        pts.push_back(low.y + ay * spacing);
        pts.push_back(low.z + az * spacing);
      }
  uint32_t id = 0;
  Container box{wall_next_, {low.x, low.y, low.z}, {high.x, high.y, high.z}};
  containers_.push_back(box);
  st = add_wall_points(pts, &id);
  if (st != SFL_OK) {
    containers_.pop_back();
    return st;
  }
  if (id_out) *id_out = id;
  return SFL_OK;
}

int32_t World::add_wall(sfl_vec3 low, sfl_vec3 high, uint32_t* id_out) {
  ++version_;
  DeviceGuard g(device_);
  int32_t st = check_box(low, high, detail_);
  if (st != SFL_OK) return st;
  auto cells = [&](float length) -> int64_t {
    float c = std::min(std::round(length / settings_.spacing), 1e9f);
    return std::max<int64_t>((int64_t)c, 1);
  };
  int64_t c[3] = {cells(high.x - low.x), cells(high.y - low.y), cells(high.z - low.z)};
  int64_t count = c[0] * c[1] * c[2];
  uint64_t have = wall_owner_.size();
  uint64_t room = settings_.max_particles > have ? settings_.max_particles - have : 0;
  if (count < 0 || (uint64_t)count > room || wall_next_ == 0xffffffffu) return SFL_FLUID_FULL;
  float spacing = settings_.spacing;
  std::vector<float> pts;
  pts.reserve((size_t)count * 3);
  for (int64_t y = 0; y < c[1]; ++y)
    for (int64_t z = 0; z < c[2]; ++z)
      for (int64_t x = 0; x < c[0]; ++x) {
        float ax = (float)x + 0.5f, ay = (float)y + 0.5f, az = (float)z + 0.5f;
        pts.push_back(low.x + ax * spacing);
        pts.push_back(low.y + ay * spacing);
        pts.push_back(low.z + az * spacing);
      }
  return add_wall_points(pts, id_out);
}

int32_t World::add_boundary_points(size_t count, const float* xyz, uint32_t* id_out) {
  ++version_;
  DeviceGuard g(device_);
  for (size_t i = 0; i < count; ++i)
    if (!finite3(xyz + 3 * i)) {
      detail_ = "fluid wall";
      return SFL_NON_FINITE;
    }
  uint64_t have = wall_owner_.size();
  uint64_t room = settings_.max_particles > have ? settings_.max_particles - have : 0;
  if (count > room || wall_next_ == 0xffffffffu) return SFL_FLUID_FULL;
  std::vector<float> pts(xyz, xyz + 3 * count);
  return add_wall_points(pts, id_out);
}

// add_wall_points (boundary.rs:139-151); the bounds follow the containers
// (boundary.rs:240-247).
int32_t World::add_wall_points(const std::vector<float>& pts, uint32_t* id_out) {
  uint32_t id = wall_next_;
  size_t before = wall_owner_.size();
  wall_pos_.insert(wall_pos_.end(), pts.begin(), pts.end());
  wall_owner_.insert(wall_owner_.end(), pts.size() / 3, id);
  int32_t st = settle_walls();
  if (st != SFL_OK) {
    // undo: the settle failed before touching the device
    std::vector<float> keep_pos;
    std::vector<uint32_t> keep_owner;
    for (size_t i = 0; i < wall_owner_.size(); ++i)
      if (wall_owner_[i] != id) {
        keep_pos.insert(keep_pos.end(), wall_pos_.begin() + 3 * i, wall_pos_.begin() + 3 * i + 3);
        keep_owner.push_back(wall_owner_[i]);
      }
    wall_pos_.swap(keep_pos);
    wall_owner_.swap(keep_owner);
    (void)before;
    (void)settle_walls();
    return st;
  }
  wall_next_ += 1;
  if (id_out) *id_out = id;
  return SFL_OK;
}

// remove_boundary (boundary.rs:94-115)
int32_t World::remove_boundary(uint32_t id, int32_t* removed) {
  ++version_;
  DeviceGuard g(device_);
  size_t before = wall_owner_.size();
  std::vector<float> keep_pos;
  std::vector<uint32_t> keep_owner;
  for (size_t i = 0; i < wall_owner_.size(); ++i)
    if (wall_owner_[i] != id) {
      keep_pos.insert(keep_pos.end(), wall_pos_.begin() + 3 * i, wall_pos_.begin() + 3 * i + 3);
      keep_owner.push_back(wall_owner_[i]);
    }
  bool had = false;
  std::vector<Container> keep_c;
  for (auto& c : containers_) {
    if (c.id == id)
      had = true;
    else
      keep_c.push_back(c);
  }
  const size_t fields = terrains_.size();
  terrains_.erase(std::remove_if(terrains_.begin(), terrains_.end(), [&](const Terrain& t) { return t.owner == id; }),
                  terrains_.end());
  if (terrains_.size() != fields) terrain_dirty_ = true;
  bool gone = keep_owner.size() != before || had || terrains_.size() != fields;
  if (gone) {
    wall_pos_.swap(keep_pos);
    wall_owner_.swap(keep_owner);
    containers_.swap(keep_c);
    int32_t st = settle_walls();
    if (st != SFL_OK) return st;
  }
  if (removed) *removed = gone ? 1 : 0;
  return SFL_OK;
}

// settle_walls (boundary.rs:155-178): sort by (cell, owner, x/y/z bits,
// index), then the dense cell table the kernels look walls up in, and the
// container bounds.
int32_t World::settle_walls() {
  size_t n = wall_owner_.size();
  struct Key {
    int32_t c[3];
    uint32_t owner, bx, by, bz;
    size_t index;
  };
  std::vector<Key> order(n);
  for (size_t i = 0; i < n; ++i) {
    const float* p = wall_pos_.data() + 3 * i;
    Key k;
    for (int a = 0; a < 3; ++a) k.c[a] = cell_coord_host(p[a], radius_);
    k.owner = wall_owner_[i];
    std::memcpy(&k.bx, p, 4);
    std::memcpy(&k.by, p + 1, 4);
    std::memcpy(&k.bz, p + 2, 4);
    k.index = i;
    order[i] = k;
  }
  auto before = [](const Key& a, const Key& b) {
    return std::tie(a.c[0], a.c[1], a.c[2], a.owner, a.bx, a.by, a.bz, a.index) <
           std::tie(b.c[0], b.c[1], b.c[2], b.owner, b.bx, b.by, b.bz, b.index);
  };
  // walls from a snapshot come settled already
  if (!std::is_sorted(order.begin(), order.end(), before)) std::sort(order.begin(), order.end(), before);
  // table box
  CellBox box{};
  uint64_t volume = 0;
  if (n) {
    int32_t lo[3] = {order[0].c[0], order[0].c[1], order[0].c[2]};
    int32_t hi[3] = {lo[0], lo[1], lo[2]};
    for (auto& k : order)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], k.c[a]);
        hi[a] = std::max(hi[a], k.c[a]);
      }
    volume = 1;
    for (int a = 0; a < 3; ++a) {
      uint64_t s = (uint64_t)((int64_t)hi[a] - (int64_t)lo[a] + 1);
      if (s > 0xffffffffull || volume * s / s != volume || volume * s >= kMaxKeyedCells) {
        detail_ = "walls spread over more cells than a 64-bit cell key addresses";
        return SFL_UNSUPPORTED;
      }
      box.lo[a] = lo[a];
      box.span[a] = (uint32_t)s;
      volume *= s;
    }
  }
  // dense while the box holds at most 64 cells a wall particle and a floor
  // (the reference keeps its dense index up to 64 per occupied cell plus
  // 4096, mod.rs:38-41, 214-217); past that the occupied cells are searched
  const bool searched = volume > std::max<uint64_t>(64 * (uint64_t)n + 4096, kDenseFloor);
  std::vector<float> pos(3 * n);
  std::vector<uint32_t> own(n);
  std::vector<float4> dev(n);
  std::vector<uint32_t> start;
  std::vector<uint64_t> keys;  // searched: the occupied cells' slots, ascending
  if (!searched) start.assign(volume + 1, 0);
  for (size_t k = 0; k < n; ++k) {
    const float* p = wall_pos_.data() + 3 * order[k].index;
    pos[3 * k] = p[0];
    pos[3 * k + 1] = p[1];
    pos[3 * k + 2] = p[2];
    own[k] = order[k].owner;
    dev[k] = make_float4(p[0], p[1], p[2], 0.0f);
    uint64_t slot = ((uint64_t)(order[k].c[0] - box.lo[0]) * box.span[1] + (uint64_t)(order[k].c[1] - box.lo[1])) *
                        box.span[2] +
                    (uint64_t)(order[k].c[2] - box.lo[2]);
    if (!searched) {
      start[slot + 1] += 1;
    } else if (keys.empty() || keys.back() != slot) {
      keys.push_back(slot);
      start.push_back((uint32_t)k);
    }
  }
  if (!searched)
    for (uint64_t c = 0; c < volume; ++c) start[c + 1] += start[c];
  else
    start.push_back((uint32_t)n);
  if (n) {
    SFL_HIP(walls_d_.reserve(n, 0, stream_), "hipMalloc(walls)");
    SFL_HIP(wall_start_d_.reserve(start.size(), 0, stream_), "hipMalloc(wall table)");
    SFL_HIP(hipMemcpyAsync(walls_d_.ptr, dev.data(), n * sizeof(float4), hipMemcpyHostToDevice, stream_),
            "upload walls");
    SFL_HIP(hipMemcpyAsync(wall_start_d_.ptr, start.data(), start.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, stream_),
            "upload wall table");
    if (searched) {
      SFL_HIP(wall_keys_d_.reserve(keys.size(), 0, stream_), "hipMalloc(wall table)");
      SFL_HIP(hipMemcpyAsync(wall_keys_d_.ptr, keys.data(), keys.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                             stream_),
              "upload wall table");
    }
    SFL_HIP(hipStreamSynchronize(stream_), "upload walls");
  }
  wall_pos_.swap(pos);
  wall_owner_.swap(own);
  wall_box_ = box;
  wall_keyed_ = searched ? (uint32_t)keys.size() : 0u;
  wall_cells_ = searched ? wall_keyed_ : (uint32_t)volume;
  // union of the containers (boundary.rs:240-247)
  has_bounds_ = !containers_.empty();
  for (size_t i = 0; i < containers_.size(); ++i)
    for (int a = 0; a < 3; ++a) {
      blo_[a] = i == 0 ? containers_[i].lo[a] : fmin_rust(blo_[a], containers_[i].lo[a]);
      bhi_[a] = i == 0 ? containers_[i].hi[a] : fmax_rust(bhi_[a], containers_[i].hi[a]);
    }
  invalidate_lists();
  return SFL_OK;
}

int32_t World::read_walls(size_t cap, float* out) const {
  size_t n = std::min(cap, wall_owner_.size());
  std::copy(wall_pos_.begin(), wall_pos_.begin() + 3 * n, out);
  return SFL_OK;
}

// ---- stepping ------------------------------------------------------------------

StepParams World::params(float dt) const {
  const sfl_settings& s = settings_;
  StepParams p{};
  p.n = (uint32_t)n_;
  p.radius = radius_;
  p.rinv = 1.0f / radius_;
  p.reach = reach_;
  p.scale = scale_;
  p.mass = mass_;
  p.rest = s.rest_density;
  float c = s.sound_speed;
  p.stiffness = s.rest_density * (c * c) / 7.0f;  // step.rs:149
  p.dt = dt;
  float r = radius_;
  p.nu = fmin_rust(s.viscosity * c * r, 0.05f * r * r / dt);         // step.rs:254-259
  p.wall_nu = fmin_rust(1.0f * c * r, 0.25f * (r * r / dt));         // step.rs:266, 282-283
  p.soften = 0.01f * r * r;                                          // step.rs:165
  p.gx = s.gravity.x;
  p.gy = s.gravity.y;
  p.gz = s.gravity.z;
  p.half_spacing = s.spacing / 2.0f;  // step.rs:308
  p.has_bounds = has_bounds_ ? 1 : 0;
  for (int a = 0; a < 3; ++a) {
    p.blo[a] = blo_[a];
    p.bhi[a] = bhi_[a];
  }
  p.walls = walls_d_.ptr;
  p.wall_start = wall_start_d_.ptr;
  p.wall_keys = wall_keyed_ ? wall_keys_d_.ptr : nullptr;
  p.wall_keyed = wall_keyed_;
  p.wall_box = wall_box_;
  p.wall_cells = wall_owner_.empty() ? 0u : wall_cells_;
  p.cell_start = searched_ ? sstart_.ptr : cell_start_.ptr;
  p.cell_keys = searched_ ? cell_keys_.ptr : nullptr;
  p.ctl = ctl_;
  p.pairs = pairs_.ptr;
  p.pair_counts = pair_counts_.ptr;
  p.pair_stride = pair_stride_;
  p.pair_cap = pair_cap_;
  p.masks = masks_.ptr;
  p.own_slot = own_slot_.ptr;
  p.near = near_.ptr;
  p.factor = factor_.ptr;
  p.format = format_;
  p.terrain = terrain_view();
  return p;
}

World::Totals World::totals() const {
  Totals t{0.0f, 0.0f, 0, 0};
  uint32_t vmax = 0, comp = 0;
  for (int k = 0; k < kStripes; ++k) {
    const Stripe& s = ctl_host_->stripes[k];
    vmax = std::max(vmax, s.vmax_bits);
    comp = std::max(comp, s.comp_bits);
    t.held += s.held;
    t.pairs_max = std::max(t.pairs_max, s.pairs_max);
  }
  std::memcpy(&t.vmax, &vmax, 4);
  std::memcpy(&t.comp, &comp, 4);
  return t;
}

// Lists of `cap` entries for each of n_ particles (column-major, stride n_).
int32_t World::ensure_pairs(uint32_t cap) {
  uint32_t stride = (uint32_t)std::max<size_t>(n_, 1);
  if (cap <= pair_cap_ && stride <= pair_stride_ && pairs_.ptr) return SFL_OK;
  cap = std::max(cap, pair_cap_);
  stride = std::max(stride, pair_stride_);
  pairs_.release();
  // kListSlack rows past the capacity: the solve's loads run ahead of a
  // list's end
  const size_t rows = (size_t)cap + kListSlack;
  SFL_HIP(pairs_.reserve(rows * stride, 0, stream_), "hipMalloc(neighbour lists)");
  SFL_HIP(pair_counts_.reserve(stride, 0, stream_), "hipMalloc(neighbour counts)");
  if (format_ == kMasks16 || format_ == kMasks32) {
    // 27 masks per particle at the widest mask (32 bits)
    SFL_HIP(masks_.reserve((size_t)27 * stride, 0, stream_), "hipMalloc(neighbour masks)");
    SFL_HIP(own_slot_.reserve(stride, 0, stream_), "hipMalloc(neighbour masks)");
  }
  if (format_ == kCompact) {
    // groups of four entries a lane: (rows rounded up to 4) * stride
    const size_t grouped = (rows + 3) / 4 * 4 * stride;
    SFL_HIP(near_.reserve(grouped, 0, stream_), "hipMalloc(compact neighbours)");
    SFL_HIP(factor_.reserve(grouped, 0, stream_), "hipMalloc(compact neighbours)");
  }
  pair_cap_ = cap;
  pair_stride_ = stride;
  return SFL_OK;
}

uint64_t World::dense_limit() const { return std::max<uint64_t>(64 * (uint64_t)n_ + 4096, kDenseFloor); }

// The searched table's buffers for the current particle count; the sorts
// from here on go through it.
int32_t World::use_searched() {
  const size_t n = std::max<size_t>(n_, 1);
  SFL_HIP(skeys_in_.reserve(n, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(skeys_.reserve(n, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(svals_in_.reserve(n, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(svals_.reserve(n, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(heads_.reserve(n, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(head_scan_.reserve(n + 1, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(cell_keys_.reserve(n, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(sstart_.reserve(n + 1, 0, stream_), "hipMalloc(searched table)");
  SFL_HIP(tile_sums_.reserve((n + 2048) / 2048 + 1, 0, stream_), "hipMalloc(tile sums)");
  size_t temp = 0;
  SFL_HIP(searched_sort_bytes((uint32_t)n, &temp), "radix sort size");
  SFL_HIP(ssort_temp_.reserve(std::max<size_t>(temp, 1), 0, stream_), "hipMalloc(radix sort)");
  searched_ = true;
  return SFL_OK;
}

int32_t World::ensure_cells(uint64_t volume) {
  if (volume <= cell_cap_ && counts_.ptr) return SFL_OK;
  if (no_alloc_) {
    detail_ = "the fluid spreads over more cells than the allocation-free step's scratch holds";
    return SFL_OUT_OF_RANGE;
  }
  if (volume == kSearchedTable) {
    detail_ = "the fluid spread over more cells than a 63-bit cell key addresses";
    return SFL_OUT_OF_RANGE;
  }
  if (volume > dense_limit() || volume > kMaxCells) {
    // too thin for a dense table: search the occupied cells, as the
    // reference does (mod.rs:214-217). The DFSPH stop sum runs over the
    // tiles of the dense box, so a DFSPH world keeps a dense table.
    if (settings_.solver == SFL_WCSPH) {
      last_box_ = volume;  // the box that sent it there, until a searched sort reports its own
      return use_searched();
    }
    if (volume > kMaxCells) {
      detail_ = "a DFSPH world spread over more cells than the dense cell table holds";
      return SFL_OUT_OF_RANGE;
    }
  }
  uint64_t cap = std::min<uint64_t>(volume + volume / 4 + 4096, kMaxCells);
  counts_.release();
  cell_start_.release();
  SFL_HIP(hipStreamSynchronize(stream_), "sync");
  SFL_HIP(counts_.reserve(cap, 0, stream_), "hipMalloc(cell counts)");
  SFL_HIP(hipMemsetAsync(counts_.ptr, 0, cap * sizeof(uint32_t), stream_), "clear cell counts");
  SFL_HIP(cell_start_.reserve(cap + 1, 0, stream_), "hipMalloc(cell table)");
  size_t tiles = (cap + 2047) / 2048 + 1;
  SFL_HIP(tile_sums_.reserve(std::max<size_t>(tiles, (n_ + 1 + 2047) / 2048), 0, stream_), "hipMalloc(tile sums)");
  // solve tiles of a box of V cells: ceil(sx/4) ceil(sy/2) ceil(sz/16), at
  // most max(1, 2V/3) (sy = 1 and sx or sz >= 2 halves a factor; sy >= 2
  // gives ceil(sy/2) <= 2sy/3)
  const size_t stop = 2 * cap / 3 + 2;
  SFL_HIP(stop_sum_.reserve(2 * stop, 0, stream_), "hipMalloc(stop sums)");
  SFL_HIP(stop_count_.reserve(2 * stop, 0, stream_), "hipMalloc(stop counts)");
  if (flow_ && format_ == kCompact) {
    SFL_HIP(flow_done_.reserve(2 * stop, 0, stream_), "hipMalloc(flow tiles)");  // one half a solve
    SFL_HIP(flow_work_.reserve(stop, 0, stream_), "hipMalloc(flow tiles)");
    SFL_HIP(flow_list_.reserve(stop, 0, stream_), "hipMalloc(flow tiles)");
  }
  SFL_HIP(hipStreamSynchronize(stream_), "sync");
  cell_cap_ = cap;
  return SFL_OK;
}

// A first table size: the containers' cells with a margin, or a small one.
// A world on the searched table goes back to a dense one once its last box
// is well within the dense limit.
int32_t World::ensure_cells_for_bounds() {
  if (searched_) {
    if (last_box_ > dense_limit() / 2) return use_searched();
    searched_ = false;
  }
  uint64_t want = std::max<uint64_t>(4096, 2 * (uint64_t)n_);
  if (has_bounds_) {
    uint64_t v = 1;
    for (int a = 0; a < 3; ++a) {
      int64_t lo = (int64_t)cell_coord_host(blo_[a], radius_) - 1;
      int64_t hi = (int64_t)cell_coord_host(bhi_[a], radius_) + 1;
      uint64_t s = (uint64_t)std::max<int64_t>(hi - lo + 1, 1);
      v = std::min<uint64_t>(v * s, kMaxCells);
    }
    want = std::max(want, v);
  }
  return ensure_cells(std::min<uint64_t>(want, std::min<uint64_t>(dense_limit(), kMaxCells)));
}

int32_t World::read_control() {
  SFL_HIP(hipMemcpyAsync(ctl_host_, ctl_, offsetof(Control, residual_log), hipMemcpyDeviceToHost, stream_),
          "read control");
  SFL_HIP(hipStreamSynchronize(stream_), "step");
  collect_profile();
  // a searched sort leaves its box's volume here
  if (searched_ && ctl_host_->need_volume != kSearchedTable) last_box_ = ctl_host_->need_volume;
  return SFL_OK;
}

int32_t World::ensure_fastest() {
  if (vmax_valid_) return SFL_OK;
  launch_step_begin(stream_, ctl_);
  launch_substep_begin(stream_, ctl_);
  launch_fastest(stream_, va_.ptr, (uint32_t)n_, ctl_);
  int32_t st = read_control();
  if (st != SFL_OK) return st;
  vmax_ = totals().vmax;
  vmax_valid_ = true;
  return SFL_OK;
}

void World::phase_begin(int phase) {
  if (!profiling_ || no_alloc_) return;
  hipEvent_t e;
  if (free_events_.empty()) {
    if (hipEventCreate(&e) != hipSuccess) return;
  } else {
    e = free_events_.back();
    free_events_.pop_back();
  }
  (void)hipEventRecord(e, stream_);
  open_ = e;
  (void)phase;
}

void World::phase_end(int phase) {
  if (!profiling_ || no_alloc_ || !open_) return;
  hipEvent_t e;
  if (free_events_.empty()) {
    if (hipEventCreate(&e) != hipSuccess) return;
  } else {
    e = free_events_.back();
    free_events_.pop_back();
  }
  (void)hipEventRecord(e, stream_);
  pending_.push_back({phase, open_, e});
  open_ = nullptr;
}

void World::collect_profile() {
  for (auto& p : pending_) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      phase_ms_[p.phase] += ms;
      phase_launches_[p.phase] += 1;
    }
    free_events_.push_back(p.a);
    free_events_.push_back(p.b);
  }
  pending_.clear();
}

// update_neighbors' sort, gathers and grid build (src/fluid/mod.rs:540-566)
// from A into B.
int32_t World::enqueue_sort() {
  uint32_t n = (uint32_t)n_;
  grid_any_ = n > 0;
  grid_dropped_ = false;
  ++sorts_;
  launch_substep_begin(stream_, ctl_);
  phase_begin(SFL_PHASE_BOUNDS);
  launch_bounds(stream_, pa_.ptr, n, radius_, ctl_);
  launch_grid_setup(stream_, ctl_, n, searched_ ? kSearchedTable : cell_cap_);
  phase_end(SFL_PHASE_BOUNDS);
  if (searched_) {
    phase_begin(SFL_PHASE_SORT);
    const SearchedSort t{skeys_in_.ptr, skeys_.ptr, svals_in_.ptr, svals_.ptr, heads_.ptr, head_scan_.ptr,
                         tile_sums_.ptr, scan_total_.ptr, cell_keys_.ptr, sstart_.ptr, ssort_temp_.ptr,
                         ssort_temp_.cap};
    hipError_t e = launch_searched_sort(stream_, t, pa_.ptr, va_.ptr, n, radius_, pb_.ptr, vb_.ptr, ctl_);
    phase_end(SFL_PHASE_SORT);
    if (e != hipSuccess) return device_error("radix sort (searched cell table)", e);
    return SFL_OK;
  }
  phase_begin(SFL_PHASE_SORT);
  launch_count(stream_, pa_.ptr, n, radius_, counts_.ptr, slot_rank_.ptr, ctl_);
  launch_scan(stream_, counts_.ptr, cell_start_.ptr, tile_sums_.ptr, scan_total_.ptr, &ctl_->volume, cell_cap_, 1,
              ctl_);
  launch_scatter(stream_, pa_.ptr, n, slot_rank_.ptr, cell_start_.ptr, scatter_.ptr, ctl_);
  phase_end(SFL_PHASE_SORT);
  phase_begin(SFL_PHASE_REORDER);
  launch_rank_gather(stream_, pa_.ptr, va_.ptr, n, radius_, scatter_.ptr, cell_start_.ptr, pb_.ptr, vb_.ptr, ctl_);
  phase_end(SFL_PHASE_REORDER);
  return SFL_OK;
}

// wcsph_substep (src/fluid/step.rs:144-224), substep `index` of the step;
// with coupled bodies in the step their points are carried to index * small
// first (step.rs:134-137), and each chunk's damping goes to the bodies after.
int32_t World::enqueue_wcsph_substep(const StepParams& base, uint32_t index, float small) {
  StepParams prm = base;
  if (movers_m_) {
    phase_begin(SFL_PHASE_BODIES);
    int32_t st = carry_movers(small * (float)index);
    phase_end(SFL_PHASE_BODIES);
    if (st != SFL_OK) return st;
    prm.movers = mover_view(prm.dt);
  }
  int32_t st = enqueue_sort();
  if (st != SFL_OK) return st;
  phase_begin(SFL_PHASE_DENSITY);
  launch_wcsph_density(stream_, pb_.ptr, vb_.ptr, aux_.ptr, prm);
  phase_end(SFL_PHASE_DENSITY);
  if (movers_m_) {
    phase_begin(SFL_PHASE_BODIES);
    launch_pairs_prepare(stream_, pair_room(), prm.n, ctl_);
    phase_end(SFL_PHASE_BODIES);
  }
  phase_begin(SFL_PHASE_FORCE);
  launch_wcsph_force(stream_, pb_.ptr, vb_.ptr, aux_.ptr, prm, pa_.ptr, va_.ptr,
                     keep_forces_ ? forces_.ptr : nullptr);
  phase_end(SFL_PHASE_FORCE);
  if (movers_m_) {
    phase_begin(SFL_PHASE_BODIES);
    hipError_t e = launch_pairs_fold(stream_, pair_room(), bodies_d_.ptr, (uint32_t)coupled_.size(), prm.dt, ctl_);
    phase_end(SFL_PHASE_BODIES);
    if (e != hipSuccess) return device_error("radix sort (body pairs)", e);
  }
  return SFL_OK;
}

// dfsph_substep (src/fluid/dfsph.rs:79-103)
int32_t World::enqueue_dfsph_substep(const StepParams& prm) {
  const sfl_settings& s = settings_;
  int32_t st = enqueue_sort();
  if (st != SFL_OK) return st;
  // the flow solve's tile set-up, once a substep, runs as one more block of
  // the density launch (a stream of its own would take one of the process's
  // few hardware queues from every other stream, a tape's replay stream
  // among them)
  const FlowBuffers flow = flow_buffers(va_.ptr);
  FlowSetup setup;
  if (dfsph_solve_tiled(prm, solve_lds_) && flow.v2) {
    setup.tile_sum = stop_sum_.ptr;
    setup.tile_count = stop_count_.ptr;
    setup.tile_cap = stop_cap();
    setup.tile_done = flow.tile_done;
    setup.tile_work = flow.tile_work;
    setup.tile_list = flow.tile_list;
  }
  phase_begin(SFL_PHASE_DENSITY);
  SFL_HIP(launch_dfsph_density(stream_, pb_.ptr, vb_.ptr, aux_.ptr, prm, setup), "density pass");
  phase_end(SFL_PHASE_DENSITY);
  phase_begin(SFL_PHASE_SOLVE);
  SFL_HIP(launch_dfsph_solve(stream_, prm, pb_.ptr, vb_.ptr, aux_.ptr, kappa_.ptr, errors_.ptr, stop_sum_.ptr,
                             stop_count_.ptr, stop_cap(), 0, s.divergence_tolerance, s.min_iterations, s.max_iterations,
                             solve_grid_, solve_lds_, nullptr, nullptr, flow_buffers(va_.ptr)),
          "cooperative launch (divergence solve)");
  phase_end(SFL_PHASE_SOLVE);
  // the tiled density solve runs the viscosity pass (VB -> VA) itself
  const bool fused = dfsph_solve_tiled(prm, solve_lds_);
  if (!fused) {
    phase_begin(SFL_PHASE_VISCOUS);
    launch_dfsph_viscous(stream_, pb_.ptr, vb_.ptr, prm, va_.ptr, keep_forces_ ? forces_.ptr : nullptr);
    phase_end(SFL_PHASE_VISCOUS);
  }
  phase_begin(SFL_PHASE_SOLVE);
  SFL_HIP(launch_dfsph_solve(stream_, prm, pb_.ptr, va_.ptr, aux_.ptr, kappa_.ptr, errors_.ptr, stop_sum_.ptr,
                             stop_count_.ptr, stop_cap(), 1, s.density_tolerance, s.min_iterations, s.max_iterations,
                             solve_grid_, solve_lds_, fused ? vb_.ptr : nullptr,
                             fused && keep_forces_ ? forces_.ptr : nullptr, flow_buffers(vb_.ptr)),
          "cooperative launch (density solve)");
  phase_end(SFL_PHASE_SOLVE);
  phase_begin(SFL_PHASE_ADVANCE);
  launch_dfsph_advance(stream_, pb_.ptr, prm, pa_.ptr, va_.ptr, vb_.ptr);
  phase_end(SFL_PHASE_ADVANCE);
  return SFL_OK;
}

int32_t World::step(float dt, sfl_report* out) {
  ++version_;
  DeviceGuard g(device_);
  sfl_report r{};
  if (!std::isfinite(dt) || dt < 0.0f) {
    detail_ = "fluid step";
    return SFL_NON_FINITE;
  }
  if (dt == 0.0f) {
    if (out) *out = r;
    return SFL_OK;
  }
  if (n_ > 0) {
    if (keep_forces_) {
      int32_t st = ensure_particles(n_);
      if (st != SFL_OK) return st;
    }
    int32_t st = settings_.solver == SFL_WCSPH ? wcsph(dt, r) : dfsph(dt, r);
    if (st != SFL_OK) return st;
    r.max_speed = vmax_;
    r.max_compression = totals().comp;
  }
  // flows, then the domain (src/fluid/step.rs:76-82)
  bool changed = false;
  uint32_t emitted = 0, drained = 0, lost = 0;
  const bool flowing = !flows_.empty() || has_domain_;
  if (flowing) phase_begin(SFL_PHASE_FLOWS);
  int32_t st = run_flows(emitted, drained, changed);
  if (st != SFL_OK) {
    vmax_valid_ = false;
    return st;
  }
  if (has_domain_) {
    st = keep_only(domain_, true, lost);
    if (st != SFL_OK) return st;
    changed = changed || lost > 0;
  }
  r.emitted = emitted;
  r.drained = drained;
  r.lost = lost;
  if (changed) {
    // the report's fastest and most squeezed over the particles now there
    // (step.rs:83-88)
    launch_flow_totals(stream_, va_.ptr, (uint32_t)n_, settings_.rest_density, ctl_);
    st = read_control();
    if (st != SFL_OK) return st;
    const Totals t = totals();
    r.max_speed = t.vmax;
    r.max_compression = t.comp;
    vmax_ = t.vmax;
    vmax_valid_ = true;
  }
  if (flowing) phase_end(SFL_PHASE_FLOWS);
  if (out) *out = r;
  return SFL_OK;
}

// The allocation-free step (sfl_step_no_alloc): World::step with the
// control block and the sort's cell table (counts, their scan, the scan's
// tile sums and total) lent from the shared scratch of this GPU. The particle
// state and its sorted copy are the world's own, sized when the particles
// were added, so nothing needs to grow; a buffer that would have to fails
// the call (t_no_alloc) instead of allocating. The last grid is copied into
// the world's own table when it has one large enough, else it is dropped
// with the lists that refer to it.
int32_t World::step_no_alloc(float dt, sfl_report* out) {
  DeviceGuard g(device_);
  if (settings_.solver != SFL_WCSPH) {
    detail_ = "the allocation-free step runs WCSPH worlds: a DFSPH step keeps neighbour lists that grow with the water";
    return SFL_UNSUPPORTED;
  }
  if (!coupled_.empty() || !terrains_.empty() || !flows_.empty() || has_domain_) {
    detail_ = "the allocation-free step runs worlds without coupled bodies, terrain, flows or a domain";
    return SFL_UNSUPPORTED;
  }
  if (keep_forces_ && forces_.cap < 3 * n_) {
    detail_ = "the kept forces need a larger buffer than the world holds";
    return SFL_UNSUPPORTED;
  }
  if (searched_) {
    detail_ = "the fluid spreads over more cells than the allocation-free step's scratch holds";
    return SFL_OUT_OF_RANGE;
  }
  if (n_ == 0 || !(dt > 0.0f) || !std::isfinite(dt)) {
    // no device work: the refusals and the empty step of World::step
    t_no_alloc = true;
    int32_t st = step(dt, out);
    t_no_alloc = false;
    return st;
  }
  unsigned char* base = nullptr;
  SFL_HIP(no_alloc_scratch(device_, &base), "hipGetSymbolAddress(scratch)");
  // layout, each piece on 256 bytes: control block, scan total, cell counts,
  // cell table (cap + 1), scan tile sums
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t head = up(sizeof(Control)) + up(sizeof(uint32_t));
  const size_t room = kNoAllocScratchBytes - head;
  auto tiles_for = [&](uint64_t c) { return std::max<uint64_t>((c + 2047) / 2048 + 1, ((uint64_t)n_ + 2048) / 2048); };
  auto need = [&](uint64_t c) { return up(4 * c) + up(4 * (c + 1)) + up(4 * tiles_for(c)); };
  uint64_t cap = room / 8;
  while (cap > 0 && need(cap) > room) cap -= std::min<uint64_t>(cap, 4096);
  if (cap < 4096) {
    detail_ = "the world holds more particles than the allocation-free step's scratch can sort";
    return SFL_OUT_OF_RANGE;
  }
  uint32_t* counts = reinterpret_cast<uint32_t*>(base + head);
  uint32_t* starts = reinterpret_cast<uint32_t*>(base + head + up(4 * cap));
  uint32_t* tiles = reinterpret_cast<uint32_t*>(base + head + up(4 * cap) + up(4 * (cap + 1)));

  struct Lease {
    World& w;
    Control* ctl;
    uint64_t cell_cap;
    DeviceBuffer<uint32_t> counts, starts, tiles, total;  // the world's own, set aside
    static void lend(DeviceBuffer<uint32_t>& member, DeviceBuffer<uint32_t>& keep, uint32_t* view, size_t cap) {
      keep.swap(member);
      member.ptr = view;
      member.cap = cap;
    }
    static void give_back(DeviceBuffer<uint32_t>& member, DeviceBuffer<uint32_t>& keep) {
      member.ptr = nullptr;  // a view into the scratch: never freed
      member.cap = 0;
      member.swap(keep);
    }
    ~Lease() {
      give_back(w.counts_, counts);
      give_back(w.cell_start_, starts);
      give_back(w.tile_sums_, tiles);
      give_back(w.scan_total_, total);
      w.ctl_ = ctl;
      w.cell_cap_ = cell_cap;
      w.no_alloc_ = false;
      t_no_alloc = false;
    }
  } lease{*this, ctl_, cell_cap_, {}, {}, {}, {}};
  Lease::lend(counts_, lease.counts, counts, cap);
  Lease::lend(cell_start_, lease.starts, starts, cap + 1);
  Lease::lend(tile_sums_, lease.tiles, tiles, tiles_for(cap));
  Lease::lend(scan_total_, lease.total, reinterpret_cast<uint32_t*>(base + up(sizeof(Control))), 1);
  ctl_ = reinterpret_cast<Control*>(base);
  cell_cap_ = cap;
  no_alloc_ = true;
  t_no_alloc = true;
  // the control block and the counts start at zero; the rest is written
  // before it is read
  launch_fill_zero(stream_, base, head + up(4 * cap));
  const uint64_t sorts = sorts_;
  int32_t st = step(dt, out);
  if (sorts_ != sorts) {
    // the last sort's grid: the cell box (lo, span, volume) and the table
    const uint64_t volume = ctl_host_->volume;
    if (st == SFL_OK && lists_valid_ && lease.counts.ptr && volume <= lease.cell_cap) {
      SFL_HIP(hipMemcpyAsync(lease.starts.ptr, starts, (volume + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                             stream_),
              "keep the cell table");
      SFL_HIP(hipMemcpyAsync(lease.ctl, ctl_, offsetof(Control, flags), hipMemcpyDeviceToDevice, stream_),
              "keep the cell box");
      SFL_HIP(hipStreamSynchronize(stream_), "keep the grid");
    } else {
      invalidate_lists();
      grid_dropped_ = true;
    }
  }
  return st;
}

// wcsph / wcsph_steps (src/fluid/step.rs:102-142): the step's sound speed
// is raised to IMPACT_SOUND (4, step.rs:20) times the fastest coupled body
// near water, and its substeps follow that body too.
int32_t World::wcsph(float dt, sfl_report& r) {
  int32_t st = ensure_fastest();
  if (st != SFL_OK) return st;
  const float configured = settings_.sound_speed;
  float mover = 0.0f;
  if (movers_m_) {
    st = fastest_near_water(dt, configured / 4.0f, mover);
    if (st != SFL_OK) return st;
  }
  struct Restore {
    sfl_settings& s;
    float c;
    ~Restore() { s.sound_speed = c; }
  } restore{settings_, configured};
  settings_.sound_speed = fmax_rust(configured, 4.0f * mover);
  const sfl_settings& s = settings_;
  r.sound_speed = s.sound_speed;
  float largest = s.courant * radius_ / (s.sound_speed + fmax_rust(vmax_, mover));
  float wanted = fmax_rust(std::ceil(dt / largest), 1.0f);
  float cap = (float)s.max_substeps;
  float take = fmin_rust(wanted, cap);
  uint32_t substeps = take >= 4294967295.0f ? 0xffffffffu : (uint32_t)take;
  r.capped = wanted > cap ? 1u : 0u;
  float small = dt / (float)substeps;
  st = ensure_cells_for_bounds();
  if (st != SFL_OK) return st;
  st = upload_terrains();
  if (st != SFL_OK) return st;
  if (!terrains_.empty()) SFL_HIP(slopes_.reserve(n_, 0, stream_), "hipMalloc(ground slopes)");
  StepParams prm = params(small);
  uint32_t done = 0, held = 0;
  lists_valid_ = false;
  if (movers_m_) {
    // per particle its body points in reach, and room for their pair records
    SFL_HIP(pair_counts_b_.reserve(n_, 0, stream_), "hipMalloc(body pairs)");
    SFL_HIP(pair_offsets_b_.reserve(n_ + 1, 0, stream_), "hipMalloc(body pairs)");
    SFL_HIP(pair_tiles_b_.reserve(n_ / 2048 + 2, 0, stream_), "hipMalloc(body pairs)");
    SFL_HIP(pair_total_b_.reserve(1, 0, stream_), "hipMalloc(body pairs)");
    st = ensure_pair_room(std::max<uint64_t>(body_pair_cap_, std::max<uint64_t>((uint64_t)movers_m_ * 16, pair_slack_)));
    if (st != SFL_OK) return st;
  }
  bool retry = false;
  while (done < substeps) {
    launch_step_begin(stream_, ctl_);
    if (retry && movers_m_) {
      // the points' order after `done` substeps, carried again from where
      // they were placed: the aborted launches left the buffers behind the
      // host's count of carries
      st = replay_movers(done, small);
      if (st != SFL_OK) return st;
    }
    retry = true;
    for (uint32_t k = done; k < substeps; ++k) {
      st = enqueue_wcsph_substep(prm, k, small);
      if (st != SFL_OK) return st;
    }
    st = read_control();
    if (st != SFL_OK) return st;
    const Control& c = *ctl_host_;
    if (movers_m_ && !(c.flags & (kFlagOverflow | kFlagPairRoom))) {
      // the next step sorts what this one used at most, and a quarter more
      const uint64_t want = (uint64_t)c.pairs_used + c.pairs_used / 4 + pair_slack_;
      body_pair_use_ = (uint32_t)std::min<uint64_t>(want, body_pair_cap_);
    }
    if (c.flags & kFlagMoverBox) {
      detail_ = "the coupled bodies' points spread over more cells than their sort key holds";
      return SFL_UNSUPPORTED;
    }
    if (c.flags & (kFlagOverflow | kFlagPairRoom)) {
      uint32_t completed = c.substeps_done ? c.substeps_done - 1 : 0;
      held += totals().held;
      done += completed;
      work_substeps_ += (uint64_t)completed * n_;
      if (completed > 0) {
        lists_valid_ = true;
        lists_on_b_ = true;
        lists_emitted_ = false;
      }
      if (c.flags & kFlagOverflow) st = ensure_cells(c.need_volume);
      if (st == SFL_OK && (c.flags & kFlagPairRoom)) {
        body_pair_use_ = 0;  // sort the whole room until a step says how much it needs
        st = ensure_pair_room(c.need_pairs + c.need_pairs / 4 + 1024);
        if (st == SFL_OK && body_pair_cap_ < c.need_pairs) {
          detail_ = "the coupled bodies' points pair with more water than 2^32 records hold";
          return SFL_UNSUPPORTED;
        }
      }
      if (st != SFL_OK) return st;
      prm = params(small);
      continue;
    }
    lists_valid_ = true;
    lists_on_b_ = true;
    lists_emitted_ = false;
    if (c.flags & kFlagNonFinite) {
      launch_wcsph_repair(stream_, pb_.ptr, vb_.ptr, aux_.ptr, prm, pa_.ptr, va_.ptr);
      SFL_HIP(hipStreamSynchronize(stream_), "repair");
      vmax_valid_ = false;
      detail_ = "fluid particle";
      return SFL_NON_FINITE;
    }
    held += totals().held;
    work_substeps_ += (uint64_t)(substeps - done) * n_;
    done = substeps;
  }
  vmax_ = totals().vmax;
  vmax_valid_ = true;
  forces_valid_ = keep_forces_;
  r.substeps = substeps;
  r.held_back = held;
  return SFL_OK;
}

// dfsph (src/fluid/dfsph.rs:46-77); ground pushes through the weakly
// compressible solver's pressure only (dfsph.rs:52-55)
int32_t World::dfsph(float dt, sfl_report& r) {
  if (!terrains_.empty()) return SFL_INVALID_SETTINGS;
  int32_t st = ensure_fastest();
  if (st != SFL_OK) return st;
  const sfl_settings& s = settings_;
  float spacing = s.spacing, courant = s.courant;
  float pull = length3(s.gravity.x, s.gravity.y, s.gravity.z);
  st = ensure_cells_for_bounds();
  if (st != SFL_OK) return st;
  // list entries keep the neighbour's index in their low 28 bits: past that
  // many particles the masks hold the fluid neighbours
  if (format_ == kLists && n_ > (size_t)kListIndexMask + 1) {
    format_ = kMasks32;
    pair_cap_ = 0;
  }
  // lists of walls only in the mask formats, of every neighbour otherwise
  st = ensure_pairs(std::max<uint32_t>(pair_cap_, format_ == kLists || format_ == kCompact ? 64 : 16));
  if (st != SFL_OK) return st;
  float left = dt;
  lists_valid_ = false;
  while (left > 0.0f) {
    float small = courant * spacing / fmax_rust(vmax_, 1e-6f);
    if (pull > 0.0f) small = fmin_rust(small, std::sqrt(2.0f * courant * spacing / pull));
    if (r.substeps + 1 >= s.max_substeps && small < left) {
      r.capped = 1;
      small = left;
    }
    small = fmin_rust(small, left);
    StepParams prm = params(small);
    for (;;) {
      launch_step_begin(stream_, ctl_);
      st = enqueue_dfsph_substep(prm);
      if (st != SFL_OK) return st;
      st = read_control();
      if (st != SFL_OK) return st;
      if (ctl_host_->flags & kFlagOverflow) {
        st = ensure_cells(ctl_host_->need_volume);
        if (st != SFL_OK) return st;
        prm = params(small);
        continue;
      }
      if (ctl_host_->flags & kFlagWide) {
        // a cell held more particles than a mask has bits: widen the masks,
        // or list every neighbour
        if (format_ == kMasks32 && n_ > (size_t)kListIndexMask + 1) {
          detail_ = "a cell holds more than 32 particles in a world too large for neighbour lists";
          return SFL_OUT_OF_RANGE;
        }
        // compact lists widen to full ones (their counts bound fluid and
        // walls apart), or to 32-bit masks past the lists' index range
        const int from = format_;
        if (from == kCompact)
          format_ = n_ > (size_t)kListIndexMask + 1 ? kMasks32 : kLists;
        else
          format_ = from == kMasks16 ? kMasks32 : kLists;
        uint32_t cap = pair_cap_;
        if (format_ == kLists)
          cap = from == kCompact ? 2 * totals().pairs_max + 16 : std::max<uint32_t>(totals().pairs_max + 64, 64);
        pair_cap_ = 0;  // reallocate with the new format
        st = ensure_pairs(cap);
        if (st != SFL_OK) return st;
        prm = params(small);
        continue;
      }
      if (ctl_host_->flags & kFlagPairs) {
        st = ensure_pairs(totals().pairs_max + 16);
        if (st != SFL_OK) return st;
        prm = params(small);
        continue;
      }
      break;
    }
    const Control& c = *ctl_host_;
    lists_valid_ = true;
    lists_on_b_ = true;
    lists_emitted_ = false;
    if (c.flags & kFlagBarrier) {
      detail_ = "a grid barrier of the persistent solve timed out";
      return SFL_DEVICE_ERROR;
    }
    if (c.flags & kFlagTiles) {
      detail_ = "the solve tiles outnumbered the room of the stop sum";
      return SFL_DEVICE_ERROR;
    }
    if (c.flags & kFlagNonFinite) {
      launch_dfsph_repair(stream_, pb_.ptr, vb_.ptr, prm, pa_.ptr, va_.ptr);
      SFL_HIP(hipStreamSynchronize(stream_), "repair");
      vmax_valid_ = false;
      detail_ = "fluid particle";
      return SFL_NON_FINITE;
    }
    Totals tot = totals();
    r.held_back += tot.held;
    r.iterations += c.iterations[0] + c.iterations[1];
    r.iterations_divergence += c.iterations[0];
    r.iterations_density += c.iterations[1];
    r.substeps += 1;
    work_substeps_ += n_;
    work_pushes_ += (uint64_t)(c.iterations[0] + c.iterations[1]) * n_;
    work_rates_ += (uint64_t)c.rates_step * n_;
    left -= small;
    vmax_ = tot.vmax;
  }
  vmax_valid_ = true;
  forces_valid_ = keep_forces_;
  return SFL_OK;
}

// ---- neighbour build and densities on their own ---------------------------

int32_t World::update_neighbors() {
  ++version_;
  DeviceGuard g(device_);
  if (n_ == 0) {
    lists_valid_ = true;
    lists_on_b_ = false;
    lists_emitted_ = false;
    grid_any_ = false;
    return SFL_OK;
  }
  int32_t st = ensure_cells_for_bounds();
  if (st != SFL_OK) return st;
  for (;;) {
    launch_step_begin(stream_, ctl_);
    st = enqueue_sort();
    if (st != SFL_OK) return st;
    st = read_control();
    if (st != SFL_OK) return st;
    if (ctl_host_->flags & kFlagOverflow) {
      st = ensure_cells(ctl_host_->need_volume);
      if (st != SFL_OK) return st;
      continue;
    }
    break;
  }
  pa_.swap(pb_);
  va_.swap(vb_);
  lists_valid_ = true;
  lists_on_b_ = false;
  lists_emitted_ = false;
  forces_valid_ = false;
  return SFL_OK;
}

int32_t World::emit_lists() {
  if (lists_emitted_) return SFL_OK;
  size_t n = n_;
  SFL_HIP(fstart_.reserve(n + 1, 0, stream_), "hipMalloc(list starts)");
  SFL_HIP(wstart_.reserve(n + 1, 0, stream_), "hipMalloc(wall list starts)");
  SFL_HIP(list_counts_.reserve(std::max<size_t>(n, 1), 0, stream_), "hipMalloc(list counts)");
  SFL_HIP(tile_sums_.reserve((n + 1 + 2047) / 2048 + 1, 0, stream_), "hipMalloc(tile sums)");
  if (n == 0) {
    uint32_t zero = 0;
    SFL_HIP(hipMemcpy(fstart_.ptr, &zero, 4, hipMemcpyHostToDevice), "list starts");
    SFL_HIP(hipMemcpy(wstart_.ptr, &zero, 4, hipMemcpyHostToDevice), "list starts");
    fitems_len_ = witems_len_ = 0;
    lists_emitted_ = true;
    return SFL_OK;
  }
  const float4* lp = lists_on_b_ ? pb_.ptr : pa_.ptr;
  StepParams prm = params(1.0f);
  for (int kind = 0; kind < 2; ++kind) {
    DeviceBuffer<uint32_t>& start = kind == 0 ? fstart_ : wstart_;
    DeviceBuffer<uint32_t>& items = kind == 0 ? fitems_ : witems_;
    launch_list_count(stream_, lp, prm, list_counts_.ptr, kind);
    launch_scan(stream_, list_counts_.ptr, start.ptr, tile_sums_.ptr, scan_total_.ptr, nullptr, n, 0, nullptr);
    uint32_t total = 0;
    SFL_HIP(hipMemcpyAsync(&total, scan_total_.ptr, 4, hipMemcpyDeviceToHost, stream_), "list total");
    SFL_HIP(hipStreamSynchronize(stream_), "list count");
    SFL_HIP(items.reserve(std::max<uint32_t>(total, 1), 0, stream_), "hipMalloc(list items)");
    launch_list_emit(stream_, lp, prm, start.ptr, items.ptr, kind);
    (kind == 0 ? fitems_len_ : witems_len_) = total;
  }
  SFL_HIP(hipStreamSynchronize(stream_), "list emit");
  lists_emitted_ = true;
  return SFL_OK;
}

// compute_densities (src/fluid/mod.rs:621-675) over the lists of the last
// build with the current positions.
int32_t World::compute_densities() {
  ++version_;
  DeviceGuard g(device_);
  if (!lists_valid_) return SFL_OUT_OF_RANGE;
  int32_t st = emit_lists();
  if (st != SFL_OK) return st;
  st = upload_terrains();
  if (st != SFL_OK) return st;
  StepParams prm = params(1.0f);
  launch_list_density(stream_, pa_.ptr, va_.ptr, prm, fstart_.ptr, fitems_.ptr, wstart_.ptr, witems_.ptr);
  SFL_HIP(hipStreamSynchronize(stream_), "densities");
  return SFL_OK;
}

// ---- readback -----------------------------------------------------------------

int32_t World::read_ids(size_t cap, uint32_t* out) {
  DeviceGuard g(device_);
  size_t n = std::min(cap, n_);
  if (!n) return SFL_OK;
  std::vector<float4> h(n);
  SFL_HIP(hipMemcpyAsync(h.data(), pa_.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "read ids");
  SFL_HIP(hipStreamSynchronize(stream_), "read ids");
  for (size_t i = 0; i < n; ++i) std::memcpy(out + i, &h[i].w, 4);
  return SFL_OK;
}

int32_t World::read_positions(size_t cap, float* out) {
  DeviceGuard g(device_);
  size_t n = std::min(cap, n_);
  if (!n) return SFL_OK;
  std::vector<float4> h(n);
  SFL_HIP(hipMemcpyAsync(h.data(), pa_.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "read positions");
  SFL_HIP(hipStreamSynchronize(stream_), "read positions");
  for (size_t i = 0; i < n; ++i) {
    out[3 * i] = h[i].x;
    out[3 * i + 1] = h[i].y;
    out[3 * i + 2] = h[i].z;
  }
  return SFL_OK;
}

int32_t World::read_velocities(size_t cap, float* out) {
  DeviceGuard g(device_);
  size_t n = std::min(cap, n_);
  if (!n) return SFL_OK;
  std::vector<float4> h(n);
  SFL_HIP(hipMemcpyAsync(h.data(), va_.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "read velocities");
  SFL_HIP(hipStreamSynchronize(stream_), "read velocities");
  for (size_t i = 0; i < n; ++i) {
    out[3 * i] = h[i].x;
    out[3 * i + 1] = h[i].y;
    out[3 * i + 2] = h[i].z;
  }
  return SFL_OK;
}

int32_t World::read_densities(size_t cap, float* out) {
  DeviceGuard g(device_);
  size_t n = std::min(cap, n_);
  if (!n) return SFL_OK;
  std::vector<float4> h(n);
  SFL_HIP(hipMemcpyAsync(h.data(), va_.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "read densities");
  SFL_HIP(hipStreamSynchronize(stream_), "read densities");
  for (size_t i = 0; i < n; ++i) out[i] = h[i].w;
  return SFL_OK;
}

int32_t World::keep_forces(bool keep) {
  DeviceGuard g(device_);
  keep_forces_ = keep;
  forces_valid_ = false;
  if (keep) {
    SFL_HIP(forces_.reserve(std::max<size_t>(3 * n_, 3), 0, stream_), "hipMalloc(forces)");
  }
  return SFL_OK;
}

int32_t World::read_forces(size_t cap, float* out) {
  DeviceGuard g(device_);
  if (!forces_valid_) return SFL_OUT_OF_RANGE;
  size_t n = std::min(cap, n_);
  if (!n) return SFL_OK;
  SFL_HIP(hipMemcpyAsync(out, forces_.ptr, 3 * n * sizeof(float), hipMemcpyDeviceToHost, stream_), "read forces");
  SFL_HIP(hipStreamSynchronize(stream_), "read forces");
  return SFL_OK;
}

int32_t World::read_neighbors(int kind, size_t start_cap, uint32_t* start, size_t items_cap, uint32_t* items,
                              size_t* items_len) {
  DeviceGuard g(device_);
  if (!lists_valid_) return SFL_OUT_OF_RANGE;
  int32_t st = emit_lists();
  if (st != SFL_OK) return st;
  size_t len = kind == 0 ? fitems_len_ : witems_len_;
  if (items_len) *items_len = len;
  if (start_cap < n_ + 1 || items_cap < len) return SFL_OUT_OF_RANGE;
  const DeviceBuffer<uint32_t>& s = kind == 0 ? fstart_ : wstart_;
  const DeviceBuffer<uint32_t>& it = kind == 0 ? fitems_ : witems_;
  SFL_HIP(hipMemcpyAsync(start, s.ptr, (n_ + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "read lists");
  if (len) SFL_HIP(hipMemcpyAsync(items, it.ptr, len * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "read lists");
  SFL_HIP(hipStreamSynchronize(stream_), "read lists");
  return SFL_OK;
}

int32_t World::neighbors_of(size_t index, size_t cap, uint32_t* out, size_t* len) {
  DeviceGuard g(device_);
  if (!lists_valid_ || index >= n_) return SFL_OUT_OF_RANGE;
  int32_t st = emit_lists();
  if (st != SFL_OK) return st;
  uint32_t range[2] = {0, 0};
  SFL_HIP(hipMemcpyAsync(range, fstart_.ptr + index, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
          "read list");
  SFL_HIP(hipStreamSynchronize(stream_), "read list");
  const size_t count = range[1] - range[0];
  if (len) *len = count;
  if (cap < count) return SFL_OUT_OF_RANGE;
  if (count) {
    SFL_HIP(hipMemcpyAsync(out, fitems_.ptr + range[0], count * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
            "read list");
    SFL_HIP(hipStreamSynchronize(stream_), "read list");
  }
  return SFL_OK;
}

uint64_t hash_particles(const float4* p, const float4* v, size_t n) {
  std::vector<std::pair<uint32_t, size_t>> order(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t id;
    std::memcpy(&id, &p[i].w, 4);
    order[i] = {id, i};
  }
  std::sort(order.begin(), order.end());
  uint64_t h = 0xcbf29ce484222325ull;
  auto feed = [&](uint32_t w) {
    for (int b = 0; b < 4; ++b) {
      h ^= (uint64_t)((w >> (8 * b)) & 0xffu);
      h *= 0x00000100000001b3ull;
    }
  };
  for (auto& o : order) {
    size_t i = o.second;
    feed(o.first);
    float vals[6] = {p[i].x, p[i].y, p[i].z, v[i].x, v[i].y, v[i].z};
    for (float f : vals) {
      uint32_t bits;
      std::memcpy(&bits, &f, 4);
      feed(bits);
    }
  }
  return h;
}

// state_hash (src/fluid/mod.rs:678-691) with FNV-1a (src/snapshot/bytes.rs:6-13)
int32_t World::state_hash(uint64_t* out) {
  DeviceGuard g(device_);
  size_t n = n_;
  std::vector<float4> p(n), v(n);
  if (n) {
    SFL_HIP(hipMemcpyAsync(p.data(), pa_.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "read state");
    SFL_HIP(hipMemcpyAsync(v.data(), va_.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, stream_), "read state");
    SFL_HIP(hipStreamSynchronize(stream_), "read state");
  }
  *out = hash_particles(p.data(), v.data(), n);
  return SFL_OK;
}

size_t World::device_bytes() const {
  auto b = [](const auto& buf) { return buf.cap * sizeof(*buf.ptr); };
  return sizeof(Control) + b(pa_) + b(va_) + b(pb_) + b(vb_) + b(aux_) + b(kappa_) + b(errors_) + b(forces_) +
         b(slot_rank_) + b(scatter_) + b(counts_) + b(cell_start_) + b(tile_sums_) + b(scan_total_) + b(pairs_) +
         b(masks_) + b(own_slot_) + b(near_) + b(factor_) +
         b(pair_counts_) + b(stop_sum_) + b(stop_count_) + b(flow_done_) + b(flow_work_) + b(flow_list_) + b(walls_d_) + b(wall_start_d_) + b(wall_keys_d_) + b(skeys_in_) + b(skeys_) + b(cell_keys_) +
         b(svals_in_) + b(svals_) + b(heads_) + b(head_scan_) + b(sstart_) + b(ssort_temp_) + b(fstart_) + b(fitems_) +
         b(wstart_) + b(witems_) + b(list_counts_) + b(flow_flags_) + b(flow_start_) + b(flow_tiles_) +
         b(flow_total_) + b(body_local_) + b(probe_local_) + b(bodies_d_) + b(ms_[0]) + b(ms_[1]) + b(mv_[0]) +
         b(mv_[1]) + b(mp_) + b(mkeys_in_) + b(mkeys_) + b(mperm_in_) + b(mperm_) + b(msort_temp_) + b(mgrid_) +
         b(mtable_) + b(pair_counts_b_) + b(pair_offsets_b_) + b(pair_tiles_b_) + b(pair_total_b_) +
         b(pair_key_[0]) + b(pair_key_[1]) + b(pair_value_[0]) + b(pair_value_[1]) + b(pair_force_) +
         b(pair_torque_) + b(pair_temp_) +
         b(extent_) + b(wet_) + b(lift_) + b(probe_arm_) + b(probe_vel_) + b(lift_near_) + b(lift_sort_) +
         b(lift_columns_) + b(lift_counts_) + b(lift_offsets_) + b(lift_tiles_) + b(lift_boxes_) +
         b(lift_ckeys_) + b(lift_ckeys2_) + b(lift_ukeys_) + b(lift_cvals_) + b(lift_cvals2_) + b(lift_sort_temp_) +
         b(terrain_desc_) + b(terrain_dist_) + b(far_d_) + b(damping_d_) + b(sites_d_) + b(slopes_);
}

int32_t World::synchronize() {
  DeviceGuard g(device_);
  SFL_HIP(hipStreamSynchronize(stream_), "synchronize");
  collect_profile();
  return SFL_OK;
}

int32_t World::profile(bool enable) {
  DeviceGuard g(device_);
  SFL_HIP(hipStreamSynchronize(stream_), "synchronize");
  collect_profile();
  profiling_ = enable;
  return SFL_OK;
}

int32_t World::profile_read(double* ms, uint64_t* launches, bool reset) {
  DeviceGuard g(device_);
  SFL_HIP(hipStreamSynchronize(stream_), "synchronize");
  collect_profile();
  for (int p = 0; p < SFL_PHASE_COUNT; ++p) {
    if (ms) ms[p] = phase_ms_[p];
    if (launches) launches[p] = phase_launches_[p];
    if (reset) {
      phase_ms_[p] = 0;
      phase_launches_[p] = 0;
    }
  }
  return SFL_OK;
}

int32_t World::read_solve_log(size_t cap, double* average, uint32_t* loaded, size_t* len) {
  DeviceGuard g(device_);
  uint32_t n = 0;
  SFL_HIP(hipMemcpyAsync(&n, &ctl_->log_len, sizeof(n), hipMemcpyDeviceToHost, stream_), "read solve log");
  SFL_HIP(hipStreamSynchronize(stream_), "read solve log");
  n = std::min<uint32_t>(n, (uint32_t)kResidualLogCap);
  if (len) *len = n;
  const size_t take = std::min<size_t>(cap, n);
  if (take && average)
    SFL_HIP(hipMemcpyAsync(average, ctl_->residual_log, take * sizeof(double), hipMemcpyDeviceToHost, stream_),
            "read solve log");
  if (take && loaded)
    SFL_HIP(hipMemcpyAsync(loaded, ctl_->loaded_log, take * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
            "read solve log");
  SFL_HIP(hipStreamSynchronize(stream_), "read solve log");
  return SFL_OK;
}

int32_t World::work_read(uint64_t* substeps, uint64_t* pushes, uint64_t* rates, bool reset) {
  if (substeps) *substeps = work_substeps_;
  if (pushes) *pushes = work_pushes_;
  if (rates) *rates = work_rates_;
  if (reset) work_substeps_ = work_pushes_ = work_rates_ = 0;
  return SFL_OK;
}

template struct DeviceBuffer<float4>;
template struct DeviceBuffer<float>;
template struct DeviceBuffer<uint2>;
template struct DeviceBuffer<uint32_t>;
template struct DeviceBuffer<double>;

}  // namespace sfl
