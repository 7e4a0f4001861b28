// TEST INFRASTRUCTURE, NOT PRODUCT CODE: C entry points over the CPU oracle
// (fluid_oracle.hpp) for the Python test harness, the smoke check and the
// CPU baseline leg of bench.py. The calls take the same shapes as the
// product's include/solunar_fluid.h with an `orc_` prefix, plus controls the
// reference's dispatcher gives its tests (thread count, a skipped job), the
// precision (32 or 64) and the DFSPH residual mode.
#include <cstdio>
#include <memory>

#include "adjoint_oracle.hpp"
#include "snapshot_oracle.hpp"
#include "fluid_oracle.hpp"

using namespace oracle;

extern "C" {

// Same layout as sfl_settings in include/solunar_fluid.h.
struct orc_settings {
  int32_t solver;
  float spacing, rest_density, support;
  uint64_t max_particles;
  float sound_speed, viscosity;
  float gx, gy, gz;
  float courant;
  uint32_t max_substeps;
  float divergence_tolerance, density_tolerance;
  uint32_t min_iterations, max_iterations;
  int32_t device;
};

struct orc_report {
  uint32_t substeps, capped;
  float max_compression, max_speed;
  uint32_t held_back, iterations, emitted, drained, lost;
  float sound_speed;
  uint32_t iterations_divergence, iterations_density;
};
}

namespace {

struct Handle {
  int precision;
  std::unique_ptr<World<float>> f;
  std::unique_ptr<World<double>> d;
  std::unique_ptr<AdjointTape<float>> tf;
  std::unique_ptr<AdjointTape<float, double>> tfd;
  std::unique_ptr<AdjointTape<double>> td;
};

Settings convert(const orc_settings* s) {
  Settings o;
  o.solver = s->solver;
  o.spacing = s->spacing;
  o.rest_density = s->rest_density;
  o.support = s->support;
  o.max_particles = s->max_particles;
  o.sound_speed = s->sound_speed;
  o.viscosity = s->viscosity;
  o.gravity = {s->gx, s->gy, s->gz};
  o.courant = s->courant;
  o.max_substeps = s->max_substeps;
  o.divergence_tolerance = s->divergence_tolerance;
  o.density_tolerance = s->density_tolerance;
  o.min_iterations = s->min_iterations;
  o.max_iterations = s->max_iterations;
  return o;
}

template <class F>
auto with(Handle* h, F&& fn) {
  if (h->precision == 64) return fn(*h->d);
  return fn(*h->f);
}

void fill_report(const Report& r, orc_report* out) {
  if (!out) return;
  out->substeps = r.substeps;
  out->capped = r.capped;
  out->max_compression = r.max_compression;
  out->max_speed = r.max_speed;
  out->held_back = r.held_back;
  out->iterations = r.iterations;
  out->emitted = r.emitted;
  out->drained = r.drained;
  out->lost = r.lost;
  out->sound_speed = r.sound_speed;
  out->iterations_divergence = r.iterations_divergence;
  out->iterations_density = r.iterations_density;
}

}  // namespace

extern "C" {

void orc_settings_default(orc_settings* out) {
  Settings d;
  out->solver = d.solver;
  out->spacing = d.spacing;
  out->rest_density = d.rest_density;
  out->support = d.support;
  out->max_particles = d.max_particles;
  out->sound_speed = d.sound_speed;
  out->viscosity = d.viscosity;
  out->gx = d.gravity.x;
  out->gy = d.gravity.y;
  out->gz = d.gravity.z;
  out->courant = d.courant;
  out->max_substeps = d.max_substeps;
  out->divergence_tolerance = d.divergence_tolerance;
  out->density_tolerance = d.density_tolerance;
  out->min_iterations = d.min_iterations;
  out->max_iterations = d.max_iterations;
  out->device = 0;
}

size_t orc_settings_size() { return sizeof(orc_settings); }
size_t orc_report_size() { return sizeof(orc_report); }

int32_t orc_create(const orc_settings* s, int32_t precision, void** out) {
  if (!s || !out) return kOutOfRange;
  Settings st = convert(s);
  if (!st.valid()) return kInvalidSettings;
  auto* h = new Handle();
  h->precision = precision == 64 ? 64 : 32;
  if (h->precision == 64)
    h->d.reset(new World<double>(st));
  else
    h->f.reset(new World<float>(st));
  *out = h;
  return kOk;
}

void orc_destroy(void* w) { delete static_cast<Handle*>(w); }

void orc_set_threads(void* w, int32_t threads) {
  with(static_cast<Handle*>(w), [&](auto& x) {
    x.dispatch.threads = threads < 1 ? 1 : threads;
    return 0;
  });
}

// Drops job `job` (-1: none) from every later dispatch, like
// dispatch::testing::SkipsJobs (src/dispatch.rs:50-65).
void orc_set_skip_job(void* w, int64_t job) {
  with(static_cast<Handle*>(w), [&](auto& x) {
    if (job < 0)
      x.dispatch.skip = nullptr;
    else
      x.dispatch.skip = [job](size_t j) { return (int64_t)j == job; };
    return 0;
  });
}

void orc_set_residual_mode(void* w, int32_t mode) {
  with(static_cast<Handle*>(w), [&](auto& x) {
    x.residual_mode = mode == 1 ? kFixedTreeF64 : kReferenceF32;
    return 0;
  });
}

void orc_set_solve_log(void* w, int32_t enable) {
  with(static_cast<Handle*>(w), [&](auto& x) {
    x.log_solves = enable != 0;
    x.solve_log.clear();
    return 0;
  });
}

// Rows of (solve, iteration, loaded, average_f32, average_f64).
size_t orc_read_solve_log(void* w, size_t cap, int32_t* solve, uint32_t* iteration, uint32_t* loaded,
                          float* avg32, double* avg64) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    size_t n = std::min(cap, x.solve_log.size());
    for (size_t i = 0; i < n; ++i) {
      solve[i] = x.solve_log[i].solve;
      iteration[i] = x.solve_log[i].iteration;
      loaded[i] = x.solve_log[i].loaded;
      avg32[i] = x.solve_log[i].average_f32;
      avg64[i] = x.solve_log[i].average_f64;
    }
    return x.solve_log.size();
  });
}

size_t orc_len(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.len(); });
}

double orc_particle_mass(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return (double)x.mass; });
}

double orc_kernel_value(void* w, double distance) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    return (double)x.kernel.value((R)distance);
  });
}

double orc_kernel_slope(void* w, double distance) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    return (double)x.kernel.slope((R)distance);
  });
}

int32_t orc_add_particles(void* w, size_t n, const float* xyz, const float* vxyz, uint32_t* ids_out) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    for (size_t i = 0; i < n; ++i) {
      Vec3<R> p{R(xyz[3 * i]), R(xyz[3 * i + 1]), R(xyz[3 * i + 2])};
      Vec3<R> v{R(0), R(0), R(0)};
      if (vxyz) v = {R(vxyz[3 * i]), R(vxyz[3 * i + 1]), R(vxyz[3 * i + 2])};
      int32_t st = x.add_particle(p, v, ids_out ? ids_out + i : nullptr);
      if (st != kOk) return st;
    }
    return (int32_t)kOk;
  });
}

int32_t orc_fill_box_jittered(void* w, float lx, float ly, float lz, float hx, float hy, float hz, float jitter,
                              uint32_t seed, size_t* count_out) {
  return with(static_cast<Handle*>(w),
              [&](auto& x) { return x.fill_box({lx, ly, lz}, {hx, hy, hz}, jitter, seed, count_out); });
}

int32_t orc_fill_box(void* w, float lx, float ly, float lz, float hx, float hy, float hz, size_t* count_out) {
  return orc_fill_box_jittered(w, lx, ly, lz, hx, hy, hz, 0.0f, 0, count_out);
}

int32_t orc_add_box_container(void* w, float lx, float ly, float lz, float hx, float hy, float hz, uint32_t* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.add_box_container({lx, ly, lz}, {hx, hy, hz}, out); });
}

int32_t orc_add_wall(void* w, float lx, float ly, float lz, float hx, float hy, float hz, uint32_t* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.add_wall({lx, ly, lz}, {hx, hy, hz}, out); });
}

// FluidWorld::set_domain (src/fluid/mod.rs:433-444); a null corner clears it
int32_t orc_set_domain(void* w, const float* low, const float* high) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    if (!low || !high) return x.set_domain(nullptr, nullptr);
    Vec3<float> lo{low[0], low[1], low[2]}, hi{high[0], high[1], high[2]};
    return x.set_domain(&lo, &hi);
  });
}

// add_emitter / add_drain / remove_flow (src/fluid/flow.rs:92-118)
int32_t orc_add_emitter(void* w, const float* low, const float* high, const float* velocity, uint32_t* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    return x.add_flow({low[0], low[1], low[2]}, {high[0], high[1], high[2]}, false,
                      {velocity[0], velocity[1], velocity[2]}, out);
  });
}

int32_t orc_add_drain(void* w, const float* low, const float* high, uint32_t* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    return x.add_flow({low[0], low[1], low[2]}, {high[0], high[1], high[2]}, true, {0.0f, 0.0f, 0.0f}, out);
  });
}

// save_snapshot (src/fluid/save.rs:17-66): out null sets *len only; a short
// buffer returns kOutOfRange
int32_t orc_save_snapshot(void* w, uint8_t* out, size_t cap, size_t* len) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    std::vector<uint8_t> bytes = save_snapshot(x);
    if (len) *len = bytes.size();
    if (!out) return (int32_t)kOk;
    if (cap < bytes.size()) return (int32_t)kOutOfRange;
    std::memcpy(out, bytes.data(), bytes.size());
    return (int32_t)kOk;
  });
}

// load_snapshot (src/fluid/save.rs:68-195): *kind 0 and a world, or the
// refusal (1..6 as SnapshotError)
int32_t orc_load_snapshot(const uint8_t* data, size_t len, const orc_settings* base, int32_t precision, void** out,
                          int32_t* kind, uint32_t* version, const char** what) {
  *out = nullptr;
  Settings b = base ? convert(base) : Settings{};
  SnapshotState st;
  static const uint8_t none = 0;
  SnapshotResult r = parse_snapshot(data ? data : &none, len, b, st);
  *kind = r.kind;
  *version = r.version;
  *what = r.what;
  if (r.kind != kSnapOk) return r.kind;
  auto* h = new Handle();
  h->precision = precision == 64 ? 64 : 32;
  if (h->precision == 64) {
    h->d.reset(new World<double>(st.settings));
    apply_snapshot(*h->d, st);
  } else {
    h->f.reset(new World<float>(st.settings));
    apply_snapshot(*h->f, st);
  }
  *out = h;
  return kOk;
}

// ---- coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) ----
// orc_shape, Pose and BodyPush have the layouts of sfl_shape, sfl_body_pose
// and sfl_body_push.
struct orc_shape {
  int32_t kind;
  float a, b, c;
};

int32_t orc_couple(void* w, uint32_t index, uint32_t generation, const orc_shape* shape) {
  Shape sh{shape->kind, shape->a, shape->b, shape->c};
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.couple(index, generation, sh); });
}

int32_t orc_uncouple(void* w, uint32_t index, uint32_t generation, int32_t* was) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    bool r = x.uncouple(index, generation);
    if (was) *was = r ? 1 : 0;
    return (int32_t)kOk;
  });
}

size_t orc_coupled_len(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.coupled.size(); });
}

// body k: handle, point and probe counts, probe spacing
int32_t orc_coupled_info(void* w, size_t k, uint32_t* index, uint32_t* generation, size_t* points, size_t* probes,
                         float* probe_spacing) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    if (k >= x.coupled.size()) return (int32_t)kOutOfRange;
    const auto& c = x.coupled[k];
    *index = c.index;
    *generation = c.generation;
    *points = c.points.size();
    *probes = c.probes.size();
    *probe_spacing = c.probe_spacing;
    return (int32_t)kOk;
  });
}

// body k's points (which 0) or probes (which 1) in body axes
int32_t orc_read_coupled_points(void* w, size_t k, int32_t which, size_t cap, float* xyz) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    if (k >= x.coupled.size()) return (int32_t)kOutOfRange;
    const auto& v = which ? x.coupled[k].probes : x.coupled[k].points;
    for (size_t i = 0; i < std::min(cap, v.size()); ++i) {
      xyz[3 * i] = v[i].x;
      xyz[3 * i + 1] = v[i].y;
      xyz[3 * i + 2] = v[i].z;
    }
    return (int32_t)kOk;
  });
}

// place_movers_at alone, for tests of the placement and the sort
int32_t orc_place_movers(void* w, const Pose* poses, size_t n) {
  std::vector<Pose> p(poses, poses + n);
  return with(static_cast<Handle*>(w), [&](auto& x) {
    x.place_movers(p);
    return (int32_t)kOk;
  });
}

size_t orc_mover_len(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.movers.positions.size(); });
}

int32_t orc_read_movers(void* w, size_t cap, float* xyz, float* vxyz, uint32_t* owners) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    size_t n = std::min(cap, x.movers.positions.size());
    for (size_t i = 0; i < n; ++i) {
      const auto& p = x.movers.positions[i];
      const auto& v = x.movers.velocities[i];
      float a[6] = {(float)p.x, (float)p.y, (float)p.z, (float)v.x, (float)v.y, (float)v.z};
      std::memcpy(xyz + 3 * i, a, 12);
      std::memcpy(vxyz + 3 * i, a + 3, 12);
      owners[i] = x.movers.owners[i];
    }
    return (int32_t)kOk;
  });
}

// lifts (src/fluid/lift.rs:25-152) over the present state, nothing given
// back: one entry per coupled body, lifted 1 where the water holds a lift
int32_t orc_lifts(void* w, float dt, const Pose* poses, size_t n, BodyPush* out, size_t cap, size_t* count) {
  std::vector<Pose> p(poses, poses + n);
  return with(static_cast<Handle*>(w), [&](auto& x) {
    *count = x.coupled.size();
    if (cap < x.coupled.size()) return (int32_t)kOutOfRange;
    for (size_t b = 0; b < x.coupled.size(); ++b) {
      out[b] = BodyPush{};
      out[b].index = x.coupled[b].index;
      out[b].generation = x.coupled[b].generation;
    }
    for (const auto& lift : x.lifts(dt, p)) {
      BodyPush& bp = out[lift.body];
      bp.lifted = 1;
      float v[6] = {(float)lift.impulse.x, (float)lift.impulse.y, (float)lift.impulse.z,
                    (float)lift.angular.x, (float)lift.angular.y, (float)lift.angular.z};
      std::memcpy(bp.lift_impulse, v, 12);
      std::memcpy(bp.lift_angular, v + 3, 12);
    }
    return (int32_t)kOk;
  });
}

int32_t orc_step_with_bodies(void* w, float dt, const Pose* poses, size_t n, orc_report* report, BodyPush* out,
                             size_t cap, size_t* count) {
  std::vector<Pose> p(poses, poses + n);
  return with(static_cast<Handle*>(w), [&](auto& x) {
    Report r;
    std::vector<BodyPush> pushes;
    int32_t st = x.step_with_bodies(dt, p, r, pushes);
    if (st != kOk) return st;
    fill_report(r, report);
    *count = pushes.size();
    for (size_t b = 0; b < std::min(cap, pushes.size()); ++b) out[b] = pushes[b];
    return (int32_t)kOk;
  });
}

// ---- terrain ground (src/fluid/terrain.rs) ----

int32_t orc_terrain_nodes(void* w, const float* low, const float* high, size_t* nodes) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    x.terrain_nodes({low[0], low[1], low[2]}, {high[0], high[1], high[2]}, nodes);
    return (int32_t)kOk;
  });
}

int32_t orc_add_terrain(void* w, const float* low, const float* high, const float* distances, size_t count,
                        uint32_t* id) {
  std::vector<float> d(distances, distances + count);
  return with(static_cast<Handle*>(w), [&](auto& x) {
    return x.add_terrain({low[0], low[1], low[2]}, {high[0], high[1], high[2]}, d, id);
  });
}

size_t orc_terrain_len(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.terrains.size(); });
}

// field k's distance and way out at `point` (TerrainField::sample); kOutOfRange outside
int32_t orc_terrain_sample(void* w, size_t k, const double* point, double* distance, double* normal) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = typename std::decay_t<decltype(x)>::Real;
    if (k >= x.terrains.size()) return (int32_t)kOutOfRange;
    R d;
    Vec3<R> n;
    if (!x.terrains[k].sample(Vec3<R>{R(point[0]), R(point[1]), R(point[2])}, d, n)) return (int32_t)kOutOfRange;
    *distance = d;
    normal[0] = n.x, normal[1] = n.y, normal[2] = n.z;
    return (int32_t)kOk;
  });
}

// ground_near (terrain.rs:307-313) summed over the fields as compute_densities sums it
int32_t orc_ground_near(void* w, const double* point, double* density, double* gradient) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = typename std::decay_t<decltype(x)>::Real;
    R total = R(0);
    Vec3<R> slope{R(0), R(0), R(0)};
    for (const auto& f : x.terrains) {
      R d;
      Vec3<R> g;
      if (x.ground_at(f, Vec3<R>{R(point[0]), R(point[1]), R(point[2])}, d, g)) {
        total = total + d;
        slope = slope + g;
      }
    }
    *density = total;
    gradient[0] = slope.x, gradient[1] = slope.y, gradient[2] = slope.z;
    return (int32_t)kOk;
  });
}

// run_flows alone (the reference's tests call it without a step)
int32_t orc_run_flows(void* w, uint32_t* emitted, uint32_t* drained) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.run_flows(*emitted, *drained); });
}

int32_t orc_remove_flow(void* w, uint32_t id) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return (int32_t)x.remove_flow(id); });
}

int32_t orc_add_boundary_points(void* w, size_t n, const float* xyz, uint32_t* out) {
  std::vector<Vec3<float>> pts(n);
  for (size_t i = 0; i < n; ++i) pts[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.add_boundary_points(pts, out); });
}

int32_t orc_remove_boundary(void* w, uint32_t id, int32_t* removed) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    bool r = x.remove_boundary(id);
    if (removed) *removed = r ? 1 : 0;
    return (int32_t)kOk;
  });
}

size_t orc_wall_len(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.walls.positions.size(); });
}

int32_t orc_read_wall_positions(void* w, size_t cap, float* xyz) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    size_t n = std::min(cap, x.walls.positions.size());
    for (size_t i = 0; i < n; ++i) {
      xyz[3 * i] = (float)x.walls.positions[i].x;
      xyz[3 * i + 1] = (float)x.walls.positions[i].y;
      xyz[3 * i + 2] = (float)x.walls.positions[i].z;
    }
    return (int32_t)kOk;
  });
}

int32_t orc_step(void* w, float dt, orc_report* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    Report r;
    int32_t st = x.step(dt, r);
    if (st == kOk) fill_report(r, out);
    return st;
  });
}

int32_t orc_update_neighbors(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.update_neighbors(); });
}

int32_t orc_compute_densities(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.compute_densities(); });
}

int32_t orc_read_ids(void* w, size_t cap, uint32_t* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    size_t n = std::min(cap, x.len());
    for (size_t i = 0; i < n; ++i) out[i] = x.ids[i];
    return (int32_t)kOk;
  });
}

// Vector fields as float (precision 32) or double (precision 64) into a
// caller buffer of the matching type: field 0 positions, 1 velocities,
// 2 forces.
int32_t orc_read_vectors(void* w, int32_t field, size_t cap, void* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    const auto& src = field == 0 ? x.positions : field == 1 ? x.velocities : x.forces;
    R* o = static_cast<R*>(out);
    size_t n = std::min(cap, src.size());
    for (size_t i = 0; i < n; ++i) {
      o[3 * i] = src[i].x;
      o[3 * i + 1] = src[i].y;
      o[3 * i + 2] = src[i].z;
    }
    return (int32_t)kOk;
  });
}

int32_t orc_read_densities(void* w, size_t cap, void* out) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    R* o = static_cast<R*>(out);
    size_t n = std::min(cap, x.densities.size());
    for (size_t i = 0; i < n; ++i) o[i] = x.densities[i];
    return (int32_t)kOk;
  });
}

// Replaces the particles (ids, positions, velocities) wholesale; densities
// reset to zero and lists are dropped. Used to feed both sides one state.
int32_t orc_set_particles(void* w, size_t n, const uint32_t* ids, const float* xyz, const float* vxyz) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    x.ids.assign(ids, ids + n);
    x.positions.resize(n);
    x.velocities.resize(n);
    x.densities.assign(n, R(0));
    uint32_t next = 0;
    for (size_t i = 0; i < n; ++i) {
      x.positions[i] = {R(xyz[3 * i]), R(xyz[3 * i + 1]), R(xyz[3 * i + 2])};
      x.velocities[i] = {R(vxyz[3 * i]), R(vxyz[3 * i + 1]), R(vxyz[3 * i + 2])};
      next = std::max(next, ids[i] + 1);
    }
    x.next_id = std::max(x.next_id, next);
    x.neighbors_built = false;
    return (int32_t)kOk;
  });
}

int32_t orc_read_neighbors(void* w, int32_t kind, size_t start_cap, uint32_t* start, size_t items_cap,
                           uint32_t* items, size_t* items_len) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    const Lists& l = kind == 0 ? x.neighbors : (kind == 1 ? x.wall_neighbors : x.mover_neighbors);
    if (items_len) *items_len = l.items.size();
    if (!x.neighbors_built) return (int32_t)kOutOfRange;
    if (start_cap < l.start.size() || items_cap < l.items.size()) return (int32_t)kOutOfRange;
    std::copy(l.start.begin(), l.start.end(), start);
    std::copy(l.items.begin(), l.items.end(), items);
    return (int32_t)kOk;
  });
}

// The DFSPH quantities of a substep's first solve iteration at the current
// state (World::dfsph_probe), lists already built, as doubles in storage
// order: factors, density rates, divergence and density stiffness and
// error (n each), and the velocity changes (3n) push_apart makes for the
// given stiffness (skipped when `stiffness` is null).
int32_t orc_dfsph_probe(void* w, float dt, const double* stiffness, double* alpha, double* rate, double* div_k,
                        double* div_e, double* den_k, double* den_e, double* change) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    using WorldT = std::decay_t<decltype(x)>;
    typename WorldT::Probe pr;
    std::vector<R> k;
    if (stiffness) k.assign(stiffness, stiffness + x.len());
    int32_t st = x.dfsph_probe(R(dt), stiffness ? &k : nullptr, pr);
    if (st != kOk) return st;
    for (size_t i = 0; i < x.len(); ++i) {
      alpha[i] = pr.alpha[i];
      rate[i] = pr.rate[i];
      div_k[i] = pr.div_stiffness[i];
      div_e[i] = pr.div_error[i];
      den_k[i] = pr.den_stiffness[i];
      den_e[i] = pr.den_error[i];
      if (stiffness) {
        change[3 * i] = pr.change[i].x;
        change[3 * i + 1] = pr.change[i].y;
        change[3 * i + 2] = pr.change[i].z;
      }
    }
    return (int32_t)kOk;
  });
}

uint64_t orc_state_hash(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.state_hash(); });
}

const char* orc_error_detail(void* w) {
  return with(static_cast<Handle*>(w), [&](auto& x) { return x.detail.c_str(); });
}

void orc_plant_mistake(int32_t which) { planted_mistake() = which; }

// FNV-1a 64 (src/snapshot/bytes.rs:6-13) over raw bytes, for known answers.
uint64_t orc_fnv1a(const uint8_t* data, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= data[i];
    h *= 0x00000100000001b3ull;
  }
  return h;
}

// The reference's grid test (src/fluid/mod.rs:803-828) on this Grid: the
// dense table and the binary search find the same ranges, a far-spread grid
// keeps no table. Returns 1 when every check holds.
int32_t orc_grid_selftest() {
  std::vector<CellKey> keys = {{-3, 0, 2}, {-3, 0, 2}, {-1, 4, 0}, {0, 0, 0}, {5, -2, 1}};
  std::sort(keys.begin(), keys.end());
  Grid dense = Grid::build(keys.begin(), keys.end());
  if (dense.dense.empty()) return 0;
  Grid search = dense;
  search.dense.clear();
  for (int x = -5; x < 7; ++x)
    for (int y = -4; y < 6; ++y)
      for (int z = -2; z < 4; ++z) {
        auto a = dense.range({x, y, z}), b = search.range({x, y, z});
        if (a.has_value() != b.has_value()) return 0;
        if (a && !(*a == *b)) return 0;
      }
  std::vector<CellKey> far_keys = {{-100000, 0, 0}, {100000, 0, 0}};
  Grid far = Grid::build(far_keys.begin(), far_keys.end());
  if (!far.dense.empty()) return 0;
  auto r = far.range({100000, 0, 0});
  if (!r || r->begin != 1 || r->end != 2) return 0;
  if (far.range({INT32_MAX, INT32_MIN, 0})) return 0;
  return 1;
}

// The fixed-shape FP64 residual sum (ResidualMode::kFixedTreeF64) over
// caller values and cells (x, y, z per value, in sorted order), so tests can
// check it against exact arithmetic.
double orc_tile_tree_sum(const float* e, const int32_t* xyz, size_t n) {
  std::vector<float> v(e, e + n);
  std::vector<CellKey> cells(n);
  for (size_t i = 0; i < n; ++i) cells[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  return tile_tree_sum(v, cells);
}

// Particles from double inputs (exact in the FP64 world), for finite
// differences that perturb below f32 resolution.
int32_t orc_add_particles_f64(void* w, size_t n, const double* xyz, const double* vxyz) {
  return with(static_cast<Handle*>(w), [&](auto& x) {
    using R = std::decay_t<decltype(x.mass)>;
    for (size_t i = 0; i < n; ++i) {
      Vec3<R> p{R(xyz[3 * i]), R(xyz[3 * i + 1]), R(xyz[3 * i + 2])};
      Vec3<R> v{R(vxyz[3 * i]), R(vxyz[3 * i + 1]), R(vxyz[3 * i + 2])};
      int32_t st = x.add_particle(p, v, nullptr);
      if (st != kOk) return st;
    }
    return (int32_t)kOk;
  });
}

// Records `steps` WCSPH steps of dt for the adjoint (adjoint_oracle.hpp).
// `reverse_precision` 64 runs the reverse pass of an f32 world in FP64.
int32_t orc_adjoint_record(void* w, float dt, uint32_t steps, int32_t reverse_precision) {
  auto* h = static_cast<Handle*>(w);
  h->tf.reset();
  h->tfd.reset();
  h->td.reset();
  if (h->precision == 64) {
    h->td.reset(new AdjointTape<double>(*h->d));
    return h->td->record(dt, steps);
  }
  if (reverse_precision == 64) {
    h->tfd.reset(new AdjointTape<float, double>(*h->f));
    return h->tfd->record(dt, steps);
  }
  h->tf.reset(new AdjointTape<float>(*h->f));
  return h->tf->record(dt, steps);
}

// Reverse pass of the recorded steps: per-id adjoints of the final x, v in,
// of the initial x, v out, and dL/d(viscosity).
int32_t orc_adjoint_reverse(void* w, const double* xbar, const double* vbar, double* xbar0, double* vbar0,
                            double* alpha_bar) {
  auto* h = static_cast<Handle*>(w);
  auto run = [&](auto& tape) {
    using R = std::decay_t<decltype(tape.substeps[0].dt)>;
    size_t n = tape.world.len();
    std::vector<Vec3<R>> xb(n), vb(n), x0, v0;
    for (size_t i = 0; i < n; ++i) {
      xb[i] = {R(xbar[3 * i]), R(xbar[3 * i + 1]), R(xbar[3 * i + 2])};
      vb[i] = {R(vbar[3 * i]), R(vbar[3 * i + 1]), R(vbar[3 * i + 2])};
    }
    R a = tape.reverse(xb, vb, x0, v0);
    for (size_t i = 0; i < n; ++i) {
      xbar0[3 * i] = x0[i].x;
      xbar0[3 * i + 1] = x0[i].y;
      xbar0[3 * i + 2] = x0[i].z;
      vbar0[3 * i] = v0[i].x;
      vbar0[3 * i + 1] = v0[i].y;
      vbar0[3 * i + 2] = v0[i].z;
    }
    if (alpha_bar) *alpha_bar = (double)a;
    return (int32_t)kOk;
  };
  if (h->td) return run(*h->td);
  if (h->tfd) return run(*h->tfd);
  if (h->tf) return run(*h->tf);
  return (int32_t)kOutOfRange;
}

}  // extern "C"
