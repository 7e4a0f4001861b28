// The C ABI of include/solunar_fluid.h over sfl::World.
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/solunar_fluid.h"
#include "adjoint.hpp"
#include "batch.hpp"
#include "world.hpp"

struct sfl_world {
  sfl::World* w;
};

struct sfl_batch {
  sfl::Batch* b;
};

struct sfl_tape {
  sfl::Tape* t;
};

namespace {

thread_local std::string g_create_detail;

template <class F>
int32_t guarded(sfl_world* h, F&& f) {
  if (!h || !h->w) return SFL_INVALID_HANDLE;
  try {
    return f(*h->w);
  } catch (const std::bad_alloc&) {
    h->w->set_detail("host allocation failed");
    return SFL_DEVICE_ERROR;
  }
}

template <class F>
int32_t guarded_batch(sfl_batch* h, F&& f) {
  if (!h || !h->b) return SFL_INVALID_HANDLE;
  try {
    return f(*h->b);
  } catch (const std::bad_alloc&) {
    h->b->set_detail("host allocation failed");
    return SFL_DEVICE_ERROR;
  }
}

}  // namespace

extern "C" {

// FluidSettings::default (src/fluid/mod.rs:80-95) and the DFSPH constants of
// src/fluid/dfsph.rs:8-16.
void sfl_settings_default(sfl_settings* out) {
  if (!out) return;
  std::memset(out, 0, sizeof(*out));
  out->solver = SFL_WCSPH;
  out->spacing = 0.05f;
  out->rest_density = 1000.0f;
  out->support = 2.0f;
  out->max_particles = uint64_t(1) << 22;
  out->sound_speed = 20.0f;
  out->viscosity = 0.15f;
  out->gravity = {0.0f, -9.81f, 0.0f};
  out->courant = 0.4f;
  out->max_substeps = 2000;
  out->divergence_tolerance = 1e-3f;
  out->density_tolerance = 5e-4f;
  out->min_iterations = 2;
  out->max_iterations = 100;
  out->device = 0;
}

int32_t sfl_create(const sfl_settings* settings, sfl_world** out) {
  if (!settings || !out) return SFL_OUT_OF_RANGE;
  *out = nullptr;
  try {
    sfl::World* w = nullptr;
    std::string why;
    int32_t st = sfl::World::create(*settings, &w, why);
    g_create_detail = why;
    if (st != SFL_OK) return st;
    *out = new sfl_world{w};
    return SFL_OK;
  } catch (const std::bad_alloc&) {
    g_create_detail = "host allocation failed";
    return SFL_DEVICE_ERROR;
  }
}

int32_t sfl_save_snapshot(sfl_world* world, uint8_t* out, size_t cap, size_t* len_out) {
  return guarded(world, [&](sfl::World& w) {
    const size_t size = w.snapshot_size();
    if (len_out) *len_out = size;
    if (!out) return (int32_t)SFL_OK;
    if (cap < size) return (int32_t)SFL_OUT_OF_RANGE;
    return w.save_snapshot(out);
  });
}

int32_t sfl_load_snapshot(const uint8_t* data, size_t len, const sfl_settings* base, sfl_world** world_out,
                          sfl_snapshot_error* error_out) {
  if (!world_out || (!data && len)) return SFL_OUT_OF_RANGE;
  *world_out = nullptr;
  if (error_out) *error_out = sfl_snapshot_error{0, 0, nullptr};
  try {
    sfl_settings b;
    if (base)
      b = *base;
    else
      sfl_settings_default(&b);
    sfl::World* w = nullptr;
    sfl::SnapshotFailure fail{0, 0, nullptr};
    std::string why;
    static const uint8_t none = 0;
    int32_t st = sfl::World::load_snapshot(data ? data : &none, len, b, &w, fail, why);
    g_create_detail = why;
    if (error_out) *error_out = sfl_snapshot_error{fail.kind, fail.version, fail.what};
    if (st != SFL_OK) return st;
    *world_out = new sfl_world{w};
    return SFL_OK;
  } catch (const std::bad_alloc&) {
    g_create_detail = "host allocation failed";
    return SFL_DEVICE_ERROR;
  }
}

void sfl_destroy(sfl_world* world) {
  if (!world) return;
  delete world->w;
  delete world;
}

int32_t sfl_get_settings(const sfl_world* world, sfl_settings* out) {
  if (!world || !world->w || !out) return SFL_INVALID_HANDLE;
  *out = world->w->settings();
  return SFL_OK;
}

float sfl_particle_mass(const sfl_world* world) { return world && world->w ? world->w->mass() : 0.0f; }

size_t sfl_len(const sfl_world* world) { return world && world->w ? world->w->len() : 0; }

int32_t sfl_add_particles(sfl_world* world, size_t count, const float* xyz, const float* vxyz, uint32_t* ids_out) {
  return guarded(world, [&](sfl::World& w) {
    if (count && !xyz) return (int32_t)SFL_OUT_OF_RANGE;
    return w.add_particles(count, xyz, vxyz, ids_out);
  });
}

int32_t sfl_fill_box(sfl_world* world, sfl_vec3 low, sfl_vec3 high, size_t* count_out) {
  return guarded(world, [&](sfl::World& w) { return w.fill_box(low, high, 0.0f, 0, count_out); });
}

int32_t sfl_fill_box_jittered(sfl_world* world, sfl_vec3 low, sfl_vec3 high, float jitter, uint32_t seed,
                              size_t* count_out) {
  return guarded(world, [&](sfl::World& w) { return w.fill_box(low, high, jitter, seed, count_out); });
}

int32_t sfl_set_domain(sfl_world* world, const sfl_vec3* low, const sfl_vec3* high) {
  return guarded(world, [&](sfl::World& w) { return w.set_domain(low, high); });
}

int32_t sfl_add_emitter(sfl_world* world, sfl_vec3 low, sfl_vec3 high, sfl_vec3 velocity, uint32_t* flow_out) {
  return guarded(world, [&](sfl::World& w) { return w.add_emitter(low, high, velocity, flow_out); });
}

int32_t sfl_add_drain(sfl_world* world, sfl_vec3 low, sfl_vec3 high, uint32_t* flow_out) {
  return guarded(world, [&](sfl::World& w) { return w.add_drain(low, high, flow_out); });
}

int32_t sfl_remove_flow(sfl_world* world, uint32_t flow, int32_t* removed_out) {
  return guarded(world, [&](sfl::World& w) { return w.remove_flow(flow, removed_out); });
}

int32_t sfl_terrain_nodes(const sfl_world* world, sfl_vec3 low, sfl_vec3 high, uint64_t* nodes_out) {
  if (!world || !world->w) return SFL_INVALID_HANDLE;
  if (!nodes_out) return SFL_OUT_OF_RANGE;
  world->w->terrain_nodes(low, high, nodes_out);
  return SFL_OK;
}

int32_t sfl_add_terrain(sfl_world* world, sfl_vec3 low, sfl_vec3 high, const float* distances, size_t count,
                        uint32_t* boundary_out) {
  return guarded(world, [&](sfl::World& w) { return w.add_terrain(low, high, distances, count, boundary_out); });
}

size_t sfl_terrain_len(const sfl_world* world) { return world && world->w ? world->w->terrain_len() : 0; }

int32_t sfl_couple(sfl_world* world, uint32_t index, uint32_t generation, const sfl_shape* shape) {
  return guarded(world, [&](sfl::World& w) {
    if (!shape) return (int32_t)SFL_INVALID_SHAPE;
    return w.couple(index, generation, *shape);
  });
}

int32_t sfl_uncouple(sfl_world* world, uint32_t index, uint32_t generation, int32_t* was_out) {
  return guarded(world, [&](sfl::World& w) { return w.uncouple(index, generation, was_out); });
}

size_t sfl_coupled_len(const sfl_world* world) { return world && world->w ? world->w->coupled_len() : 0; }

int32_t sfl_step_with_bodies(sfl_world* world, float dt, const sfl_body_pose* poses, size_t pose_count,
                             sfl_report* report_out, sfl_body_push* pushes, size_t push_cap,
                             size_t* push_count_out) {
  return guarded(world, [&](sfl::World& w) {
    if ((pose_count && !poses) || (push_cap && !pushes)) return (int32_t)SFL_OUT_OF_RANGE;
    return w.step_with_bodies(dt, poses, pose_count, report_out, pushes, push_cap, push_count_out);
  });
}

int32_t sfl_add_box_container(sfl_world* world, sfl_vec3 low, sfl_vec3 high, uint32_t* boundary_out) {
  return guarded(world, [&](sfl::World& w) { return w.add_box_container(low, high, boundary_out); });
}

int32_t sfl_add_wall(sfl_world* world, sfl_vec3 low, sfl_vec3 high, uint32_t* boundary_out) {
  return guarded(world, [&](sfl::World& w) { return w.add_wall(low, high, boundary_out); });
}

int32_t sfl_add_boundary_points(sfl_world* world, size_t count, const float* xyz, uint32_t* boundary_out) {
  return guarded(world, [&](sfl::World& w) {
    if (count && !xyz) return (int32_t)SFL_OUT_OF_RANGE;
    return w.add_boundary_points(count, xyz, boundary_out);
  });
}

int32_t sfl_remove_boundary(sfl_world* world, uint32_t boundary, int32_t* removed_out) {
  return guarded(world, [&](sfl::World& w) { return w.remove_boundary(boundary, removed_out); });
}

size_t sfl_wall_len(const sfl_world* world) { return world && world->w ? world->w->wall_len() : 0; }

int32_t sfl_read_wall_positions(const sfl_world* world, size_t cap, float* xyz) {
  if (!world || !world->w) return SFL_INVALID_HANDLE;
  if (cap && !xyz) return SFL_OUT_OF_RANGE;
  return world->w->read_walls(cap, xyz);
}

int32_t sfl_step(sfl_world* world, float dt, sfl_report* report_out) {
  return guarded(world, [&](sfl::World& w) { return w.step(dt, report_out); });
}

int32_t sfl_get_domain(const sfl_world* world, int32_t* has_out, sfl_vec3* low_out, sfl_vec3* high_out) {
  if (!world || !world->w) return SFL_INVALID_HANDLE;
  const bool has = world->w->domain(low_out, high_out);
  if (has_out) *has_out = has ? 1 : 0;
  return SFL_OK;
}

float sfl_kernel_value(const sfl_world* world, float distance) {
  return world && world->w ? world->w->kernel_value_at(distance) : 0.0f;
}

float sfl_kernel_slope(const sfl_world* world, float distance) {
  return world && world->w ? world->w->kernel_slope_at(distance) : 0.0f;
}

int32_t sfl_kernel_gradient(const sfl_world* world, sfl_vec3 offset, sfl_vec3* out) {
  if (!world || !world->w) return SFL_INVALID_HANDLE;
  if (!out) return SFL_OUT_OF_RANGE;
  *out = world->w->kernel_gradient_at(offset);
  return SFL_OK;
}

int32_t sfl_neighbors_of(sfl_world* world, size_t index, size_t cap, uint32_t* out, size_t* len_out) {
  return guarded(world, [&](sfl::World& w) { return w.neighbors_of(index, cap, out, len_out); });
}

int32_t sfl_step_no_alloc(sfl_world* world, float dt, sfl_report* report_out) {
  return guarded(world, [&](sfl::World& w) { return w.step_no_alloc(dt, report_out); });
}

int32_t sfl_update_neighbors(sfl_world* world) {
  return guarded(world, [&](sfl::World& w) { return w.update_neighbors(); });
}

int32_t sfl_compute_densities(sfl_world* world) {
  return guarded(world, [&](sfl::World& w) { return w.compute_densities(); });
}

int32_t sfl_read_ids(const sfl_world* world, size_t cap, uint32_t* out) {
  return guarded(const_cast<sfl_world*>(world), [&](sfl::World& w) {
    if (cap && !out) return (int32_t)SFL_OUT_OF_RANGE;
    return w.read_ids(cap, out);
  });
}

int32_t sfl_read_positions(const sfl_world* world, size_t cap, float* xyz) {
  return guarded(const_cast<sfl_world*>(world), [&](sfl::World& w) {
    if (cap && !xyz) return (int32_t)SFL_OUT_OF_RANGE;
    return w.read_positions(cap, xyz);
  });
}

int32_t sfl_read_velocities(const sfl_world* world, size_t cap, float* xyz) {
  return guarded(const_cast<sfl_world*>(world), [&](sfl::World& w) {
    if (cap && !xyz) return (int32_t)SFL_OUT_OF_RANGE;
    return w.read_velocities(cap, xyz);
  });
}

int32_t sfl_read_densities(const sfl_world* world, size_t cap, float* out) {
  return guarded(const_cast<sfl_world*>(world), [&](sfl::World& w) {
    if (cap && !out) return (int32_t)SFL_OUT_OF_RANGE;
    return w.read_densities(cap, out);
  });
}

int32_t sfl_keep_forces(sfl_world* world, int32_t keep) {
  return guarded(world, [&](sfl::World& w) { return w.keep_forces(keep != 0); });
}

int32_t sfl_read_forces(const sfl_world* world, size_t cap, float* xyz) {
  return guarded(const_cast<sfl_world*>(world), [&](sfl::World& w) {
    if (cap && !xyz) return (int32_t)SFL_OUT_OF_RANGE;
    return w.read_forces(cap, xyz);
  });
}

int32_t sfl_read_neighbors(sfl_world* world, int32_t kind, size_t start_cap, uint32_t* start, size_t items_cap,
                           uint32_t* items, size_t* items_len_out) {
  return guarded(world, [&](sfl::World& w) {
    if (kind != 0 && kind != 1) return (int32_t)SFL_OUT_OF_RANGE;
    return w.read_neighbors(kind, start_cap, start, items_cap, items, items_len_out);
  });
}

int32_t sfl_state_hash(const sfl_world* world, uint64_t* out) {
  return guarded(const_cast<sfl_world*>(world), [&](sfl::World& w) {
    if (!out) return (int32_t)SFL_OUT_OF_RANGE;
    return w.state_hash(out);
  });
}

const char* sfl_error_detail(const sfl_world* world) {
  if (!world || !world->w) return g_create_detail.c_str();
  return world->w->detail();
}

// Display of PhysicsError (src/error.rs:24-45).
const char* sfl_status_text(int32_t status) {
  switch (status) {
    case SFL_OK: return "ok";
    case SFL_INVALID_HANDLE: return "invalid body handle";
    case SFL_INVALID_FRAME: return "invalid frame";
    case SFL_FRAME_TOO_DEEP: return "frame hierarchy too deep";
    case SFL_NON_FINITE: return "non finite value";
    case SFL_INVALID_SHAPE: return "invalid shape";
    case SFL_INVALID_MATERIAL: return "invalid material";
    case SFL_INVALID_SETTINGS: return "invalid settings";
    case SFL_NOT_DYNAMIC: return "body is not dynamic";
    case SFL_NOT_KINEMATIC: return "body is not kinematic";
    case SFL_OUT_OF_RANGE: return "position out of range";
    case SFL_DISPATCH_INCOMPLETE: return "dispatcher did not run every job";
    case SFL_INVALID_JOINT: return "invalid joint";
    case SFL_BODY_HAS_JOINTS: return "body has joints";
    case SFL_SHAPE_IN_USE: return "a body still uses the shape";
    case SFL_TERRAIN_BUSY: return "terrain edits are still waiting";
    case SFL_FLUID_FULL: return "fluid has no room for more particles";
    case SFL_DEVICE_ERROR: return "device error";
    case SFL_UNSUPPORTED: return "unsupported on the device path";
    default: return "unknown status";
  }
}

int32_t sfl_profile(sfl_world* world, int32_t enable) {
  return guarded(world, [&](sfl::World& w) { return w.profile(enable != 0); });
}

int32_t sfl_profile_read(sfl_world* world, double* ms_out, uint64_t* launches_out, int32_t reset) {
  return guarded(world, [&](sfl::World& w) { return w.profile_read(ms_out, launches_out, reset != 0); });
}

int32_t sfl_work_read(sfl_world* world, uint64_t* particle_substeps, uint64_t* particle_pushes,
                      uint64_t* particle_rates, int32_t reset) {
  return guarded(world,
                 [&](sfl::World& w) { return w.work_read(particle_substeps, particle_pushes, particle_rates, reset != 0); });
}

int32_t sfl_read_solve_log(sfl_world* world, size_t cap, double* average, uint32_t* loaded, size_t* len_out) {
  return guarded(world, [&](sfl::World& w) { return w.read_solve_log(cap, average, loaded, len_out); });
}

size_t sfl_memory_bytes(const sfl_world* world) { return world && world->w ? world->w->device_bytes() : 0; }

int32_t sfl_solve_grid(const sfl_world* world) { return world && world->w ? world->w->solve_grid() : 0; }

int32_t sfl_synchronize(sfl_world* world) {
  return guarded(world, [&](sfl::World& w) { return w.synchronize(); });
}

// ---- batched WCSPH ------------------------------------------------------------

int32_t sfl_batch_create(sfl_world* const* worlds, size_t count, sfl_batch** out) {
  if (!out) return SFL_OUT_OF_RANGE;
  *out = nullptr;
  if (!worlds) return SFL_OUT_OF_RANGE;
  try {
    std::vector<sfl::World*> ws(count);
    for (size_t s = 0; s < count; ++s) {
      if (!worlds[s] || !worlds[s]->w) return SFL_INVALID_HANDLE;
      ws[s] = worlds[s]->w;
    }
    sfl::Batch* b = nullptr;
    std::string why;
    int32_t st = sfl::Batch::create(ws.data(), count, &b, why);
    g_create_detail = why;
    if (st != SFL_OK) return st;
    *out = new sfl_batch{b};
    return SFL_OK;
  } catch (const std::bad_alloc&) {
    g_create_detail = "host allocation failed";
    return SFL_DEVICE_ERROR;
  }
}

void sfl_batch_destroy(sfl_batch* batch) {
  if (!batch) return;
  delete batch->b;
  delete batch;
}

size_t sfl_batch_scenes(const sfl_batch* batch) { return batch && batch->b ? batch->b->scenes() : 0; }

size_t sfl_batch_len(const sfl_batch* batch, size_t scene) {
  return batch && batch->b ? batch->b->len(scene) : 0;
}

int32_t sfl_batch_step(sfl_batch* batch, float dt, uint32_t steps, sfl_report* reports) {
  return guarded_batch(batch, [&](sfl::Batch& b) { return b.step(dt, steps, reports); });
}

int32_t sfl_batch_prepare(sfl_batch* batch, uint32_t steps) {
  return guarded_batch(batch, [&](sfl::Batch& b) { return b.prepare(steps); });
}

int32_t sfl_batch_scene_failed(const sfl_batch* batch, size_t scene, int32_t* failed) {
  if (!batch || !batch->b) return SFL_INVALID_HANDLE;
  if (!failed) return SFL_OUT_OF_RANGE;
  return batch->b->scene_failed(scene, failed);
}

int32_t sfl_batch_autotune(sfl_batch* batch, float dt, uint32_t repeats, sfl_batch_tuning* out) {
  return guarded_batch(batch, [&](sfl::Batch& b) { return b.autotune(dt, repeats, out); });
}

int32_t sfl_batch_set_tuning(sfl_batch* batch, const sfl_batch_tuning* tuning) {
  return guarded_batch(batch, [&](sfl::Batch& b) {
    if (!tuning) return (int32_t)SFL_OUT_OF_RANGE;
    return b.set_tuning(*tuning);
  });
}

int32_t sfl_batch_get_tuning(const sfl_batch* batch, sfl_batch_tuning* out) {
  if (!batch || !batch->b) return SFL_INVALID_HANDLE;
  if (!out) return SFL_OUT_OF_RANGE;
  return batch->b->get_tuning(out);
}

int32_t sfl_batch_use_graph(sfl_batch* batch, int32_t on) {
  return guarded_batch(batch, [&](sfl::Batch& b) { return b.use_graph(on != 0); });
}

int32_t sfl_batch_read_state(sfl_batch* batch, size_t scene, size_t cap, uint32_t* ids, float* xyz, float* vxyz,
                             float* densities) {
  return guarded_batch(batch, [&](sfl::Batch& b) { return b.read_state(scene, cap, ids, xyz, vxyz, densities); });
}

int32_t sfl_batch_state_hash(sfl_batch* batch, size_t scene, uint64_t* out) {
  return guarded_batch(batch, [&](sfl::Batch& b) {
    if (!out) return (int32_t)SFL_OUT_OF_RANGE;
    return b.state_hash(scene, out);
  });
}

int32_t sfl_batch_store(sfl_batch* batch, sfl_world* const* worlds, size_t count) {
  return guarded_batch(batch, [&](sfl::Batch& b) {
    if (!worlds) return (int32_t)SFL_OUT_OF_RANGE;
    std::vector<sfl::World*> ws(count);
    for (size_t s = 0; s < count; ++s) {
      if (!worlds[s] || !worlds[s]->w) return (int32_t)SFL_INVALID_HANDLE;
      ws[s] = worlds[s]->w;
    }
    return b.store(ws.data(), count);
  });
}

const char* sfl_batch_error_detail(const sfl_batch* batch) {
  if (!batch || !batch->b) return g_create_detail.c_str();
  return batch->b->detail();
}

size_t sfl_batch_memory_bytes(const sfl_batch* batch) { return batch && batch->b ? batch->b->device_bytes() : 0; }

uint64_t sfl_batch_graph_nodes(const sfl_batch* batch) { return batch && batch->b ? batch->b->graph_nodes() : 0; }

int32_t sfl_batch_synchronize(sfl_batch* batch) {
  return guarded_batch(batch, [&](sfl::Batch& b) { return b.synchronize(); });
}

// ---- WCSPH adjoint ---------------------------------------------------------------

int32_t sfl_wcsph_record(sfl_world* world, float dt, uint32_t steps, uint32_t every, sfl_tape** out) {
  if (!out) return SFL_OUT_OF_RANGE;
  *out = nullptr;
  return guarded(world, [&](sfl::World& w) {
    sfl::Tape* t = nullptr;
    std::string why;
    int32_t st = sfl::Tape::record(&w, dt, steps, every, &t, why);
    if (st != SFL_OK) {
      w.set_detail(why);
      g_create_detail = why;
      return st;
    }
    *out = new sfl_tape{t};
    return (int32_t)SFL_OK;
  });
}

int32_t sfl_wcsph_adjoint(sfl_tape* tape, const float* xbar, const float* vbar, float* xbar0, float* vbar0,
                          float* viscosity_bar) {
  if (!tape || !tape->t) return SFL_INVALID_HANDLE;
  if (!xbar || !vbar || !xbar0 || !vbar0) return SFL_OUT_OF_RANGE;
  try {
    return tape->t->adjoint(xbar, vbar, xbar0, vbar0, viscosity_bar);
  } catch (const std::bad_alloc&) {
    tape->t->set_detail("host allocation failed");
    return SFL_DEVICE_ERROR;
  }
}

void sfl_tape_destroy(sfl_tape* tape) {
  if (!tape) return;
  delete tape->t;
  delete tape;
}

uint64_t sfl_tape_substeps(const sfl_tape* tape) { return tape && tape->t ? tape->t->substeps() : 0; }

size_t sfl_tape_memory_bytes(const sfl_tape* tape) { return tape && tape->t ? tape->t->device_bytes() : 0; }

const char* sfl_tape_error_detail(const sfl_tape* tape) {
  if (!tape || !tape->t) return g_create_detail.c_str();
  return tape->t->detail();
}

}  // extern "C"
