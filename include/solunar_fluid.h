/*
 * solunar_fluid.h: C ABI of the MI355X (gfx950) SPH particle-fluid step.
 *
 * One handle (`sfl_world`) stands for one reference `FluidWorld`
 * (src/fluid/mod.rs:310-339). Every call returns an `int32_t` status whose
 * non-zero values map one to one onto the reference's `PhysicsError`
 * variants (src/error.rs:5-22), plus SFL_DEVICE_ERROR and SFL_UNSUPPORTED.
 *
 * Contracts
 *  - Buffers are caller owned host memory. They are copied in or out during
 *    the call; no pointer is kept.
 *  - A world owns one HIP stream on the device named in its settings. A
 *    handle may move between threads but must not be used by two at once
 *    (the reference's `&mut self`).
 *  - Bad input changes nothing (src/fluid/mod.rs:915-972).
 *  - Vectors are three packed floats (x, y, z); arrays of them are
 *    interleaved xyz, 12 bytes per element.
 */
#ifndef SOLUNAR_FLUID_H
#define SOLUNAR_FLUID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the PhysicsError variants in declaration order
 * (src/error.rs:5-22), offset by one so zero is success. */
enum sfl_status {
  SFL_OK = 0,
  SFL_INVALID_HANDLE = 1,
  SFL_INVALID_FRAME = 2,
  SFL_FRAME_TOO_DEEP = 3,
  SFL_NON_FINITE = 4, /* detail string from sfl_error_detail */
  SFL_INVALID_SHAPE = 5,
  SFL_INVALID_MATERIAL = 6,
  SFL_INVALID_SETTINGS = 7,
  SFL_NOT_DYNAMIC = 8,
  SFL_NOT_KINEMATIC = 9,
  SFL_OUT_OF_RANGE = 10,
  SFL_DISPATCH_INCOMPLETE = 11,
  SFL_INVALID_JOINT = 12,
  SFL_BODY_HAS_JOINTS = 13,
  SFL_SHAPE_IN_USE = 14,
  SFL_TERRAIN_BUSY = 15,
  SFL_FLUID_FULL = 16,
  /* No reference counterpart: a HIP call failed (detail names it). */
  SFL_DEVICE_ERROR = 100,
  /* No reference counterpart: a feature the device path does not carry. */
  SFL_UNSUPPORTED = 101,
  /* A snapshot was refused; sfl_snapshot_error says why (SnapshotError). */
  SFL_SNAPSHOT_REFUSED = 102
};

enum sfl_solver { SFL_WCSPH = 0, SFL_DFSPH = 1 };

typedef struct sfl_vec3 {
  float x, y, z;
} sfl_vec3;

/* Mirrors FluidSettings (src/fluid/mod.rs:55-78) field for field, then the
 * DFSPH constants of src/fluid/dfsph.rs:8-16 made settable, then the device.
 * sfl_settings_default gives the reference defaults (mod.rs:80-95,
 * dfsph.rs:8-16). */
typedef struct sfl_settings {
  int32_t solver; /* enum sfl_solver */
  float spacing;
  float rest_density;
  float support;
  uint64_t max_particles;
  float sound_speed;
  float viscosity;
  sfl_vec3 gravity;
  float courant;
  uint32_t max_substeps;
  /* extensions, defaults = reference constants */
  float divergence_tolerance; /* DIVERGENCE_TOLERANCE, dfsph.rs:11 */
  float density_tolerance;    /* DENSITY_TOLERANCE, dfsph.rs:9 */
  uint32_t min_iterations;    /* MIN_ITERATIONS, dfsph.rs:12 */
  uint32_t max_iterations;    /* MAX_ITERATIONS, dfsph.rs:13 */
  int32_t device;             /* HIP device ordinal */
} sfl_settings;

/* Mirrors FluidReport (src/fluid/step.rs:28-49); the last two fields split
 * `iterations` by solve. Booleans are 0/1. */
typedef struct sfl_report {
  uint32_t substeps;
  uint32_t capped;
  float max_compression;
  float max_speed;
  uint32_t held_back;
  uint32_t iterations;
  uint32_t emitted;
  uint32_t drained;
  uint32_t lost;
  float sound_speed;
  uint32_t iterations_divergence;
  uint32_t iterations_density;
} sfl_report;

/* FluidSettings::default (src/fluid/mod.rs:80-95). */
void sfl_settings_default(sfl_settings* out);

/* FluidWorld::new (src/fluid/mod.rs:344). SFL_INVALID_SETTINGS on settings
 * out of range (mod.rs:98-117); SFL_DEVICE_ERROR if the device is absent. */
typedef struct sfl_world sfl_world;
int32_t sfl_create(const sfl_settings* settings, sfl_world** out);
/* Drop for FluidWorld. Null is ignored. */
void sfl_destroy(sfl_world* world);

/* FluidWorld::settings / kernel radius / particle_mass / len
 * (src/fluid/mod.rs:372-390). */
int32_t sfl_get_settings(const sfl_world* world, sfl_settings* out);
float sfl_particle_mass(const sfl_world* world);
size_t sfl_len(const sfl_world* world);

/* FluidWorld::add_particle (src/fluid/mod.rs:412-426), one call per
 * particle in order; stops at the first failure, earlier ones stay added
 * (as a loop of add_particle calls would). `ids_out` may be null. */
int32_t sfl_add_particles(sfl_world* world, size_t count, const float* xyz,
                          const float* vxyz, uint32_t* ids_out);

/* FluidWorld::fill_box (src/fluid/mod.rs:493-524). */
int32_t sfl_fill_box(sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                     size_t* count_out);

/* fill_box with every lattice site nudged by a seeded xorshift32 offset of up
 * to +-jitter*spacing per axis, drawn x, y, z per particle in fill order.
 * jitter = 0 is fill_box. The generator is the one the reference's tests use
 * (src/fluid/mod.rs:755-761): u = state / 4294967295, offset =
 * (u - 0.5) * (2 * jitter * spacing). A seed of 0 is replaced by 0x9E3779B9. */
int32_t sfl_fill_box_jittered(sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                              float jitter, uint32_t seed, size_t* count_out);

/* FluidWorld::set_domain (src/fluid/mod.rs:433-444): water outside the box
 * (inclusive bounds) is taken out after each step, after the flows, and
 * counted in sfl_report.lost (keep_in_domain, mod.rs:451-465). Null clears
 * it. */
int32_t sfl_set_domain(sfl_world* world, const sfl_vec3* low,
                       const sfl_vec3* high);

/* FluidWorld::domain (src/fluid/mod.rs:446-448): *has_out 1 and the box
 * when a domain is set, else 0 (the box is left alone). */
int32_t sfl_get_domain(const sfl_world* world, int32_t* has_out,
                       sfl_vec3* low_out, sfl_vec3* high_out);

/* FluidWorld::kernel (mod.rs:376-378): the world's Wendland C2 kernel,
 * Kernel::value, slope and gradient (mod.rs:140-164) in the f32 arithmetic
 * of the device passes. A null world gives 0. */
float sfl_kernel_value(const sfl_world* world, float distance);
float sfl_kernel_slope(const sfl_world* world, float distance);
int32_t sfl_kernel_gradient(const sfl_world* world, sfl_vec3 offset,
                            sfl_vec3* out);

/* FluidWorld::add_emitter / add_drain (src/fluid/flow.rs:92-111). After each
 * step an emitter fills every lattice slot of its box (spacing apart, half a
 * spacing in from its low corner) that no water is nearer to than 0.9
 * spacing, the new water moving at `velocity`; a drain takes out the water
 * in its box. Drains run first, then emitters, each in the order added
 * (run_flows, flow.rs:144-223), and the report counts them in emitted and
 * drained. A full world stops an emitter with SFL_FLUID_FULL after the
 * particles that fit. An emitter box of more than 2^28 slots is refused at
 * the step with SFL_OUT_OF_RANGE. */
int32_t sfl_add_emitter(sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                        sfl_vec3 velocity, uint32_t* flow_out);
int32_t sfl_add_drain(sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                      uint32_t* flow_out);

/* FluidWorld::remove_flow (src/fluid/flow.rs:113-118): *removed_out is 0
 * for an id that is not there. */
int32_t sfl_remove_flow(sfl_world* world, uint32_t flow, int32_t* removed_out);

/* ---- snapshots (src/fluid/save.rs) ---- */

/* SnapshotError (src/snapshot/mod.rs:56-63): `version` is set for
 * UNSUPPORTED_VERSION, `what` (a static string) for INVALID. */
enum sfl_snapshot_kind {
  SFL_SNAPSHOT_TOO_LARGE = 1,
  SFL_SNAPSHOT_TRUNCATED = 2,
  SFL_SNAPSHOT_BAD_MAGIC = 3,
  SFL_SNAPSHOT_UNSUPPORTED_VERSION = 4,
  SFL_SNAPSHOT_BAD_CHECKSUM = 5,
  SFL_SNAPSHOT_INVALID = 6
};
typedef struct sfl_snapshot_error {
  int32_t kind; /* enum sfl_snapshot_kind, 0 when not refused */
  uint32_t version;
  const char* what;
} sfl_snapshot_error;

/* FluidWorld::save_snapshot (src/fluid/save.rs:17-66): the SLFL version 3
 * bytes of the world (settings, particles in order, next id, walls, flows,
 * coupled bodies, terrain regions, domain, FNV-1a checksum), the same bytes
 * the reference writes. With `out` null, only *len_out is set; with `cap`
 * too small, *len_out is set and SFL_OUT_OF_RANGE returned. */
int32_t sfl_save_snapshot(sfl_world* world, uint8_t* out, size_t cap,
                          size_t* len_out);

/* FluidWorld::load_snapshot (src/fluid/save.rs:68-195): a new world from
 * untrusted bytes, checked as the reference checks them (size, magic,
 * checksum, version 1..3, settings, counts against the bytes left, ids,
 * walls, flows, coupled bodies, terrain regions, domain, no trailing
 * bytes). A refused snapshot returns SFL_SNAPSHOT_REFUSED with *error_out
 * filled. Settings the snapshot does not hold
 * (the extensions, device included) come from `base` (defaults when null).
 * Densities start at the rest density, as the reference's do. */
int32_t sfl_load_snapshot(const uint8_t* data, size_t len,
                          const sfl_settings* base, sfl_world** world_out,
                          sfl_snapshot_error* error_out);

/* ---- terrain ground (src/fluid/terrain.rs) ---- */

/* The lattice FluidWorld::add_terrain samples over the box low..high
 * (terrain.rs:280-286): nodes_out[a] = ceil(extent_a / step) + 1 nodes per
 * axis, step = spacing / 2, node (x, y, z) at low + (x, y, z) * step in f32.
 * The caller evaluates its signed-distance field (the reference's
 * crate::sdf volume, outside this library) there. */
int32_t sfl_terrain_nodes(const sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                          uint64_t* nodes_out);
/* FluidWorld::add_terrain (terrain.rs:262-305): water rests on and flows
 * over the ground whose signed distance (positive above it) `distances`
 * holds at the nodes of sfl_terrain_nodes, in the reference's order (loops
 * y, z, x: index (y * nz + z) * nx + x); `count` is their number. The
 * ground enters the WCSPH density, force, wall viscosity and advance; a
 * DFSPH step of a world with terrain returns SFL_INVALID_SETTINGS, as the
 * reference's does. SFL_NON_FINITE for a non-finite box or distance,
 * SFL_OUT_OF_RANGE for an inside-out box or a count that is not the
 * lattice's, SFL_FLUID_FULL past 2^24 nodes. *boundary_out is the id
 * sfl_remove_boundary takes it away with. */
int32_t sfl_add_terrain(sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                        const float* distances, size_t count,
                        uint32_t* boundary_out);
size_t sfl_terrain_len(const sfl_world* world);

/* ---- coupled bodies (src/fluid/couple.rs, src/fluid/lift.rs) ---- */

/* The shapes a body couples with (src/shape.rs:8-49): a cuboid (a, b, c its
 * half extents), a sphere (a its radius), a capsule along local y (a its
 * radius, b the half height of its cylinder). */
enum sfl_shape_kind { SFL_SHAPE_CUBOID = 0, SFL_SHAPE_SPHERE = 1, SFL_SHAPE_CAPSULE = 2 };
typedef struct sfl_shape {
  int32_t kind;
  float a, b, c;
} sfl_shape;

/* Where the rigid world has a coupled body for one step (SubstepPose,
 * src/substep.rs): the body's handle (BodyHandle::to_raw, src/body.rs:18-22),
 * its centre in f64 (rounded to f32 as the reference rounds it), rotation as
 * a quaternion (x, y, z, w), velocities, and whether the body is dynamic,
 * i.e. takes the water's pushes (PhysicsWorld::apply_force fails on fixed
 * and kinematic bodies). */
typedef struct sfl_body_pose {
  uint32_t index, generation;
  double position[3];
  float rotation[4];
  float linear_velocity[3];
  float angular_velocity[3];
  int32_t dynamic;
} sfl_body_pose;

/* What the water gave one coupled body over a step, as impulses: the lift of
 * src/fluid/lift.rs:25-152 (lifted = 1 when the body is dynamic and the water
 * held a lift for it, which the water then gave back, lift.rs:154-166) and
 * the damping summed by mover_push (src/fluid/couple.rs:368-408). The
 * reference applies them as forces, impulse / dt, in step_with_bodies
 * (couple.rs:236-251); a substep hook (couple.rs:128-158) pushes them as
 * impulses. */
typedef struct sfl_body_push {
  uint32_t index, generation;
  int32_t lifted;
  float lift_impulse[3];
  float lift_angular[3];
  float impulse[3];
  float angular[3];
} sfl_body_push;

/* FluidWorld::couple (src/fluid/couple.rs:165-198): water and the body push
 * each other in sfl_step_with_bodies. Its wall points are the shape's
 * lattice at the water's spacing as deep as the kernel reaches, its probes
 * the lattice through its whole volume, coarser by 1.25 a time until at most
 * 4096. SFL_INVALID_SHAPE for another kind, a non-finite or non-positive
 * size, or a shape too small for a point; SFL_FLUID_FULL past max_particles
 * body points. Coupling a coupled body again replaces its shape. */
int32_t sfl_couple(sfl_world* world, uint32_t index, uint32_t generation,
                   const sfl_shape* shape);
/* FluidWorld::uncouple (couple.rs:200-205): *was_out = 0 for a body that was
 * not coupled. */
int32_t sfl_uncouple(sfl_world* world, uint32_t index, uint32_t generation,
                     int32_t* was_out);
/* Coupled bodies, in handle order. */
size_t sfl_coupled_len(const sfl_world* world);

/* FluidWorld::step_with_bodies (src/fluid/couple.rs:207-252) with the rigid
 * world's part handed in: the coupled bodies stand where `poses` has them
 * (a coupled body no pose names is uncoupled, as a body gone from the world
 * is), their points move with them through the substeps, the step's sound
 * speed follows the fastest body near water (step.rs:102-116), and the water
 * feels them as walls. pushes[k] gets coupled body k's impulses, in handle
 * order; *push_count_out the count (SFL_OUT_OF_RANGE, nothing stepped, when
 * `push_cap` is smaller than the coupled bodies). SFL_INVALID_SETTINGS for a
 * DFSPH world with coupled bodies, SFL_NON_FINITE for a non-finite pose,
 * SFL_INVALID_HANDLE for a body named twice; otherwise as sfl_step. */
int32_t sfl_step_with_bodies(sfl_world* world, float dt,
                             const sfl_body_pose* poses, size_t pose_count,
                             sfl_report* report_out, sfl_body_push* pushes,
                             size_t push_cap, size_t* push_count_out);

/* add_box_container / add_wall (src/fluid/boundary.rs:41-91). */
int32_t sfl_add_box_container(sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                              uint32_t* boundary_out);
int32_t sfl_add_wall(sfl_world* world, sfl_vec3 low, sfl_vec3 high,
                     uint32_t* boundary_out);
/* add_wall_points (src/fluid/boundary.rs:139-151) for arbitrary points, such
 * as spheres sampled as boundary particles. Points must be finite. */
int32_t sfl_add_boundary_points(sfl_world* world, size_t count,
                                const float* xyz, uint32_t* boundary_out);
/* remove_boundary (src/fluid/boundary.rs:94-115); *removed_out = 0/1. */
int32_t sfl_remove_boundary(sfl_world* world, uint32_t boundary,
                            int32_t* removed_out);
size_t sfl_wall_len(const sfl_world* world);
/* wall_positions (boundary.rs:118-120), in their settled order. */
int32_t sfl_read_wall_positions(const sfl_world* world, size_t cap,
                                float* xyz);

/* FluidWorld::step (src/fluid/step.rs:58-90). The dispatcher argument of the
 * reference has no counterpart: the device path has one schedule.
 * Like the reference's Grid (src/fluid/mod.rs:214-236), the cell table is
 * dense while the cells' box is small enough (at most max(64 n + 4096, 2^26)
 * cells) and searches the occupied cells past that; either gives the same
 * neighbours in the same order. A DFSPH world keeps a dense table and
 * returns SFL_OUT_OF_RANGE, its state untouched, when its water spreads over
 * more than 2^31 cells. */
int32_t sfl_step(sfl_world* world, float dt, sfl_report* report_out);

/* No counterpart in the reference: sfl_step for callers that cannot
 * allocate device memory (HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same results and report as sfl_step, and it calls
 * no hipMalloc or hipFree: the step's control block and cell table live in a
 * 32 MiB scratch array that each GPU holds from the first call until the
 * process ends and that every call on that GPU shares (calls overlapping on
 * one GPU race; each returns after its last kernel). WCSPH worlds without
 * coupled bodies, terrain, flows or a domain (else SFL_UNSUPPORTED); a sort
 * needing more cells than the scratch holds (about 4.19M, less 1 per 2,048
 * particles) returns SFL_OUT_OF_RANGE with the last finished substep's state.
 * The last grid is kept when the world's own cell table holds it (it does
 * after any sfl_step or sfl_update_neighbors at least as large); otherwise
 * the neighbour lists go stale (sfl_compute_densities and sfl_read_neighbors
 * return SFL_OUT_OF_RANGE until sfl_update_neighbors), and a coupled step
 * whose fast body needs that grid returns SFL_UNSUPPORTED. */
int32_t sfl_step_no_alloc(sfl_world* world, float dt, sfl_report* report_out);

/* update_neighbors / compute_densities (src/fluid/mod.rs:540-675): the
 * neighbour build and the density pass alone, for parity checks. */
int32_t sfl_update_neighbors(sfl_world* world);
int32_t sfl_compute_densities(sfl_world* world);

/* Readback of ids / positions / velocities / densities
 * (src/fluid/mod.rs:393-408), `cap` elements at most. */
int32_t sfl_read_ids(const sfl_world* world, size_t cap, uint32_t* out);
int32_t sfl_read_positions(const sfl_world* world, size_t cap, float* xyz);
int32_t sfl_read_velocities(const sfl_world* world, size_t cap, float* xyz);
int32_t sfl_read_densities(const sfl_world* world, size_t cap, float* out);
/* Per-particle acceleration of the last WCSPH substep (pressure, viscosity,
 * walls and gravity: the vector step.rs:207 returns) or of the last DFSPH
 * viscosity pass (dfsph.rs:94). Needs sfl_keep_forces(world, 1) before the
 * step. */
int32_t sfl_keep_forces(sfl_world* world, int32_t keep);
int32_t sfl_read_forces(const sfl_world* world, size_t cap, float* xyz);

/* Neighbour lists of the last neighbour build (src/fluid/mod.rs:279-308,
 * 610-614): CSR `start` (len+1 entries) and `items`. kind 0 = fluid list,
 * 1 = wall list. *items_len_out receives the item count even if `items` is
 * too small (then nothing is copied and SFL_OUT_OF_RANGE is returned). */
int32_t sfl_read_neighbors(sfl_world* world, int32_t kind, size_t start_cap,
                           uint32_t* start, size_t items_cap, uint32_t* items,
                           size_t* items_len_out);

/* FluidWorld::neighbors_of (src/fluid/mod.rs:612-614): the fluid list of
 * particle `index` (storage order) from the last build. *len_out receives
 * its length even if `cap` is too small (then SFL_OUT_OF_RANGE). */
int32_t sfl_neighbors_of(sfl_world* world, size_t index, size_t cap,
                         uint32_t* out, size_t* len_out);

/* FluidWorld::state_hash (src/fluid/mod.rs:678-691): FNV-1a 64
 * (src/snapshot/bytes.rs:6-13) over (id, x, v) in id order. */
int32_t sfl_state_hash(const sfl_world* world, uint64_t* out);

/* Detail of the last non-zero status on this world (NonFinite's `what`, or
 * the HIP call that failed). Never null. Valid until the next call. */
const char* sfl_error_detail(const sfl_world* world);
/* Display text of a status code (src/error.rs:24-45). */
const char* sfl_status_text(int32_t status);

/* ---- measurement (no reference counterpart) ---------------------------- */

/* Per-kernel timing with HIP events recorded on the world's stream around
 * every launch. Off by default; enabling adds one event pair per launch. */
enum sfl_phase {
  SFL_PHASE_BOUNDS = 0,   /* cell bounds of the particles */
  SFL_PHASE_SORT = 1,     /* cell count, scan, scatter, in-cell id sort */
  SFL_PHASE_REORDER = 2,  /* gather of x, v, id into sorted order */
  SFL_PHASE_DENSITY = 3,  /* density (+ DFSPH factors) */
  SFL_PHASE_FORCE = 4,    /* WCSPH force + integrate */
  SFL_PHASE_SOLVE = 5,    /* DFSPH persistent Jacobi solves */
  SFL_PHASE_VISCOUS = 6,  /* DFSPH viscosity + gravity */
  SFL_PHASE_ADVANCE = 7,  /* DFSPH advance */
  SFL_PHASE_FLOWS = 8,    /* drains, emitters and the domain after a step */
  SFL_PHASE_BODIES = 9,   /* coupled bodies: carry, damping sums, lifts */
  SFL_PHASE_COUNT = 10
};
int32_t sfl_profile(sfl_world* world, int32_t enable);
/* ms_out[SFL_PHASE_COUNT] summed milliseconds, launches_out[SFL_PHASE_COUNT]
 * launch counts since the last reset; either may be null. */
int32_t sfl_profile_read(sfl_world* world, double* ms_out,
                         uint64_t* launches_out, int32_t reset);
/* Fluid particle-substeps and Jacobi pushes since the last reset, for the
 * algorithmic byte count. */
int32_t sfl_work_read(sfl_world* world, uint64_t* particle_substeps,
                      uint64_t* particle_pushes, uint64_t* particle_rates,
                      int32_t reset);

/* Device memory the world holds, in bytes (particle state, sort and cell
 * tables, neighbour lists, walls). */
size_t sfl_memory_bytes(const sfl_world* world);

/* No counterpart in the reference (a parity hook, like the oracle's solve
 * log): the average error every DFSPH stop test of the last step saw, in
 * order, positive for a divergence solve and negated for a density solve,
 * and how many particles it averaged over (dfsph.rs:263-268). At most 256
 * entries; *len_out gets the count held. */
int32_t sfl_read_solve_log(sfl_world* world, size_t cap, double* average,
                           uint32_t* loaded, size_t* len_out);

/* Blocks of 1024 threads the persistent DFSPH solve runs on (every one
 * resident for its grid barrier). */
int32_t sfl_solve_grid(const sfl_world* world);

/* Blocks until every launch of the world has finished. */
int32_t sfl_synchronize(sfl_world* world);

/* ---- Batched WCSPH ---------------------------------------------------------
 *
 * No reference counterpart as one object: the reference steps many scenes as
 * many `FluidWorld`s, one `step` (src/fluid/step.rs:58) per world. A batch
 * takes a copy of the state of `count` WCSPH worlds with equal settings and
 * steps all of them together; scene s's results are bit-equal to world s
 * stepping alone. The run of `steps` steps is one HIP graph launch (captured
 * on first use and kept while the step count, substep slots, block shapes
 * and cell table stay the same); block shapes come from sfl_batch_autotune.
 */
typedef struct sfl_batch sfl_batch;

enum { SFL_BATCH_CANDIDATES = 5 };

typedef struct sfl_batch_tuning {
  uint32_t sort_block;     /* threads per block: bounds, count, scatter, gather */
  uint32_t density_block;  /* density pass */
  uint32_t force_block;    /* force and advance pass */
  uint32_t candidates;     /* entries used below (sfl_batch_autotune) */
  uint32_t blocks[SFL_BATCH_CANDIDATES];   /* block sizes tried */
  float sort_ms[SFL_BATCH_CANDIDATES];     /* best time of each, milliseconds */
  float density_ms[SFL_BATCH_CANDIDATES];
  float force_ms[SFL_BATCH_CANDIDATES];
} sfl_batch_tuning;

/* Copies the worlds' particles, walls and containers. Every world must be
 * WCSPH, on one device, with equal settings apart from max_particles. The
 * worlds are not changed. SFL_INVALID_SETTINGS otherwise; SFL_UNSUPPORTED
 * for a world whose walls or water went onto a searched cell table (see
 * sfl_step): the batch keeps dense tables. */
int32_t sfl_batch_create(sfl_world* const* worlds, size_t count,
                         sfl_batch** out);
void sfl_batch_destroy(sfl_batch* batch);
size_t sfl_batch_scenes(const sfl_batch* batch);
size_t sfl_batch_len(const sfl_batch* batch, size_t scene);

/* `steps` steps of `dt` for every scene; reports[scene] (count entries, may
 * be null) gets the last step's FluidReport of each scene. A non-finite
 * advance finishes that step for every scene and returns SFL_NON_FINITE
 * ("fluid particle"); sfl_batch_scene_failed names the scenes. */
int32_t sfl_batch_step(sfl_batch* batch, float dt, uint32_t steps,
                       sfl_report* reports);
int32_t sfl_batch_scene_failed(const sfl_batch* batch, size_t scene,
                               int32_t* failed);
/* Captures the graph a call of `steps` steps launches, ahead of that call
 * (graphs are kept per step count until the shapes or the cell table
 * change). */
int32_t sfl_batch_prepare(sfl_batch* batch, uint32_t steps);

/* Times every candidate block size of the three kernel groups on the
 * batch's current state (one substep of `dt`, best of `repeats` runs) and
 * keeps the fastest. The state is not changed. */
int32_t sfl_batch_autotune(sfl_batch* batch, float dt, uint32_t repeats,
                           sfl_batch_tuning* out);
int32_t sfl_batch_set_tuning(sfl_batch* batch, const sfl_batch_tuning* tuning);
int32_t sfl_batch_get_tuning(const sfl_batch* batch, sfl_batch_tuning* out);
/* 1 (default): run steps as one graph launch; 0: plain stream launches. */
int32_t sfl_batch_use_graph(sfl_batch* batch, int32_t on);

/* Scene state in the scene's current order (any pointer may be null). */
int32_t sfl_batch_read_state(sfl_batch* batch, size_t scene, size_t cap,
                             uint32_t* ids, float* xyz, float* vxyz,
                             float* densities);
int32_t sfl_batch_state_hash(sfl_batch* batch, size_t scene, uint64_t* out);
/* Writes the scenes back into `count` worlds (the ones the batch was made
 * from, or any with the same particle counts and settings). */
int32_t sfl_batch_store(sfl_batch* batch, sfl_world* const* worlds,
                        size_t count);
const char* sfl_batch_error_detail(const sfl_batch* batch);
size_t sfl_batch_memory_bytes(const sfl_batch* batch);
/* Kernel nodes of the captured graph (0 before the first graph launch). */
uint64_t sfl_batch_graph_nodes(const sfl_batch* batch);
int32_t sfl_batch_synchronize(sfl_batch* batch);

/* ---- WCSPH adjoint ---------------------------------------------------------
 *
 * No reference counterpart: the reference has no differentiation. A tape
 * records `steps` calls of sfl_step(world, dt) (the world advances as it
 * would), keeping the particle state every `every` steps. sfl_wcsph_adjoint
 * then runs the hand-written reverse pass of those steps (density, Tait
 * pressure, pressure and viscosity forces, walls, symplectic Euler with the
 * container clamp), replaying each segment from its checkpoint. Neighbour
 * sets, substep counts and clamp decisions are those of the recorded run.
 * The world must outlive the tape and must not change between recording
 * and the reverse pass (SFL_OUT_OF_RANGE otherwise); its ids must be
 * 0..len-1 (a world that only ever gained particles). Recording stops with
 * SFL_UNSUPPORTED when the water goes onto a searched cell table (see
 * sfl_step): segments replay on dense tables.
 */
typedef struct sfl_tape sfl_tape;

int32_t sfl_wcsph_record(sfl_world* world, float dt, uint32_t steps,
                         uint32_t every, sfl_tape** out);
/* xbar, vbar: dL/dx, dL/dv of the final state, 3 floats per particle id.
 * xbar0, vbar0: dL/dx, dL/dv of the state at recording start, per id.
 * viscosity_bar: dL/d(settings.viscosity), the artificial-viscosity alpha
 * (may be null). */
int32_t sfl_wcsph_adjoint(sfl_tape* tape, const float* xbar, const float* vbar,
                          float* xbar0, float* vbar0, float* viscosity_bar);
void sfl_tape_destroy(sfl_tape* tape);
uint64_t sfl_tape_substeps(const sfl_tape* tape);
size_t sfl_tape_memory_bytes(const sfl_tape* tape);
const char* sfl_tape_error_detail(const sfl_tape* tape);

#ifdef __cplusplus
}
#endif

#endif /* SOLUNAR_FLUID_H */
