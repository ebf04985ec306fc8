/*
 * avbd_capi.h - C ABI of the MI355X-native AVBD (augmented vertex block
 * descent) step.
 *
 * The library keeps a world of one or more independent scenes resident in
 * HBM and advances it with HIP kernels.  The records below are the exchange
 * format across the boundary (plain C structs, array-of-records); the device
 * keeps its own structure-of-arrays layout and converts on upload/download.
 *
 * Every entry point returns 0 on success and a negative AVBD_E_* code on
 * failure; avbd_last_error() returns a message for the calling thread.  No
 * exception crosses the boundary.
 *
 * Reference interface each entry point replaces (SkyePhyX, src/physics):
 *   avbd_create/avbd_destroy  -> Solver::Solver/~Solver, Solver::Start/Clear
 *                                (solver.hpp:24-41, solver.cpp:18-47)
 *   avbd_upload               -> Solver::AddBody/AddParticle/AddForce/AddEnergy
 *                                (solver.hpp:43-46, solver.cpp:50-164): the host
 *                                facade builds the world, this call moves it
 *                                into device memory
 *   avbd_step                 -> Solver::Step (solver.hpp:41, solver.cpp:199-627)
 *   avbd_download             -> reads of bodyPtrs[i]->transform/velocity,
 *                                Force::constraintPoints, Manifold contacts
 *                                (solver.hpp:32-34, geometry.hpp:148-256,
 *                                force.hpp:9-24,167-198)
 *   avbd_get_status           -> Solver::emergencyStop (solver.hpp:86),
 *                                CheckExplosion (solver.cpp:182-196)
 *   avbd_last_removed         -> the RemoveBody calls of Step's bounds check
 *                                (solver.cpp:520-534)
 *   avbd_remove_bodies,
 *   avbd_add_bodies           -> Solver::RemoveBody / AddBody between steps
 *                                (solver.cpp:50-124), on the resident world
 *   avbd_download_bodies,
 *   avbd_upload_bodies        -> per-frame reads/writes of Mesh state
 *                                (geometry.hpp:148-256)
 *   avbd_set_drag             -> Mesh::isDragged / addedDragVelocity
 *                                (geometry.hpp:193-194), read by Step's
 *                                velocity update (solver.cpp:506-510)
 *   avbd_get_timings          -> Solver::timings / SolverTimings
 *                                (solver.hpp:10-19, solver.cpp:599-626)
 *   avbd_step_scenes          -> (no counterpart: batched/served stepping of a
 *                                subset of resident scenes)
 *   avbd_server_*             -> (no counterpart: the served request loop)
 *   avbd_snapshot_*           -> (no counterpart: world snapshot files)
 *   avbd_get_debug_lines/_points,
 *   avbd_set_body_colours     -> Solver::lineData / debugPointData
 *                                (solver.hpp:36-37), Mesh::color
 *   avbd_get_instances        -> RenderEngine::UpdateInstanceBuffer
 *                                (RenderEngine.cpp:951-1029)
 *   avbd_footprint, avbd_get_memory, avbd_set_option, avbd_tune_*,
 *   avbd_kernel_stats         -> (no counterpart: device memory, tuning and
 *                                measurement)
 */
#ifndef AVBD_CAPI_H
#define AVBD_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVBD_ABI_VERSION 2

/* status codes */
#define AVBD_OK 0
#define AVBD_E_ARG (-1)       /* bad argument / inconsistent world */
#define AVBD_E_HIP (-2)       /* HIP runtime error */
#define AVBD_E_CAPACITY (-3)  /* a device capacity was exceeded */
#define AVBD_E_STATE (-4)     /* call not valid in the current state */
#define AVBD_E_NOGPU (-5)     /* no usable gfx950 device */
#define AVBD_E_UNSUPPORTED (-6)
#define AVBD_E_IO (-7)        /* file missing, unreadable, corrupt or of another format */

/* body flags */
#define AVBD_BODY_STATIC 1u
#define AVBD_BODY_PARTICLE 2u
#define AVBD_BODY_SPHERE 4u  /* Mesh::modelType == ModelType_Sphere (geometry.hpp:17-22) */

/* force flags */
#define AVBD_FORCE_DISABLED 1u

/* eigenvalue projection modes (math.hpp:99-104) */
#define AVBD_PROJ_CLAMP 0
#define AVBD_PROJ_ABSOLUTE 1
#define AVBD_PROJ_ADAPTIVE 2

/* Scene options: the two features the reference declares but leaves
 * unfinished, off by default (the reference's behaviour).
 * AVBD_SCENE_TRUST_REGION: with projection_mode ADAPTIVE, trust_region_rho is
 *   updated after every primal sweep (the empty step 5.2 of solver.cpp:451-455)
 *   to the trust-region ratio of the sweep's FEM-only block updates:
 *   sum (E(x) - E(x - dx)) / sum (r . dx) / 2 over the particles whose block
 *   has energies and no forces, E = the particle's inertia term plus the
 *   elastic energy of its elements; ADAPTIVE then projects by CLAMP when
 *   |rho - 1| <= 0.01, else by ABSOLUTE (math.cpp:152-160).  A sweep whose
 *   predicted decrease is below 3e-5 of the summed magnitude of its terms (a
 *   nearly converged solve, where the ratio is rounding noise) keeps rho.
 *   Without the option rho keeps its value (0 in the reference, so ADAPTIVE
 *   acts as ABSOLUTE).
 * AVBD_SCENE_SPHERE_COLLISION: bodies flagged AVBD_BODY_SPHERE collide as
 *   spheres of radius scale[0] / 2 (the radius their inertia uses,
 *   solver.cpp:84-87); the reference collides every pair as boxes
 *   (manifold.cpp:57). */
#define AVBD_SCENE_TRUST_REGION 1u
#define AVBD_SCENE_SPHERE_COLLISION 2u

/* Per-scene solver parameters and the scene's sticky status word.
 * Mirrors Solver's public tunables (solver.hpp:69-86); gravity and world
 * bounds are constants in the reference (solver.cpp:7, constants.hpp:11-12)
 * and are per-scene parameters here (defaults (0,-9.81,0) and +-50). */
typedef struct avbd_scene {
    float dt;                 /* stepValue */
    int32_t iterations;       /* numIterations */
    float alpha;
    float beta;
    float gamma;
    float on_penetration_penalty;
    int32_t post_stabilization; /* bool */
    int32_t projection_mode;    /* AVBD_PROJ_* */
    float gravity[3];
    float bounds_min[3];
    float bounds_max[3];
    int32_t emergency_stop;   /* state: nonzero once CheckExplosion fired */
    int32_t n_removed;        /* state: bodies removed out of bounds (cumulative) */
    uint32_t options;         /* AVBD_SCENE_* */
    float trust_region_rho;   /* state: trustRegionRho (solver.hpp:80) */
} avbd_scene;

/* One rigid body or particle (Mesh, geometry.hpp:148-256).
 * Quaternions are stored (w, x, y, z).  `inertia` is the body-frame diagonal
 * inertia; `radius` the broad-phase detection radius (|scale|/2,
 * solver.cpp:65).  Bodies of one scene are contiguous and in solverIndex
 * order. */
typedef struct avbd_body {
    float pos[3];
    float rot[4];
    float scale[3];
    float vel[3];
    float ang_vel[3];
    float prev_vel[3];
    float prev_ang_vel[3];
    float mass;
    float inertia[3];
    float friction;
    float radius;
    uint32_t flags;   /* AVBD_BODY_* */
    int32_t scene;
} avbd_body;

/* Joint (force.hpp:116-143, joint.cpp).  body_a == -1 is the world anchor;
 * body indices are global (into the body array).  `serial` orders forces by
 * creation: a body visits its forces in increasing serial order
 * (Mesh::forces append order, force.cpp:5-9). */
typedef struct avbd_joint {
    int32_t body_a;
    int32_t body_b;
    int64_t serial;
    uint32_t flags;           /* AVBD_FORCE_DISABLED */
    float r_a[3];
    float r_b[3];
    float rest_rot[4];        /* (w, x, y, z) */
    float torque_arm;
    float stiffness[6];
    float fracture[6];
    float fmin[6];
    float fmax[6];
    float lambda[6];
    float penalty[6];
} avbd_joint;

/* Spring (force.hpp:92-113, spring.cpp). */
typedef struct avbd_spring {
    int32_t body_a;
    int32_t body_b;
    int64_t serial;
    uint32_t flags;
    float r_a[3];
    float r_b[3];
    float rest_length;
    float stiffness;
    float fracture;
    float fmin;
    float fmax;
    float lambda;
    float penalty;
} avbd_spring;

/* One contact of a manifold (collision.hpp:8-18 plus the warm-start row
 * state of its three constraint rows: normal, tangent1, tangent2). */
typedef struct avbd_contact {
    uint32_t id;
    int32_t stick;
    float r_a[3];             /* body-A local, unscaled */
    float r_b[3];             /* body-B local, unscaled */
    float normal[3];          /* A -> B */
    float penetration;
    float lambda[3];
    float penalty[3];
} avbd_contact;

/* Contact manifold between two boxes (force.hpp:149-198, manifold.cpp). */
typedef struct avbd_manifold {
    int32_t body_a;
    int32_t body_b;
    int64_t serial;
    int32_t n_contacts;
    float friction;
    avbd_contact contacts[8];
} avbd_manifold;

/* Linear tetrahedron with Neo-Hookean energy (energy.hpp:48-85).
 * dm_inv is row-major; the per-vertex shape gradients follow from it
 * (neohookeanfem.cpp:26-36).  `serial` orders a vertex's energies. */
typedef struct avbd_tet {
    int32_t v[4];
    int32_t serial;
    float dm_inv[9];
    float rest_volume;
    float mu;
    float lambda;
} avbd_tet;

/* Triangle membrane with StVK energy (energy.hpp:88-121). */
typedef struct avbd_tri {
    int32_t v[3];
    int32_t serial;
    float dm_inv[4];          /* row-major 2x2 */
    float rest_area;
    float mu;
    float lambda;
} avbd_tri;

/* A whole world: pointers to caller-owned arrays.  For avbd_upload they are
 * read; for avbd_download they are written (sized from avbd_get_counts). */
typedef struct avbd_world {
    int32_t n_scenes;
    int32_t n_bodies;
    int32_t n_joints;
    int32_t n_springs;
    int32_t n_manifolds;
    int32_t n_tets;
    int32_t n_tris;
    avbd_scene* scenes;
    avbd_body* bodies;
    avbd_joint* joints;
    avbd_spring* springs;
    avbd_manifold* manifolds;
    avbd_tet* tets;
    avbd_tri* tris;
} avbd_world;

/* Per-phase wall time of the last step, milliseconds (SolverTimings), from
 * HIP events on the library's stream; a replayed step graph (option
 * "step_graphs") records the events of phases 4-7 itself, so every phase is
 * timed either way. */
typedef struct avbd_timings {
    float broad_phase_ms;
    float warmstart_ms;
    float prediction_ms;
    float primal_dual_ms;
    float velocity_update_ms;
    float post_stab_ms;
    float total_ms;
    int32_t n_colours;        /* max colour count over scenes, last step */
} avbd_timings;

typedef struct avbd_ctx avbd_ctx;

/* Library identification; returns AVBD_ABI_VERSION. */
int avbd_abi_version(void);
/* sizeof of each record, for bindings to check their layouts: scene, body,
 * joint, spring, contact, manifold, tet, tri, world, timings, request_info,
 * server_stats.  Returns the number of records. */
int avbd_record_sizes(int32_t* out, int32_t n);

/* Message describing the last failure on the calling thread. */
const char* avbd_last_error(void);

/* Create a context on HIP device `device`. */
int avbd_create(int32_t device, avbd_ctx** out);
int avbd_destroy(avbd_ctx* ctx);

/* Replace the context's world with `w` (all counts may be zero except
 * n_scenes).  Allocates device storage sized for the world plus headroom for
 * contact manifolds (`manifold_capacity` manifolds; 0 picks a default). */
int avbd_upload(avbd_ctx* ctx, const avbd_world* w, int64_t manifold_capacity);

/* Advance every scene by `n_steps` steps (Solver::Step, n times). */
int avbd_step(avbd_ctx* ctx, int32_t n_steps);

/* Advance only the listed scenes by `n_steps`; others keep their state. */
int avbd_step_scenes(avbd_ctx* ctx, const int32_t* scene_ids, int32_t n_ids, int32_t n_steps);

/* Block until all work queued on the context's stream has finished. */
int avbd_synchronize(avbd_ctx* ctx);

/* Current entity counts (n_* fields of `out`; pointers untouched). */
int avbd_get_counts(avbd_ctx* ctx, avbd_world* out);

/* Copy the world back into caller arrays sized by avbd_get_counts. */
int avbd_download(avbd_ctx* ctx, avbd_world* out);

/* World snapshots (no reference counterpart: the on-disk form of the records
 * above, for fixtures, reproducible configurations and server preload).
 * One binary file: a 128-byte header (magic "AVBDSNAP", format version, ABI
 * version, the size and count of each record kind, a 64-bit checksum of the
 * record bytes) and the scenes, bodies, joints, springs, manifolds (contacts
 * inline), tets and tris as arrays of the records of this header, each
 * starting at a 64-byte aligned offset.  A snapshot is read only by a build
 * with the same record layout; AVBD_E_IO for a missing, truncated, corrupt
 * or foreign file. */
#define AVBD_SNAPSHOT_VERSION 1
int avbd_snapshot_write(const char* path, const avbd_world* w);
/* Record counts of a snapshot (the n_* fields; pointers untouched). */
int avbd_snapshot_counts(const char* path, avbd_world* counts);
/* Reads a snapshot into caller arrays sized from avbd_snapshot_counts. */
int avbd_snapshot_read(const char* path, avbd_world* out);
/* The context's world (as avbd_download gives it) to a snapshot file. */
int avbd_snapshot_save(avbd_ctx* ctx, const char* path);
/* A snapshot file into the context (as avbd_upload with that world).  The
 * loaded context steps as the saved one would have; manifolds it creates
 * later may carry other creation serials (numbering restarts after the
 * largest serial in the file), in the same order. */
int avbd_snapshot_load(avbd_ctx* ctx, const char* path, int64_t manifold_capacity);

/* Per-scene status words: emergency_stop and n_removed for `n` scenes. */
int avbd_get_status(avbd_ctx* ctx, int32_t* emergency_stop, int32_t* n_removed, int32_t n);

/* Solver::trustRegionRho (solver.hpp:80) of `n` scenes, as the last primal
 * sweep left it (it changes only under AVBD_SCENE_TRUST_REGION). */
int avbd_get_trust_region_rho(avbd_ctx* ctx, float* rho, int32_t n);

/* Debug/render output of a step (option "debug_output"): the line and point
 * data Solver::Step collects for the renderer (Solver::lineData /
 * debugPointData, solver.hpp:36-37, filled at solver.cpp:266-313 by
 * Manifold/Joint/Spring/NeoHookeanFEM/STVKFEM::AddLineData and
 * Manifold::AddDebugPointData), as GPULineData / GPUDebugPointData records
 * (geometry.hpp:48-60).  Made on the device from the step-start poses after
 * each step (before its bounds removal); for scenes that ran.  Lines come
 * grouped as joints, springs, contacts, tet edges, triangle edges (the
 * reference holds the same set in the order of its force and energy
 * vectors).  An element edge takes the mean colour of its two bodies
 * (Mesh::color, set with avbd_set_body_colours after an upload; white
 * otherwise) and alpha 1. */
typedef struct avbd_line {
    float start[3];
    float end[3];
    float color[4];
} avbd_line;
typedef struct avbd_debug_point {
    float position[3];
    float pad;
    float color[4];
} avbd_debug_point;
/* Renderer instance data (RenderEngine::UpdateInstanceBuffer,
 * RenderEngine.cpp:951-1029; GPUInstanceData, RenderEngine.hpp:43-49), one
 * per live body in body order from the current poses: column-major model
 * matrix [R diag(scale) | position] (Transform::GetModelMatrix), normal
 * matrix R diag(1/scale) in three 4-float columns (GetNormalMatrix), colour
 * (Mesh::color from avbd_set_body_colours, else white) with alpha 1.  The
 * reference batches instances by model type; the caller, which knows the
 * types, does that. */
typedef struct avbd_instance {
    float model[16];
    float normal[12];
    float color[4];
} avbd_instance;
/* Copies at most cap instances; *n = the live body count. */
int avbd_get_instances(avbd_ctx* ctx, avbd_instance* out, int32_t cap, int32_t* n);

/* avbd_get_instances without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (a render
 * loop under a frame budget, a serving loop, code that forbids hipMalloc).
 * Same arguments, results and errors as avbd_get_instances, but the
 * instance, flag and offset staging lives in a file-scope device array of
 * the library (one per GPU, AVBD_NOALLOC_MAX_BODIES body slots, about
 * 17 MiB, held until the process ends; its address is looked up by
 * avbd_create).  Worlds with more body slots (dead and spare slots count)
 * get AVBD_E_CAPACITY.  Every call on one GPU shares that array: calls that
 * overlap on one GPU, from any context, race. */
#define AVBD_NOALLOC_MAX_BODIES 131072
int avbd_get_instances_noalloc(avbd_ctx* ctx, avbd_instance* out, int32_t cap, int32_t* n);
/* rgb: 3 floats per live body (avbd_download's order) */
int avbd_set_body_colours(avbd_ctx* ctx, const float* rgb, int32_t n);
/* The last step's lines / points: copies at most cap, *n = the full count. */
int avbd_get_debug_lines(avbd_ctx* ctx, avbd_line* out, int64_t cap, int64_t* n);
int avbd_get_debug_points(avbd_ctx* ctx, avbd_debug_point* out, int64_t cap, int64_t* n);

/* Solver::RemoveBody (solver.cpp:104-116, ~Mesh geometry.cpp:11-24) on the
 * device-resident world, without a re-upload: bodies[k] are live body
 * indices (the order avbd_download gives) naming bodies before this call;
 * they are removed in the listed order, each swapped with the last live
 * body of its scene, with their forces and energies (tets/tris that lose a
 * vertex).  Other scenes keep every index.  avbd_last_removed then reports
 * these removals (scene by scene in ascending scene order, each scene's in
 * the listed order). */
int avbd_remove_bodies(avbd_ctx* ctx, const int32_t* bodies, int32_t n);

/* Solver::AddBody / AddParticle (solver.cpp:50-124) on the device-resident
 * world, without a re-upload: each record (its .scene names the scene) is
 * appended to its scene's body order, into a free slot: one a removal
 * freed or a spare one (option "spare_bodies").  AVBD_E_CAPACITY when a
 * scene has no free slot left (upload the world again).  Forces and
 * energies are added with a full upload. */
int avbd_add_bodies(avbd_ctx* ctx, const avbd_body* bodies, int32_t n);

/* Bodies the last avbd_step / avbd_step_scenes call removed for leaving the
 * world bounds (Solver::RemoveBody from Step, solver.cpp:520-534,
 * 104-116), or the last avbd_remove_bodies call removed, in removal order:
 * the scene and the body's index within its scene at the moment of its
 * removal (its solverIndex then).  Applying RemoveBody (swap with last) in
 * this order to a host copy of the scene reproduces the device's body
 * order.  Writes at most cap entries; *n_out receives the full count. */
int avbd_last_removed(avbd_ctx* ctx, int32_t* scene, int32_t* index, int32_t cap, int32_t* n_out);

/* Body state only, n = the live body count (avbd_get_counts), in body order:
 * the per-step read of bodyPtrs[i]->transform/velocity by the renderer and
 * game loop (GameManager.cpp:92-99, RenderEngine.cpp:951-1029) without the
 * forces and contacts. */
int avbd_download_bodies(avbd_ctx* ctx, avbd_body* out, int32_t n);

/* Overwrite the dynamic state of every live body (pos, rot, vel, ang_vel,
 * prev_vel, prev_ang_vel; other fields of `in` are ignored): a host facade's
 * direct writes to Mesh::transform/velocity between steps (e.g. dragging,
 * geometry.hpp:193-194).  Mass, inertia, scale and flags change only
 * through avbd_upload. */
int avbd_upload_bodies(avbd_ctx* ctx, const avbd_body* in, int32_t n);

/* Mesh::isDragged and Mesh::addedDragVelocity (geometry.hpp:193-194) of the
 * listed bodies (indices among the live bodies, as avbd_download_bodies
 * orders them; added_velocity: 3 floats per body; dragged: 0 or 1).  The
 * next velocity update adds a dragged body's added velocity to its new
 * velocity once and clears it (solver.cpp:506-510); the flag stays until set
 * again.  Bodies not listed keep their state.  Removed and added bodies
 * start undragged.  Not kept in snapshots. */
int avbd_set_drag(avbd_ctx* ctx, const int32_t* bodies, const float* added_velocity, const int32_t* dragged,
                  int32_t n);

int avbd_get_timings(avbd_ctx* ctx, avbd_timings* out);

/* Per-body colour of the last step (-1 for static bodies). */
int avbd_get_colours(avbd_ctx* ctx, int32_t* out, int32_t n);

/* Tuning knobs (no reference counterpart):
 *   "lanes"          lanes cooperating on one block solve of a body with
 *                    forces: 1, 2, 4, 8, 16 or 32 (1 = the reference's
 *                    sequential accumulation order), 0 = autotuned (default)
 *   "fem_lanes"      the same for particles whose only terms are FEM
 *                    energies
 *   "kernel_timing"  1 = bracket every primal-sweep and dual launch with HIP
 *                    events on the context's stream (see avbd_kernel_stats);
 *                    steps then run without graphs
 *   "step_graphs"    1 (default) = phases 4-7 of a step (prediction, the
 *                    iterations, velocities, post-stabilisation) are captured
 *                    into a HIP graph the second time their launch shape
 *                    (colours, iterations, lane widths and grids) occurs and
 *                    replayed from then on; 0 = always launched one by one.
 *                    Results are identical either way.
 *   "fused_scenes"   1 = scenes of at most "fused_max_bodies" bodies (default
 *                    512) without breakable joints or springs run phases 4-7
 *                    in one workgroup each (one launch for all of them, on
 *                    its own stream beside the bucket launches) when "lanes"
 *                    equals "fem_lanes"; 0 (default) = every scene on the
 *                    per-colour bucket launches.  Results are identical.
 *   "fused_single_scene" 1 = a world of one such scene runs it in the
 *                    per-scene kernel as well; 0 (default) = by
 *                    "fused_scenes" alone.
 *   "scene_lds_bytes" 0 to 65536 (default 61440): LDS per workgroup of the
 *                    per-scene kernel; a scene whose body state and incidence
 *                    lists fit (256 bytes per body plus 4 per incidence
 *                    entry) runs on an LDS copy of them, others on global
 *                    memory; 0 = every scene on global memory.
 *   "spare_bodies"   k >= 0 (default 0): the next avbd_upload gives every
 *                    scene k free body slots after its bodies, for
 *                    avbd_add_bodies (bodies added between steps without a
 *                    re-upload); they take device memory, not body indices.
 *   "debug_output"   1 = after every step, make its line and point data
 *                    (avbd_get_debug_lines / _points); 0 (default) = none.
 *   "colour_updates" 1 (default) = a scene whose constraint graph gained or
 *                    lost edges since its last colouring re-evaluates only
 *                    the bodies the change can reach; 0 = every recolouring
 *                    from scratch.  The colouring is the same either way. */
int avbd_set_option(avbd_ctx* ctx, const char* name, int64_t value);

/* Autotuner table (no reference counterpart): the launch shape chosen per
 * (kernel class, log2 of bodies per launch) of the primal sweep, lanes per
 * body and workgroup size, as JSON {"key": [lanes, workgroup], ...}; keys
 * still being tuned are left out.  Writes at most len bytes (NUL-terminated)
 * and returns the full length. */
int avbd_tune_table(avbd_ctx* ctx, char* out, int32_t len);
/* Seed the autotuner with a table written by avbd_tune_table (a bare
 * width, {"key": width}, takes workgroup size 256). */
int avbd_tune_load(avbd_ctx* ctx, const char* json);

/* Accumulated launch count and device time (ms) of one kernel class since
 * the last reset: kernel = "primal_forces" / "primal_fem" (each primal
 * bucket launch: bodies with forces / FEM-only particles, timed on the
 * stream it runs on), "primal" (both together) or "dual".  reset != 0 clears
 * after reading.  Requires the "kernel_timing" option.  kernel =
 * "step_graph": launches = step graph replays, total_ms = graphs captured. */
int avbd_kernel_stats(avbd_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms, int32_t reset);

/* Device memory (no reference counterpart).  avbd_footprint: the bytes an
 * avbd_upload of `w` with this manifold_capacity allocates (host-only, needs
 * no device).  avbd_get_memory: the bytes the context holds now. */
int avbd_footprint(const avbd_world* w, int64_t manifold_capacity, int64_t* bytes);
int avbd_get_memory(avbd_ctx* ctx, int64_t* bytes);

/* ------------------------------------------------------------------------
 * Served form (no reference counterpart): requests for k steps of a
 * registered scene, continuous batching.
 *
 * Scenes are registered once (the server keeps a host copy).  A scene is
 * resident when its state lives in the server's device world; residency is
 * bounded by an HBM budget in bytes (the exact allocation of the device
 * world, avbd_footprint) and idle scenes are evicted least recently used
 * when another must come in.  Each avbd_server_tick admits queued requests
 * in arrival order into free slots (at most max_slots in flight, at most one
 * per scene), makes their scenes resident, advances every in-flight scene by
 * one step in a single batched device step and retires the requests whose
 * steps are done.  Requests submitted between ticks join at the next tick.
 * A scene's result does not depend on which other scenes share the device
 * world or its ticks: the server's context starts with fixed lane widths
 * (options "lanes" = "fem_lanes" = 4), which makes it bit-identical.  With
 * autotuned widths (0, set through avbd_server_context) the accumulation
 * order follows the launch size and results agree only to rounding.  One
 * thread drives a server.
 */
typedef struct avbd_server avbd_server;

typedef struct avbd_request_info {
    int64_t ticket;
    int32_t scene;
    int32_t steps;
    int32_t state;            /* 0 queued, 1 running, 2 done */
    int32_t pad_;
    double t_arrival;         /* seconds on the server clock (avbd_server_now) */
    double t_start;           /* start of the tick of the first step, -1 before */
    double t_done;            /* end of the tick of the last step, -1 before */
} avbd_request_info;

typedef struct avbd_server_stats {
    int64_t ticks;            /* ticks that stepped at least one scene */
    int64_t scene_steps;      /* scenes advanced by one step, summed over ticks */
    int64_t completed;
    int64_t layout_rebuilds;  /* residency changes (download + upload) */
    int64_t resident_bytes;   /* device bytes of the resident world */
    int32_t resident_scenes;
    int32_t queued;
    int32_t running;
    int32_t pad_;
} avbd_server_stats;

int avbd_server_create(int32_t device, int64_t hbm_budget_bytes, int32_t max_slots, avbd_server** out);
int avbd_server_destroy(avbd_server* s);
/* Register every scene of `w` (the server copies them); they get the
 * consecutive ids first_id, first_id + 1, ...  Fails with AVBD_E_CAPACITY if
 * a scene alone exceeds the budget (then none is registered). */
int avbd_server_add_scenes(avbd_server* s, const avbd_world* w, int32_t* first_id);
/* Make registered scenes resident now, in id order, while they fit. */
int avbd_server_preload(avbd_server* s);
/* Seconds since the server was created. */
double avbd_server_now(avbd_server* s);
/* Queue a request for n_steps steps of a scene; t_arrival < 0 means now. */
int avbd_server_submit(avbd_server* s, int32_t scene_id, int32_t n_steps, double t_arrival, int64_t* ticket);
/* One scheduling round: admit, one batched step, retire.  n_running gets the
 * number of scenes stepped (0 when nothing was runnable). */
int avbd_server_tick(avbd_server* s, int32_t* n_running);
/* State of a request that has not been polled yet. */
int avbd_server_request(avbd_server* s, int64_t ticket, avbd_request_info* out);
/* Move up to cap finished requests (completion order) into out. */
int avbd_server_poll(avbd_server* s, avbd_request_info* out, int32_t cap, int32_t* n_out);
int avbd_server_get_stats(avbd_server* s, avbd_server_stats* out);
/* Current state of one scene (local body indices, force serials renumbered
 * by rank): call with every out->n_* = 0 to get the counts, then with arrays
 * of at least those sizes. */
int avbd_server_download_scene(avbd_server* s, int32_t scene_id, avbd_world* out);
/* The server's solver context (options, autotuner, timings). */
avbd_ctx* avbd_server_context(avbd_server* s);

#ifdef __cplusplus
}
#endif

#endif /* AVBD_CAPI_H */
