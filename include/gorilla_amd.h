/*
 * gorilla_amd.h — C ABI of the MI355X batched rigid-body step.
 *
 * The reference crate (one-for-all/gorilla-physics, Rust) exposes this path
 * in-process through MechanismState / step / simulate / dynamics_continuous.
 * Each entry point below replaces one of those and names the reference
 * interface it stands in for (file:line inside the reference crate).
 *
 * Conventions (identical to the reference):
 *  - body i is the child of joint i; parent[i] < i; parent -1 is the world
 *    (mechanism.rs:37-46, 98-116);
 *  - x_tree is the joint's initial transform successor -> predecessor
 *    (init_iso: revolute.rs:18, prismatic.rs:23, floating.rs:9);
 *  - spatial vectors are [angular; linear]; floating-joint velocity is the body
 *    twist relative to the parent, expressed in the body frame
 *    (joint/mod.rs:295-296);
 *  - flattening of q per instance: revolute/prismatic 1 value, floating 7
 *    values [qx,qy,qz,qw, tx,ty,tz], fixed 0 (joint/mod.rs:208-224);
 *    v / tau / vdot: 1, 6 ([w; v]) and 0 values (joint/mod.rs:286-303);
 *  - gravity is a base acceleration +g z (dynamics.rs:212-218, lib.rs:39).
 *
 * Threading: a handle is not thread-safe; distinct handles may be used from
 * distinct threads. Every call returns an int status (GP_OK = 0); no C++
 * exception or abort crosses this boundary. gp_last_error() returns a
 * thread-local message for the last failing call on the calling thread.
 */
#ifndef GORILLA_AMD_H
#define GORILLA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_ABI_VERSION 5

/* ---- status codes (replace the reference's panic!/expect sites) ---------- */
enum {
  GP_OK = 0,
  GP_ERR_INVALID = 1,     /* bad argument (NULL, size, non-unit axis ...)     */
  GP_ERR_TOPOLOGY = 2,    /* parent >= own index, unknown joint type          */
  GP_ERR_NOMEM = 3,       /* host or device allocation failed                 */
  GP_ERR_DEVICE = 4,      /* HIP runtime error / no device                    */
  GP_ERR_UNSUPPORTED = 5, /* option not available for this batch              */
  GP_ERR_FAULT = 6,       /* an instance produced a non-finite value / D <= 0 */
  GP_ERR_BUSY = 7         /* server full (no slot / memory for the request)   */
};

/* ---- joint types (joint/mod.rs:22) ---------------------------------------- */
enum {
  GP_JOINT_FIXED = 0,
  GP_JOINT_REVOLUTE = 1,
  GP_JOINT_PRISMATIC = 2,
  GP_JOINT_FLOATING = 3
};

/* ---- integrators (integrators.rs:17) -------------------------------------- */
enum {
  GP_SEMI_IMPLICIT_EULER = 0,
  GP_RUNGE_KUTTA_2 = 1,
  GP_RUNGE_KUTTA_4 = 2,
  /* velocity stepping (integrators.rs:42-49, dynamics.rs:366-742): free velocity
   * without contact forces, then the contact impulses of a second-order-cone
   * program (friction cone mu = 0.95, dynamics.rs:674) over the touching
   * contact points and sphere colliders; q advanced with the new velocity */
  GP_VELOCITY_STEPPING = 3,
  /* the same with continuous collision detection (integrators.rs:51-175): the
   * step is cut at the earliest time a separated contact point or sphere would
   * reach a half-space, and the remainder stepped again */
  GP_CCD_VELOCITY_STEPPING = 4
};

/* ---- arithmetic of the batched step --------------------------------------- */
enum {
  /* everything in FP64 (the reference's Float = f64, types.rs:3) */
  GP_PREC_F64 = 0,
  /* FP32 kinematics, bias and contact forces and both ABA sweeps; FP64 for the
   * D = S^T I^A S solve and for the q/v accumulation of the integrator */
  GP_PREC_MIXED = 1
};

/* ---- kernel family --------------------------------------------------------- */
enum {
  GP_KERNEL_AUTO = 0,  /* FAST when the bucket topology has a compiled kernel;
                          velocity stepping: the register-resident kernel for
                          a lone floating body with at most 8 contact
                          candidates, the runtime-topology one otherwise     */
  GP_KERNEL_PLAIN = 1, /* runtime-topology FP64 kernel, per-body
                          intermediates in device scratch (also for velocity
                          stepping)                                          */
  GP_KERNEL_FAST = 2   /* compile-time topology, register resident, several
                          steps per launch                                     */
};

/* One body and the joint that connects it to its parent. */
typedef struct gp_body_desc {
  int32_t parent;          /* -1 = world, else index of the parent body (< own) */
  int32_t jtype;           /* GP_JOINT_*                                        */
  double axis[3];          /* joint axis, successor frame (revolute/prismatic)  */
  double x_tree_quat[4];   /* init_iso rotation as quaternion [x,y,z,w]         */
  double x_tree_trans[3];  /* init_iso translation                              */
  double moment[6];        /* J about the body-frame origin: xx,yy,zz,xy,xz,yz   */
  double mc[3];            /* mass * centre of mass (cross_part, inertia.rs:32) */
  double mass;
  double spring_k;         /* prismatic joint spring (k = 0: none), prismatic.rs:13 */
  double spring_l;         /* spring rest length                                */
} gp_body_desc;

/* Contact point on a body (contact.rs:17); default k is 50e3 (contact.rs:28). */
typedef struct gp_contact_desc {
  int32_t body;
  int32_t _pad;
  double p[3];  /* location in the body frame */
  double k;
} gp_contact_desc;

/* Half-space (collision/halfspace.rs:6); defaults alpha 0.9, mu 0.5. */
typedef struct gp_halfspace_desc {
  double point[3];
  double normal[3]; /* unit outward normal */
  double alpha;
  double mu;
} gp_halfspace_desc;

/* Stateful spring contact (SpringContact, contact.rs:74-94, 131-186): a
 * massless leg from the body's origin along `direction` (body frame, unit) of
 * rest length l_rest and stiffness k. While unregistered, its tip is tested
 * against the half-spaces; the first one containing the tip registers that
 * world point. While registered, a compressed leg pushes the body away from
 * the point (force k (l_rest - l) along the leg, at the body's origin); once
 * it is no longer shorter than l_rest it detaches and its direction becomes
 * the leg's last world direction. Semi-implicit Euler only (simulate.rs:56-71). */
typedef struct gp_spring_contact_desc {
  int32_t body;
  int32_t _pad;
  double direction[3];
  double l_rest;
  double k;
} gp_spring_contact_desc;

/* Sphere collider of a body (collision/sphere.rs:6-39, rigid_body.rs:11-37):
 * centre at `center` in the body frame (Sphere::base_translation), radius.
 * Touches a half-space when the centre is at most `radius` above it; used by
 * the velocity-stepping integrators only, as in the reference. A body has at
 * most one collider; enabled = 0 switches it off (Collider::enabled). */
typedef struct gp_sphere_desc {
  int32_t body;
  int32_t enabled;
  double center[3];
  double radius;
} gp_sphere_desc;

typedef struct gp_model_desc {
  int32_t n_bodies;
  int32_t n_contacts;
  int32_t n_halfspaces;
  int32_t _pad;
  const gp_body_desc* bodies;
  const gp_contact_desc* contacts;
  const gp_halfspace_desc* halfspaces;
  double gravity; /* 9.81 in the reference (lib.rs:39) */
  int32_t n_spring_contacts;  /* zero-initialised descriptions have none */
  int32_t _pad2;
  const gp_spring_contact_desc* spring_contacts;
  int32_t n_spheres;          /* sphere colliders, at most one per body */
  int32_t _pad3;
  const gp_sphere_desc* spheres;
} gp_model_desc;

/* Options of a batch. Zero-initialised = defaults. */
typedef struct gp_batch_opts {
  int32_t precision;         /* GP_PREC_*                                     */
  int32_t kernel;            /* GP_KERNEL_*                                   */
  int32_t steps_per_launch;  /* steps fused in one launch by every kernel
                                (0 = 100)                                     */
  int32_t block_size;        /* threads per workgroup, 64/128/256 (0 = 256;
                                grouped launches 128; the single-body
                                velocity-stepping kernel 64, or 16 at one
                                lane per body)                                */
  int32_t lanes_per_mech;    /* FAST free bodies with contacts: lanes sharing
                                one instance's contact points, 1/2/4/8 (0 = 4);
                                velocity stepping of a lone floating body with
                                at most 8 contact candidates: 1 = one lane
                                per body, else 8 lanes per body             */
  int32_t use_graph;         /* capture multi-launch rollouts in a hipGraph   */
                             /* (a chunk of one kernel is launched directly)  */
  int32_t separate_launches; /* 1: one launch per topology bucket instead of
                                one grouped launch over all FAST buckets      */
  int32_t lanes_per_tree;    /* FAST floating base with 4 three-joint legs:
                                1 = one lane per instance, 4 = one lane per leg
                                (0 = 4)                                         */
} gp_batch_opts;

typedef struct gp_model gp_model;
typedef struct gp_batch gp_batch;

/* ---- errors ------------------------------------------------------------------ */
const char* gp_last_error(void);
int gp_abi_version(void);

/* ---- model --------------------------------------------------------------------
 * Replaces MechanismState::new (mechanism.rs:62) with add_halfspace (:379) and
 * add_contact_point (:384). Validation that panics in the reference
 * (mechanism.rs:91-115) returns GP_ERR_TOPOLOGY here. The model is immutable. */
int gp_model_create(const gp_model_desc* desc, gp_model** out);
void gp_model_destroy(gp_model* model);
/* nq, nv of one instance (flattening of joint/mod.rs:208-303) */
int gp_model_dims(const gp_model* model, int32_t* nq, int32_t* nv, int32_t* n_bodies);
/* Topology signature used for bucketing (parents, dof classes, contacts). */
int gp_model_signature(const gp_model* model, uint64_t* sig);

/* ---- batch ----------------------------------------------------------------------
 * counts[k] instances of models[k]; instances are numbered in that order. Every
 * instance starts at the reference's zero state (q = 0, floating q = identity,
 * v = 0; mechanism.rs:68-88) with zero held torque. State lives in device memory
 * as structure-of-arrays per topology bucket. opts may be NULL. */
int gp_batch_create(gp_model* const* models, const int64_t* counts, int32_t n_models,
                    const gp_batch_opts* opts, gp_batch** out);
void gp_batch_destroy(gp_batch* batch);
int gp_batch_info(const gp_batch* batch, int64_t* n_instances, int64_t* nq_total,
                  int64_t* nv_total, int32_t* n_buckets);
/* per-instance offsets into the flattened host arrays (n_instances + 1 each) */
int gp_batch_offsets(const gp_batch* batch, int64_t* q_offsets, int64_t* v_offsets);
int gp_batch_set_opts(gp_batch* batch, const gp_batch_opts* opts);
int gp_batch_get_opts(const gp_batch* batch, gp_batch_opts* opts);

/* MechanismState::update(q, v) (mechanism.rs:209) for every instance; host
 * arrays, instance-major flattening. */
int gp_batch_set_state(gp_batch* batch, const double* q, const double* v);
int gp_batch_get_state(gp_batch* batch, double* q, double* v);
/* Torques held over the following steps (the tau argument of step(),
 * simulate.rs:20); NULL = zeros, as an empty tau in simulate.rs:27-37. */
int gp_batch_set_tau(gp_batch* batch, const double* tau);

/* A spring contact's direction (body frame, unit) and rest length for one
 * instance, as the reference's controllers set them between steps
 * (SLIP_control.rs:120-121, contact.rs:887-888); and its registration state
 * (contact.rs:78-79): registered (0/1), the registered half-space (its index in
 * the model's description, -1 when unregistered), the registered world point,
 * the current direction and rest length. index counts the model's spring
 * contacts in description order. Any output pointer may be NULL. */
int gp_batch_set_spring_contact(gp_batch* batch, int64_t instance, int32_t index, const double* direction,
                                double l_rest);
int gp_batch_get_spring_contact(gp_batch* batch, int64_t instance, int32_t index, int32_t* registered,
                                int32_t* halfspace, double* point, double* direction, double* l_rest);

/* The same transfers with DEVICE arrays in the same instance-major flattening,
 * asynchronous on `stream` (a hipStream_t; NULL = the batch's stream): the
 * transposition to and from the buckets' structure-of-arrays runs on the GPU.
 * For callers whose state or torques already live in HBM (a policy network's
 * output, a tensor), so a control loop never crosses PCIe. No reference
 * counterpart (the reference's state is host memory). */
int gp_batch_set_state_dev(gp_batch* batch, const double* q, const double* v, void* stream);
int gp_batch_get_state_dev(gp_batch* batch, double* q, double* v, void* stream);
int gp_batch_set_tau_dev(gp_batch* batch, const double* tau, void* stream);
int gp_batch_get_vdot_dev(gp_batch* batch, double* vdot, void* stream);

/* step(state, dt, tau, integrator) (simulate.rs:20) applied n_steps times with
 * the held torques. Asynchronous on `stream` (a hipStream_t; NULL = the batch's
 * own stream). */
int gp_step(gp_batch* batch, double dt, int32_t integrator, int32_t n_steps, void* stream);
/* Do now the one-time work that gp_step(batch, dt, integrator, n_steps) would
 * do on its first call: the grouped launch's table, bucket scratch and, with
 * use_graph, the hipGraphs it replays (one per chunk of steps_per_launch steps
 * and one for a shorter last chunk; a chunk whose graph holds a single kernel
 * is launched directly, which costs less than the graph's launch), so that
 * later gp_step calls only launch.
 * Nothing is stepped. No reference counterpart (setup of the batched launch). */
int gp_step_prepare(gp_batch* batch, double dt, int32_t integrator, int32_t n_steps, void* stream);
/* gp_step without any allocation, for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same results as
 * gp_step, bit for bit; no hipGraph and no grouped launch (one launch per
 * bucket and chunk). The runtime-topology and velocity-stepping kernels take
 * their per-lane scratch from one static device array of the library
 * (64 MiB per GPU, held until the process ends) instead of the batch's own
 * buffers: a batch whose largest such bucket needs more returns
 * GP_ERR_UNSUPPORTED and nothing runs. Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (other streams, other threads)
 * race on it: serialise them. Not timed (gp_batch_last_timing refers to the
 * last gp_step). */
int gp_step_noalloc(gp_batch* batch, double dt, int32_t integrator, int32_t n_steps, void* stream);
/* Block until the batch's work on its stream is done; when a call since the
 * last gp_batch_sync was given a stream of the caller's own, until all work
 * on the device is done. */
int gp_batch_sync(gp_batch* batch);

/* Joint accelerations of the last step's first dynamics evaluation (the vdot
 * the reference's semi_implicit_euler computes, integrators.rs:30), host array
 * of nv_total values. */
int gp_batch_get_vdot(gp_batch* batch, double* vdot);

/* dynamics_continuous(state, tau) (dynamics.rs:322): vdot for every instance at
 * the current state. tau NULL = held torques. Host arrays, synchronous. */
int gp_dynamics(gp_batch* batch, const double* tau, double* vdot);

/* mass_matrix(state, bodies_to_root) (mechanism.rs:637-696): the joint-space
 * inertia M(q) of every instance at the batch's current state, by the
 * composite-rigid-body algorithm on the GPU (FP64). M receives, instance after
 * instance, an nv_i x nv_i row-major block (sum of nv_i^2 values; a floating
 * joint's rows and columns are [w; v] of its body twist, as the reference
 * orders them). Host array, synchronous. */
int gp_mass_matrix(gp_batch* batch, double* M);

/* dynamics_bias(state, ...) (dynamics.rs:233-251): c(q, v) - tau_c of every
 * instance at the batch's current state, the joint torques of the
 * velocity-product and gravity terms minus those of the contact wrenches
 * (contact_dynamics, contact.rs:97-129), by recursive Newton-Euler with zero
 * joint accelerations on the GPU (FP64). Prismatic springs are not included
 * (dynamics_continuous adds them to tau, dynamics.rs:298-315); spring
 * contacts act with their current registration, which this call does not
 * change. c has nv_total values in the v flattening. Host array,
 * synchronous. */
int gp_dynamics_bias(gp_batch* batch, double* c);

/* simulate(state, final_time, dt, control_fn, integrator) (simulate.rs:87).
 * The loop is the reference's float-accumulated `while t < final_time`
 * (simulate.rs:97-109). control_fn (may be NULL = zero torque) is called once
 * per step on the host with the current state and fills tau (nv_total values).
 * If qs / vs are non-NULL they receive the initial state and every step's state,
 * (max_records rows of nq_total / nv_total values); n_records gets the number
 * of rows the loop produced (steps + 1). */
typedef int (*gp_control_fn)(void* user, const double* q, const double* v, double t,
                             double* tau_out);
int gp_simulate(gp_batch* batch, double final_time, double dt, gp_control_fn control_fn,
                void* user, int32_t integrator, double* qs, double* vs, int64_t max_records,
                int64_t* n_records);

/* MechanismState::kinetic_energy and gravitational_energy (mechanism.rs:334-362)
 * of every instance at the batch's current state: kinetic[i] = sum over
 * bodies of 1/2 (w.J w + v.(m v + 2 w x c)) in world coordinates
 * (inertia.rs:182-202); gravitational[i] = sum of m g z of each body frame's
 * origin, as the reference computes it. Either pointer may be NULL. FP64. */
int gp_batch_energy(gp_batch* batch, double* kinetic, double* gravitational);

/* Per-instance fault flags (1 = a non-finite q/v/vdot, a non-positive D or a
 * spring contact pulling into its half-space was seen since the last clear).
 * No reference counterpart: the reference panics (dynamics.rs:267-274,
 * contact.rs:161-163). */
int gp_batch_faults(gp_batch* batch, uint8_t* faults);
int gp_batch_clear_faults(gp_batch* batch);

/* Rounds of the collision loop in each instance's last velocity-stepping step
 * (the loop of ccd_velocity_stepping, integrators.rs:56-174: 1 = the step
 * was not cut; always 1 without continuous collision detection); 0 for an
 * instance that has taken no such step since the batch was built. A
 * diagnostic with no counterpart in the reference, whose loop count is
 * internal. */
int gp_batch_collision_rounds(gp_batch* batch, int32_t* rounds);

/* Time the batch's last gp_step on its own stream (HIP events around every
 * launch): total milliseconds and number of kernel launches. */
int gp_batch_last_timing(const gp_batch* batch, double* ms, int64_t* launches);

/* Autotuner. No counterpart in the reference (one mechanism per process, no
 * launch shape to choose). Times short trial rollouts of this batch (trial_steps
 * steps each, 0 = 200; the fastest of three runs, after one untimed warm-up)
 * while it walks, one option at a time, lanes_per_mech (1/2/4/8, if a FAST
 * free-body bucket exists), lanes_per_tree (1/4, if a FAST quadruped bucket
 * exists), block_size (64/128/256) and steps_per_launch (25/50/100/200),
 * keeping a value only if it is at least 2 % faster (zero options are replaced
 * by the defaults they stand for first). The batch's state, vdot and fault flags are restored afterwards and
 * the chosen options are applied (use_graph kept as set) and written to
 * *chosen; *best_ms is the chosen trial's time. Workgroup size and steps per
 * launch do not change any result bit; the lane counts change the order of
 * the contact / leg sums, i.e. results at the level of rounding. Blocking. */
int gp_batch_autotune(gp_batch* batch, double dt, int32_t integrator, int32_t trial_steps,
                      gp_batch_opts* chosen, double* best_ms);

/* ---- serving ----------------------------------------------------------------
 * No counterpart in the reference, which steps one mechanism in-process
 * (simulate.rs:87). A request is one rollout: a model, its initial state, held
 * torques, a horizon in steps and dt. The server keeps every admitted request's
 * state resident in device memory, in a pool per (topology, dt) whose lanes are
 * slots; between chunks of chunk_steps steps it admits queued requests into
 * free slots (continuous batching), creating a pool only while the memory
 * budget allows, and runs one grouped launch per chunk over every pool. Each
 * request's result is its final (q, v) and its latency = completion - arrival
 * on the server clock. Topologies without a FAST kernel are rejected
 * (GP_ERR_UNSUPPORTED). A model must outlive its requests. */
typedef struct gp_server_opts {
  int32_t precision;        /* GP_PREC_*                                      */
  int32_t chunk_steps;      /* steps per round (0 = 100)                      */
  int32_t slots_per_pool;   /* lanes of one pool (0 = 4096)                   */
  int32_t block_size;       /* threads per workgroup (0 = 128)                */
  int64_t mem_budget_bytes; /* pools' device memory (0 = half the free memory) */
  int32_t lanes_per_mech;   /* as gp_batch_opts (0 = 4)                       */
  int32_t lanes_per_tree;   /* as gp_batch_opts (0 = 4)                       */
  int32_t _reserved[2];
} gp_server_opts;
typedef struct gp_server gp_server;

int gp_server_create(const gp_server_opts* opts, gp_server** out);
void gp_server_destroy(gp_server* server);
/* seconds since the server was created */
int gp_server_clock(const gp_server* server, double* now_s);
/* Queue a request that arrives at arrival_s on the server clock (< 0: now).
 * The first request of a model uploads the model's parameters to the device
 * once (a registry the server keeps until it is destroyed); admissions then
 * copy only each request's state, torques and horizon. */
int gp_server_submit(gp_server* server, const gp_model* model, const double* q0, const double* v0,
                     const double* tau, int64_t horizon_steps, double dt, double arrival_s,
                     int64_t* request_id);
/* Run rounds until nothing is queued or running (or max_rounds > 0 rounds). */
int gp_server_run(gp_server* server, int32_t max_rounds, int32_t* rounds_done);
int gp_server_pending(const gp_server* server, int64_t* queued, int64_t* running);
/* Ids of requests completed since the last call (up to max_ids). */
int gp_server_completed(gp_server* server, int64_t* ids, int64_t max_ids, int64_t* n_ids);
/* Final state of a completed request (then forgotten); fault = 1 if the
 * instance produced a non-finite value. */
int gp_server_result(gp_server* server, int64_t request_id, double* q, double* v, double* latency_s,
                     int32_t* fault);
/* 10 values: [rounds, env_steps, busy_seconds, admitted, completed, pools,
 *  pool_bytes, budget_bytes, admit_seconds, retire_seconds]; busy_seconds is
 *  the host time from each round's launch to its results, admit_seconds the
 *  host time spent admitting (packing, one transfer and one launch per 32
 *  pools per round, its wait), retire_seconds the time spent packing and
 *  copying back finished slots after the round's steps ended (inside
 *  busy_seconds). */
int gp_server_stats(const gp_server* server, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* GORILLA_AMD_H */
