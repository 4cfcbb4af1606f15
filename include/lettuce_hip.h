/* lettuce HIP backend: C ABI of the MI355X (gfx950) force backend library liblettuce_hip.so.
 *
 * Every entry point mirrors one member of the reference's Backend interface (src/backend.hpp) or
 * its factory functions, with plain pointers and sizes. C++ callers use the HipBackend class
 * (lettuce_simulate_physics_amd/host/hip_backend.cpp), which implements Backend on top of this ABI
 * and rethrows errors as the reference's exception types.
 *
 * Status codes: LHIP_OK on success; LHIP_RUNTIME for device/runtime failures (reference:
 * std::runtime_error), LHIP_LOGIC for protocol misuse (reference: std::logic_error), etc. The
 * message of the last failure on a context is lhip_last_error(ctx).
 *
 * Threading: one caller thread per context (as the reference backends). Bodies live on the device
 * between calls; kick and drift are asynchronous; compute_forces(_with), diagnostics, download,
 * body_costs and times synchronise.
 */
#ifndef LETTUCE_HIP_H
#define LETTUCE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LHIP_OK = 0,
    LHIP_RUNTIME = 1,
    LHIP_LOGIC = 2,
    LHIP_UNAVAILABLE = 3,
    LHIP_OUT_OF_MEMORY = 4,
    LHIP_INVALID = 5,
    LHIP_NOT_PROVIDED = 6
};

enum { LHIP_METHOD_DIRECT = 0, LHIP_METHOD_TREE = 1 };

typedef struct lhip_ctx lhip_ctx;

/* ForceParams (src/backend.hpp:14-19). method: LHIP_METHOD_DIRECT or LHIP_METHOD_TREE. */
typedef struct lhip_force_params {
    double G;
    double eps;
    double theta;
    int32_t method;
    int32_t reserved;
} lhip_force_params;

/* Diag (src/backend.hpp:30-38). */
typedef struct lhip_diag {
    double mass, ke, pe;
    int32_t has_pe, pad;
    double px, py, pz;
    double lx, ly, lz;
    double cx, cy, cz;
} lhip_diag;

/* PhaseTimes (src/backend.hpp:40-49) plus split counters: p2p (body-body) and p2c (body-node)
 * interactions, and `listed`, the reference CPU backend's count (list entries per body, the body's
 * own entry included). interactions = p2p + p2c. p2c_mono: the body-node interactions evaluated as
 * monopoles by the error-budgeted far field (part of p2c; 0 with the reference's walk options). */
typedef struct lhip_phase_times {
    double force_ms, sort_ms, build_ms, walk_ms, step_ms;
    uint64_t interactions;
    double comm_ms, balance_ms;
    uint64_t p2p, p2c, listed;
    uint64_t p2c_mono;
} lhip_phase_times;

/* ---- factory (src/backend.hpp:143-148; make_cuda_backend src/cuda_backend.cu:2026-2048) ---- */

/* Devices visible to this process (cuda_device_count, src/backend.hpp:148); 0 without a driver. */
int lhip_device_count(void);
/* 1: this library was built with its device code (cuda_compiled, src/backend.hpp:146). */
int lhip_compiled(void);
/* make_cuda_backend(device, why): on failure returns LHIP_UNAVAILABLE (or another code), *out =
 * NULL and the reason in why[0..why_len). Never throws. */
int lhip_create(int device, lhip_ctx** out, char* why, size_t why_len);
void lhip_destroy(lhip_ctx* ctx);
/* Backend::name (src/backend.hpp:93), e.g. "hip: AMD Instinct MI355X (gfx950, 256 CUs, 288 GB)". */
int lhip_name(const lhip_ctx* ctx, char* buf, size_t len);
const char* lhip_last_error(const lhip_ctx* ctx);

/* ---- state (src/backend.hpp:95-104) ---- */

/* Backend::upload: copies n bodies (SoA, reference Bodies src/bodies.hpp:8-13), zeroes a and phi,
 * synchronises. Refuses n >= 2^31 - 1 (src/cuda_backend.cu:1292-1293). */
int lhip_upload(lhip_ctx* ctx, uint64_t n, const float* x, const float* y, const float* z, const float* vx,
                const float* vy, const float* vz, const float* m, const uint32_t* id);
uint64_t lhip_size(const lhip_ctx* ctx);
/* Backend::download: fills n = lhip_size() entries per array, in the backend's current order. */
int lhip_download(lhip_ctx* ctx, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* m,
                  uint32_t* id);
/* Backend::compute_forces = build_tree + compute_forces_with(no remote trees). */
int lhip_compute_forces(lhip_ctx* ctx, const lhip_force_params* p, int potential);
/* Backend::kick: v += a dt (asynchronous). LHIP_LOGIC if the accelerations are stale (a
 * build_tree since the last force call, src/cuda_backend.cu:1567-1571). */
int lhip_kick(lhip_ctx* ctx, double dt);
/* Backend::drift: x += v dt (asynchronous); invalidates the tree and the potential. */
int lhip_drift(lhip_ctx* ctx, double dt);

/* ---- reductions (src/backend.hpp:106-110) ---- */

int lhip_diagnostics(lhip_ctx* ctx, lhip_diag* out);
/* Backend::exact_potential_energy: O(N^2) pair sum in double. */
int lhip_exact_potential_energy(lhip_ctx* ctx, double G, double eps, double* out);
int lhip_times(lhip_ctx* ctx, lhip_phase_times* out);

/* ---- distributed hooks (src/backend.hpp:114-140) ---- */

/* Backend::local_bounds: float bounding box of the current positions. */
int lhip_local_bounds(lhip_ctx* ctx, float lo[3], float hi[3]);
/* Backend::build_tree: sort + tree for the current positions (tree method; no-op for direct). */
int lhip_build_tree(lhip_ctx* ctx, const lhip_force_params* p);
/* Backend::compute_forces_with with an empty remote list (the single-rank case). LHIP_LOGIC for
 * the tree method without a current tree (src/cuda_backend.cu:1399-1400). */
int lhip_compute_forces_with(lhip_ctx* ctx, const lhip_force_params* p, int potential);

/* LetNode (src/backend.hpp:63-76): one node of a locally essential tree (LET), 60 bytes, the
 * backend-neutral wire format LETs travel in between ranks. Children of a node are contiguous at
 * first_child (-1: none), its bodies at body_first (-1: none) in the tree's body arrays. A node with
 * neither is terminal (accepted by any target in the box it was exported for). rcrit2 = 3.0e38:
 * always opened. */
typedef struct lhip_let_node {
    float cx, cy, cz, mass;
    float qxx, qxy, qxz, qyy, qyz, qzz;
    float rcrit2;
    int32_t first_child, child_count, body_first, body_count;
} lhip_let_node;

/* LetTree (src/backend.hpp:84-87) as plain arrays: n_nodes records, n_bodies bodies (SoA). */
typedef struct lhip_let {
    const lhip_let_node* nodes;
    uint64_t n_nodes;
    const float *bx, *by, *bz, *bm;
    uint64_t n_bodies;
} lhip_let;

/* Backend::compute_forces_with (src/backend.hpp:131-136) with remote LETs: the forces of the local
 * bodies plus every remote tree. Each walk group also walks every remote tree with the receiver
 * rules of src/backend.hpp:78-83 (bodies enter with no self term); the direct method walks them
 * per block of 128 consecutive bodies (src/cpu_backend.cpp:691-713). Costs and interaction counts
 * include the remote entries. The trees are validated before they reach the device: every body
 * range in bounds, children after their parent (as both reference exporters emit them) and in
 * bounds, at most 2^31 - 1 nodes and bodies in all, at most 33554431 bodies per node; LHIP_INVALID
 * otherwise. LHIP_RUNTIME if the remote trees are too deep for the walk's stack. Empty trees
 * (n_nodes = 0) are skipped. */
int lhip_compute_forces_remote(lhip_ctx* ctx, const lhip_force_params* p, int potential, const lhip_let* remote,
                               int n_remote);
/* Backend::export_let (src/backend.hpp:125-130): the part of the local tree a target anywhere in
 * the box [lo, hi] needs at p->theta, built on the device. Returns its sizes; lhip_export_let_copy
 * then fills the caller's arrays (n_nodes records, n_bodies entries per body array). Tree method:
 * needs a current tree (build_tree after the last upload or drift; LHIP_LOGIC otherwise). Direct
 * method: one always-open root holding every local body in the current order. n = 0 or a root of
 * zero mass: an empty LET. */
int lhip_export_let(lhip_ctx* ctx, const float lo[3], const float hi[3], const lhip_force_params* p,
                    uint64_t* n_nodes, uint64_t* n_bodies);
/* Copies the LET of the last lhip_export_let; LHIP_LOGIC if there is none. */
int lhip_export_let_copy(lhip_ctx* ctx, lhip_let_node* nodes, float* bx, float* by, float* bz, float* bm);
/* Backend::memory_need (src/backend.hpp:137): device and host bytes a force call on n bodies
 * reserves in the plain mode, buffer by buffer as the reserves size them, plus 64 MiB for allocator
 * rounding; lhip_memory_need_mode (below) for either mode. Tree node arrays are counted for n/4 + 4096
 * nodes, which smooth distributions stay below (a tree with more grows them). Not included: grown
 * node arrays, the key-restart buffers (~76 B/body, reserved when a build first restarts keys in
 * over-full cells), remote LETs and their exports (sized by the other ranks) and the differentiable
 * simulations. */
int lhip_memory_need(uint64_t n, int32_t method, uint64_t* device_bytes, uint64_t* host_bytes);
/* Backend::device_memory (src/backend.hpp:138-143): free device bytes counting every buffer this
 * context's force path holds (lhip_device_bytes_held), and the total. */
int lhip_device_memory(lhip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* Backend::body_costs: interactions per body of the last force call (self excluded), in download
 * order; LHIP_NOT_PROVIDED when there is none. */
int lhip_body_costs(lhip_ctx* ctx, float* cost);

/* ---- mixed-precision form (no reference counterpart: the reference's state is FP32,
 *      src/bodies.hpp:8-13) ----
 *
 * LHIP_MODE_MIXED keeps positions and velocities in FP64 on the device: FP64 kick and drift, Morton
 * keys quantised from the FP64 positions (the reference recipe, src/cpu_backend.cpp:396-426), FP64
 * multipole accumulation at every tree level, pair arithmetic in FP32 on coordinates relative to the
 * walk group's centre (tree) or on FP64 separations (direct), FP32 partial sums carried into FP64,
 * and FP64 diagnostics. LHIP_MODE_PLAIN is the reference's FP32 state (the default). Switching
 * converts the resident bodies, accelerations and potential and invalidates the tree.
 *
 * Walk options (both modes): group_max, the walk group size in [16, 128] (the reference's
 * kGroupMax = 128, src/cpu_backend.cpp:27); small_leaf, the largest accepted leaf summed body by
 * body, in [0, 16] (the reference's kSmallLeaf = 4, src/cpu_backend.cpp:29); quad_tau >= 0, the
 * error-budgeted far field: an accepted cell whose quadrupole term is below quad_tau times its
 * monopole term at the group box's distance (1.5 |Q|_F / M < quad_tau d^2) is evaluated as a point
 * mass (0 keeps every quadrupole, as the reference). NULL options select the mode's defaults:
 * plain {128, 4, 0} (the reference walk), mixed {128, 1, 0.05}. */
enum { LHIP_MODE_PLAIN = 0, LHIP_MODE_MIXED = 1 };
typedef struct lhip_walk_options {
    int32_t group_max;
    int32_t small_leaf;
    double quad_tau;
} lhip_walk_options;
int lhip_set_precision(lhip_ctx* ctx, int32_t mode, const lhip_walk_options* options);
int lhip_get_precision(const lhip_ctx* ctx, int32_t* mode, lhip_walk_options* options);
/* Backend::memory_need for the given mode (LHIP_MODE_PLAIN or LHIP_MODE_MIXED). */
int lhip_memory_need_mode(uint64_t n, int32_t method, int32_t mode, uint64_t* device_bytes, uint64_t* host_bytes);
/* FP64 initial conditions and state (the extra entry Bodies' FP32 fields cannot carry). Mixed
 * mode only (LHIP_LOGIC otherwise). Masses and ids as in lhip_upload. */
int lhip_upload_f64(lhip_ctx* ctx, uint64_t n, const double* x, const double* y, const double* z, const double* vx,
                    const double* vy, const double* vz, const float* m, const uint32_t* id);
int lhip_download_f64(lhip_ctx* ctx, double* x, double* y, double* z, double* vx, double* vy, double* vz, float* m,
                      uint32_t* id);
/* FP64 accelerations and potential of the last force call (mixed mode only). Any pointer may be
 * NULL. */
int lhip_accelerations_f64(lhip_ctx* ctx, double* ax, double* ay, double* az, double* pot);

/* ---- snapshots (host code; the reference's LTCSNAP3 files, src/snapshot.hpp:1-27) ----
 *
 * The reference's layout: a 256-byte little-endian header, then planar arrays on 4096-byte
 * boundaries (f32 x, y, z, vx, vy, vz, m; u32 id). With FP64 state (the mixed-precision form) the
 * file also carries f64 x, y, z, vx, vy, vz arrays after those, marked in spare header bytes (no
 * reference counterpart: its snapshots are f32 only). The f32 arrays then hold the FP64 values
 * rounded to nearest, so the reference reads such a file as an ordinary snapshot, and these
 * readers restore the FP64 state bit for bit. Writes go to <path>.tmp, are fsync'ed and renamed
 * into place. Failures return LHIP_RUNTIME (the reference: std::runtime_error naming the path),
 * with the message in lhip_snapshot_error(); LHIP_INVALID for bad arguments. */

/* SnapshotHeader (src/snapshot.hpp:9-20). scenario: at most 31 bytes are stored. fp64 (read only):
 * 1 when the file holds the FP64 state arrays. */
typedef struct lhip_snapshot_header {
    uint64_t n, step;
    double time, G, eps, dt;
    uint64_t seed;
    char scenario[32];
    uint32_t pieces;
    uint32_t fp64;
} lhip_snapshot_header;

/* Message of the last failed snapshot call on this thread. */
const char* lhip_snapshot_error(void);
/* read_snapshot_header (src/snapshot.hpp:26): the header, validated against the file size. */
int lhip_snapshot_read_header(const char* path, lhip_snapshot_header* h);
/* write_snapshot (src/snapshot.hpp:23) of n bodies from host arrays; n is the arrays' length
 * (h->n is not used). xd..vzd: the FP64 state, or all NULL for an f32-only file. */
int lhip_snapshot_save(const char* path, const lhip_snapshot_header* h, uint64_t n, const float* x, const float* y,
                       const float* z, const float* vx, const float* vy, const float* vz, const float* m,
                       const uint32_t* id, const double* xd, const double* yd, const double* zd, const double* vxd,
                       const double* vyd, const double* vzd);
/* read_snapshot (src/snapshot.hpp:24) into host arrays of `capacity` entries (LHIP_INVALID when
 * the file holds more bodies). xd..vzd are filled when the file has FP64 arrays (NULL: skipped). */
int lhip_snapshot_load(const char* path, lhip_snapshot_header* h, uint64_t capacity, float* x, float* y, float* z,
                       float* vx, float* vy, float* vz, float* m, uint32_t* id, double* xd, double* yd, double* zd,
                       double* vxd, double* vyd, double* vzd);
/* The context's bodies (download order) to a snapshot; in the mixed-precision mode with the FP64
 * state. h->n is not used. */
int lhip_snapshot_write(lhip_ctx* ctx, const char* path, const lhip_snapshot_header* h);
/* A snapshot into the context (as upload): the FP64 state when the file has it and the context is
 * in the mixed-precision mode, else the f32 arrays (promoted in the mixed mode). *h: its header. */
int lhip_snapshot_restore(lhip_ctx* ctx, const char* path, lhip_snapshot_header* h);

/* ---- initial conditions (host code) ----
 *
 * The reference's Plummer scenario (src/scenarios.cpp:94-132; keys plummer.a, plummer.cut,
 * plummer.mass, seed): float outputs are bit-exact with the reference generator; the _f64 form
 * returns the same values before the final rounding. Any output pointer may be NULL. */
int lhip_scenario_plummer(uint64_t n, uint64_t seed, double G, double M, double a, double cut, float* x, float* y,
                          float* z, float* vx, float* vy, float* vz, float* m, uint32_t* id);
int lhip_scenario_plummer_f64(uint64_t n, uint64_t seed, double G, double M, double a, double cut, double* x,
                              double* y, double* z, double* vx, double* vy, double* vz, double* m, uint32_t* id);
/* Rotating exponential disk (scale rd, cut at rmax, total stellar mass `mass`) with a sech^2
 * vertical profile of scale z0 around a central point mass `core` (body 0; nstars + 1 bodies),
 * circular speed from the enclosed mass with Gaussian dispersion sigma * vc per component, spin
 * +1 or -1 about the disk axis; the disk is then rotated by incl about the x axis and moved to
 * (pos, vel). No reference counterpart: the reference disk (src/scenarios.cpp:173-236) is Gaussian
 * in z; radii are drawn the same way (src/scenarios.cpp:197-205). */
typedef struct lhip_disk_params {
    uint64_t nstars, seed;
    double G, core, mass, rd, rmax, z0, sigma, incl, spin;
    double pos[3], vel[3];
} lhip_disk_params;
int lhip_scenario_disk_sech2(const lhip_disk_params* p, int32_t com_frame, float* x, float* y, float* z, float* vx,
                             float* vy, float* vz, float* m, uint32_t* id);
int lhip_scenario_disk_sech2_f64(const lhip_disk_params* p, int32_t com_frame, double* x, double* y, double* z,
                                 double* vx, double* vy, double* vz, double* m, uint32_t* id);

/* Two-disk encounter (no reference counterpart: the reference's `collision` tilts one disk and
 * aims by an escape-speed fraction, src/scenarios.cpp:248-271). Two sech^2 disks as above (seeds
 * seed and seed + 1, spin +1), tilted by incl1 and incl2 about the x axis, their cores on a
 * parabolic orbit in the xy plane with pericentre `pericentre`, `separation` apart and
 * approaching, in the cores' centre-of-mass frame. Disk 1 is bodies [0, nstars + 1) with its core
 * first, disk 2 follows. Outputs in FP64: positions and velocities as n x 3 (AoS), masses, ids. */
typedef struct lhip_collision_params {
    uint64_t nstars, seed;
    double G, core, mass, rd, rmax, z0, sigma;
    double pericentre, separation, incl1, incl2;
} lhip_collision_params;
int lhip_scenario_collision(const lhip_collision_params* p, double* x, double* v, double* m, uint32_t* id);
/* Vector-Jacobian product of the encounter: given dL/dx and dL/dv (n x 3 each), returns
 * out = (dL/dpericentre, dL/dincl1, dL/dincl2). */
int lhip_scenario_collision_vjp(const lhip_collision_params* p, const double* gx, const double* gv, double out[3]);

/* ---- inspection (no reference counterpart; for parity tests and tools) ---- */

/* Device bytes this context's force path holds (the buffers device_memory counts as free). */
int lhip_device_bytes_held(lhip_ctx* ctx, uint64_t* held);

/* Accelerations and potential of the last force call, in the backend's current order (rounded to
 * float in the mixed mode). */
int lhip_accelerations(lhip_ctx* ctx, float* ax, float* ay, float* az, float* pot);
/* FP64 direct-sum accelerations and potentials (a = G sum m_j d/s^3, phi = -G sum m_j/s, the
 * target skipped by index) of the bodies at the given backend-order indices, against every body
 * at the current positions (FP32 or FP64 state), in double pair arithmetic on the device: the
 * FP64 direct oracle for force errors at sizes a CPU sum cannot reach. Any output may be NULL. */
int lhip_direct_f64_targets(lhip_ctx* ctx, uint64_t nt, const uint32_t* targets, double G, double eps, double* ax,
                            double* ay, double* az, double* pot);
/* Morton keys of the current tree in body order (n entries); a body of a restarted cell has the key
 * of its deepest frame, as the reference's key array. */
int lhip_tree_keys(lhip_ctx* ctx, uint64_t* keys);
/* Node count, walk-group count and the number of bodies in over-full finest cells: cells whose
 * bodies coincide (exact leaves, as the reference keeps them), or restarted beyond the fourth key
 * frame. Over-full cells whose bodies differ restart their keys (src/cpu_backend.cpp:501-544). */
int lhip_tree_counts(lhip_ctx* ctx, uint64_t* nodes, uint64_t* groups, uint64_t* overfull);
/* Node records in level order (children of a node contiguous): link = (lo, hi, info, first child)
 * per node with info = level | leaf << 7 | children << 8 (level = depth from the root, up to 84 with
 * key restarts) and first = -1 for a leaf; prefix = key prefix within the node's key frame; mom = 16
 * floats per node (cx, cy, cz, m, qxx, qxy, qxz, qyy, qyz, qzz, rcrit2, qrel, gx, gy, gz, half). Any
 * pointer may be NULL. */
int lhip_tree_nodes(lhip_ctx* ctx, int32_t* link, uint64_t* prefix, float* mom);
/* Walk-group starts (groups + 1 entries; the last is n). */
int lhip_tree_groups(lhip_ctx* ctx, uint32_t* gstart);

/* ---- allocation-free direct sum (no counterpart in the reference) ----
 *
 * For callers that cannot allocate device memory (code under HIP stream capture, tight serving
 * loops, harnesses that forbid hipMalloc): the direct sum of lhip_compute_forces with method
 * LHIP_METHOD_DIRECT (its kernels and split count, so the results are the same bit for bit) on
 * caller-owned device arrays, enqueued on `stream` (a hipStream_t; NULL: the null stream) of the
 * current device without synchronising. posm: n float4 records (x, y, z, m), 16-byte aligned;
 * ax, ay, az (and pot when potential != 0): n floats each. a_i = G sum_j m_j d_ij / s^3 and
 * phi_i = -G sum_j m_j / s with s^2 = d^2 + eps^2, the target skipped by index. The partial sums of
 * the split source range live in a 64 MiB device array the library holds per device for the
 * process's lifetime (4,194,304 float4; every n below 2^31 - 1 fits). Nothing is allocated or freed.
 * Returns LHIP_INVALID for n >= 2^31 - 1, NULL or misaligned arrays; LHIP_RUNTIME for launch
 * failures. Calls that overlap on one device share that array and race: serialise them (one
 * stream, or events). */
int lhip_direct_forces_noalloc(const float* posm, uint64_t n, double G, double eps, int potential, float* ax,
                               float* ay, float* az, float* pot, void* stream);

/* ---- differentiable simulation (no reference counterpart: the reference has no differentiation) ----
 *
 * FP64 state (positions, velocities, masses) integrated with the reference's KDK leapfrog
 * (src/simulation.cpp:406-438) for `steps` steps of dt, direct-sum forces (mixed precision: FP32 pair
 * arithmetic on FP64 separations, FP64 accumulation; or all-FP64). The forward keeps (x, v)
 * checkpoints every `checkpoint_every` steps on the device; the backward recomputes each segment from
 * its checkpoint and runs the adjoint sweep. Arrays are n x 3 (row-major) or n, host memory. */
enum { LHIP_PRECISION_MIXED = 0, LHIP_PRECISION_F64 = 1 };
/* Tree method: LHIP_TREE_BUILD builds a tree per force call; LHIP_TREE_RECORD also keeps each
 * call's topology (sorted order and positions); LHIP_TREE_REPLAY runs the forward from new initial
 * conditions on the recorded topology (the frozen-topology forward whose finite differences the
 * adjoint matches) and keeps the recorded forward's checkpoints for the backward. */
enum { LHIP_TREE_BUILD = 0, LHIP_TREE_RECORD = 1, LHIP_TREE_REPLAY = 2 };

typedef struct lhip_sim_params {
    double G, eps, dt, theta;
    int32_t steps;
    int32_t checkpoint_every;
    int32_t method;     /* LHIP_METHOD_DIRECT, or LHIP_METHOD_TREE (mixed-precision walk, theta) */
    int32_t precision;  /* direct: LHIP_PRECISION_MIXED or LHIP_PRECISION_F64; tree: mixed */
    int32_t tree_mode;  /* tree: LHIP_TREE_BUILD, LHIP_TREE_RECORD or LHIP_TREE_REPLAY */
    int32_t reserved;
    lhip_walk_options walk;  /* tree: walk options; group_max 0 selects the mixed mode's defaults */
} lhip_sim_params;

/* Runs the forward from (x0, v0, m); writes the final state (xT, vT may be NULL). */
int lhip_sim_forward(lhip_ctx* ctx, const lhip_sim_params* p, uint64_t n, const double* x0, const double* v0,
                     const double* m, double* xT, double* vT);
/* Given dL/dxT and dL/dvT, writes dL/dx0, dL/dv0, dL/dm (any output may be NULL). LHIP_LOGIC without
 * a forward. The tree method runs on the context's bodies (the forward replaces them), keeps its
 * checkpoints per body id, and gives no mass gradient (dL/dm must be NULL). */
int lhip_sim_backward(lhip_ctx* ctx, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gm);
/* Generation of the forward whose adjoint lhip_sim_backward runs (0 before the first forward). Every
 * lhip_sim_forward other than a tree replay starts a new generation: the checkpoints of the previous
 * one are gone. A replay keeps the recorded forward's checkpoints, parameters and generation. */
int lhip_sim_generation(const lhip_ctx* ctx, uint64_t* generation);
/* lhip_sim_backward for the forward of the given generation: LHIP_LOGIC if another forward has run
 * on this context since (a caller holding several forwards, e.g. an autograd graph, gets an error
 * instead of another trajectory's gradients). generation 0 skips the check (lhip_sim_backward). */
int lhip_sim_backward_gen(lhip_ctx* ctx, uint64_t generation, const double* gxT, const double* gvT, double* gx0,
                          double* gv0, double* gm);
/* Device time of the last forward and backward (HIP events on the context's stream). */
int lhip_sim_times(lhip_ctx* ctx, double* forward_ms, double* backward_ms);

#ifdef __cplusplus
}
#endif
#endif /* LETTUCE_HIP_H */
