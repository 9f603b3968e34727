/*
 * sevc.h -- C ABI of the MI355X-native vertex-model time step.
 *
 * Drop-in boundary for the body of the reference's main loop
 * (src/main.cu:120-394).  The reference keeps its state in three device
 * arrays of pointer-linked structs (geometry.h:39-86) built by
 * host_to_device (translation.h:140-155) from the loader output
 * (io.h:66-264), and drives it from struct options (options.h:15-46).
 * This ABI keeps the same options and the same vertex / boundary / grain
 * arrays -- same slots, same ids (id == slot), same per-vertex boundary
 * and grain order, same per-grain vertex order -- laid out as
 * structure-of-arrays with int32 indices instead of pointers.
 *
 * Errors are returned as negative codes (never exit(), unlike
 * HERR/CERR/CNULL in utils.h:17-44); sevc_last_error() gives the text.
 * One state per host thread; distinct states may be used concurrently.
 */
#ifndef SEVC_H
#define SEVC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- */
/* Status codes                                                       */
/* ---------------------------------------------------------------- */
#define SEVC_OK              0
#define SEVC_E_INVALID_ARG  -1   /* bad argument / option                      */
#define SEVC_E_IO           -2   /* file missing or unreadable (io.h:78-81)    */
#define SEVC_E_TOPOLOGY     -3   /* illegal grains / count mismatch (io.h:155-186) */
#define SEVC_E_HIP          -4   /* HIP runtime error                          */
#define SEVC_E_NOMEM        -5   /* allocation failed                          */
#define SEVC_E_GRAIN_FULL   -6   /* grain list would exceed max_vrt_per_grn    */
#define SEVC_E_STATE        -7   /* inconsistent topology met during an event  */
#define SEVC_E_NO_DEVICE    -8   /* no HIP device                              */
#define SEVC_E_CHECK        -9   /* check_data failed with debug_lvl >= 1      */

/* Reference fixed-size limits (geometry.h:9-11). */
#define SEVC_REF_MAX_VRT_PER_GRN 100
#define SEVC_MAX_T_EXT 100000.0
#define SEVC_GRAD_EPS 1e-8          /* calculus.h:7 */

/* Nucleation models. */
#define SEVC_NUCL_MC   0   /* reference: one Metropolis site per gap (nucleation2.cu:239-443) */
#define SEVC_NUCL_RATE 1   /* extension: per-grain probability rate*SE per step              */

/* Velocity models. */
#define SEVC_VEL_FD       0   /* reference forward difference (calculus.h:19-73,143-154) */
#define SEVC_VEL_ANALYTIC 1   /* extension: exact -grad E (smooth, used by the adjoint)  */

#define SEVC_FNAME 256

/* ---------------------------------------------------------------- */
/* Options: every key of options.h:15-46 plus extensions.            */
/* ---------------------------------------------------------------- */
typedef struct sevc_options {
    /* --- reference keys (options.h:15-46, defaults options.h:71-85) --- */
    double dt;                 /* fixed dt, or the upper bound when adaptive_dt */
    double dtau;               /* nucleation period; nucleation_gap = (int)(dtau/dt) */
    int32_t nucleation_gap;    /* derived (options.h:172)                      */
    double tmax;               /* -1 = unset                                    */
    int32_t nmax;              /* -1 = unset                                    */
    int32_t snap_gap;
    double leave_percent;      /* stop when n_grains <= floor(leave_percent*n_grains0) */
    int32_t n_grains;          /* expected initial grain count (io.h:182)       */
    char vertex_fname[SEVC_FNAME];
    char boundary_fname[SEVC_FNAME];
    char orientation_fname[SEVC_FNAME];
    char SE_fname[SEVC_FNAME];
    double GB_eps;
    double GB_scaling;
    double SE_eps;
    double SE_min;
    double SE_max;
    double nucleation_scaling;
    double domain_bound;
    int32_t do_nucleation;
    double MC_k;
    int32_t n_blocks;          /* accepted and ignored: launch shapes are the library's */
    int32_t n_threads;
    int32_t debug_lvl;
    char output_folder[SEVC_FNAME];

    /* --- extensions (absent from the reference; defaults reproduce it) --- */
    double mobility;           /* mu, multiplies -dE/eps (1.0)                  */
    int32_t adaptive_dt;       /* 0: fixed dt (reference)                       */
    double dt_safety;          /* dt_n = clamp(dt_safety * tau_min, dt_min, dt) */
    double dt_min;
    double flip_threshold;     /* edges of length in (0, threshold) that shrink also flip (0 = off) */
    int32_t nucleation_mode;   /* SEVC_NUCL_MC / SEVC_NUCL_RATE                 */
    double nucleation_rate;    /* RATE: p_g = nucleation_rate * SE_g per step   */
    uint64_t seed;             /* Philox key for nucleation                     */
    int32_t max_vrt_per_grn;   /* grain list stride (capacity per grain)        */
    double extra_capacity;     /* spare disabled slots appended, fraction of size */
    int32_t velocity_model;    /* SEVC_VEL_FD / SEVC_VEL_ANALYTIC               */
    int32_t events;            /* 0: topology-fixed (no T1/T2/nucleation)       */
    int32_t nucleation_gate;   /* MC mode opens once enabled grains <= this count;
                                  -1: 5*n_grains0/6 (main.cu:137-139)            */
} sevc_options;

/* ---------------------------------------------------------------- */
/* Host view of one microstructure (structure of arrays).            */
/* Slot i holds the entity with id i, as in the reference.           */
/* ---------------------------------------------------------------- */
typedef struct sevc_mesh {
    int32_t n_vertices;        /* slots, enabled or not */
    int32_t n_boundaries;
    int32_t n_grains;
    int32_t max_vrt_per_grn;   /* stride of g_vtx */
    double domain_bound;

    /* vertices (geometry.h:39-59) */
    double*  pos;              /* 2*n_vertices, x y interleaved */
    double*  vel;              /* 2*n_vertices                  */
    int32_t* v_bnd;            /* 3*n_vertices, order as vertex.boundaries[3] */
    int32_t* v_grn;            /* 3*n_vertices, order as vertex.grains[3]     */
    uint8_t* v_enabled;
    double*  v_energy;         /* optional (NULL): vertex energy diagnostic    */
    double*  v_vel_parts;      /* optional (NULL): 4*n_vertices vel_bnd_x vel_grn_x vel_bnd_y vel_grn_y */

    /* boundaries (geometry.h:61-72) */
    int32_t* b_vtx;            /* 2*n_boundaries: ini, end */
    double*  b_energy;
    double*  b_arclength;
    double*  b_t_ext;
    uint8_t* b_enabled;
    uint8_t* b_candidate;

    /* grains (geometry.h:74-86) */
    int32_t* g_vtx;            /* n_grains*max_vrt_per_grn, -1 padded */
    int32_t* g_vlen;
    double*  g_area;
    double*  g_dAdt;           /* optional (NULL) */
    double*  g_orientation;
    double*  g_SE;
    uint8_t* g_enabled;
    uint8_t* g_type;           /* 0 COMMON, 1 NUCLEATED (geometry.h:19) */
} sevc_mesh;

/* Per-call statistics of sevc_step (the printf lines of main.cu:122,
 * 231, 324-325 become fields). */
typedef struct sevc_step_stats {
    int32_t steps_done;        /* steps executed by this call               */
    int32_t n;                 /* step counter after the call (main.cu:388) */
    double  t;                 /* simulation time after the call            */
    double  dt_last;           /* dt of the last executed step              */
    int32_t n_grains;          /* enabled grains (main.cu:317)              */
    int32_t n_vertices;        /* enabled vertices                          */
    double  nucleated_area_pct;/* main.cu:321                               */
    int32_t done;              /* stop rule reached                         */
    int32_t status;            /* SEVC_OK or the first device-side error    */
    int64_t n_flips;           /* T1 events during the call                 */
    int64_t n_removals;        /* T2 events                                 */
    int64_t n_two_grain;       /* two-triangle collapses                    */
    int64_t n_nucleations;
    int64_t n_vertex_steps;    /* sum over steps of enabled vertices        */
} sevc_step_stats;

typedef struct sevc_state sevc_state;

/* ---------------------------------------------------------------- */
/* Options                                                           */
/* ---------------------------------------------------------------- */

/* Defaults of load_config_file (options.h:71-85) plus extension defaults. */
void sevc_options_default(sevc_options* o);

/* Parse an .ini file with the semantics of load_config_file
 * (options.h:69-206): the defaults first, then the file's keys;
 * '#'/'[' lines skipped, exact "key=value" match,
 * do_nucleation true only for "true", nucleation_gap = (int)(dtau/dt),
 * tmax/nmax defaulting (options.h:173-181).  Unknown keys are ignored;
 * the extension keys are accepted too. */
int sevc_options_from_ini(const char* path, sevc_options* o);

/* ---------------------------------------------------------------- */
/* Host meshes                                                       */
/* ---------------------------------------------------------------- */

/* Allocate every array of a host mesh (calloc, -1 in index arrays). */
int  sevc_mesh_alloc(sevc_mesh* m, int32_t n_vertices, int32_t n_boundaries,
                     int32_t n_grains, int32_t max_vrt_per_grn);
void sevc_mesh_free(sevc_mesh* m);

/* Build the initial mesh from the reference's input files with the ids
 * and orderings of load_from_files (io.h:66-264): vertices in file order,
 * boundaries in file order, grains by the face walk (io.h:137-180), then
 * counter-clockwise boundary order (io.h:188-192), id-sorted vertex grain
 * lists (io.h:196-200), orientations * 2pi (io.h:214-218) and the SE
 * transform (io.h:223-248).  Positions are scaled by domain_bound. */
int sevc_mesh_load_files(const sevc_options* o, sevc_mesh* out);

/* Same, from arrays already in memory (positions in [0,1)^2, ridges as
 * 2*n_ridges vertex pairs, one orientation and one SE value per grain). */
int sevc_mesh_build(const sevc_options* o, int32_t n_vertices, const double* unit_pos,
                    int32_t n_ridges, const int32_t* ridges, int32_t n_grains,
                    const double* orientations01, const double* se_raw, sevc_mesh* out);

/* Periodic Voronoi tessellation of n_seeds uniform seeds on the unit
 * torus (the role of voronoi/voronoi.py:12-109), written as the arrays
 * sevc_mesh_build takes.  Seeds are ordered along a space-filling curve so
 * that neighbouring grains get neighbouring ids.  Arrays are malloc'd;
 * free them with sevc_free. */
int sevc_voronoi_generate(int32_t n_seeds, uint64_t seed,
                          int32_t* n_vertices, double** unit_pos,
                          int32_t* n_ridges, int32_t** ridges);
void sevc_free(void* p);

/* Host port of check_data (utils.h:85-336).  Returns the number of
 * violations found (0 = valid) and writes the summed grain area.  An id
 * outside [0, n) counts once for the entity holding it (the checks that
 * would follow it are skipped); a grain list is read up to
 * min(vlen, max_vrt_per_grn) entries. */
int sevc_check_mesh(const sevc_mesh* m, double* total_area);

/* The same conditions counted on the GPU (check_data, utils.h:85-336, the
 * checks main.cu:185-196,237-248,269-280,297-308 run with debug_lvl >= 1),
 * one lane per vertex, edge and grain.  sevc_check_state checks the live
 * device state of every member in place (nothing is copied to the host)
 * and writes each member's violation count to bad[0..B-1].
 * sevc_check_mesh_device copies a host mesh (valid or not) to the device,
 * checks it there and returns its count in *bad.  The summed grain area of
 * check_data is a printed warning in the reference and is not formed. */
int sevc_check_state(sevc_state* s, int32_t* bad);
int sevc_check_mesh_device(const sevc_mesh* m, int32_t* bad);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): the count of
 * sevc_check_mesh_device for n_meshes meshes whose topology arrays (v_bnd,
 * v_grn, v_enabled, b_vtx, b_enabled, g_vtx, g_vlen, g_enabled) are already
 * in device memory; the other array pointers are not read.  bad[m] (device
 * memory, n_meshes ints) receives mesh m's violation count.  The call only
 * enqueues kernels on `stream` (a hipStream_t, NULL for the default
 * stream) and returns without waiting; it calls no hipMalloc, memset or
 * copy.  Its scratch (a parameter record and a counter per mesh) is a
 * static device array of 11.5 KB per GPU, held until the process ends:
 * at most 64 meshes per call (more: SEVC_E_INVALID_ARG, the reason in
 * sevc_last_error(NULL)), and calls that overlap on one GPU race. */
int sevc_check_meshes_device_noalloc(const sevc_mesh* meshes, int32_t n_meshes, int32_t* bad, void* stream);

/* Snapshot in the format of export_data (io.h:280-377) to
 * <folder>/<index as %06d>.txt.  The format holds 100 vertices per grain
 * (MAX_VRT_PER_GRN, geometry.h:9): a grain with a longer list returns
 * SEVC_E_GRAIN_FULL and no file is written. */
int sevc_export_snapshot(const sevc_mesh* m, const sevc_options* o, const char* folder,
                         int32_t index, double time);

/* Read a snapshot written by sevc_export_snapshot back into a newly
 * allocated mesh (grain list stride max_vrt_per_grn), for resuming a run;
 * *time receives its TIME.  Values are the snapshot's %.16f text, so
 * positions come back rounded to 16 decimals.  The format lists the other
 * end of each vertex boundary, not its id, so vertex.boundaries of enabled
 * vertices is rebuilt from the enabled boundaries in that neighbour order;
 * disabled vertices get none. */
int sevc_load_snapshot(const char* path, int32_t max_vrt_per_grn, sevc_mesh* out, double* time);

/* ---------------------------------------------------------------- */
/* Device state                                                      */
/* ---------------------------------------------------------------- */

/* Upload one mesh (host_to_device, translation.h:140-155; scratch of
 * main.cu:68-107).  The library copies; the caller keeps ownership. */
int sevc_create(const sevc_options* o, const sevc_mesh* m, sevc_state** out);

/* Upload n_members independent meshes as one batched state: every kernel
 * of a step covers all members in one launch.  velocity_model and
 * nucleation_mode must be known values and domain_bound positive
 * (SEVC_E_INVALID_ARG, checked before any device call). */
int sevc_batch_create(const sevc_options* opts, const sevc_mesh* meshes, int32_t n_members,
                      sevc_state** out);

/* Run up to n_steps steps of main.cu:124-389 on every active member;
 * a member stops at its own stop rule (main.cu:120,391).  stats may be
 * NULL; for batches it receives member 0's statistics.  An event that meets
 * an inconsistent topology or a full grain list stops only its own member
 * (status SEVC_E_STATE / SEVC_E_GRAIN_FULL in sevc_member_stats); the call
 * in which a member stops so returns that code and names the member, the
 * other members keep stepping.  A one-member state returns the code again
 * on every later call. */
int sevc_step(sevc_state* s, int32_t n_steps, sevc_step_stats* stats);

/* Same, capturing one step into a HIP graph on first use and replaying
 * it (no host synchronisation inside the steps). */
int sevc_step_graph(sevc_state* s, int32_t n_steps, sevc_step_stats* stats);

/* Per-member statistics after the last call. */
int sevc_member_stats(sevc_state* s, int32_t member, sevc_step_stats* stats);

/* Slot counts of a member (they include spare capacity). */
int sevc_member_sizes(const sevc_state* s, int32_t member, int32_t* n_vertices,
                      int32_t* n_boundaries, int32_t* n_grains, int32_t* max_vrt_per_grn);
int32_t sevc_n_members(const sevc_state* s);

/* Copy a member back to host arrays sized by sevc_member_sizes
 * (device_to_host, translation.h:184-199, without freeing the device copy). */
int sevc_get_mesh(sevc_state* s, int32_t member, sevc_mesh* out);

/* Pick the FD velocity's path and tile size for this state by timing each
 * candidate on the current positions (extension; no counterpart in the
 * reference): path 0, two passes (grain tiles of 16..64 grains, then one
 * lane per vertex); path 1, one pass over tiles of 32 or 64 vertex slots;
 * path 2, two passes with one lane per grain, then per vertex (paths 1 and
 * 2 when the list stride is a multiple of 4, at least 8).  Results do
 * not depend on the choice, and the state is left unchanged (velocities
 * and grain areas written by the timed launches are put back).
 * tile_grains (may be NULL) receives the chosen tile size: grains per tile
 * for path 0, vertices per tile for path 1. */
int sevc_autotune(sevc_state* s, int32_t* tile_grains);

/* The velocity path and tile size in use (see sevc_autotune), and a way to
 * set them (path 0: tile 16..64 grains; path 1: tile 32 or 64 vertices;
 * path 2: tile ignored, reported as 0).
 * A new state uses path 0 with 64 grains per tile. */
int sevc_velocity_config(const sevc_state* s, int32_t* path, int32_t* tile);

/* How the vote rounds (polling_system, topological.h:549-578) and the
 * conflict check run, by the length of the step's candidate list: lists up
 * to max_fast (0..2048, default 2048) with every candidate's mapped
 * neighbours in registers, lists up to max_lds (0..16384, default 16384)
 * with the flags and extinction times in LDS, longer lists by the general
 * rounds over global memory.  Results do not depend on it (extension; a
 * tuning and test knob). */
int sevc_set_poll_limit(sevc_state* s, int32_t max_fast, int32_t max_lds);

/* The nucleation free-slot search (the first free grain, boundary and
 * vertex slots in index order, nucleation2.cu:33-72) reads each member's
 * per-block free counts (256 slots a block) in rounds: by default 4096,
 * 12288 and 8192 blocks (one round for members up to about a million
 * grains).  blocks = 1..1024 makes each round read that many blocks of each
 * kind, so that small states search in several rounds; 0 restores the
 * default.  Results do not depend on it (extension; a test knob). */
int sevc_set_search_round(sevc_state* s, int32_t blocks);
int sevc_set_velocity_config(sevc_state* s, int32_t path, int32_t tile);

/* Copy only the vertex positions of a member (2 per vertex slot). */
int sevc_get_positions(sevc_state* s, int32_t member, double* pos);

/* Replace the positions / stored energies of a member (adjoint driver,
 * parameter sweeps).  NULL arrays are left unchanged. */
int sevc_set_member_params(sevc_state* s, int32_t member, const double* pos,
                           const double* g_SE, double gb_scaling, double mobility);

/* Replace member m's whole mesh (positions, topology, stored energies,
 * orientations, flags) by a host mesh of the member's slot counts and
 * stride, e.g. one taken with sevc_get_mesh earlier: a checkpoint restore
 * (extension; the reference has no resume, io.h:280-377 only writes).  The
 * mesh must pass check_data (SEVC_E_TOPOLOGY otherwise).  Velocities and
 * t_ext are taken when given; the member's step counter and time go on. */
int sevc_set_member_mesh(sevc_state* s, int32_t member, const sevc_mesh* mesh);

/* Device checkpoints of the whole state (every member's mesh, scratch and
 * loop state; not the parameters of sevc_set_member_params): slot 0..63,
 * kept in device memory until sevc_destroy.  sevc_restore puts the state
 * back as it was at sevc_checkpoint, so that a later step repeats exactly
 * (extension; e.g. the windows of a calibration horizon). */
int sevc_checkpoint(sevc_state* s, int32_t slot);
int sevc_restore(sevc_state* s, int32_t slot);

/* Gradients of one member from sevc_adjoint_window (extension). */
typedef struct sevc_grads {
    double* d_pos;           /* 2 per vertex slot: dL/dx at the window's start (NULL: not wanted) */
    double* d_SE;            /* 1 per grain slot: dL/dSE_g (NULL: not wanted) */
    double  d_gb_scaling;    /* dL/dGB_scaling */
    double  d_mobility;      /* dL/dmobility */
} sevc_grads;

/* Forward over a topology-fixed window (extension; no counterpart in the
 * reference, whose step is calculus.h:143-154 + :391-404 inside
 * main.cu:124-389): `steps` steps of x <- dom_mod(x + dt*mobility*F(x)),
 * F = -grad E the analytic force, with T1/T2/nucleation off and dt the
 * fixed opts.dt, on every member.  Positions are checkpointed every
 * ckpt_every steps (<= 0: every step is kept, no recomputation later).
 * The state ends at the window's end; the next sevc_step or
 * sevc_set_member_params invalidates the window. */
int sevc_window_forward(sevc_state* s, int32_t steps, int32_t ckpt_every);

/* Reverse pass over the last window: dL_dpos[m] = dL/dx at the window's
 * end for member m (2 per vertex slot, host memory), or dL_dpos == NULL to
 * use the device seed left by sevc_area_histogram; fills grads[m].
 * Segments are recomputed from their checkpoints; results are
 * deterministic (fixed-order sums). */
int sevc_adjoint_window(sevc_state* s, const double* const* dL_dpos, sevc_grads* grads);

/* Smooth histogram of grain areas and its loss, on the device, at the
 * current positions (extension, for calibration with the adjoint window;
 * no counterpart in the reference).  Member m with ranges[3m..3m+2] =
 * (a_min, a_max, width) and n_bins bins of centres
 * c_k = a_min + (k + 1/2)(a_max - a_min)/n_bins:
 *   h_k = (1/G) sum_g exp(-((|A_g| - c_k)/width)^2 / 2)
 * over its enabled grains with >= 3 vertices, and
 *   loss[m] = sum_k (h_k - targets[m*n_bins + k])^2.
 * hist (n_bins per member), loss and dL_dpos (per-member host arrays of
 * 2 per vertex slot) may be NULL.  With seed_adjoint != 0, dL/dx stays on
 * the device as the seed of the next sevc_adjoint_window call made with
 * dL_dpos == NULL.  n_bins <= 64. */
int sevc_area_histogram(sevc_state* s, int32_t n_bins, const double* ranges, const double* targets,
                        double* hist, double* loss, double* const* dL_dpos, int32_t seed_adjoint);

/* Compute the diagnostics of the next step (vertex energies,
 * dA/dt and velocity parts, calculus.h:123-133,243-253,62-69). */
int sevc_request_diagnostics(sevc_state* s);

/* Device time (ms) of the kernels of the last call, by name; returns the
 * number of entries written.  Used by the benchmark for roofline figures. */
int sevc_kernel_times(sevc_state* s, int32_t max_entries, char (*names)[48], double* ms,
                      int64_t* launches);
/* on: 0 off, 1 every kernel, 2 only the velocity ("vel_fused" for the
 * one-pass path, "vel_grain" and "velocity" for the two passes),
 * "translate" and "text". */
int sevc_enable_kernel_timing(sevc_state* s, int32_t on);
/* With timing on, sevc_step times the kernels of one step in `every`
 * (default 1; counted from the last sevc_enable_kernel_timing or this
 * call), so that the events disturb the timed run less. */
int sevc_set_timing_sample(sevc_state* s, int32_t every);

/* Device self-test of the shared arithmetic: evaluates op on n inputs on
 * the GPU (0 atan2(y,x), 1 sin(x), 2 cos(x), 3 exp(x), 4 x/y, 5 sqrt(x),
 * 6 |(x,y)| as geometry.h:217, 7 wrap_dist(x,y,D=1), 8 Philox word 0 of
 * counter (x,y,0,0) key (1,2)).  Lets tests compare device and host bits. */
int sevc_selftest_math(int32_t op, int32_t n, const double* x, const double* y, double* out);

void sevc_destroy(sevc_state* s);
/* Message of the last failure of state s; with s == NULL, of the last
 * failed call on this thread that takes no state. */
const char* sevc_last_error(const sevc_state* s);
const char* sevc_strerror(int code);
int sevc_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* SEVC_H */
