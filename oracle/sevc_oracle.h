/*
 * sevc_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial FP64 CPU restatement of the reference's vertex-model time step
 * (src/main.cu:120-394 and the kernels it calls).  It is the checker for
 * the HIP implementation: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product links
 * or calls it.
 *
 * The state uses the sevc_mesh layout of include/sevc.h so that a GPU
 * state and an oracle state can be compared slot by slot.
 */
#ifndef SEVC_ORACLE_H
#define SEVC_ORACLE_H

#include "../include/sevc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* Deep copy of an initial mesh. */
orc* orc_create(const sevc_options* o, const sevc_mesh* m);
void orc_destroy(orc* s);

/* Run up to n_steps steps of main.cu:124-389. */
int orc_step(orc* s, int32_t n_steps, sevc_step_stats* st);

/* Copy the current state into caller arrays sized like the input mesh. */
int orc_get_mesh(const orc* s, sevc_mesh* out);

/* Compute the per-step quantities of main.cu:124-134,200-206 only
 * (arclengths, areas, energies, velocities, t_ext) without moving or
 * changing topology.  For unit tests of the velocity. */
int orc_compute_velocities(orc* s);

/* Request the export diagnostics (vertex energy, dA/dt, velocity parts)
 * on the next step. */
void orc_request_diagnostics(orc* s);

/* Topology-fixed window (extension): `steps` steps of the analytic model
 * with events and nucleation off and dt fixed; the trajectory is kept. */
int orc_window_forward(orc* s, int32_t steps);
/* Reverse pass over the last window (see sevc_adjoint_window). */
int orc_adjoint_window(orc* s, const double* dL_dpos, double* d_pos, double* d_SE, double* d_gb, double* d_mu);
/* Loader restatement of io.h:66-264 (independent of the product loader). */
int orc_load_files(const sevc_options* o, sevc_mesh* out);
int orc_build_mesh(const sevc_options* o, int32_t nv, const double* unit_pos, int32_t nr,
                   const int32_t* ridges, int32_t ng, const double* ori01, const double* se_raw,
                   sevc_mesh* out);

/* Host mesh helpers for the oracle loader output. */
int  orc_mesh_alloc(sevc_mesh* m, int32_t nv, int32_t ne, int32_t ng, int32_t S);
void orc_mesh_free(sevc_mesh* m);

/* check_data restatement (utils.h:85-336): number of violations. */
int orc_check_mesh(const sevc_mesh* m, double* total_area);

/* Single stages of the step, for the tests that pin the oracle. */
int orc_poll(orc* s, const double* t_ext, uint8_t* cand, int32_t* conf, int32_t stages);
int orc_nucl_geometry(orc* s, int32_t A, double r, double ori, int32_t do_scale, double* xy, double* dE,
                      int32_t* ids);
int orc_nucleate_vertex(orc* s, int32_t A, double r, double ori, int32_t* slots);
int orc_mc_trial(orc* s, int32_t step, int32_t* site, double* dE, double* nfac);
int orc_rate_draws(orc* s, int32_t step, int32_t* grains, int32_t* sites, int32_t max);

/* Geometry primitives exposed for closed-form tests. */
double orc_wrap_dist(double dest, double orig, double D);     /* geometry.h:147 */
double orc_dom_mod(double a, double D);                       /* geometry.h:133 */
double orc_mag(double x, double y);                           /* geometry.h:217 */
double orc_polygon_area(const double* x, const double* y, int32_t n, double D);  /* geometry.h:757 */
double orc_sum_chunked(const double* v, const uint8_t* mask, int32_t n);

/* Elementary functions of include/sevc_math.h compiled on the host. */
double orc_atan2(double y, double x);
double orc_sin(double x);
double orc_cos(double x);
double orc_exp(double x);
void   orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
