/*
 * sevc_oracle.c -- TEST INFRASTRUCTURE ONLY (see sevc_oracle.h).
 *
 * Serial FP64 restatement of the reference time step.  Every function names
 * the reference lines it follows.  Build with -O2 -ffp-contract=off so that
 * no multiply-add is fused: the operation order below is the reference's.
 *
 * Decisions where the reference is undefined or racy (the HIP path
 * follows the same ones):
 *  - vote loop: Jacobi rounds until a round changes no candidate flag
 *    (the reference's convergence test ORs inside blocks and reads an
 *    uninitialised buffer, topological.h:486-537, main.cu:102);
 *  - apply_flippings on adjacent candidates (a data race in the reference,
 *    topological.h:1016-1026) is defined as serial edge-index order;
 *  - cuRAND XORWOW is replaced by Philox4x32-10 keyed by (seed, step,
 *    entity) (nucleation2.cu:12-16,243,298,392,421); the random vertex is
 *    the floor(u*n_alive)-th enabled vertex (replaces the launch-shape
 *    dependent pick of nucleation2.cu:292-402);
 *  - libm atan2/cos/sin/exp/pow are replaced by include/sevc_math.h
 *    (pow(c,3) -> c*c*c);
 *  - nucleation needs all 1 grain + 3 boundary + 2 vertex free slots (the
 *    reference proceeds with index -1 when only some are free,
 *    nucleation2.cu:486-499);
 *  - a grain list longer than max_vrt_per_grn is an error (the reference
 *    writes past the array, geometry.h:615-619);
 *  - the hard-coded min_grains = 5000 (main.cu:49) is not reproduced: the
 *    stop rule uses leave_percent.
 *
 * ORC_MUTANT (0 by default) builds a deliberately broken copy of the oracle
 * for the tests that show its pins catch plausible mistakes: each value
 * plants one mistake (see the `#if ORC_MUTANT == k` sites).  The normal
 * builds never define it.
 */
#include "sevc_oracle.h"
#include "../include/sevc_math.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

enum { FIX_ADD = 0, FIX_REMOVE = 1, FIX_REPLACE = 2, FIX_NONE = 3 };   /* geometry.h:17 */
enum { T_COMMON = 0, T_NUCLEATED = 1 };                                  /* geometry.h:19 */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define STREAM_MC_SITE 1u
#define STREAM_MC_PICK 2u
#define STREAM_RATE    3u

struct orc {
    sevc_options o;
    int nv, ne, ng, S;
    double D;
    /* vertices */
    double *pos, *vel, *venergy, *vparts;
    int *vb, *vg;
    uint8_t *ven;
    /* boundaries */
    int *bvx;
    double *ben, *blen, *bt;
    uint8_t *benab, *bcand;
    /* grains */
    int *gvx, *glen;
    double *garea, *gdadt, *gori, *gse;
    uint8_t *genab, *gtype;
    /* scratch of the reference structs */
    uint8_t *mfix;
    int *mgrn, *mvrt, *mprev;          /* 4 per vertex, geometry.h:26-31,47 */
    int *voted, *nvotes, *conf;
    uint8_t *gcand, *gfix, *vnuc, *cand_prev;
    double *nucl_r, *nucl_ori, *dE, *nfac;
    /* topology-fixed window (extension): trajectory x_0..x_N, grain areas
     * at x_0..x_{N-1}, d gamma_b / d GB_scaling */
    int wN;
    double *wtraj, *warea, *webase;
    /* loop state (main.cu:30-113) */
    int n;
    double t, dt_last;
    int curr_grains, n_grains0, min_grains, allow_nucl, done, status, diag;
    long long n_flips, n_removals, n_two, n_nucl;
};

/* The per-entity loops of the physics kernels below (each iteration writes
 * only its own entity and reads nothing the loop writes) run in parallel in
 * the OpenMP build, liboracle_omp.so, which serves as a multi-core timing
 * baseline only.  The serial build is the checker; both give the same bits. */
#ifdef _OPENMP
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PARALLEL_FOR
#endif

/* ------------------------------------------------------------------ */
/* geometry.h primitives                                               */
/* ------------------------------------------------------------------ */

/* geometry.h:133-137 */
double orc_dom_mod(double a, double D) {
    if (a < 0) return a + D;
    else if (a >= D) return a - D;
    return a;
}

/* geometry.h:147-152 */
double orc_wrap_dist(double dest, double orig, double D) {
    double delta = dest - orig;
    if (delta > +(D / 2.0)) return (delta - D);
    if (delta < -(D / 2.0)) return (D + delta);
    return delta;
}

/* geometry.h:217-223 */
double orc_mag(double x, double y) {
    double ax = fabs(x), ay = fabs(y);
    double p = ax > ay ? ax : ay;
    if (p == 0) return 0;
    double q = ax < ay ? ax : ay;
    double r = q / p;
    return p * sqrt(1.0 + r * r);
}

/* geometry.h:328-330: |wrap(ini - end)| */
static double arclen(double ax, double ay, double bx, double by, double D) {
    return orc_mag(orc_wrap_dist(ax, bx, D), orc_wrap_dist(ay, by, D));
}

/* geometry.h:908-931, in place */
static void adjust_origin(double* X, double* Y, int n, double D) {
    if (n > 0) {
        double px = X[0], py = Y[0];
        X[0] = 0; Y[0] = 0;
        for (int i = 1; i < n; i++) {
            double cx = X[i], cy = Y[i];
            double dx = orc_wrap_dist(cx, px, D);
            double dy = orc_wrap_dist(cy, py, D);
            px = X[i]; py = Y[i];
            X[i] = X[i - 1] + dx;
            Y[i] = Y[i - 1] + dy;
        }
    }
}

/* geometry.h:757-765 (destroys X, Y) */
static double area_inplace(double* X, double* Y, int n, double D) {
    double area = 0;
    adjust_origin(X, Y, n, D);
    for (int k = 0; k < n; k++) {
        int k1 = (k + 1) % n;
        area += X[k] * Y[k1] - Y[k] * X[k1];
    }
    area *= 0.5;
    return area;
}

double orc_polygon_area(const double* x, const double* y, int32_t n, double D) {
    double X[4096], Y[4096];
    if (n > 4096) return NAN;
    for (int i = 0; i < n; i++) { X[i] = x[i]; Y[i] = y[i]; }
    return area_inplace(X, Y, n, D);
}

/* geometry.h:779-787 (destroys X, Y) */
static double dadt_inplace(double* X, double* Y, const double* Vx, const double* Vy, int n, double D) {
    double d = 0;
    adjust_origin(X, Y, n, D);
    for (int k = 0; k < n; k++) {
        int k1 = (k + 1) % n;
        d += (Vx[k] * Y[k1] + X[k] * Vy[k1]);
        d -= (Vy[k] * X[k1] + Y[k] * Vx[k1]);
    }
    return d;
}

/* geometry.h:548-550 with pow(c,3) -> c*c*c */
static double gb_energy_func(double d_alpha, double eps) {
    double c = sevc_cos(4 * d_alpha);
    return (1 + 0.5 * eps * (1 - c * c * c));
}

/* geometry.h:274-287 */
static void unitary(double vx, double vy, double* ux, double* uy) {
    double m = orc_mag(vx, vy);
    if (m == 0) { *ux = 0; *uy = 0; }
    else {
        double a = sevc_atan2(vy, vx);
        *ux = sevc_cos(a);
        *uy = sevc_sin(a);
    }
}

double orc_atan2(double y, double x) { return sevc_atan2(y, x); }
double orc_sin(double x) { return sevc_sin(x); }
double orc_cos(double x) { return sevc_cos(x); }
double orc_exp(double x) { return sevc_exp(x); }
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    sevc_philox4x32_10(ctr, key, out);
}

/* Deterministic chunked sum: chunks of 256 consecutive selected values
 * summed left to right, then the chunk sums reduced the same way.  The
 * HIP reduction uses the identical order. */
double orc_sum_chunked(const double* v, const uint8_t* mask, int32_t n) {
    if (n <= 0) return 0.0;
    int nc = (n + 255) / 256;
    double* part = (double*)malloc(sizeof(double) * (size_t)nc);
    for (int c = 0; c < nc; c++) {
        double s = 0.0;
        int e = (c + 1) * 256 < n ? (c + 1) * 256 : n;
        for (int i = c * 256; i < e; i++)
            if (!mask || mask[i]) s += v[i];
        part[c] = s;
    }
    while (nc > 1) {
        int nn = (nc + 255) / 256;
        for (int c = 0; c < nn; c++) {
            double s = 0.0;
            int e = (c + 1) * 256 < nc ? (c + 1) * 256 : nc;
            for (int i = c * 256; i < e; i++) s += part[i];
            part[c] = s;
        }
        nc = nn;
    }
    double r = part[0];
    free(part);
    return r;
}

/* ------------------------------------------------------------------ */
/* accessors                                                            */
/* ------------------------------------------------------------------ */
#define VB(s, v, i) ((s)->vb[3 * (v) + (i)])
#define VG(s, v, i) ((s)->vg[3 * (v) + (i)])
#define BI(s, b) ((s)->bvx[2 * (b)])
#define BE(s, b) ((s)->bvx[2 * (b) + 1])
#define GV(s, g, k) ((s)->gvx[(size_t)(g) * (s)->S + (k)])
#define PX(s, v) ((s)->pos[2 * (v)])
#define PY(s, v) ((s)->pos[2 * (v) + 1])
#define VX(s, v) ((s)->vel[2 * (v)])
#define VY(s, v) ((s)->vel[2 * (v) + 1])

static void set_status(orc* s, int code) {
    if (s->status == SEVC_OK) s->status = code;
}

/* geometry.h:369-376 */
static void reset_mod_list(orc* s, int v) {
    for (int i = 0; i < 4; i++) {
        s->mvrt[4 * v + i] = -1;
        s->mprev[4 * v + i] = -1;
        s->mgrn[4 * v + i] = -1;
        s->mfix[4 * v + i] = FIX_NONE;
    }
}

/* geometry.h:837-848; returns how many shared grains were found */
static int grains_intersect(const orc* s, int a, int b, int out[2]) {
    int k = 0;
    out[0] = out[1] = -1;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (VG(s, a, i) == VG(s, b, j)) {
                out[k++] = VG(s, a, i);
                if (k == 2) return 2;
            }
    return k;
}

/* geometry.h:863-875 */
static void grains_union(const orc* s, int a, int b, int out[4]) {
    out[3] = -1;
    for (int i = 0; i < 3; i++) out[i] = VG(s, a, i);
    for (int i = 0; i < 3; i++) {
        int g = VG(s, b, i);
        if (g != out[0] && g != out[1] && g != out[2]) { out[3] = g; break; }
    }
}

/* geometry.h:886-898 */
static void grains_symdiff(const orc* s, int a, int b, int out[2]) {
    int gic[2], gun[4], k = 0;
    out[0] = out[1] = -1;
    grains_intersect(s, a, b, gic);
    grains_union(s, a, b, gun);
    for (int i = 0; i < 4; i++)
        if (gun[i] != gic[0] && gun[i] != gic[1]) {
            out[k++] = gun[i];
            if (k == 2) return;
        }
}

/* geometry.h:703-710 */
static int grain_contains_vertex(const orc* s, int g, int v) {
    for (int i = 0; i < s->glen[g]; i++)
        if (GV(s, g, i) == v) return i;
    return -1;
}

/* geometry.h:724-734 */
static int grain_contains_consecutive(const orc* s, int g, int a, int b) {
    int n = s->glen[g];
    for (int i = 0; i < n; i++) {
        if (GV(s, g, i) == a && GV(s, g, (i + 1) % n) == b) return 1;
        if (GV(s, g, i) == b && GV(s, g, (i + 1) % n) == a) return 2;
    }
    return -1;
}

/* geometry.h:614-620 */
static void grain_push_vertex(orc* s, int g, int v) {
    if (s->glen[g] >= s->S) { set_status(s, SEVC_E_GRAIN_FULL); return; }
    GV(s, g, s->glen[g]) = v;
    s->glen[g]++;
}

/* geometry.h:629-645 */
static void grain_add_vertex_after(orc* s, int g, int v, int prev) {
    int n = s->glen[g], pos = -1;
    for (int i = 0; i < n; i++)
        if (GV(s, g, i) == prev) { pos = i + 1; break; }
    if (pos < 0) { set_status(s, SEVC_E_STATE); return; }
    if (n >= s->S) { set_status(s, SEVC_E_GRAIN_FULL); return; }
    for (int i = n; i > pos; i--) GV(s, g, i) = GV(s, g, i - 1);
    GV(s, g, pos) = v;
    s->glen[g]++;
}

/* geometry.h:653-666 */
static void grain_remove_vertex(orc* s, int g, int v) {
    for (int i = 0; i < s->glen[g]; i++)
        if (GV(s, g, i) == v) {
            for (int j = i; j < s->glen[g] - 1; j++) GV(s, g, j) = GV(s, g, j + 1);
            s->glen[g]--;
            GV(s, g, s->glen[g]) = -1;
            return;
        }
}

/* geometry.h:686-694 */
static void grain_replace_vertex(orc* s, int g, int old_v, int new_v) {
    for (int i = 0; i < s->glen[g]; i++)
        if (GV(s, g, i) == old_v) { GV(s, g, i) = new_v; return; }
}

/* geometry.h:385-396 */
static void vertex_add_boundary(orc* s, int v, int b) {
    for (int i = 0; i < 3; i++)
        if (VB(s, v, i) < 0) { VB(s, v, i) = b; return; }
    set_status(s, SEVC_E_STATE);
}

/* geometry.h:406-417 */
static void vertex_add_grain(orc* s, int v, int g) {
    for (int i = 0; i < 3; i++)
        if (VG(s, v, i) < 0) { VG(s, v, i) = g; return; }
    set_status(s, SEVC_E_STATE);
}

/* geometry.h:424-429 */
static void vertex_invert_boundary_order(orc* s, int v) {
    int b = VB(s, v, 1);
    VB(s, v, 1) = VB(s, v, 0);
    VB(s, v, 0) = b;
}

/* geometry.h:513-517, extended with the flip threshold (flip_threshold=0
 * gives the reference rule). */
static int boundary_is_candidate(const orc* s, int b, double dt) {
    double te = s->bt[b];
    if (te >= 0 && te <= dt) return 1;
    /* a zero-length edge (just flipped; its t_ext is the no-motion value)
     * is not shrinking, so the length arm needs L > 0 */
    if (s->o.flip_threshold > 0 && te >= 0 && s->blen[b] > 0 && s->blen[b] < s->o.flip_threshold) return 1;
    return 0;
}

/* geometry.h:559-564 */
static double boundary_compute_energy(const orc* s, int b, double eps) {
    int gic[2];
    grains_intersect(s, BI(s, b), BE(s, b), gic);
    if (gic[0] < 0 || gic[1] < 0) return 1.0;
    double d_alpha = s->gori[gic[0]] - s->gori[gic[1]];
    return gb_energy_func(d_alpha, eps);
}

/* geometry.h:472-479 */
static double vertex_compute_energy(const orc* s, int v) {
    double e = 0.0;
    for (int i = 0; i < 3; i++) {
        int b = VB(s, v, i), g = VG(s, v, i);
        e += (s->blen[b] * s->ben[b] * 0.5);
        e += (s->gse[g] * s->garea[g] / s->glen[g]);
    }
    return e;
}

/* ------------------------------------------------------------------ */
/* calculus.h kernels                                                  */
/* ------------------------------------------------------------------ */

/* calculus.h:83-93 */
static void compute_boundary_arclengths(orc* s) {
    ORC_PARALLEL_FOR
    for (int b = 0; b < s->ne; b++)
        if (s->benab[b])
            s->blen[b] = arclen(PX(s, BI(s, b)), PY(s, BI(s, b)), PX(s, BE(s, b)), PY(s, BE(s, b)), s->D);
}

/* geometry.h:797-804 */
static double grain_area(const orc* s, int g) {
    double X[SEVC_REF_MAX_VRT_PER_GRN * 4], Y[SEVC_REF_MAX_VRT_PER_GRN * 4];
    int n = s->glen[g];
    for (int i = 0; i < n; i++) { X[i] = PX(s, GV(s, g, i)); Y[i] = PY(s, GV(s, g, i)); }
    return area_inplace(X, Y, n, s->D);
}

/* calculus.h:224-234 */
static void compute_grain_areas(orc* s) {
    ORC_PARALLEL_FOR
    for (int g = 0; g < s->ng; g++)
        if (s->genab[g]) s->garea[g] = grain_area(s, g);
}

/* calculus.h:104-114 */
static void compute_boundary_energies(orc* s) {
    ORC_PARALLEL_FOR
    for (int b = 0; b < s->ne; b++)
        if (s->benab[b]) s->ben[b] = boundary_compute_energy(s, b, s->o.GB_eps) * s->o.GB_scaling;
}

/* calculus.h:123-133 */
static void compute_vertex_energies(orc* s) {
    if (!s->venergy) return;
    ORC_PARALLEL_FOR
    for (int v = 0; v < s->nv; v++)
        if (s->ven[v]) s->venergy[v] = vertex_compute_energy(s, v);
}

/* calculus.h:19-73 */
static double energy_diff(orc* s, int comp, int v) {
    double nx = PX(s, v), ny = PY(s, v);
    if (comp == 0) nx += SEVC_GRAD_EPS;
    else ny += SEVC_GRAD_EPS;
    nx = orc_dom_mod(nx, s->D);
    ny = orc_dom_mod(ny, s->D);

    double arcl[3] = {0, 0, 0};
    for (int j = 0; j < 3; j++) {
        int b = VB(s, v, j);
        int nb = (BI(s, b) == v) ? BE(s, b) : BI(s, b);
        arcl[j] = arclen(nx, ny, PX(s, nb), PY(s, nb), s->D);
    }
    double areas[3] = {0, 0, 0};
    for (int j = 0; j < 3; j++) {
        double X[SEVC_REF_MAX_VRT_PER_GRN * 4], Y[SEVC_REF_MAX_VRT_PER_GRN * 4];
        int g = VG(s, v, j);
        int n = s->glen[g];
        for (int k = 0; k < n; k++) {
            int w = GV(s, g, k);
            X[k] = PX(s, w); Y[k] = PY(s, w);
            if (w == v) { X[k] = nx; Y[k] = ny; }
        }
        areas[j] = area_inplace(X, Y, n, s->D);
    }
    double bnd_term = 0.0, grn_term = 0.0;
    for (int j = 0; j < 3; j++) {
        int b = VB(s, v, j), g = VG(s, v, j);
        bnd_term += (s->ben[b] * (arcl[j] - s->blen[b]));
        grn_term += (s->gse[g] * (fabs(areas[j]) - fabs(s->garea[g])));
    }
    if (s->vparts && s->diag) {
        if (comp == 0) {
            s->vparts[4 * v + 0] = -bnd_term / SEVC_GRAD_EPS;
            s->vparts[4 * v + 1] = -grn_term / SEVC_GRAD_EPS;
        } else {
            s->vparts[4 * v + 2] = -bnd_term / SEVC_GRAD_EPS;
            s->vparts[4 * v + 3] = -grn_term / SEVC_GRAD_EPS;
        }
    }
    return bnd_term + grn_term;
}

/* Analytic -grad E at vertex v (extension used by the adjoint): edges
 * contribute -gamma_b * wrap(p_v - p_nb)/L_b, grains -SE_g sign(A_g) dA_g/dp_v
 * with dA/dx_k = (dy_prev + dy_next)/2, dA/dy_k = -(dx_prev + dx_next)/2
 * where d_prev = wrap(p_k - p_{k-1}), d_next = wrap(p_{k+1} - p_k). */
static void analytic_force(const orc* s, int v, double* fx, double* fy) {
    double x = PX(s, v), y = PY(s, v);
    double ex = 0.0, ey = 0.0, gx = 0.0, gy = 0.0;
    for (int j = 0; j < 3; j++) {
        int b = VB(s, v, j);
        int nb = (BI(s, b) == v) ? BE(s, b) : BI(s, b);
        double dx = orc_wrap_dist(x, PX(s, nb), s->D);
        double dy = orc_wrap_dist(y, PY(s, nb), s->D);
        double L = orc_mag(dx, dy);
        if (L > 0) {
            ex += s->ben[b] * (dx / L);
            ey += s->ben[b] * (dy / L);
        }
    }
    for (int j = 0; j < 3; j++) {
        int g = VG(s, v, j);
        int n = s->glen[g];
        int k = grain_contains_vertex(s, g, v);
        if (k < 0 || n < 3) continue;
        int wp = GV(s, g, (k + n - 1) % n), wn = GV(s, g, (k + 1) % n);
        double dpx = orc_wrap_dist(x, PX(s, wp), s->D), dpy = orc_wrap_dist(y, PY(s, wp), s->D);
        double dnx = orc_wrap_dist(PX(s, wn), x, s->D), dny = orc_wrap_dist(PY(s, wn), y, s->D);
        double sg = s->garea[g] < 0 ? -1.0 : 1.0;
        double c = s->gse[g] * sg * 0.5;
        gx += c * (dpy + dny);
        gy += c * (-(dpx + dnx));
    }
    *fx = -(ex + gx);
    *fy = -(ey + gy);
}

/* calculus.h:143-154 (+ mobility, + analytic model) */
static void compute_vertex_velocities(orc* s) {
    double mu = s->o.mobility;
    ORC_PARALLEL_FOR
    for (int v = 0; v < s->nv; v++) {
        if (!s->ven[v]) continue;
        if (s->o.velocity_model == SEVC_VEL_ANALYTIC) {
            double fx, fy;
            analytic_force(s, v, &fx, &fy);
            VX(s, v) = fx * mu;
            VY(s, v) = fy * mu;
        } else {
            VX(s, v) = (-energy_diff(s, 0, v) / SEVC_GRAD_EPS) * mu;
            VY(s, v) = (-energy_diff(s, 1, v) / SEVC_GRAD_EPS) * mu;
        }
    }
}

/* calculus.h:243-253 */
static void compute_grain_dadts(orc* s) {
    if (!s->gdadt) return;
    ORC_PARALLEL_FOR
    for (int g = 0; g < s->ng; g++) {
        if (!s->genab[g]) continue;
        double X[SEVC_REF_MAX_VRT_PER_GRN * 4], Y[SEVC_REF_MAX_VRT_PER_GRN * 4];
        double U[SEVC_REF_MAX_VRT_PER_GRN * 4], W[SEVC_REF_MAX_VRT_PER_GRN * 4];
        int n = s->glen[g];
        for (int i = 0; i < n; i++) {
            int w = GV(s, g, i);
            X[i] = PX(s, w); U[i] = VX(s, w);
            Y[i] = PY(s, w); W[i] = VY(s, w);
        }
        s->gdadt[g] = dadt_inplace(X, Y, U, W, n, s->D);
    }
}

/* calculus.h:391-404 */
static void update_vertices_positions(orc* s, double dt) {
    ORC_PARALLEL_FOR
    for (int v = 0; v < s->nv; v++) {
        if (!s->ven[v]) continue;
        PX(s, v) += VX(s, v) * dt;
        PY(s, v) += VY(s, v) * dt;
        PX(s, v) = orc_dom_mod(PX(s, v), s->D);
        PY(s, v) = orc_dom_mod(PY(s, v), s->D);
    }
}

/* ------------------------------------------------------------------ */
/* topological.h                                                        */
/* ------------------------------------------------------------------ */

/* topological.h:303-334 */
static void compute_t_ext(orc* s) {
    ORC_PARALLEL_FOR
    for (int b = 0; b < s->ne; b++) {
        if (!s->benab[b]) continue;
        s->bcand[b] = 0;
        int ini = BI(s, b), end = BE(s, b);
        double xp[2] = {PX(s, ini), PX(s, end)}, yp[2] = {PY(s, ini), PY(s, end)};
        adjust_origin(xp, yp, 2, s->D);
        double dx = xp[1] - xp[0], dy = yp[1] - yp[0];
        double norm = orc_mag(dx, dy);
        double tx, ty;
        unitary(dx, dy, &tx, &ty);
        double px = VX(s, end) - VX(s, ini);
        double py = VY(s, end) - VY(s, ini);
        double aux = tx * px + ty * py;
        if (aux == 0) s->bt[b] = 100;
        else s->bt[b] = -norm / aux;
    }
}

/* Adaptive dt (extension): dt_n = clamp(dt_safety * tau, dt_min, dt) with
 * tau the smallest positive t_ext over enabled edges not shorter than the
 * flip threshold; dt when there is no such edge or adaptive_dt = 0. */
static double step_dt(const orc* s) {
    if (!s->o.adaptive_dt) return s->o.dt;
    double tau = INFINITY;
    for (int b = 0; b < s->ne; b++) {
        if (!s->benab[b]) continue;
        double te = s->bt[b];
        if (te > 0 && !(s->blen[b] < s->o.flip_threshold) && te < tau) tau = te;
    }
    if (!(tau < INFINITY)) return s->o.dt;
    double d = s->o.dt_safety * tau;
    if (d < s->o.dt_min) d = s->o.dt_min;
    if (d > s->o.dt) d = s->o.dt;
    return d;
}

/* topological.h:343-353 (all boundaries, enabled or not: quirk A-6) */
static void label_boundaries(orc* s, double dt) {
    ORC_PARALLEL_FOR
    for (int b = 0; b < s->ne; b++) {
        s->bcand[b] = (uint8_t)boundary_is_candidate(s, b, dt);
        s->conf[b] = -1;
    }
}

/* polling_system, topological.h:549-578 with the kernels of :362-457, run
 * as Jacobi rounds until a round changes nothing. */
static void polling_system(orc* s) {
    for (;;) {
        for (int b = 0; b < s->ne; b++) if (s->benab[b]) s->nvotes[b] = 0;       /* :426 */
        for (int v = 0; v < s->nv; v++) if (s->ven[v]) s->voted[v] = -1;         /* :388 */
        for (int v = 0; v < s->nv; v++) {                                        /* :362 */
            if (!s->ven[v]) continue;
            double min_t = 99999999;
            int arg = 0;
            for (int i = 0; i < 3; i++) {
                int b = VB(s, v, i);
#if ORC_MUTANT == 1
                if (s->bt[b] < min_t && s->bcand[b]) { min_t = s->bt[b]; arg = i; }
#else
                if (s->bt[b] <= min_t && s->bcand[b]) { min_t = s->bt[b]; arg = i; }
#endif
            }
            s->voted[v] = VB(s, v, arg);
        }
        for (int b = 0; b < s->ne; b++) {                                        /* :406 */
            if (s->benab[b] && s->bcand[b]) {
                if (s->voted[BI(s, b)] == b) s->nvotes[b]++;
                if (s->voted[BE(s, b)] == b) s->nvotes[b]++;
            }
        }
        for (int b = 0; b < s->ne; b++) s->cand_prev[b] = s->benab[b] ? s->bcand[b] : 0;
        for (int b = 0; b < s->ne; b++) {                                        /* :441 */
            if (s->benab[b] && s->bcand[b] && s->nvotes[b] == 2) {
                for (int i = 0; i < 3; i++) {
                    int x = VB(s, BI(s, b), i);
                    if (x != b) s->bcand[x] = 0;
                    x = VB(s, BE(s, b), i);
                    if (x != b) s->bcand[x] = 0;
                }
            }
        }
        int changed = 0;
        for (int b = 0; b < s->ne; b++)
            if (s->benab[b] && s->bcand[b] != s->cand_prev[b]) { changed = 1; break; }
#if ORC_MUTANT == 2
        changed = 0;
#endif
        if (!changed) break;
    }
}

/* topological.h:83-121 */
static void check_conflictive_boundaries(orc* s) {
    for (int b = 0; b < s->ne; b++) {
        if (!(s->benab[b] && !s->bcand[b])) continue;
        int bc_ini = -1, bc_end = -1;
        int ini = BI(s, b), end = BE(s, b);
        for (int i = 0; i < 3; i++) {
            int x = VB(s, ini, i);
            if (x != b && s->bcand[x]) { bc_ini = x; break; }
        }
        for (int i = 0; i < 3; i++) {
            int x = VB(s, end, i);
            if (x != b && s->bcand[x]) { bc_end = x; break; }
        }
        if (bc_ini >= 0 && bc_end >= 0) {
#if ORC_MUTANT == 3
            if (s->bt[bc_ini] == s->bt[bc_end])
                s->conf[b] = bc_ini < bc_end ? bc_ini : bc_end;
            else
                s->conf[b] = s->bt[bc_ini] > s->bt[bc_end] ? bc_ini : bc_end;
#elif ORC_MUTANT == 4
            if (s->bt[bc_ini] == s->bt[bc_end])
                s->conf[b] = bc_ini > bc_end ? bc_ini : bc_end;
            else
                s->conf[b] = s->bt[bc_ini] < s->bt[bc_end] ? bc_ini : bc_end;
#else
            if (s->bt[bc_ini] == s->bt[bc_end])
                s->conf[b] = bc_ini < bc_end ? bc_ini : bc_end;
            else
                s->conf[b] = s->bt[bc_ini] < s->bt[bc_end] ? bc_ini : bc_end;
#endif
        }
    }
}

/* topological.h:129-155 */
static void unlabel_conflictive_boundaries(orc* s) {
    for (int b = 0; b < s->ne; b++) {
        if (!(s->benab[b] && s->bcand[b])) continue;
        for (int i = 0; i < 3; i++) {
            int x = VB(s, BI(s, b), i);
            if (x != b && s->conf[x] != -1 && s->conf[x] != b) { s->bcand[b] = 0; break; }
        }
        for (int i = 0; i < 3; i++) {
            int x = VB(s, BE(s, b), i);
            if (x != b && s->conf[x] != -1 && s->conf[x] != b) { s->bcand[b] = 0; break; }
        }
    }
}

/* topological.h:22-73 */
static int two_three_sided_grains(const orc* s) {
    int cnt = 0;
    for (int b = 0; b < s->ne; b++) {
        if (!s->benab[b]) continue;
        int gic[2];
        if (grains_intersect(s, BI(s, b), BE(s, b), gic) < 2) continue;
        if (s->glen[gic[0]] == 3 && s->glen[gic[1]] == 3) cnt++;
    }
    return cnt;
}

static void set_mod(orc* s, int v, int slot, int fix, int grn, int vrt, int prev, int set_prev) {
    s->mfix[4 * v + slot] = (uint8_t)fix;
    s->mgrn[4 * v + slot] = grn;
    s->mvrt[4 * v + slot] = vrt;
    if (set_prev) s->mprev[4 * v + slot] = prev;
}

/* topological.h:840-955 */
static void remove_two_grains(orc* s, int bnd, const int gnic[2]) {
    int A = BI(s, bnd), B = BE(s, bnd);
    int AC = -1, BD = -1, CE = -1, DF = -1;
    for (int i = 0; i < 3; i++)
        if (VB(s, A, i) == bnd) { AC = VB(s, A, (i + 1) % 3); break; }
    for (int j = 0; j < 3; j++)
        if (VB(s, B, j) == bnd) { BD = VB(s, B, (j + 1) % 3); break; }
    if (AC < 0 || BD < 0) { set_status(s, SEVC_E_STATE); return; }
    int C = (BI(s, AC) == A) ? BE(s, AC) : BI(s, AC);
    int D = (BI(s, BD) == B) ? BE(s, BD) : BI(s, BD);
    for (int k = 0; k < 3; k++)
        if (VB(s, C, k) == AC) { CE = VB(s, C, (k + 2) % 3); break; }
    for (int l = 0; l < 3; l++)
        if (VB(s, D, l) == BD) { DF = VB(s, D, (l + 2) % 3); break; }
    if (CE < 0 || DF < 0) { set_status(s, SEVC_E_STATE); return; }
    int F = (BI(s, DF) == D) ? BE(s, DF) : BI(s, DF);
    if (BI(s, CE) == C) BI(s, CE) = F;
    else BE(s, CE) = F;
    for (int p = 0; p < 3; p++)
        if (VB(s, F, p) == DF) { VB(s, F, p) = CE; break; }
    int G3, G4;
    if (grain_contains_vertex(s, gnic[0], A) >= 0) { G3 = gnic[0]; G4 = gnic[1]; }
    else { G3 = gnic[1]; G4 = gnic[0]; }
    set_mod(s, A, 0, FIX_REMOVE, G3, A, 0, 0); s->gfix[G3] = 1;
    set_mod(s, A, 1, FIX_REMOVE, G3, C, 0, 0);
    set_mod(s, A, 2, FIX_REMOVE, G3, D, 0, 0);
    set_mod(s, B, 0, FIX_REMOVE, G4, B, 0, 0); s->gfix[G4] = 1;
    set_mod(s, B, 1, FIX_REMOVE, G4, C, 0, 0);
    set_mod(s, B, 2, FIX_REMOVE, G4, D, 0, 0);
    for (int i = 0; i < 3; i++) {
        s->benab[VB(s, A, i)] = 0;
        s->benab[VB(s, B, i)] = 0;
    }
    s->benab[DF] = 0;
    s->benab[bnd] = 0;
    s->ven[A] = 0; s->ven[D] = 0; s->ven[C] = 0; s->ven[B] = 0;
}

/* topological.h:965-985 (serial <<<1,1>>> scan) */
static int remove_two_grains_configurations(orc* s) {
    int done = 0;
    for (int b = 0; b < s->ne; b++) {
        if (!s->benab[b]) continue;
        int gic[2], gnic[2];
        if (grains_intersect(s, BI(s, b), BE(s, b), gic) < 2) continue;
        grains_symdiff(s, BI(s, b), BE(s, b), gnic);
        if (s->glen[gic[0]] == 3 && s->glen[gic[1]] == 3) {
            if (gnic[0] < 0 || gnic[1] < 0) { set_status(s, SEVC_E_STATE); continue; }
            remove_two_grains(s, b, gnic);
            s->genab[gic[0]] = 0;
            s->genab[gic[1]] = 0;
            done++;
        }
    }
    return done;
}

/* fix_grain_lists, topological.h:1209-1227 with :1035-1163 */
static void fix_grain_lists(orc* s) {
    for (int g = 0; g < s->ng; g++) {                                         /* ADD :1035 */
        if (!(s->genab[g] && s->gfix[g])) continue;
        for (int i = 0; i < s->glen[g]; i++) {
            int v = GV(s, g, i);
            for (int j = 0; j < 4; j++)
                if (s->mfix[4 * v + j] == FIX_ADD && s->mgrn[4 * v + j] == g)
                    grain_add_vertex_after(s, g, s->mvrt[4 * v + j], s->mprev[4 * v + j]);
        }
    }
    for (int g = 0; g < s->ng; g++) {                                         /* REMOVE :1081 */
        if (!(s->genab[g] && s->gfix[g])) continue;
        int n = s->glen[g];
        int rv[4 * SEVC_REF_MAX_VRT_PER_GRN * 4];
        int nr = 0;
        for (int i = 0; i < n; i++) {
            int v = GV(s, g, i);
            for (int j = 0; j < 4; j++)
                if (s->mfix[4 * v + j] == FIX_REMOVE && s->mgrn[4 * v + j] == g) rv[nr++] = s->mvrt[4 * v + j];
        }
        for (int k = 0; k < nr; k++) grain_remove_vertex(s, g, rv[k]);
    }
    for (int g = 0; g < s->ng; g++) {                                         /* REPLACE :1119 */
        if (!(s->genab[g] && s->gfix[g])) continue;
        for (int i = 0; i < s->glen[g]; i++) {
            int v = GV(s, g, i);
            for (int j = 0; j < 4; j++)
                if (s->mfix[4 * v + j] == FIX_REPLACE && s->mgrn[4 * v + j] == g)
                    grain_replace_vertex(s, g, s->mprev[4 * v + j], s->mvrt[4 * v + j]);
        }
    }
    for (int g = 0; g < s->ng; g++) {                                         /* clear :1150 */
        if (!s->genab[g]) continue;
        s->gfix[g] = 0;
        for (int i = 0; i < s->glen[g]; i++) reset_mod_list(s, GV(s, g, i));
    }
}

/* topological.h:614-623 + :634-668 */
static void label_3sided_grains(orc* s, double dt) {
    for (int g = 0; g < s->ng; g++) if (s->genab[g]) s->gcand[g] = 0;
    for (int g = 0; g < s->ng; g++) {
        if (!(s->genab[g] && s->glen[g] == 3)) continue;
        int ans = 0;
        for (int i = 0; i < 3; i++) {
            int v = GV(s, g, i);
            for (int k = 0; k < 3; k++) {
                int b = VB(s, v, k);
                int ip = grain_contains_vertex(s, g, BI(s, b));
                int ep = grain_contains_vertex(s, g, BE(s, b));
                if (ip >= 0 && ep >= 0 && s->bcand[b] && s->benab[b]) { s->gcand[g] = 1; break; }
                if (ip >= 0 && ep >= 0 && boundary_is_candidate(s, b, dt) && s->benab[b]) ans++;
            }
        }
        ans /= 2;
        if (ans == 3) s->gcand[g] = 1;
    }
}

/* topological.h:687-822 */
static void remove_grain(orc* s, int g) {
    int A = GV(s, g, 0), B = GV(s, g, 1), C = GV(s, g, 2);
    double D = s->D;
    double xp[3] = {PX(s, A), PX(s, B), PX(s, C)};
    double yp[3] = {PY(s, A), PY(s, B), PY(s, C)};
    double nAB = arclen(PX(s, A), PY(s, A), PX(s, B), PY(s, B), D);
    double nBC = arclen(PX(s, B), PY(s, B), PX(s, C), PY(s, C), D);
    double nCA = arclen(PX(s, C), PY(s, C), PX(s, A), PY(s, A), D);
    double denom = nAB + nBC + nCA;
    adjust_origin(xp, yp, 3, D);
    PX(s, A) += (nBC * xp[0] + nCA * xp[1] + nAB * xp[2]) / denom;
    PY(s, A) += (nBC * yp[0] + nCA * yp[1] + nAB * yp[2]) / denom;
    PX(s, A) = orc_dom_mod(PX(s, A), D);
    PY(s, A) = orc_dom_mod(PY(s, A), D);
    int gic[2], G1, G2, G3;
    grains_intersect(s, B, C, gic); G3 = (gic[0] == g) ? gic[1] : gic[0];
    grains_intersect(s, A, B, gic); G2 = (gic[0] == g) ? gic[1] : gic[0];
    grains_intersect(s, A, C, gic); G1 = (gic[0] == g) ? gic[1] : gic[0];
    if (G1 < 0 || G2 < 0 || G3 < 0) { set_status(s, SEVC_E_STATE); return; }
    VX(s, A) = 0.0;
    VY(s, A) = 0.0;
    for (int i = 0; i < 3; i++) {
        int e = VB(s, C, i);
        if ((BI(s, e) == B && BE(s, e) == C) || (BE(s, e) == B && BI(s, e) == C)) { s->benab[e] = 0; break; }
    }
    int BEb = -1, CF = -1;
    for (int i = 0; i < 3; i++) {
        int e = VB(s, B, i);
        if (!(BI(s, e) == A || BE(s, e) == A) && !(BI(s, e) == C || BE(s, e) == C)) { BEb = e; break; }
    }
    for (int i = 0; i < 3; i++) {
        int e = VB(s, C, i);
        if (!(BI(s, e) == A || BE(s, e) == A) && !(BI(s, e) == B || BE(s, e) == B)) { CF = e; break; }
    }
    if (BEb < 0 || CF < 0) { set_status(s, SEVC_E_STATE); return; }
    if (BI(s, BEb) == B) BI(s, BEb) = A; else BE(s, BEb) = A;
    if (BI(s, CF) == C) BI(s, CF) = A; else BE(s, CF) = A;
    for (int i = 0; i < 3; i++) {
        int e = VB(s, A, i);
        if (BI(s, e) == B || BE(s, e) == B) { s->benab[e] = 0; VB(s, A, i) = BEb; break; }
    }
    for (int i = 0; i < 3; i++) {
        int e = VB(s, A, i);
        if (BI(s, e) == C || BE(s, e) == C) { s->benab[e] = 0; VB(s, A, i) = CF; break; }
    }
    s->ven[B] = 0;
    s->ven[C] = 0;
    s->genab[g] = 0;
    set_mod(s, B, 0, FIX_REMOVE, G2, B, 0, 0); s->gfix[G2] = 1;
    set_mod(s, C, 0, FIX_REMOVE, G1, C, 0, 0); s->gfix[G1] = 1;
    set_mod(s, B, 1, FIX_REMOVE, G3, B, 0, 0); s->gfix[G3] = 1;
    set_mod(s, C, 1, FIX_REPLACE, G3, A, C, 1);
    for (int i = 0; i < 3; i++)
        if (VG(s, A, i) == g) VG(s, A, i) = G3;
}

/* topological.h:994-1004 (serial <<<1,1>>>, main.cu:259) */
static int remove_grains(orc* s) {
    int cnt = 0;
    for (int g = 0; g < s->ng; g++)
        if (s->genab[g] && s->gcand[g]) { remove_grain(s, g); cnt++; }
    return cnt;
}

/* topological.h:588-606 */
static void unlabel_boundaries_st2(orc* s) {
    for (int b = 0; b < s->ne; b++) {
        if (!(s->benab[b] && s->bcand[b])) continue;
        int gic[2];
        if (grains_intersect(s, BI(s, b), BE(s, b), gic) < 2) { s->bcand[b] = 0; continue; }
        if (s->glen[gic[0]] == 3 || s->glen[gic[1]] == 3) s->bcand[b] = 0;
    }
}

/* topological.h:174-291 */
static void flip(orc* s, int bnd) {
    int A = BI(s, bnd), B = BE(s, bnd);
    int gic[2], gnic[2];
    grains_intersect(s, A, B, gic);
    grains_symdiff(s, A, B, gnic);
    int i = -1, j = -1, AC = -1, BEb = -1;
    for (int k = 0; k < 3; k++)
        if (VB(s, A, k) == bnd) { i = (k + 1) % 3; AC = VB(s, A, i); break; }
    for (int k = 0; k < 3; k++)
        if (VB(s, B, k) == bnd) { j = (k + 1) % 3; BEb = VB(s, B, j); break; }
    if (i < 0 || j < 0 || gic[1] < 0 || gnic[1] < 0) { set_status(s, SEVC_E_STATE); return; }
    VB(s, A, i) = BEb;
    VB(s, B, j) = AC;
    if (BI(s, AC) == A) BI(s, AC) = B; else BE(s, AC) = B;
    if (BI(s, BEb) == B) BI(s, BEb) = A; else BE(s, BEb) = A;
    int g1, g2, g3, g4;
    for (g2 = 0; g2 < 2; g2++)
        if (grain_contains_consecutive(s, gic[g2], A, B) == 1) {
            set_mod(s, A, 0, FIX_REMOVE, gic[g2], A, 0, 0);
            s->gfix[gic[g2]] = 1;
            break;
        }
    for (g4 = 0; g4 < 2; g4++)
        if (grain_contains_consecutive(s, gic[g4], B, A) == 1) {
            set_mod(s, B, 0, FIX_REMOVE, gic[g4], B, 0, 0);
            s->gfix[gic[g4]] = 1;
            break;
        }
    for (g3 = 0; g3 < 2; g3++)
        if (grain_contains_vertex(s, gnic[g3], A) >= 0) {
            set_mod(s, A, 1, FIX_ADD, gnic[g3], B, A, 1);
            s->gfix[gnic[g3]] = 1;
            break;
        }
    for (g1 = 0; g1 < 2; g1++)
        if (grain_contains_vertex(s, gnic[g1], B) >= 0) {
            set_mod(s, B, 1, FIX_ADD, gnic[g1], A, B, 1);
            s->gfix[gnic[g1]] = 1;
            break;
        }
    if (g1 > 1 || g2 > 1 || g3 > 1 || g4 > 1) { set_status(s, SEVC_E_STATE); return; }
    for (int k = 0; k < 3; k++) {
        if (VG(s, A, k) == gic[g2]) VG(s, A, k) = gnic[g1];
        if (VG(s, B, k) == gic[g4]) VG(s, B, k) = gnic[g3];
    }
    /* vector2_mean, geometry.h:311-317 */
    double dx = orc_wrap_dist(PX(s, B), PX(s, A), s->D);
    double dy = orc_wrap_dist(PY(s, B), PY(s, A), s->D);
    double cx = PX(s, A) + dx / 2, cy = PY(s, A) + dy / 2;
    cx = orc_dom_mod(cx, s->D);
    cy = orc_dom_mod(cy, s->D);
    PX(s, A) = cx; PY(s, A) = cy;
    PX(s, B) = cx; PY(s, B) = cy;
    vertex_invert_boundary_order(s, A);
    vertex_invert_boundary_order(s, B);
    s->ben[bnd] = boundary_compute_energy(s, bnd, s->o.GB_eps);   /* no GB_scaling: quirk A-5 */
    VX(s, A) = 0; VY(s, A) = 0;
    VX(s, B) = 0; VY(s, B) = 0;
}

/* topological.h:1016-1026, serial edge-index order */
static int apply_flippings(orc* s) {
    int cnt = 0;
    for (int b = 0; b < s->ne; b++)
        if (s->benab[b] && s->bcand[b]) { flip(s, b); cnt++; }
    return cnt;
}

/* ------------------------------------------------------------------ */
/* nucleation2.cu                                                       */
/* ------------------------------------------------------------------ */

/* geometry.h:346-361 */
static void init_vertex(orc* s, int v, double x, double y) {
    PX(s, v) = x; PY(s, v) = y;
    VX(s, v) = 0.0; VY(s, v) = 0.0;
    if (s->venergy) s->venergy[v] = 0.0;
    s->ven[v] = 1;
    s->vnuc[v] = 0;
    for (int i = 0; i < 3; i++) { VG(s, v, i) = -1; VB(s, v, i) = -1; }
    s->voted[v] = -1;
    reset_mod_list(s, v);
}

/* geometry.h:491-504 */
static void init_boundary(orc* s, int b, int ini, int end) {
    BI(s, b) = ini;
    vertex_add_boundary(s, ini, b);
    BE(s, b) = end;
    vertex_add_boundary(s, end, b);
    s->blen[b] = 0;
    s->bt[b] = SEVC_MAX_T_EXT;
    s->benab[b] = 1;
    s->bcand[b] = 0;
    s->nvotes[b] = 0;
    s->conf[b] = -1;
}

/* geometry.h:596-606 */
static void nucleate_grain(orc* s, int g) {
    for (int k = 0; k < s->S; k++) GV(s, g, k) = -1;
    s->glen[g] = 0;
    s->garea[g] = 0;
    s->gse[g] = 0;
    s->gori[g] = 0;
    s->genab[g] = 1;
    s->gcand[g] = 0;
    s->gfix[g] = 0;
    s->gtype[g] = T_NUCLEATED;
}

typedef struct {
    double Apos[2], Bpos[2], Cpos[2];
    int D, E, F, AD, AE, AF, g1, g2, g3;
    double dE;
} nucl_geom;

/* nucleation2.cu:116-212 */
static int generate_grain_for_nucleation(const orc* s, int A, double eps, double scaling, int do_scale,
                                         nucl_geom* r) {
    double Dm = s->D;
    r->AD = VB(s, A, 0); r->AE = VB(s, A, 1); r->AF = VB(s, A, 2);
    r->D = (A == BI(s, r->AD)) ? BE(s, r->AD) : BI(s, r->AD);
    r->E = (A == BI(s, r->AE)) ? BE(s, r->AE) : BI(s, r->AE);
    r->F = (A == BI(s, r->AF)) ? BE(s, r->AF) : BI(s, r->AF);
    int gAD[2], gAE[2], gAF[2];
    grains_intersect(s, A, r->D, gAD);
    grains_intersect(s, A, r->E, gAE);
    grains_intersect(s, A, r->F, gAF);
    if (gAD[0] == gAE[0] || gAD[0] == gAE[1]) { r->g2 = gAD[0]; r->g1 = gAD[1]; }
    else if (gAD[1] == gAE[0] || gAD[1] == gAE[1]) { r->g2 = gAD[1]; r->g1 = gAD[0]; }
    else return -1;
    if (gAE[0] == r->g2) r->g3 = gAE[1];
    else if (gAE[1] == r->g2) r->g3 = gAE[0];
    else return -1;
    if (r->g1 < 0 || r->g2 < 0 || r->g3 < 0) return -1;
    double ax = PX(s, A), ay = PY(s, A), rr = s->nucl_r[A];
    const int nb[3] = {r->D, r->E, r->F};
    double* out[3] = {r->Apos, r->Bpos, r->Cpos};
    for (int k = 0; k < 3; k++) {
        double tx = orc_wrap_dist(PX(s, nb[k]), ax, Dm);
        double ty = orc_wrap_dist(PY(s, nb[k]), ay, Dm);
        double w = rr / orc_mag(tx, ty);
        if (do_scale) w *= 0.99;
        double px = w * tx + ax, py = w * ty + ay;
        out[k][0] = orc_dom_mod(px, Dm);
        out[k][1] = orc_dom_mod(py, Dm);
    }
    double LAB = arclen(r->Apos[0], r->Apos[1], r->Bpos[0], r->Bpos[1], Dm);
    double LBC = arclen(r->Bpos[0], r->Bpos[1], r->Cpos[0], r->Cpos[1], Dm);
    double LCA = arclen(r->Cpos[0], r->Cpos[1], r->Apos[0], r->Apos[1], Dm);
    double xx[3], yy[3];
    xx[0] = ax; xx[1] = r->Cpos[0]; xx[2] = r->Apos[0];
    yy[0] = ay; yy[1] = r->Cpos[1]; yy[2] = r->Apos[1];
    double A1 = area_inplace(xx, yy, 3, Dm);
    xx[0] = ax; xx[1] = r->Apos[0]; xx[2] = r->Bpos[0];
    yy[0] = ay; yy[1] = r->Apos[1]; yy[2] = r->Bpos[1];
    double A2 = area_inplace(xx, yy, 3, Dm);
    xx[0] = ax; xx[1] = r->Bpos[0]; xx[2] = r->Cpos[0];
    yy[0] = ay; yy[1] = r->Bpos[1]; yy[2] = r->Cpos[1];
    double A3 = area_inplace(xx, yy, 3, Dm);
    double E0 = 0.0, E1;
    for (int i = 0; i < 3; i++) {
        if (do_scale) E0 += 0.99 * rr * s->ben[VB(s, A, i)];
        else E0 += rr * s->ben[VB(s, A, i)];
    }
#if ORC_MUTANT == 5
    E0 += A1 * s->gse[r->g2] + A2 * s->gse[r->g1] + A3 * s->gse[r->g3];
#elif ORC_MUTANT == 6
    E0 -= A1 * s->gse[r->g1] + A2 * s->gse[r->g2] + A3 * s->gse[r->g3];
#else
    E0 += A1 * s->gse[r->g1] + A2 * s->gse[r->g2] + A3 * s->gse[r->g3];
#endif
    double ori = s->nucl_ori[A];
#if ORC_MUTANT == 11
    E1 = gb_energy_func(ori - s->gori[r->g1], eps) * scaling * LAB;
    E1 += gb_energy_func(ori - s->gori[r->g2], eps) * scaling * LCA;
#else
    E1 = gb_energy_func(ori - s->gori[r->g1], eps) * scaling * LCA;
    E1 += gb_energy_func(ori - s->gori[r->g2], eps) * scaling * LAB;
#endif
    E1 += gb_energy_func(ori - s->gori[r->g3], eps) * scaling * LBC;
    r->dE = E1 - E0;
    return 0;
}

/* nucleation2.cu:239-282 with Philox(seed, (n, v, 1, 0)) */
static void compute_nucleation_factor(orc* s) {
    for (int v = 0; v < s->nv; v++) {
        if (!s->ven[v]) continue;
        s->dE[v] = 0;
        s->nfac[v] = 0;
        double ALs[3], minAL;
        for (int i = 0; i < 3; i++) ALs[i] = s->blen[VB(s, v, i)];
        minAL = ALs[0];
        for (int i = 0; i < 3; i++) if (minAL > ALs[i]) minAL = ALs[i];
        double u0, u1;
        sevc_uniform2(s->o.seed, (uint32_t)s->n, (uint32_t)v, STREAM_MC_SITE, 0u, &u0, &u1);
        s->nucl_r[v] = 0.8 * minAL * u0 + 0.1;
        s->nucl_ori[v] = 2 * SEVC_PI * u1;
        nucl_geom g0, g1;
        double de = 0.0, de_ = 0.0;
        if (generate_grain_for_nucleation(s, v, s->o.GB_eps, s->o.GB_scaling, 0, &g0) == 0) de = g0.dE;
        s->dE[v] = de;
        if (generate_grain_for_nucleation(s, v, s->o.GB_eps, s->o.GB_scaling, 1, &g1) == 0) de_ = g1.dE;
        s->nfac[v] = de - de_;
    }
}

/* nucleation2.cu:474-676, after the candidate checks; slots given. */
static int nucleate_at(orc* s, int A, int B, int C, int AB, int BC, int CA, int g) {
    nucl_geom r;
    /* nucleation2.cu:526 passes (scaling, eps) swapped; only dE differs */
    if (generate_grain_for_nucleation(s, A, s->o.GB_scaling, s->o.GB_eps, 1, &r) != 0) return -1;
#if ORC_MUTANT == 9
    init_vertex(s, B, r.Cpos[0], r.Cpos[1]);
    init_vertex(s, C, r.Bpos[0], r.Bpos[1]);
#else
    init_vertex(s, B, r.Bpos[0], r.Bpos[1]);
    init_vertex(s, C, r.Cpos[0], r.Cpos[1]);
#endif
    PX(s, A) = r.Apos[0]; PY(s, A) = r.Apos[1];
    for (int i = 0; i < 3; i++) VB(s, A, i) = -1;
    init_boundary(s, AB, A, B);
    init_boundary(s, BC, B, C);
    init_boundary(s, CA, C, A);
    vertex_add_boundary(s, A, r.AD);
    nucleate_grain(s, g);
    s->gori[g] = s->nucl_ori[A];
    grain_push_vertex(s, g, A);
    grain_push_vertex(s, g, B);
    grain_push_vertex(s, g, C);
#if ORC_MUTANT == 8
    for (int i = 0; i < 3; i++)
        if (VG(s, A, i) == r.g1) { VG(s, A, i) = g; break; }
#else
    for (int i = 0; i < 3; i++)
        if (VG(s, A, i) == r.g3) { VG(s, A, i) = g; break; }
#endif
    vertex_add_grain(s, B, g); vertex_add_grain(s, B, r.g2); vertex_add_grain(s, B, r.g3);
    vertex_add_grain(s, C, g); vertex_add_grain(s, C, r.g3); vertex_add_grain(s, C, r.g1);
    grain_add_vertex_after(s, r.g1, C, A);
    if (BI(s, r.AE) == A) grain_add_vertex_after(s, r.g2, B, BE(s, r.AE));
    else if (BE(s, r.AE) == A) grain_add_vertex_after(s, r.g2, B, BI(s, r.AE));
    grain_add_vertex_after(s, r.g3, C, A);
    grain_add_vertex_after(s, r.g3, B, C);
    grain_remove_vertex(s, r.g3, A);
    vertex_invert_boundary_order(s, B);
    vertex_invert_boundary_order(s, C);
    if (BI(s, r.AE) == A) BI(s, r.AE) = B;
    else if (BE(s, r.AE) == A) BE(s, r.AE) = B;
    VB(s, B, 2) = r.AE;
    if (BI(s, r.AF) == A) BI(s, r.AF) = C;
    else BE(s, r.AF) = C;
    VB(s, C, 2) = r.AF;
    return 0;
}

/* rank-th (0-based) disabled slot of an enable array, or -1 */
static int kth_free(const uint8_t* en, int n, int rank) {
    for (int i = 0; i < n; i++)
        if (!en[i]) { if (rank == 0) return i; rank--; }
    return -1;
}

/* nucleation2.cu:292-440: the random alive vertex and the Metropolis rule
 * (accept if dE < 0, or if d2E < 0 and u < exp(-MC_k dE)), after
 * compute_nucleation_factor.  Returns the site or -1; vnuc[site] holds the
 * decision. */
static int mc_pick_and_accept(orc* s) {
    int n_alive = 0;
    for (int v = 0; v < s->nv; v++) n_alive += s->ven[v] ? 1 : 0;
    if (n_alive == 0) return -1;
    double u0, u1;
    sevc_uniform2(s->o.seed, (uint32_t)s->n, 0u, STREAM_MC_PICK, 0u, &u0, &u1);
    int k = (int)(u0 * (double)n_alive);
    if (k >= n_alive) k = n_alive - 1;
    int A = -1;
    for (int v = 0; v < s->nv; v++)
        if (s->ven[v]) { if (k == 0) { A = v; break; } k--; }
    if (A < 0) return -1;
    s->vnuc[A] = 0;
    if (s->dE[A] < 0) s->vnuc[A] = 1;
    else if (s->nfac[A] < 0) {
        double thr = sevc_exp(-s->o.MC_k * s->dE[A]);
#if ORC_MUTANT == 7
        if (u1 > thr) s->vnuc[A] = 1;
#else
        if (u1 < thr) s->vnuc[A] = 1;
#endif
    }
    return A;
}

/* nucleation2.cu:292-443 + :33-72 + :474 (reference MC mode). */
static int nucleation_mc(orc* s) {
    compute_nucleation_factor(s);
    int A = mc_pick_and_accept(s);
    if (A < 0) return 0;
    if (!s->vnuc[A]) return 0;
    int g = kth_free(s->genab, s->ng, 0);
    int b0 = kth_free(s->benab, s->ne, 0), b1 = kth_free(s->benab, s->ne, 1), b2 = kth_free(s->benab, s->ne, 2);
    int v0 = kth_free(s->ven, s->nv, 0), v1 = kth_free(s->ven, s->nv, 1);
    if (g < 0 || b0 < 0 || b1 < 0 || b2 < 0 || v0 < 0 || v1 < 0) return 0;
    if (nucleate_at(s, A, v0, v1, b0, b1, b2, g) != 0) return 0;
    return 1;
}

/* Rate mode (extension): each enabled grain g nucleates with probability
 * nucleation_rate * SE_g at one of its vertices chosen uniformly; radius
 * nucleation_scaling * min(L) * u, orientation 2 pi u.  An event is kept
 * when no earlier candidate (lower grain index) touches any of its three
 * grains; kept events take the free slots in rank order. */
/* The draws of the rate mode: every enabled grain draws with probability
 * rate * SE_g; a drawing grain names one of its vertices, the radius and
 * the orientation.  Returns the number of candidate sites; ev_src[e] is the
 * grain that drew candidate e (may be NULL). */
static int rate_draws(orc* s, int* ev_v, double* ev_r, double* ev_o, int* ev_g, int* ev_src) {
    int ne = 0;
    for (int g = 0; g < s->ng; g++) {
        if (!s->genab[g] || s->glen[g] < 1) continue;
#if ORC_MUTANT == 10
        double p = s->o.nucleation_rate;
#else
        double p = s->o.nucleation_rate * s->gse[g];
#endif
        double u0, u1, u2, u3;
        sevc_uniform2(s->o.seed, (uint32_t)s->n, (uint32_t)g, STREAM_RATE, 0u, &u0, &u1);
        if (!(u0 < p)) continue;
        sevc_uniform2(s->o.seed, (uint32_t)s->n, (uint32_t)g, STREAM_RATE, 1u, &u2, &u3);
        int k = (int)(u1 * (double)s->glen[g]);
        if (k >= s->glen[g]) k = s->glen[g] - 1;
        int A = GV(s, g, k);
        double minAL = s->blen[VB(s, A, 0)];
        for (int i = 0; i < 3; i++) if (minAL > s->blen[VB(s, A, i)]) minAL = s->blen[VB(s, A, i)];
        /* a vertex on a zero-length boundary (both ends of a fresh flip sit
         * at one point) cannot host the nucleus: it is not a candidate */
        if (!(minAL > 0)) continue;
        ev_v[ne] = A;
        ev_r[ne] = s->o.nucleation_scaling * minAL * u2;
        ev_o[ne] = 2 * SEVC_PI * u3;
        for (int i = 0; i < 3; i++) ev_g[3 * ne + i] = VG(s, A, i);
        if (ev_src) ev_src[ne] = g;
        ne++;
    }
    return ne;
}

static int nucleation_rate(orc* s) {
    int cap = 0;
    for (int g = 0; g < s->ng; g++) cap += s->genab[g] ? 1 : 0;
    int* ev_v = (int*)malloc(sizeof(int) * (size_t)(cap + 1));
    double* ev_r = (double*)malloc(sizeof(double) * (size_t)(cap + 1));
    double* ev_o = (double*)malloc(sizeof(double) * (size_t)(cap + 1));
    int* ev_g = (int*)malloc(sizeof(int) * (size_t)(cap + 1) * 3);
    int ne = rate_draws(s, ev_v, ev_r, ev_o, ev_g, NULL);
    int kept = 0, made = 0;
    /* free slots ranked before any insertion of this step */
    int nfg = 0, nfb = 0, nfv = 0;
    for (int g = 0; g < s->ng; g++) nfg += !s->genab[g];
    for (int b = 0; b < s->ne; b++) nfb += !s->benab[b];
    for (int v = 0; v < s->nv; v++) nfv += !s->ven[v];
    int* fg = (int*)malloc(sizeof(int) * (size_t)(nfg + 1));
    int* fb = (int*)malloc(sizeof(int) * (size_t)(nfb + 1));
    int* fv = (int*)malloc(sizeof(int) * (size_t)(nfv + 1));
    nfg = nfb = nfv = 0;
    for (int g = 0; g < s->ng; g++) if (!s->genab[g]) fg[nfg++] = g;
    for (int b = 0; b < s->ne; b++) if (!s->benab[b]) fb[nfb++] = b;
    for (int v = 0; v < s->nv; v++) if (!s->ven[v]) fv[nfv++] = v;
    for (int e = 0; e < ne; e++) {
        int clash = 0;
        for (int f = 0; f < e && !clash; f++)
            for (int i = 0; i < 3 && !clash; i++)
                for (int j = 0; j < 3; j++)
                    if (ev_g[3 * e + i] == ev_g[3 * f + j]) { clash = 1; break; }
        if (clash) continue;
        int r = kept++;
        if (r + 1 > nfg || 3 * r + 3 > nfb || 2 * r + 2 > nfv) continue;
        int A = ev_v[e];
        s->nucl_r[A] = ev_r[e];
        s->nucl_ori[A] = ev_o[e];
        if (nucleate_at(s, A, fv[2 * r], fv[2 * r + 1], fb[3 * r], fb[3 * r + 1], fb[3 * r + 2], fg[r]) == 0)
            made++;
    }
    free(ev_v); free(ev_r); free(ev_o); free(ev_g); free(fg); free(fb); free(fv);
    return made;
}

/* ------------------------------------------------------------------ */
/* the step (main.cu:120-394)                                          */
/* ------------------------------------------------------------------ */

static int count_enabled(const uint8_t* en, int n) {
    int c = 0;
    for (int i = 0; i < n; i++) c += en[i] ? 1 : 0;
    return c;
}

static double nucleated_pct(const orc* s) {
    uint8_t* m = (uint8_t*)malloc((size_t)s->ng + 1);
    for (int g = 0; g < s->ng; g++) m[g] = s->genab[g] && s->gtype[g] == T_NUCLEATED;
    double a = orc_sum_chunked(s->garea, m, s->ng);
    free(m);
    return 100 * a / (s->D * s->D);
}

static int time_ok(const orc* s) {
    double tmax = s->o.tmax;
    int nmax = s->o.nmax;
    if (tmax <= 0 && nmax <= 0) return 1;
    return (s->t <= tmax && tmax > 0) || (s->n <= nmax && nmax > 0);
}

static void one_step(orc* s) {
    const sevc_options* o = &s->o;
    compute_boundary_arclengths(s);                         /* main.cu:124 */
    compute_grain_areas(s);                                 /* :127 */
    compute_boundary_energies(s);                           /* :130 */
    if (s->diag) compute_vertex_energies(s);                /* :133 */

    if (o->events && o->do_nucleation) {
        int gap = o->nucleation_gap > 0 ? o->nucleation_gap : 1;
        int made = 0, ran = 0;
        if (o->nucleation_mode == SEVC_NUCL_MC) {
            int gate = s->o.nucleation_gate >= 0 ? s->o.nucleation_gate : 5 * s->n_grains0 / 6;
            if (s->curr_grains <= gate) s->allow_nucl = 1;   /* :137 */
            if ((s->n % gap) == 0 && s->allow_nucl) { made = nucleation_mc(s); ran = 1; }
        } else if ((s->n % gap) == 0) {
            made = nucleation_rate(s);
            ran = 1;
        }
        s->n_nucl += made;
        if (ran) {                                          /* :171-181 */
            compute_boundary_arclengths(s);
            compute_grain_areas(s);
            compute_boundary_energies(s);
            if (s->diag) compute_vertex_energies(s);
        }
    }

    compute_vertex_velocities(s);                           /* :203 */
    if (s->diag) compute_grain_dadts(s);                    /* :206 */
    compute_t_ext(s);                                       /* :212 */
    double dt = step_dt(s);
    s->dt_last = dt;

    if (o->events) {
        label_boundaries(s, dt);                            /* :216 */
        polling_system(s);                                  /* :219 */
        check_conflictive_boundaries(s);                    /* :222 */
        unlabel_conflictive_boundaries(s);                  /* :224 */
        for (int it = 0; it < 20; it++) {                   /* :227-250 */
            if (two_three_sided_grains(s) == 0) break;
            s->n_two += remove_two_grains_configurations(s);
            fix_grain_lists(s);
        }
        label_3sided_grains(s, dt);                         /* :252-255 */
        s->n_removals += remove_grains(s);                  /* :259 */
        fix_grain_lists(s);                                 /* :262 */
        compute_grain_areas(s);                             /* :266 */
        unlabel_boundaries_st2(s);                          /* :282 */
        s->n_flips += apply_flippings(s);                   /* :289 */
        fix_grain_lists(s);                                 /* :294 */
    }
    update_vertices_positions(s, dt);                       /* :314 */
    s->curr_grains = count_enabled(s->genab, s->ng);        /* :317 */
    s->n++;                                                 /* :388-389 */
    if (o->adaptive_dt) s->t += dt;
    else s->t = s->n * o->dt;
    s->diag = 0;
}

int orc_step(orc* s, int32_t n_steps, sevc_step_stats* st) {
    long long vsteps = 0;
    long long f0 = s->n_flips, r0 = s->n_removals, t0 = s->n_two, n0 = s->n_nucl;
    int done_steps = 0;
    for (int k = 0; k < n_steps; k++) {
        if (s->done) break;
        if (!time_ok(s) || !(s->min_grains <= 0 || s->curr_grains >= s->min_grains)) { s->done = 1; break; }
        vsteps += count_enabled(s->ven, s->nv);
        one_step(s);
        done_steps++;
        if (s->status != SEVC_OK) break;
        if (s->curr_grains <= s->min_grains && s->min_grains > 0) s->done = 1;   /* :391 */
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->steps_done = done_steps;
        st->n = s->n;
        st->t = s->t;
        st->dt_last = s->dt_last;
        st->n_grains = s->curr_grains;
        st->n_vertices = count_enabled(s->ven, s->nv);
        st->nucleated_area_pct = nucleated_pct(s);
        st->done = s->done;
        st->status = s->status;
        st->n_flips = s->n_flips - f0;
        st->n_removals = s->n_removals - r0;
        st->n_two_grain = s->n_two - t0;
        st->n_nucleations = s->n_nucl - n0;
        st->n_vertex_steps = vsteps;
    }
    return s->status;
}

int orc_compute_velocities(orc* s) {
    compute_boundary_arclengths(s);
    compute_grain_areas(s);
    compute_boundary_energies(s);
    compute_vertex_velocities(s);
    compute_t_ext(s);
    return s->status;
}

void orc_request_diagnostics(orc* s) { s->diag = 1; }

/* ------------------------------------------------------------------ */
/* create / copy                                                        */
/* ------------------------------------------------------------------ */

#define ALLOC(p, n) do { (p) = calloc((size_t)(n) + 1, sizeof(*(p))); if (!(p)) goto fail; } while (0)
#define COPY(dst, src, n) do { if (src) memcpy((dst), (src), sizeof(*(dst)) * (size_t)(n)); } while (0)

void orc_destroy(orc* s) {
    if (!s) return;
    free(s->pos); free(s->vel); free(s->venergy); free(s->vparts);
    free(s->vb); free(s->vg); free(s->ven);
    free(s->bvx); free(s->ben); free(s->blen); free(s->bt); free(s->benab); free(s->bcand);
    free(s->gvx); free(s->glen); free(s->garea); free(s->gdadt); free(s->gori); free(s->gse);
    free(s->genab); free(s->gtype);
    free(s->mfix); free(s->mgrn); free(s->mvrt); free(s->mprev);
    free(s->voted); free(s->nvotes); free(s->conf); free(s->gcand); free(s->gfix);
    free(s->vnuc); free(s->cand_prev);
    free(s->nucl_r); free(s->nucl_ori); free(s->dE); free(s->nfac);
    free(s->wtraj); free(s->warea); free(s->webase);
    free(s);
}

orc* orc_create(const sevc_options* o, const sevc_mesh* m) {
    orc* s = (orc*)calloc(1, sizeof(orc));
    if (!s) return NULL;
    s->o = *o;
    s->nv = m->n_vertices; s->ne = m->n_boundaries; s->ng = m->n_grains;
    s->S = m->max_vrt_per_grn;
    s->D = m->domain_bound;
    if (s->S > 128) goto fail;
    ALLOC(s->pos, 2 * s->nv); ALLOC(s->vel, 2 * s->nv); ALLOC(s->venergy, s->nv);
    ALLOC(s->vparts, 4 * s->nv);
    ALLOC(s->vb, 3 * s->nv); ALLOC(s->vg, 3 * s->nv); ALLOC(s->ven, s->nv);
    ALLOC(s->bvx, 2 * s->ne); ALLOC(s->ben, s->ne); ALLOC(s->blen, s->ne); ALLOC(s->bt, s->ne);
    ALLOC(s->benab, s->ne); ALLOC(s->bcand, s->ne);
    ALLOC(s->gvx, (size_t)s->ng * s->S); ALLOC(s->glen, s->ng); ALLOC(s->garea, s->ng);
    ALLOC(s->gdadt, s->ng); ALLOC(s->gori, s->ng); ALLOC(s->gse, s->ng);
    ALLOC(s->genab, s->ng); ALLOC(s->gtype, s->ng);
    ALLOC(s->mfix, 4 * s->nv); ALLOC(s->mgrn, 4 * s->nv); ALLOC(s->mvrt, 4 * s->nv); ALLOC(s->mprev, 4 * s->nv);
    ALLOC(s->voted, s->nv); ALLOC(s->nvotes, s->ne); ALLOC(s->conf, s->ne);
    ALLOC(s->gcand, s->ng); ALLOC(s->gfix, s->ng); ALLOC(s->vnuc, s->nv); ALLOC(s->cand_prev, s->ne);
    ALLOC(s->nucl_r, s->nv); ALLOC(s->nucl_ori, s->nv); ALLOC(s->dE, s->nv); ALLOC(s->nfac, s->nv);
    COPY(s->pos, m->pos, 2 * s->nv); COPY(s->vel, m->vel, 2 * s->nv);
    COPY(s->venergy, m->v_energy, s->nv); COPY(s->vparts, m->v_vel_parts, 4 * s->nv);
    COPY(s->vb, m->v_bnd, 3 * s->nv); COPY(s->vg, m->v_grn, 3 * s->nv); COPY(s->ven, m->v_enabled, s->nv);
    COPY(s->bvx, m->b_vtx, 2 * s->ne); COPY(s->ben, m->b_energy, s->ne); COPY(s->blen, m->b_arclength, s->ne);
    COPY(s->bt, m->b_t_ext, s->ne); COPY(s->benab, m->b_enabled, s->ne); COPY(s->bcand, m->b_candidate, s->ne);
    COPY(s->gvx, m->g_vtx, (size_t)s->ng * s->S); COPY(s->glen, m->g_vlen, s->ng);
    COPY(s->garea, m->g_area, s->ng); COPY(s->gdadt, m->g_dAdt, s->ng);
    COPY(s->gori, m->g_orientation, s->ng); COPY(s->gse, m->g_SE, s->ng);
    COPY(s->genab, m->g_enabled, s->ng); COPY(s->gtype, m->g_type, s->ng);
    for (int v = 0; v < s->nv; v++) { reset_mod_list(s, v); s->voted[v] = -1; }
    for (int b = 0; b < s->ne; b++) s->conf[b] = -1;
    s->n = 0; s->t = 0; s->dt_last = o->dt;
    s->n_grains0 = count_enabled(s->genab, s->ng);
    s->curr_grains = s->n_grains0;
    s->min_grains = o->leave_percent < 0 ? -1 : (int)floor(o->leave_percent * s->n_grains0);
    s->status = SEVC_OK;
    return s;
fail:
    orc_destroy(s);
    return NULL;
}

int orc_get_mesh(const orc* s, sevc_mesh* m) {
    if (m->n_vertices != s->nv || m->n_boundaries != s->ne || m->n_grains != s->ng || m->max_vrt_per_grn != s->S)
        return SEVC_E_INVALID_ARG;
    m->domain_bound = s->D;
    COPY(m->pos, s->pos, 2 * s->nv); COPY(m->vel, s->vel, 2 * s->nv);
    if (m->v_energy) COPY(m->v_energy, s->venergy, s->nv);
    if (m->v_vel_parts) COPY(m->v_vel_parts, s->vparts, 4 * s->nv);
    COPY(m->v_bnd, s->vb, 3 * s->nv); COPY(m->v_grn, s->vg, 3 * s->nv); COPY(m->v_enabled, s->ven, s->nv);
    COPY(m->b_vtx, s->bvx, 2 * s->ne);
    if (m->b_energy) COPY(m->b_energy, s->ben, s->ne);
    if (m->b_arclength) COPY(m->b_arclength, s->blen, s->ne);
    if (m->b_t_ext) COPY(m->b_t_ext, s->bt, s->ne);
    COPY(m->b_enabled, s->benab, s->ne);
    if (m->b_candidate) COPY(m->b_candidate, s->bcand, s->ne);
    COPY(m->g_vtx, s->gvx, (size_t)s->ng * s->S); COPY(m->g_vlen, s->glen, s->ng);
    if (m->g_area) COPY(m->g_area, s->garea, s->ng);
    if (m->g_dAdt) COPY(m->g_dAdt, s->gdadt, s->ng);
    COPY(m->g_orientation, s->gori, s->ng); COPY(m->g_SE, s->gse, s->ng);
    COPY(m->g_enabled, s->genab, s->ng); COPY(m->g_type, s->gtype, s->ng);
    return SEVC_OK;
}

/* ------------------------------------------------------------------ */
/* topology-fixed window and its adjoint (extension)                    */
/* ------------------------------------------------------------------ */

/* Forward over a window of `steps` steps with the topology held fixed:
 * the analytic model v = mobility * F(x), F = -grad E, explicit Euler with
 * the fixed dt and the wrap of update_vertices_positions
 * (calculus.h:391-404); no events, no nucleation.  Boundary energies are
 * evaluated once (they depend on orientations only).  Each step is the
 * analytic-model step of orc_step with events off: areas, velocities,
 * translation.  The trajectory is kept for the reverse pass. */
int orc_window_forward(orc* s, int32_t steps) {
    if (steps < 0) return SEVC_E_INVALID_ARG;
    free(s->wtraj); free(s->warea); free(s->webase);
    const size_t n2 = 2 * (size_t)s->nv;
    s->wtraj = (double*)malloc(sizeof(double) * n2 * ((size_t)steps + 1));
    s->warea = (double*)malloc(sizeof(double) * (size_t)s->ng * ((size_t)steps + 1));
    s->webase = (double*)calloc((size_t)s->ne + 1, sizeof(double));
    if (!s->wtraj || !s->warea || !s->webase) return SEVC_E_NOMEM;
    s->wN = steps;
    compute_boundary_energies(s);
    for (int b = 0; b < s->ne; b++)
        if (s->benab[b]) s->webase[b] = boundary_compute_energy(s, b, s->o.GB_eps);
    memcpy(s->wtraj, s->pos, sizeof(double) * n2);
    const double mu = s->o.mobility, dt = s->o.dt;
    for (int n = 0; n < steps; n++) {
        compute_grain_areas(s);
        memcpy(s->warea + (size_t)n * s->ng, s->garea, sizeof(double) * (size_t)s->ng);
        for (int v = 0; v < s->nv; v++) {
            if (!s->ven[v]) continue;
            double fx, fy;
            analytic_force(s, v, &fx, &fy);
            VX(s, v) = fx * mu;
            VY(s, v) = fy * mu;
        }
        update_vertices_positions(s, dt);
        memcpy(s->wtraj + (size_t)(n + 1) * n2, s->pos, sizeof(double) * n2);
        s->n += 1;
        if (s->o.adaptive_dt) s->t += dt;
        else s->t = s->n * dt;
    }
    return SEVC_OK;
}

/* One reverse step at vertex v: with lambda = dL/dx_{n+1} and the state
 * x_n (positions X, grain areas A),
 *   lambda_n = lambda - dt mu H lambda          (H = Hessian of E at x_n)
 *   acc_mu  += dt (lambda . F)
 *   acc_s   += dt mu (lambda . dF/dGB_scaling),  dF/ds = -sum_b e_b u_b
 *   acc_se_j += dt mu (lambda . dF/dSE_g),       dF/dSE_g = -sign(A_g) dA_g/dp_v
 * Edge b (other end w, d = wrap(p_v - p_w), L = |d|, u = d/L) adds
 * (gamma_b/L)(I - u u^T)(lambda_v - lambda_w) to (H lambda)_v; grain g
 * (list neighbours prev, next) adds
 * SE_g sign(A_g)/2 (lambda_next.y - lambda_prev.y, lambda_prev.x - lambda_next.x). */
static void adjoint_vertex(const orc* s, const double* X, const double* A, const double* lam, double* lam2,
                           double* acc_mu, double* acc_s, double* acc_se, int v) {
    const double D = s->D, mu = s->o.mobility, dt = s->o.dt, dtm = dt * mu;
    const double x = X[2 * v], y = X[2 * v + 1];
    const double lx = lam[2 * v], ly = lam[2 * v + 1];
    double ex = 0.0, ey = 0.0, gx = 0.0, gy = 0.0;
    double hex = 0.0, hey = 0.0, hgx = 0.0, hgy = 0.0, sx = 0.0, sy = 0.0;
    for (int j = 0; j < 3; j++) {
        int b = VB(s, v, j);
        int w = (BI(s, b) == v) ? BE(s, b) : BI(s, b);
        double dx = orc_wrap_dist(x, X[2 * w], D);
        double dy = orc_wrap_dist(y, X[2 * w + 1], D);
        double L = orc_mag(dx, dy);
        if (L > 0) {
            double ux = dx / L, uy = dy / L;
            ex += s->ben[b] * ux;
            ey += s->ben[b] * uy;
            double wx = lx - lam[2 * w], wy = ly - lam[2 * w + 1];
            double uw = ux * wx + uy * wy;
            double c = s->ben[b] / L;
            hex += c * (wx - ux * uw);
            hey += c * (wy - uy * uw);
            sx += s->webase[b] * ux;
            sy += s->webase[b] * uy;
        }
    }
    for (int j = 0; j < 3; j++) {
        int g = VG(s, v, j);
        int n = s->glen[g];
        int k = grain_contains_vertex(s, g, v);
        if (k < 0 || n < 3) continue;
        int wp = GV(s, g, (k + n - 1) % n), wn = GV(s, g, (k + 1) % n);
        double dpx = orc_wrap_dist(x, X[2 * wp], D), dpy = orc_wrap_dist(y, X[2 * wp + 1], D);
        double dnx = orc_wrap_dist(X[2 * wn], x, D), dny = orc_wrap_dist(X[2 * wn + 1], y, D);
        double sg = A[g] < 0 ? -1.0 : 1.0;
        double c = s->gse[g] * sg * 0.5;
        gx += c * (dpy + dny);
        gy += c * (-(dpx + dnx));
        hgx += c * (lam[2 * wn + 1] - lam[2 * wp + 1]);
        hgy += c * (lam[2 * wp] - lam[2 * wn]);
        double q = sg * 0.5;
        double tse = -(lx * (q * (dpy + dny)) + ly * (q * (-(dpx + dnx))));
        acc_se[3 * v + j] += dtm * tse;
    }
    double fx = -(ex + gx), fy = -(ey + gy);
    double hx = hex + hgx, hy = hey + hgy;
    acc_mu[v] += dt * (lx * fx + ly * fy);
    acc_s[v] += dtm * (-(lx * sx + ly * sy));
    lam2[2 * v] = lx - dtm * hx;
    lam2[2 * v + 1] = ly - dtm * hy;
}

/* Reverse pass over the last window: dL_dpos = dL/dx_N (2 per vertex
 * slot).  Outputs dL/dx_0 (d_pos, 2 per vertex slot), dL/dSE (per grain
 * slot: the grain's list order sum of its vertices' terms), dL/dGB_scaling
 * and dL/dmobility (orc_sum_chunked over the vertex slots). */
int orc_adjoint_window(orc* s, const double* dL_dpos, double* d_pos, double* d_SE, double* d_gb, double* d_mu) {
    if (!s->wtraj) return SEVC_E_STATE;
    const int nv = s->nv;
    double* lam = (double*)calloc(2 * (size_t)nv + 2, sizeof(double));
    double* lam2 = (double*)calloc(2 * (size_t)nv + 2, sizeof(double));
    double* acc_mu = (double*)calloc((size_t)nv + 1, sizeof(double));
    double* acc_s = (double*)calloc((size_t)nv + 1, sizeof(double));
    double* acc_se = (double*)calloc(3 * (size_t)nv + 3, sizeof(double));
    if (!lam || !lam2 || !acc_mu || !acc_s || !acc_se) {
        free(lam); free(lam2); free(acc_mu); free(acc_s); free(acc_se);
        return SEVC_E_NOMEM;
    }
    for (int v = 0; v < nv; v++)
        if (s->ven[v]) { lam[2 * v] = dL_dpos[2 * v]; lam[2 * v + 1] = dL_dpos[2 * v + 1]; }
    for (int n = s->wN - 1; n >= 0; n--) {
        const double* X = s->wtraj + (size_t)n * 2 * nv;
        const double* A = s->warea + (size_t)n * s->ng;
        for (int v = 0; v < nv; v++) {
            if (!s->ven[v]) { lam2[2 * v] = 0.0; lam2[2 * v + 1] = 0.0; continue; }
            adjoint_vertex(s, X, A, lam, lam2, acc_mu, acc_s, acc_se, v);
        }
        double* t = lam; lam = lam2; lam2 = t;
    }
    if (d_pos) memcpy(d_pos, lam, sizeof(double) * 2 * (size_t)nv);
    if (d_mu) *d_mu = orc_sum_chunked(acc_mu, s->ven, nv);
    if (d_gb) *d_gb = orc_sum_chunked(acc_s, s->ven, nv);
    if (d_SE)
        for (int g = 0; g < s->ng; g++) {
            double r = 0.0;
            if (s->genab[g])
                for (int k = 0; k < s->glen[g]; k++) {
                    int v = GV(s, g, k);
                    for (int j = 0; j < 3; j++)
                        if (VG(s, v, j) == g) { r += acc_se[3 * v + j]; break; }
                }
            d_SE[g] = r;
        }
    free(lam); free(lam2); free(acc_mu); free(acc_s); free(acc_se);
    return SEVC_OK;
}

/* ------------------------------------------------------------------ */
/* test hooks: single stages of the step on a state, for the tests that  */
/* pin the oracle (brute force, closed forms, statistics)               */
/* ------------------------------------------------------------------ */

/* The vote rounds (stages & 1: polling_system, topological.h:549-578) and
 * the conflict filter (stages & 2: topological.h:83-155) on given
 * extinction times and candidate flags (one per boundary slot); cand
 * receives the flags after the stages, conf the conflict records. */
int orc_poll(orc* s, const double* t_ext, uint8_t* cand, int32_t* conf, int32_t stages) {
    for (int b = 0; b < s->ne; b++) {
        s->bt[b] = t_ext[b];
        s->bcand[b] = cand[b] ? 1 : 0;
        s->conf[b] = -1;
    }
    if (stages & 1) polling_system(s);
    if (stages & 2) {
        check_conflictive_boundaries(s);
        unlabel_conflictive_boundaries(s);
    }
    for (int b = 0; b < s->ne; b++) {
        cand[b] = s->bcand[b];
        if (conf) conf[b] = s->conf[b];
    }
    return 0;
}

/* generate_grain_for_nucleation (nucleation2.cu:116-212) at vertex A with
 * radius r and orientation ori, on the current lengths and energies:
 * xy = (A', B, C) positions, *dE, ids = (D, E, F, AD, AE, AF, g1, g2, g3). */
int orc_nucl_geometry(orc* s, int32_t A, double r, double ori, int32_t do_scale, double* xy, double* dE,
                      int32_t* ids) {
    if (A < 0 || A >= s->nv || !s->ven[A]) return SEVC_E_INVALID_ARG;
    compute_boundary_arclengths(s);
    compute_boundary_energies(s);
    s->nucl_r[A] = r;
    s->nucl_ori[A] = ori;
    nucl_geom g;
    if (generate_grain_for_nucleation(s, A, s->o.GB_eps, s->o.GB_scaling, do_scale, &g) != 0) return SEVC_E_STATE;
    for (int k = 0; k < 2; k++) { xy[k] = g.Apos[k]; xy[2 + k] = g.Bpos[k]; xy[4 + k] = g.Cpos[k]; }
    *dE = g.dE;
    const int v[9] = {g.D, g.E, g.F, g.AD, g.AE, g.AF, g.g1, g.g2, g.g3};
    for (int k = 0; k < 9; k++) ids[k] = v[k];
    return 0;
}

/* nucleate (nucleation2.cu:474-676) at vertex A with radius r and
 * orientation ori into the first free slots (one grain, three boundaries,
 * two vertices); slots = (g, AB, BC, CA, B, C) used. */
int orc_nucleate_vertex(orc* s, int32_t A, double r, double ori, int32_t* slots) {
    if (A < 0 || A >= s->nv || !s->ven[A]) return SEVC_E_INVALID_ARG;
    compute_boundary_arclengths(s);
    compute_boundary_energies(s);
    s->nucl_r[A] = r;
    s->nucl_ori[A] = ori;
    int g = kth_free(s->genab, s->ng, 0);
    int b0 = kth_free(s->benab, s->ne, 0), b1 = kth_free(s->benab, s->ne, 1), b2 = kth_free(s->benab, s->ne, 2);
    int v0 = kth_free(s->ven, s->nv, 0), v1 = kth_free(s->ven, s->nv, 1);
    if (g < 0 || b0 < 0 || b1 < 0 || b2 < 0 || v0 < 0 || v1 < 0) return SEVC_E_GRAIN_FULL;
    if (nucleate_at(s, A, v0, v1, b0, b1, b2, g) != 0) return SEVC_E_STATE;
    compute_boundary_arclengths(s);
    compute_grain_areas(s);
    const int u[6] = {g, b0, b1, b2, v0, v1};
    if (slots) for (int k = 0; k < 6; k++) slots[k] = u[k];
    return s->status;
}

/* One Metropolis trial of the MC mode at step counter `step`
 * (compute_nucleation_factor, the random alive vertex, the acceptance
 * rule): returns 1 if accepted, 0 if not; *site, *dE and *nfac (Delta E and
 * Delta^2 E of the site) are written.  Nothing is inserted. */
int orc_mc_trial(orc* s, int32_t step, int32_t* site, double* dE, double* nfac) {
    const int n0 = s->n;
    s->n = step;
    compute_boundary_arclengths(s);
    compute_boundary_energies(s);
    compute_nucleation_factor(s);
    const int A = mc_pick_and_accept(s);
    s->n = n0;
    *site = A;
    if (A < 0) return 0;
    *dE = s->dE[A];
    *nfac = s->nfac[A];
    return s->vnuc[A] ? 1 : 0;
}

/* The rate mode's draws at step counter `step`: the grain that drew each
 * candidate site and the site (at most `max`); returns their number. */
int orc_rate_draws(orc* s, int32_t step, int32_t* grains, int32_t* sites, int32_t max) {
    const int n0 = s->n;
    s->n = step;
    compute_boundary_arclengths(s);
    int* ev_v = (int*)malloc(sizeof(int) * (size_t)(s->ng + 1));
    double* ev_r = (double*)malloc(sizeof(double) * (size_t)(s->ng + 1));
    double* ev_o = (double*)malloc(sizeof(double) * (size_t)(s->ng + 1));
    int* ev_g = (int*)malloc(sizeof(int) * (size_t)(s->ng + 1) * 3);
    int* ev_s = (int*)malloc(sizeof(int) * (size_t)(s->ng + 1));
    const int ne = rate_draws(s, ev_v, ev_r, ev_o, ev_g, ev_s);
    for (int e = 0; e < ne && e < max; e++) { grains[e] = ev_s[e]; sites[e] = ev_v[e]; }
    free(ev_v); free(ev_r); free(ev_o); free(ev_g); free(ev_s);
    s->n = n0;
    return ne;
}
