/*
 * sevc_oracle_io.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Restatements of the reference loader (io.h:66-264) and of the
 * consistency checker check_data (utils.h:85-336), written independently
 * of the product loader so that the two can be compared.
 */
#include "sevc_oracle.h"
#include "../include/sevc_math.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

static double wrapd(double dest, double orig, double D) { return orc_wrap_dist(dest, orig, D); }

int orc_mesh_alloc(sevc_mesh* m, int nv, int ne, int ng, int S);
void orc_mesh_free(sevc_mesh* m);

int orc_mesh_alloc(sevc_mesh* m, int nv, int ne, int ng, int S) {
    memset(m, 0, sizeof(*m));
    m->n_vertices = nv; m->n_boundaries = ne; m->n_grains = ng; m->max_vrt_per_grn = S;
    m->pos = calloc((size_t)2 * nv + 1, sizeof(double));
    m->vel = calloc((size_t)2 * nv + 1, sizeof(double));
    m->v_bnd = malloc(sizeof(int32_t) * ((size_t)3 * nv + 1));
    m->v_grn = malloc(sizeof(int32_t) * ((size_t)3 * nv + 1));
    m->v_enabled = calloc((size_t)nv + 1, 1);
    m->v_energy = calloc((size_t)nv + 1, sizeof(double));
    m->v_vel_parts = calloc((size_t)4 * nv + 1, sizeof(double));
    m->b_vtx = malloc(sizeof(int32_t) * ((size_t)2 * ne + 1));
    m->b_energy = calloc((size_t)ne + 1, sizeof(double));
    m->b_arclength = calloc((size_t)ne + 1, sizeof(double));
    m->b_t_ext = calloc((size_t)ne + 1, sizeof(double));
    m->b_enabled = calloc((size_t)ne + 1, 1);
    m->b_candidate = calloc((size_t)ne + 1, 1);
    m->g_vtx = malloc(sizeof(int32_t) * ((size_t)ng * S + 1));
    m->g_vlen = calloc((size_t)ng + 1, sizeof(int32_t));
    m->g_area = calloc((size_t)ng + 1, sizeof(double));
    m->g_dAdt = calloc((size_t)ng + 1, sizeof(double));
    m->g_orientation = calloc((size_t)ng + 1, sizeof(double));
    m->g_SE = calloc((size_t)ng + 1, sizeof(double));
    m->g_enabled = calloc((size_t)ng + 1, 1);
    m->g_type = calloc((size_t)ng + 1, 1);
    if (!m->pos || !m->vel || !m->v_bnd || !m->v_grn || !m->v_enabled || !m->b_vtx || !m->g_vtx) return -1;
    for (size_t i = 0; i < (size_t)3 * nv; i++) { m->v_bnd[i] = -1; m->v_grn[i] = -1; }
    for (size_t i = 0; i < (size_t)2 * ne; i++) m->b_vtx[i] = -1;
    for (size_t i = 0; i < (size_t)ng * S; i++) m->g_vtx[i] = -1;
    for (int i = 0; i < ne; i++) m->b_t_ext[i] = SEVC_MAX_T_EXT;
    return 0;
}

void orc_mesh_free(sevc_mesh* m) {
    free(m->pos); free(m->vel); free(m->v_bnd); free(m->v_grn); free(m->v_enabled);
    free(m->v_energy); free(m->v_vel_parts);
    free(m->b_vtx); free(m->b_energy); free(m->b_arclength); free(m->b_t_ext);
    free(m->b_enabled); free(m->b_candidate);
    free(m->g_vtx); free(m->g_vlen); free(m->g_area); free(m->g_dAdt);
    free(m->g_orientation); free(m->g_SE); free(m->g_enabled); free(m->g_type);
    memset(m, 0, sizeof(*m));
}

/* geometry.h:437-460: sort the three boundaries by descending atan2 of the
 * direction to the neighbour (bubble sort as written). */
static void set_clockwise(sevc_mesh* m, int v) {
    double D = m->domain_bound, ang[3];
    int* vb = &m->v_bnd[3 * v];
    for (int i = 0; i < 3; i++) {
        int b = vb[i];
        int nb = (m->b_vtx[2 * b] == v) ? m->b_vtx[2 * b + 1] : m->b_vtx[2 * b];
        double dx = wrapd(m->pos[2 * nb], m->pos[2 * v], D);
        double dy = wrapd(m->pos[2 * nb + 1], m->pos[2 * v + 1], D);
        ang[i] = sevc_atan2(dy, dx);
    }
    for (int i = 0; i < 2; i++)
        for (int p = 1; p < 3 - i; p++)
            if (ang[p - 1] < ang[p]) {
                double a = ang[p - 1]; ang[p - 1] = ang[p]; ang[p] = a;
                int t = vb[p - 1]; vb[p - 1] = vb[p]; vb[p] = t;
            }
}

int orc_build_mesh(const sevc_options* o, int32_t nv, const double* upos, int32_t nr,
                   const int32_t* ridges, int32_t ng_expect, const double* ori01, const double* se_raw,
                   sevc_mesh* out) {
    double D = o->domain_bound;
    int S = o->max_vrt_per_grn;
    int xv = (int)(o->extra_capacity * nv), xe = (int)(o->extra_capacity * nr);
    int ngrains_walk_cap = 2 * nv + 8;
    /* first pass into a temporary grain table to learn the grain count */
    int* gtmp = malloc(sizeof(int) * (size_t)ngrains_walk_cap * S);
    int* glen = calloc((size_t)ngrains_walk_cap, sizeof(int));
    sevc_mesh m;
    if (orc_mesh_alloc(&m, nv + xv, nr + xe, 1, S) != 0) return SEVC_E_NOMEM;
    m.domain_bound = D;
    for (int i = 0; i < nv; i++) {
        m.pos[2 * i] = upos[2 * i] * D;            /* io.h:89-90 */
        m.pos[2 * i + 1] = upos[2 * i + 1] * D;
        m.v_enabled[i] = 1;
    }
    for (int i = 0; i < nr; i++) {                   /* io.h:124-126 */
        int a = ridges[2 * i], b = ridges[2 * i + 1];
        if (a < 0 || a >= nv || b < 0 || b >= nv) goto bad;
        m.b_vtx[2 * i] = a; m.b_vtx[2 * i + 1] = b;
        m.b_enabled[i] = 1;
        int k;
        for (k = 0; k < 3 && m.v_bnd[3 * a + k] >= 0; k++) {}
        if (k == 3) goto bad;
        m.v_bnd[3 * a + k] = i;
        for (k = 0; k < 3 && m.v_bnd[3 * b + k] >= 0; k++) {}
        if (k == 3) goto bad;
        m.v_bnd[3 * b + k] = i;
    }
    for (int j = 0; j < nv; j++) {
        for (int k = 0; k < 3; k++) if (m.v_bnd[3 * j + k] < 0) goto bad;
        set_clockwise(&m, j);                        /* io.h:132-134 */
    }
    uint8_t* seen = calloc((size_t)3 * nv + 1, 1);
    int ng = 0;
    for (int j = 0; j < nv; j++) {                   /* io.h:143-180 */
        for (int g = 0; g < 3; g++) {
            if (seen[3 * j + g]) continue;
            if (ng >= ngrains_walk_cap) { free(seen); goto bad; }
            int cur = j, side = g;
            do {
                if (seen[3 * cur + side]) { free(seen); goto bad; }
                seen[3 * cur + side] = 1;
                if (glen[ng] >= S) { free(seen); free(gtmp); free(glen); orc_mesh_free(&m); return SEVC_E_GRAIN_FULL; }
                gtmp[(size_t)ng * S + glen[ng]++] = cur;
                int b = m.v_bnd[3 * cur + side];
                cur = (m.b_vtx[2 * b] == cur) ? m.b_vtx[2 * b + 1] : m.b_vtx[2 * b];
                int i = 0;
                while (i < 3 && m.v_bnd[3 * cur + i] != b) i++;
                if (i == 3) { free(seen); goto bad; }
                side = (i + 1) % 3;
            } while (cur != j);
            ng++;
        }
    }
    free(seen);
    if (ng_expect > 0 && ng != ng_expect) goto bad;   /* io.h:182-186 */
    for (int j = 0; j < nv; j++) {                     /* io.h:189-192 */
        set_clockwise(&m, j);
        int t = m.v_bnd[3 * j + 1]; m.v_bnd[3 * j + 1] = m.v_bnd[3 * j]; m.v_bnd[3 * j] = t;
    }
    {
        int xg = (int)(o->extra_capacity * ng);
        sevc_mesh full;
        if (orc_mesh_alloc(&full, nv + xv, nr + xe, ng + xg, S) != 0) return SEVC_E_NOMEM;
        full.domain_bound = D;
        memcpy(full.pos, m.pos, sizeof(double) * 2 * (size_t)(nv + xv));
        memcpy(full.v_bnd, m.v_bnd, sizeof(int32_t) * 3 * (size_t)(nv + xv));
        memcpy(full.v_enabled, m.v_enabled, (size_t)(nv + xv));
        memcpy(full.b_vtx, m.b_vtx, sizeof(int32_t) * 2 * (size_t)(nr + xe));
        memcpy(full.b_enabled, m.b_enabled, (size_t)(nr + xe));
        orc_mesh_free(&m);
        for (int g = 0; g < ng; g++) {
            full.g_vlen[g] = glen[g];
            full.g_enabled[g] = 1;
            for (int k = 0; k < glen[g]; k++) full.g_vtx[(size_t)g * S + k] = gtmp[(size_t)g * S + k];
            for (int k = 0; k < glen[g]; k++) {        /* io.h:196-200 */
                int v = gtmp[(size_t)g * S + k], q;
                for (q = 0; q < 3 && full.v_grn[3 * v + q] >= 0; q++) {}
                if (q == 3) { free(gtmp); free(glen); orc_mesh_free(&full); return SEVC_E_TOPOLOGY; }
                full.v_grn[3 * v + q] = g;
            }
            full.g_orientation[g] = ori01[g] * 2 * SEVC_PI;   /* io.h:214-218 */
            double se = se_raw[g];                            /* io.h:223-248 */
            if (o->SE_min == 999 && o->SE_max == 999) se *= o->SE_eps;
            else if (o->SE_min == o->SE_max) se = o->SE_max * o->SE_eps;
            else se = (o->SE_max - o->SE_min) * se + o->SE_min;
            full.g_SE[g] = se;
        }
        free(gtmp); free(glen);
        *out = full;
        return SEVC_OK;
    }
bad:
    free(gtmp); free(glen);
    orc_mesh_free(&m);
    return SEVC_E_TOPOLOGY;
}

static int read_doubles(const char* path, int per_line, double** out, int* n) {
    FILE* f = fopen(path, "r");
    if (!f) return SEVC_E_IO;
    int cap = 1024, k = 0;
    double* a = malloc(sizeof(double) * (size_t)cap);
    double v;
    while (fscanf(f, "%lf", &v) == 1) {
        if (k == cap) { cap *= 2; a = realloc(a, sizeof(double) * (size_t)cap); }
        a[k++] = v;
    }
    fclose(f);
    *out = a;
    *n = k / per_line;
    return SEVC_OK;
}

int orc_load_files(const sevc_options* o, sevc_mesh* out) {
    double *vp = NULL, *rr = NULL, *ori = NULL, *se = NULL;
    int nv, nr, no, ns;
    int rc = read_doubles(o->vertex_fname, 2, &vp, &nv);
    if (rc == SEVC_OK) rc = read_doubles(o->boundary_fname, 2, &rr, &nr);
    if (rc == SEVC_OK) rc = read_doubles(o->orientation_fname, 1, &ori, &no);
    if (rc == SEVC_OK) rc = read_doubles(o->SE_fname, 1, &se, &ns);
    if (rc == SEVC_OK) {
        int32_t* ridges = malloc(sizeof(int32_t) * (size_t)(2 * nr + 1));
        for (int i = 0; i < 2 * nr; i++) ridges[i] = (int32_t)rr[i];
        if (no < o->n_grains || ns < o->n_grains) rc = SEVC_E_IO;
        else rc = orc_build_mesh(o, nv, vp, nr, ridges, o->n_grains, ori, se, out);
        free(ridges);
    }
    free(vp); free(rr); free(ori); free(se);
    return rc;
}

/* check_data, utils.h:85-336: counts the conditions that set
 * is_valid = false (warnings are not counted).  The reference holds
 * pointers; here an id outside [0, n) counts once for the entity holding
 * it and the checks that would follow it are skipped, and a grain list is
 * read up to min(vlen, S) entries. */
int orc_check_mesh(const sevc_mesh* m, double* total_area) {
    int bad = 0;
    const int S = m->max_vrt_per_grn;
    const int nv = m->n_vertices, ne = m->n_boundaries, ng = m->n_grains;
#define BINI(b) (m->b_vtx[2 * (b)])
#define BEND(b) (m->b_vtx[2 * (b) + 1])
#define VBND(v, i) (m->v_bnd[3 * (v) + (i)])
#define VGRN(v, i) (m->v_grn[3 * (v) + (i)])
    for (int b = 0; b < m->n_boundaries; b++) {
        if (!m->b_enabled[b]) continue;
        int a = BINI(b), e = BEND(b);
        if (a < 0 || e < 0 || a >= nv || e >= nv) { bad++; continue; }
        if (a == e) bad++;
        if (!m->v_enabled[a] || !m->v_enabled[e]) bad++;
        if (VBND(a, 0) != b && VBND(a, 1) != b && VBND(a, 2) != b) bad++;
        if (VBND(e, 0) != b && VBND(e, 1) != b && VBND(e, 2) != b) bad++;
        for (int i = 0; i < 3; i++) {
            if (VBND(e, i) < 0 || VBND(e, i) >= ne || !m->b_enabled[VBND(e, i)]) { bad++; break; }
        }
        for (int i = 0; i < 3; i++) {
            if (VBND(a, i) < 0 || VBND(a, i) >= ne || !m->b_enabled[VBND(a, i)]) { bad++; break; }
        }
        int gic[2] = {-1, -1}, k = 0;
        for (int i = 0; i < 3 && k < 2; i++)
            for (int j = 0; j < 3; j++)
                if (VGRN(a, i) == VGRN(e, j)) { gic[k++] = VGRN(a, i); if (k == 2) break; }
        if (k < 2 || gic[0] == gic[1]) bad++;
    }
    for (int v = 0; v < m->n_vertices; v++) {
        if (!m->v_enabled[v]) continue;
        int ok = 1;
        for (int j = 0; j < 3; j++)
            if (VBND(v, j) < 0 || VBND(v, j) >= ne || VGRN(v, j) < 0 || VGRN(v, j) >= ng) ok = 0;
        if (!ok) { bad++; continue; }
        if (VBND(v, 0) == VBND(v, 1) || VBND(v, 1) == VBND(v, 2) || VBND(v, 2) == VBND(v, 0)) bad++;
        if (VGRN(v, 0) == VGRN(v, 1) || VGRN(v, 1) == VGRN(v, 2) || VGRN(v, 2) == VGRN(v, 0)) bad++;
        for (int j = 0; j < 3; j++) if (!m->b_enabled[VBND(v, j)]) bad++;
        for (int j = 0; j < 3; j++) if (!m->g_enabled[VGRN(v, j)]) bad++;
        for (int j = 0; j < 3; j++) if (BINI(VBND(v, j)) == BEND(VBND(v, j))) bad++;
        int arr[3];
        for (int j = 0; j < 3; j++) arr[j] = BINI(VBND(v, j)) == v ? BEND(VBND(v, j)) : BINI(VBND(v, j));
        if (arr[0] == arr[1] || arr[1] == arr[2] || arr[2] == arr[0]) bad++;
        for (int j = 0; j < 3; j++)
            if (BINI(VBND(v, j)) != v && BEND(VBND(v, j)) != v) bad++;
        for (int j = 0; j < 3; j++) {
            int g = VGRN(v, j), found = 0;
            int n = m->g_vlen[g] < S ? m->g_vlen[g] : S;
            for (int k = 0; k < n; k++) if (m->g_vtx[(size_t)g * S + k] == v) found = 1;
            if (!found) bad++;
        }
    }
    double total = 0;
    for (int g = 0; g < m->n_grains; g++) {
        if (!m->g_enabled[g]) continue;
        total += m->g_area ? m->g_area[g] : 0.0;
        const int n = m->g_vlen[g] < S ? m->g_vlen[g] : S;
        for (int j = 0; j < n; j++) {
            int v = m->g_vtx[(size_t)g * S + j];
            if (v < 0 || v >= nv || !m->v_enabled[v]) bad++;
        }
        for (int j = 0; j < n; j++) {
            int v = m->g_vtx[(size_t)g * S + j];
            if (v < 0 || v >= nv) continue;
            if (VGRN(v, 0) != g && VGRN(v, 1) != g && VGRN(v, 2) != g) { bad++; break; }
        }
        int cnt = 0;
        for (int j = 0; j < S; j++) if (m->g_vtx[(size_t)g * S + j] >= 0) cnt++;
        if (cnt != m->g_vlen[g]) bad++;
    }
    if (total_area) *total_area = total;
    return bad;
#undef BINI
#undef BEND
#undef VBND
#undef VGRN
}
