// sevc_host.cpp -- host side of the C ABI: options, ini parsing, host
// meshes, the initial-condition loader, the consistency checker and the
// snapshot writer.  No device code here.
#include "../../include/sevc.h"
#include "../../include/sevc_math.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <sys/types.h>

namespace {

void copy_str(char* dst, const char* src) {
    std::strncpy(dst, src, SEVC_FNAME - 1);
    dst[SEVC_FNAME - 1] = '\0';
}

// wrap_dist of geometry.h:147-152 (minimum-image difference).
inline double wrap(double dest, double orig, double D) {
    double d = dest - orig;
    if (d > +(D / 2.0)) return d - D;
    if (d < -(D / 2.0)) return D + d;
    return d;
}

}  // namespace

extern "C" {

// Defaults of load_config_file, options.h:71-85.  Keys the reference leaves
// uninitialised (dt, dtau, snap_gap, n_grains, n_blocks, n_threads) get the
// values of config/example.ini; extensions default to reference behaviour.
void sevc_options_default(sevc_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->dt = 1e-4;
    o->dtau = 1e-4;
    o->nucleation_gap = 1;
    o->tmax = -1;
    o->nmax = -1;
    o->snap_gap = 100;
    o->leave_percent = -1;
    o->n_grains = 0;
    o->GB_eps = 0.02;
    o->GB_scaling = 1.0;
    o->SE_eps = 1.0;
    o->SE_min = 999;
    o->SE_max = 999;
    o->MC_k = 1.0;
    o->nucleation_scaling = 0.9;
    o->domain_bound = 1.0;
    o->do_nucleation = 1;
    o->n_blocks = 32;
    o->n_threads = 256;
    o->debug_lvl = 0;
    o->mobility = 1.0;
    o->adaptive_dt = 0;
    o->dt_safety = 0.5;
    o->dt_min = 0.0;
    o->flip_threshold = 0.0;
    o->nucleation_mode = SEVC_NUCL_MC;
    o->nucleation_rate = 0.0;
    o->seed = 32;               // curand_init(32, ...) in nucleation2.cu:15
    o->max_vrt_per_grn = 32;
    o->extra_capacity = 0.0;
    o->velocity_model = SEVC_VEL_FD;
    o->events = 1;
    o->nucleation_gate = -1;
}

// load_config_file, options.h:69-206.  Lines starting with '#' or '[' are
// skipped; the key is the text before the first '=' matched exactly (no
// trimming); numbers go through atof/atoi; string values lose their line
// terminator; do_nucleation is true only for the value "true".
int sevc_options_from_ini(const char* path, sevc_options* o) {
    if (!path || !o) return SEVC_E_INVALID_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return SEVC_E_IO;
    sevc_options_default(o);                      // load_config_file starts from its defaults (options.h:71-85)
    char line[1000];
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '[') continue;
        char* eq = std::strchr(line, '=');
        if (!eq) continue;
        std::string key(line, static_cast<size_t>(eq - line));
        const char* val = eq + 1;
        std::string sval(val);
        while (!sval.empty() && (sval.back() == '\n' || sval.back() == '\r')) sval.pop_back();
        const char* v = sval.c_str();
        if (key == "dt") o->dt = std::atof(v);
        else if (key == "dtau") o->dtau = std::atof(v);
        else if (key == "tmax") o->tmax = std::atof(v);
        else if (key == "nmax") o->nmax = std::atoi(v);
        else if (key == "snap_gap") o->snap_gap = std::atoi(v);
        else if (key == "leave_percent") o->leave_percent = std::atof(v);
        else if (key == "n_grains") o->n_grains = std::atoi(v);
        else if (key == "vertex_fname") copy_str(o->vertex_fname, v);
        else if (key == "boundary_fname") copy_str(o->boundary_fname, v);
        else if (key == "orientation_fname") copy_str(o->orientation_fname, v);
        else if (key == "SE_fname") copy_str(o->SE_fname, v);
        else if (key == "n_blocks") o->n_blocks = std::atoi(v);
        else if (key == "n_threads") o->n_threads = std::atoi(v);
        else if (key == "debug_lvl") o->debug_lvl = std::atoi(v);
        else if (key == "folder") copy_str(o->output_folder, v);
        else if (key == "GB_eps") o->GB_eps = std::atof(v);
        else if (key == "GB_scaling") o->GB_scaling = std::atof(v);
        else if (key == "SE_eps") o->SE_eps = std::atof(v);
        else if (key == "SE_min") o->SE_min = std::atof(v);
        else if (key == "SE_max") o->SE_max = std::atof(v);
        else if (key == "MC_k") o->MC_k = std::atof(v);
        else if (key == "nucleation_scaling") o->nucleation_scaling = std::atof(v);
        else if (key == "domain_bound") o->domain_bound = std::atof(v);
        else if (key == "do_nucleation") o->do_nucleation = (std::strcmp(v, "true") == 0) ? 1 : 0;
        // extension keys
        else if (key == "mobility") o->mobility = std::atof(v);
        else if (key == "adaptive_dt") o->adaptive_dt = (std::strcmp(v, "true") == 0 || std::atoi(v) != 0) ? 1 : 0;
        else if (key == "dt_safety") o->dt_safety = std::atof(v);
        else if (key == "dt_min") o->dt_min = std::atof(v);
        else if (key == "flip_threshold") o->flip_threshold = std::atof(v);
        else if (key == "nucleation_mode") o->nucleation_mode = (std::strcmp(v, "rate") == 0) ? SEVC_NUCL_RATE : SEVC_NUCL_MC;
        else if (key == "nucleation_rate") o->nucleation_rate = std::atof(v);
        else if (key == "seed") o->seed = std::strtoull(v, nullptr, 10);
        else if (key == "max_vrt_per_grn") o->max_vrt_per_grn = std::atoi(v);
        else if (key == "extra_capacity") o->extra_capacity = std::atof(v);
        else if (key == "velocity_model") o->velocity_model = (std::strcmp(v, "analytic") == 0) ? SEVC_VEL_ANALYTIC : SEVC_VEL_FD;
        else if (key == "events") o->events = std::atoi(v);
        else if (key == "nucleation_gate") o->nucleation_gate = std::atoi(v);
    }
    std::fclose(f);
    // options.h:172-181
    if (o->dt > 0) o->nucleation_gap = static_cast<int32_t>(o->dtau / o->dt);
    if (o->tmax > 0 && o->nmax == -1) {
    } else if (o->tmax == -1 && o->nmax > 0) {
    } else {
        o->tmax = 1.0;
        o->nmax = -1;
    }
    return SEVC_OK;
}

int sevc_mesh_alloc(sevc_mesh* m, int32_t nv, int32_t ne, int32_t ng, int32_t S) {
    if (!m || nv < 0 || ne < 0 || ng < 0 || S < 3) return SEVC_E_INVALID_ARG;
    std::memset(m, 0, sizeof(*m));
    m->n_vertices = nv;
    m->n_boundaries = ne;
    m->n_grains = ng;
    m->max_vrt_per_grn = S;
    m->domain_bound = 1.0;
    auto dz = [](size_t n) { return static_cast<double*>(std::calloc(n + 1, sizeof(double))); };
    auto iz = [](size_t n) {
        int32_t* p = static_cast<int32_t*>(std::malloc((n + 1) * sizeof(int32_t)));
        if (p) for (size_t i = 0; i < n; ++i) p[i] = -1;
        return p;
    };
    auto bz = [](size_t n) { return static_cast<uint8_t*>(std::calloc(n + 1, 1)); };
    size_t V = static_cast<size_t>(nv), E = static_cast<size_t>(ne), G = static_cast<size_t>(ng);
    m->pos = dz(2 * V); m->vel = dz(2 * V);
    m->v_bnd = iz(3 * V); m->v_grn = iz(3 * V); m->v_enabled = bz(V);
    m->v_energy = dz(V); m->v_vel_parts = dz(4 * V);
    m->b_vtx = iz(2 * E); m->b_energy = dz(E); m->b_arclength = dz(E); m->b_t_ext = dz(E);
    m->b_enabled = bz(E); m->b_candidate = bz(E);
    m->g_vtx = iz(G * static_cast<size_t>(S));
    m->g_vlen = static_cast<int32_t*>(std::calloc(G + 1, sizeof(int32_t)));
    m->g_area = dz(G); m->g_dAdt = dz(G); m->g_orientation = dz(G); m->g_SE = dz(G);
    m->g_enabled = bz(G); m->g_type = bz(G);
    if (!m->pos || !m->vel || !m->v_bnd || !m->v_grn || !m->v_enabled || !m->b_vtx || !m->g_vtx ||
        !m->g_vlen || !m->b_enabled || !m->g_enabled) {
        sevc_mesh_free(m);
        return SEVC_E_NOMEM;
    }
    for (size_t i = 0; i < E; ++i) m->b_t_ext[i] = SEVC_MAX_T_EXT;   // init_boundary, geometry.h:499
    return SEVC_OK;
}

void sevc_mesh_free(sevc_mesh* m) {
    if (!m) return;
    void* ptrs[] = {m->pos, m->vel, m->v_bnd, m->v_grn, m->v_enabled, m->v_energy, m->v_vel_parts,
                    m->b_vtx, m->b_energy, m->b_arclength, m->b_t_ext, m->b_enabled, m->b_candidate,
                    m->g_vtx, m->g_vlen, m->g_area, m->g_dAdt, m->g_orientation, m->g_SE,
                    m->g_enabled, m->g_type};
    for (void* p : ptrs) std::free(p);
    std::memset(m, 0, sizeof(*m));
}

void sevc_free(void* p) { std::free(p); }

// Mesh construction with the ids and orderings of load_from_files,
// io.h:66-264.  Vertex i of the file is slot i, ridge i is boundary i.
// Each vertex lists its boundaries in descending-angle ("clockwise") order
// for the face walk; a grain is traced by leaving a vertex along side s,
// arriving at the far vertex, and continuing on the side after the arrival
// edge.  Grains are numbered in the order the walk discovers them
// (vertex-major, side-minor).  Finally the boundary order is made
// counter-clockwise by swapping the first two entries, and every vertex
// lists its grains by increasing grain id.
int sevc_mesh_build(const sevc_options* o, int32_t nv, const double* upos, int32_t nr,
                    const int32_t* ridges, int32_t ng_expect, const double* ori01,
                    const double* se_raw, sevc_mesh* out) {
    if (!o || !upos || !ridges || !ori01 || !se_raw || !out || nv <= 0 || nr <= 0) return SEVC_E_INVALID_ARG;
    const double D = o->domain_bound;
    const int S = o->max_vrt_per_grn;
    if (S < 3) return SEVC_E_INVALID_ARG;

    std::vector<double> px(static_cast<size_t>(nv)), py(static_cast<size_t>(nv));
    for (int i = 0; i < nv; ++i) { px[i] = upos[2 * i] * D; py[i] = upos[2 * i + 1] * D; }

    // incidence in file order (init_boundary -> vertex_add_boundary)
    std::vector<int> inc(static_cast<size_t>(3) * nv, -1);
    std::vector<int> deg(static_cast<size_t>(nv), 0);
    for (int e = 0; e < nr; ++e) {
        const int a = ridges[2 * e], b = ridges[2 * e + 1];
        if (a < 0 || a >= nv || b < 0 || b >= nv) return SEVC_E_TOPOLOGY;
        if (deg[a] == 3 || deg[b] == 3) return SEVC_E_TOPOLOGY;
        inc[3 * a + deg[a]++] = e;
        inc[3 * b + deg[b]++] = e;
    }
    for (int v = 0; v < nv; ++v) if (deg[v] != 3) return SEVC_E_TOPOLOGY;

    auto other = [&](int e, int v) { return ridges[2 * e] == v ? ridges[2 * e + 1] : ridges[2 * e]; };
    // vertex_set_boundaries_clockwise, geometry.h:437-460
    auto sort_desc = [&](int v) {
        double ang[3];
        int* b = &inc[3 * v];
        for (int i = 0; i < 3; ++i) {
            const int w = other(b[i], v);
            ang[i] = sevc_atan2(wrap(py[w], py[v], D), wrap(px[w], px[v], D));
        }
        for (int pass = 0; pass < 2; ++pass)
            for (int p = 1; p < 3 - pass; ++p)
                if (ang[p - 1] < ang[p]) { std::swap(ang[p - 1], ang[p]); std::swap(b[p - 1], b[p]); }
    };
    for (int v = 0; v < nv; ++v) sort_desc(v);

    // face walk, io.h:137-180
    std::vector<uint8_t> used(static_cast<size_t>(3) * nv, 0);
    std::vector<int> glist;
    std::vector<int> gstart;
    for (int v0 = 0; v0 < nv; ++v0) {
        for (int s0 = 0; s0 < 3; ++s0) {
            if (used[3 * v0 + s0]) continue;
            gstart.push_back(static_cast<int>(glist.size()));
            int v = v0, s = s0, len = 0;
            do {
                if (used[3 * v + s]) return SEVC_E_TOPOLOGY;   // io.h:155-160
                used[3 * v + s] = 1;
                if (++len > S) return SEVC_E_GRAIN_FULL;
                glist.push_back(v);
                const int e = inc[3 * v + s];
                const int w = other(e, v);
                int k = 0;
                while (k < 3 && inc[3 * w + k] != e) ++k;
                if (k == 3) return SEVC_E_TOPOLOGY;
                v = w;
                s = (k + 1) % 3;
            } while (v != v0);
        }
    }
    const int ng = static_cast<int>(gstart.size());
    gstart.push_back(static_cast<int>(glist.size()));
    if (ng_expect > 0 && ng != ng_expect) return SEVC_E_TOPOLOGY;    // io.h:182-186

    // io.h:189-192: clockwise again, then swap the first two (counter-clockwise)
    for (int v = 0; v < nv; ++v) {
        sort_desc(v);
        std::swap(inc[3 * v], inc[3 * v + 1]);
    }

    const int xv = static_cast<int>(o->extra_capacity * nv);
    const int xe = static_cast<int>(o->extra_capacity * nr);
    const int xg = static_cast<int>(o->extra_capacity * ng);
    int rc = sevc_mesh_alloc(out, nv + xv, nr + xe, ng + xg, S);
    if (rc != SEVC_OK) return rc;
    out->domain_bound = D;
    for (int v = 0; v < nv; ++v) {
        out->pos[2 * v] = px[v];
        out->pos[2 * v + 1] = py[v];
        out->v_enabled[v] = 1;
        for (int k = 0; k < 3; ++k) out->v_bnd[3 * v + k] = inc[3 * v + k];
    }
    for (int e = 0; e < nr; ++e) {
        out->b_vtx[2 * e] = ridges[2 * e];
        out->b_vtx[2 * e + 1] = ridges[2 * e + 1];
        out->b_enabled[e] = 1;
    }
    std::vector<int> gdeg(static_cast<size_t>(nv), 0);
    for (int g = 0; g < ng; ++g) {
        const int b = gstart[g], n = gstart[g + 1] - gstart[g];
        out->g_vlen[g] = n;
        out->g_enabled[g] = 1;
        out->g_type[g] = 0;
        for (int k = 0; k < n; ++k) {
            const int v = glist[b + k];
            out->g_vtx[static_cast<size_t>(g) * S + k] = v;
            if (gdeg[v] == 3) { sevc_mesh_free(out); return SEVC_E_TOPOLOGY; }
            out->v_grn[3 * v + gdeg[v]++] = g;          // io.h:196-200
        }
        out->g_orientation[g] = ori01[g] * 2 * SEVC_PI;  // fix_orientation, geometry.h:536-538
        double se = se_raw[g];                           // io.h:226-248
        if (o->SE_min == 999 && o->SE_max == 999) se *= o->SE_eps;
        else if (o->SE_min == o->SE_max) se = o->SE_max * o->SE_eps;
        else se = (o->SE_max - o->SE_min) * se + o->SE_min;
        out->g_SE[g] = se;
    }
    return SEVC_OK;
}

namespace {
int read_numbers(const char* path, std::vector<double>& out) {
    FILE* f = std::fopen(path, "r");
    if (!f) return SEVC_E_IO;
    double v;
    while (std::fscanf(f, "%lf", &v) == 1) out.push_back(v);
    std::fclose(f);
    return SEVC_OK;
}
}  // namespace

// load_from_files, io.h:66-264 (file reading part).
int sevc_mesh_load_files(const sevc_options* o, sevc_mesh* out) {
    if (!o || !out) return SEVC_E_INVALID_ARG;
    std::vector<double> vtx, rid, ori, se;
    int rc;
    if ((rc = read_numbers(o->vertex_fname, vtx)) != SEVC_OK) return rc;
    if ((rc = read_numbers(o->boundary_fname, rid)) != SEVC_OK) return rc;
    if ((rc = read_numbers(o->orientation_fname, ori)) != SEVC_OK) return rc;
    if ((rc = read_numbers(o->SE_fname, se)) != SEVC_OK) return rc;
    const int nv = static_cast<int>(vtx.size() / 2), nr = static_cast<int>(rid.size() / 2);
    std::vector<int32_t> r(rid.size());
    for (size_t i = 0; i < rid.size(); ++i) r[i] = static_cast<int32_t>(rid[i]);
    const int ng = o->n_grains > 0 ? o->n_grains : 2 * nv;   // upper bound when unknown
    if (static_cast<int>(ori.size()) < (o->n_grains > 0 ? o->n_grains : 0) ||
        static_cast<int>(se.size()) < (o->n_grains > 0 ? o->n_grains : 0))
        return SEVC_E_IO;
    ori.resize(static_cast<size_t>(ng), 0.0);
    se.resize(static_cast<size_t>(ng), 0.0);
    return sevc_mesh_build(o, nv, vtx.data(), nr, r.data(), o->n_grains, ori.data(), se.data(), out);
}

// check_data, utils.h:85-336: counts the conditions that make the
// reference report invalid data (its warnings are not counted).  Where the
// reference holds pointers, an id outside [0, n) counts once for the entity
// holding it and the checks that would follow it are skipped; a grain list
// is read up to min(vlen, S) entries.
int sevc_check_mesh(const sevc_mesh* m, double* total_area) {
    if (!m) return -1;
    const int S = m->max_vrt_per_grn;
    const int nv = m->n_vertices, ne = m->n_boundaries, ng = m->n_grains;
    auto vb = [&](int v, int i) { return m->v_bnd[3 * v + i]; };
    auto vg = [&](int v, int i) { return m->v_grn[3 * v + i]; };
    auto bi = [&](int b) { return m->b_vtx[2 * b]; };
    auto be = [&](int b) { return m->b_vtx[2 * b + 1]; };
    int bad = 0;
    for (int b = 0; b < ne; ++b) {
        if (!m->b_enabled[b]) continue;
        const int a = bi(b), e = be(b);
        if (a < 0 || e < 0 || a >= nv || e >= nv) { ++bad; continue; }
        if (a == e) ++bad;
        if (!m->v_enabled[a] || !m->v_enabled[e]) ++bad;
        if (vb(a, 0) != b && vb(a, 1) != b && vb(a, 2) != b) ++bad;
        if (vb(e, 0) != b && vb(e, 1) != b && vb(e, 2) != b) ++bad;
        bool dis = false;
        for (int i = 0; i < 3; ++i) if (vb(e, i) < 0 || vb(e, i) >= ne || !m->b_enabled[vb(e, i)]) dis = true;
        if (dis) ++bad;
        dis = false;
        for (int i = 0; i < 3; ++i) if (vb(a, i) < 0 || vb(a, i) >= ne || !m->b_enabled[vb(a, i)]) dis = true;
        if (dis) ++bad;
        int gic[2] = {-1, -1}, k = 0;
        for (int i = 0; i < 3 && k < 2; ++i)
            for (int j = 0; j < 3 && k < 2; ++j)
                if (vg(a, i) == vg(e, j)) gic[k++] = vg(a, i);
        if (k < 2 || gic[0] == gic[1]) ++bad;
    }
    for (int v = 0; v < nv; ++v) {
        if (!m->v_enabled[v]) continue;
        bool ok = true;
        for (int j = 0; j < 3; ++j)
            if (vb(v, j) < 0 || vb(v, j) >= ne || vg(v, j) < 0 || vg(v, j) >= ng) ok = false;
        if (!ok) { ++bad; continue; }
        if (vb(v, 0) == vb(v, 1) || vb(v, 1) == vb(v, 2) || vb(v, 2) == vb(v, 0)) ++bad;
        if (vg(v, 0) == vg(v, 1) || vg(v, 1) == vg(v, 2) || vg(v, 2) == vg(v, 0)) ++bad;
        for (int j = 0; j < 3; ++j) if (!m->b_enabled[vb(v, j)]) ++bad;
        for (int j = 0; j < 3; ++j) if (!m->g_enabled[vg(v, j)]) ++bad;
        for (int j = 0; j < 3; ++j) if (bi(vb(v, j)) == be(vb(v, j))) ++bad;
        int nb[3];
        for (int j = 0; j < 3; ++j) nb[j] = bi(vb(v, j)) == v ? be(vb(v, j)) : bi(vb(v, j));
        if (nb[0] == nb[1] || nb[1] == nb[2] || nb[2] == nb[0]) ++bad;
        for (int j = 0; j < 3; ++j) if (bi(vb(v, j)) != v && be(vb(v, j)) != v) ++bad;
        for (int j = 0; j < 3; ++j) {
            const int g = vg(v, j);
            bool found = false;
            const int n = std::min(m->g_vlen[g], S);
            for (int k = 0; k < n; ++k) if (m->g_vtx[static_cast<size_t>(g) * S + k] == v) found = true;
            if (!found) ++bad;
        }
    }
    double total = 0;
    for (int g = 0; g < ng; ++g) {
        if (!m->g_enabled[g]) continue;
        total += m->g_area ? m->g_area[g] : 0.0;
        const int32_t* L = &m->g_vtx[static_cast<size_t>(g) * S];
        const int n = std::min(m->g_vlen[g], S);
        for (int j = 0; j < n; ++j) if (L[j] < 0 || L[j] >= nv || !m->v_enabled[L[j]]) ++bad;
        for (int j = 0; j < n; ++j) {
            const int v = L[j];
            if (v < 0 || v >= nv) continue;
            if (vg(v, 0) != g && vg(v, 1) != g && vg(v, 2) != g) { ++bad; break; }
        }
        int cnt = 0;
        for (int j = 0; j < S; ++j) if (L[j] >= 0) ++cnt;
        if (cnt != m->g_vlen[g]) ++bad;
    }
    if (total_area) *total_area = total;
    return bad;
}

// export_data, io.h:280-377: same header lines, tables and number formats.
// The grain table always has SEVC_REF_MAX_VRT_PER_GRN vertex columns padded
// with -1, as the reference's fixed-size grain struct does; a grain with a
// longer list (possible with max_vrt_per_grn above 100) does not fit the
// format and is an error, nothing is written.
int sevc_export_snapshot(const sevc_mesh* m, const sevc_options* o, const char* folder,
                         int32_t index, double time) {
    if (!m || !o || !folder) return SEVC_E_INVALID_ARG;
    for (int i = 0; i < m->n_grains; ++i)
        if (m->g_vlen[i] > SEVC_REF_MAX_VRT_PER_GRN || m->g_vlen[i] > m->max_vrt_per_grn) return SEVC_E_GRAIN_FULL;
    char fname[2 * SEVC_FNAME + 16];
    std::snprintf(fname, sizeof fname, "%s/%06d.txt", folder, index);
    FILE* f = std::fopen(fname, "w");
    if (!f) return SEVC_E_IO;
    const int S = m->max_vrt_per_grn;
    int cv = 0, cb = 0, cg = 0;
    for (int i = 0; i < m->n_vertices; ++i) cv += m->v_enabled[i] ? 1 : 0;
    for (int i = 0; i < m->n_boundaries; ++i) cb += m->b_enabled[i] ? 1 : 0;
    for (int i = 0; i < m->n_grains; ++i) cg += m->g_enabled[i] ? 1 : 0;
    std::fprintf(f, "TIME %.16f\n", time);
    std::fprintf(f, "DELTA T %.16f\n", o->dt);
    std::fprintf(f, "DELTA TAU %.16f\n", o->dtau);
    std::fprintf(f, "MAX TIME %.16f\n", o->tmax);
    std::fprintf(f, "NUCLEATION %d\n", o->do_nucleation);
    std::fprintf(f, "DOMAIN BOUND %.16f\n", o->domain_bound);
    std::fprintf(f, "GB EPSILON %.16f\n", o->GB_eps);
    std::fprintf(f, "GB SCALING %.16f\n", o->GB_scaling);
    std::fprintf(f, "SE EPSILON %.16f\n", o->SE_eps);
    std::fprintf(f, "SE MIN %.16f\n", o->SE_min);
    std::fprintf(f, "SE MAX %.16f\n", o->SE_max);
    std::fprintf(f, "NUCLEATION SCALING %.16f\n", o->nucleation_scaling);
    std::fprintf(f, "MAX VERTICES PER GRN %d\n", SEVC_REF_MAX_VRT_PER_GRN);
    std::fprintf(f, "VERTICES %d\n", m->n_vertices);
    std::fprintf(f, "BOUNDARIES %d\n", m->n_boundaries);
    std::fprintf(f, "GRAINS %d\n", m->n_grains);
    std::fprintf(f, "CURR VERTICES %d\n", cv);
    std::fprintf(f, "CURR BOUNDARIES %d\n", cb);
    std::fprintf(f, "CURR GRAINS %d\n", cg);
    std::fprintf(f, "#VERTEX ENABLED ENERGY X Y VX VY VXbnd VXgrn VYbnd VYgrn VG1 G2 G3 N1 N2 N3\n");
    for (int i = 0; i < m->n_vertices; ++i) {
        const double* vp = m->v_vel_parts ? &m->v_vel_parts[4 * i] : nullptr;
        std::fprintf(f, "%d %d %.16f %.16f %.16f %.16f %.16f %.16f %.16f %.16f %.16f %d %d %d ", i,
                     m->v_enabled[i], m->v_energy ? m->v_energy[i] : 0.0, m->pos[2 * i], m->pos[2 * i + 1],
                     m->vel[2 * i], m->vel[2 * i + 1], vp ? vp[0] : 0.0, vp ? vp[1] : 0.0, vp ? vp[2] : 0.0,
                     vp ? vp[3] : 0.0, m->v_grn[3 * i], m->v_grn[3 * i + 1], m->v_grn[3 * i + 2]);
        for (int j = 0; j < 3; ++j) {
            const int b = m->v_bnd[3 * i + j];
            const int nb = b < 0 ? -1 : (m->b_vtx[2 * b] == i ? m->b_vtx[2 * b + 1] : m->b_vtx[2 * b]);
            std::fprintf(f, j == 2 ? "%d\n" : "%d ", nb);
        }
    }
    std::fprintf(f, "#BOUNDARY ENABLED INI END ENERGY ARCLEN EXTTIME CANDIDATE\n");
    for (int i = 0; i < m->n_boundaries; ++i)
        std::fprintf(f, "%d %d %d %d %.16f %.16f %.16f %d\n", i, m->b_enabled[i], m->b_vtx[2 * i],
                     m->b_vtx[2 * i + 1], m->b_energy ? m->b_energy[i] : 0.0,
                     m->b_arclength ? m->b_arclength[i] : 0.0, m->b_t_ext ? m->b_t_ext[i] : 0.0,
                     m->b_candidate ? m->b_candidate[i] : 0);
    std::fprintf(f, "#GRAIN ID ENABLED NUCLEATED ORIENT SE AREA dAdt VLEN VERTICES\n");
    for (int i = 0; i < m->n_grains; ++i) {
        std::fprintf(f, "%d %d %d %.16f %.16f %.16f %.16f %d ", i, m->g_enabled[i], m->g_type[i],
                     m->g_orientation[i], m->g_SE[i], m->g_area ? m->g_area[i] : 0.0,
                     m->g_dAdt ? m->g_dAdt[i] : 0.0, m->g_vlen[i]);
        for (int j = 0; j < SEVC_REF_MAX_VRT_PER_GRN; ++j) {
            const int v = (j < m->g_vlen[i] && j < S) ? m->g_vtx[static_cast<size_t>(i) * S + j] : -1;
            if (j < m->g_vlen[i]) std::fprintf(f, "%d ", v);
            else if (j == SEVC_REF_MAX_VRT_PER_GRN - 1) std::fprintf(f, "%d", -1);
            else std::fprintf(f, "%d ", -1);
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return SEVC_OK;
}

int sevc_load_snapshot(const char* path, int32_t max_vrt_per_grn, sevc_mesh* out, double* time) {
    if (!path || !out) return SEVC_E_INVALID_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return SEVC_E_IO;
    char line[4096];
    double t = 0, D = 1;
    int nv = -1, ne = -1, ng = -1;
    // header (io.h:300-317): "KEY value" lines up to the vertex section
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') break;
        if (std::sscanf(line, "TIME %lf", &t) == 1) continue;
        if (std::sscanf(line, "DOMAIN BOUND %lf", &D) == 1) continue;
        if (std::sscanf(line, "VERTICES %d", &nv) == 1) continue;
        if (std::sscanf(line, "BOUNDARIES %d", &ne) == 1) continue;
        std::sscanf(line, "GRAINS %d", &ng);
    }
    if (nv < 0 || ne < 0 || ng < 0) { std::fclose(f); return SEVC_E_IO; }
    const int S = max_vrt_per_grn > 0 ? max_vrt_per_grn : 32;
    int rc = sevc_mesh_alloc(out, nv, ne, ng, S);
    if (rc != SEVC_OK) { std::fclose(f); return rc; }
    out->domain_bound = D;
    std::vector<int> nbr(3 * static_cast<size_t>(nv), -1);
    auto fail_io = [&](int code) { std::fclose(f); sevc_mesh_free(out); return code; };
    for (int i = 0; i < nv; ++i) {                       // #VERTEX section
        if (!std::fgets(line, sizeof line, f)) return fail_io(SEVC_E_IO);
        int id, en, g[3], n[3];
        double e, x, y, vx, vy, p0, p1, p2, p3;
        if (std::sscanf(line, "%d %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %d %d %d %d %d %d", &id, &en, &e, &x, &y,
                        &vx, &vy, &p0, &p1, &p2, &p3, &g[0], &g[1], &g[2], &n[0], &n[1], &n[2]) != 17 || id != i)
            return fail_io(SEVC_E_IO);
        out->v_enabled[i] = static_cast<uint8_t>(en);
        out->pos[2 * i] = x; out->pos[2 * i + 1] = y;
        out->vel[2 * i] = vx; out->vel[2 * i + 1] = vy;
        if (out->v_energy) out->v_energy[i] = e;
        if (out->v_vel_parts) {
            out->v_vel_parts[4 * i] = p0; out->v_vel_parts[4 * i + 1] = p1;
            out->v_vel_parts[4 * i + 2] = p2; out->v_vel_parts[4 * i + 3] = p3;
        }
        for (int j = 0; j < 3; ++j) { out->v_grn[3 * i + j] = g[j]; nbr[3 * static_cast<size_t>(i) + j] = n[j]; }
    }
    if (!std::fgets(line, sizeof line, f) || line[0] != '#') return fail_io(SEVC_E_IO);
    for (int b = 0; b < ne; ++b) {                       // #BOUNDARY section
        if (!std::fgets(line, sizeof line, f)) return fail_io(SEVC_E_IO);
        int id, en, ini, end, cand;
        double e, L, te;
        if (std::sscanf(line, "%d %d %d %d %lf %lf %lf %d", &id, &en, &ini, &end, &e, &L, &te, &cand) != 8 || id != b)
            return fail_io(SEVC_E_IO);
        out->b_enabled[b] = static_cast<uint8_t>(en);
        out->b_vtx[2 * b] = ini; out->b_vtx[2 * b + 1] = end;
        if (out->b_energy) out->b_energy[b] = e;
        if (out->b_arclength) out->b_arclength[b] = L;
        if (out->b_t_ext) out->b_t_ext[b] = te;
        if (out->b_candidate) out->b_candidate[b] = static_cast<uint8_t>(cand);
    }
    if (!std::fgets(line, sizeof line, f) || line[0] != '#') return fail_io(SEVC_E_IO);
    for (int g = 0; g < ng; ++g) {                       // #GRAIN section
        if (!std::fgets(line, sizeof line, f)) return fail_io(SEVC_E_IO);
        char* q = line;
        auto next_long = [&q]() { char* e; long v = std::strtol(q, &e, 10); q = e; return v; };
        auto next_double = [&q]() { char* e; double v = std::strtod(q, &e); q = e; return v; };
        if (next_long() != g) return fail_io(SEVC_E_IO);
        out->g_enabled[g] = static_cast<uint8_t>(next_long());
        const long typ = next_long();
        if (out->g_type) out->g_type[g] = static_cast<uint8_t>(typ);
        out->g_orientation[g] = next_double();
        out->g_SE[g] = next_double();
        const double a = next_double(), dadt = next_double();
        if (out->g_area) out->g_area[g] = a;
        if (out->g_dAdt) out->g_dAdt[g] = dadt;
        const long n = next_long();
        if (n < 0 || n > SEVC_REF_MAX_VRT_PER_GRN) return fail_io(SEVC_E_IO);
        if (n > S) return fail_io(SEVC_E_GRAIN_FULL);
        out->g_vlen[g] = static_cast<int32_t>(n);
        for (long k = 0; k < n; ++k) out->g_vtx[static_cast<size_t>(g) * S + k] = static_cast<int32_t>(next_long());
    }
    std::fclose(f);
    // v_bnd from the neighbour columns: for an enabled vertex, its enabled
    // boundaries ordered as its listed neighbours (export_data writes the
    // other end of vertex.boundaries[j] as N_j)
    std::vector<std::vector<int>> inc(static_cast<size_t>(nv));
    for (int b = 0; b < ne; ++b)
        if (out->b_enabled[b])
            for (int e = 0; e < 2; ++e) {
                const int v = out->b_vtx[2 * b + e];
                if (v >= 0 && v < nv) inc[static_cast<size_t>(v)].push_back(b);
            }
    for (int i = 0; i < nv; ++i) {
        if (!out->v_enabled[i]) continue;
        std::vector<int>& cand = inc[static_cast<size_t>(i)];
        std::vector<char> used(cand.size(), 0);
        for (int j = 0; j < 3; ++j) {
            const int want = nbr[3 * static_cast<size_t>(i) + j];
            for (size_t c = 0; c < cand.size(); ++c) {
                const int b = cand[c];
                const int other = out->b_vtx[2 * b] == i ? out->b_vtx[2 * b + 1] : out->b_vtx[2 * b];
                if (!used[c] && other == want) {
                    out->v_bnd[3 * i + j] = b;
                    used[c] = 1;
                    break;
                }
            }
            if (out->v_bnd[3 * i + j] < 0) { sevc_mesh_free(out); return SEVC_E_TOPOLOGY; }
        }
    }
    if (time) *time = t;
    return SEVC_OK;
}

const char* sevc_strerror(int code) {
    switch (code) {
        case SEVC_OK: return "ok";
        case SEVC_E_INVALID_ARG: return "invalid argument";
        case SEVC_E_IO: return "file missing or unreadable";
        case SEVC_E_TOPOLOGY: return "illegal topology in input";
        case SEVC_E_HIP: return "HIP runtime error";
        case SEVC_E_NOMEM: return "out of memory";
        case SEVC_E_GRAIN_FULL: return "grain has more vertices than max_vrt_per_grn";
        case SEVC_E_STATE: return "inconsistent topology met during an event";
        case SEVC_E_NO_DEVICE: return "no HIP device";
        case SEVC_E_CHECK: return "check_data failed (debug_lvl >= 1)";
        default: return "unknown error";
    }
}

}  // extern "C"
