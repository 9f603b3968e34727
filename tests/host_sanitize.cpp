// Host code of the library and the oracle built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_host_sanitizers.py compiles and
// runs it): the .ini parser, the loader, the periodic Voronoi generator,
// both check_data ports on valid and corrupted meshes (ids in and out of
// range), a snapshot round trip and oracle steps with every event kind.
// Exit status 0 when every check holds; a sanitizer report aborts.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/sevc.h"
#include "../oracle/sevc_oracle.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

// A deep copy of the topology arrays the checkers read (positions shared).
struct Copy {
    sevc_mesh m{};
    std::vector<int32_t> vb, vg, bvx, gvx, glen;
    std::vector<uint8_t> ven, ben, gen;
    explicit Copy(const sevc_mesh& a) {
        m = a;
        const size_t nv = a.n_vertices, ne = a.n_boundaries, ng = a.n_grains, S = a.max_vrt_per_grn;
        vb.assign(a.v_bnd, a.v_bnd + 3 * nv); vg.assign(a.v_grn, a.v_grn + 3 * nv);
        bvx.assign(a.b_vtx, a.b_vtx + 2 * ne); gvx.assign(a.g_vtx, a.g_vtx + ng * S);
        glen.assign(a.g_vlen, a.g_vlen + ng);
        ven.assign(a.v_enabled, a.v_enabled + nv); ben.assign(a.b_enabled, a.b_enabled + ne);
        gen.assign(a.g_enabled, a.g_enabled + ng);
        m.v_bnd = vb.data(); m.v_grn = vg.data(); m.b_vtx = bvx.data(); m.g_vtx = gvx.data();
        m.g_vlen = glen.data(); m.v_enabled = ven.data(); m.b_enabled = ben.data(); m.g_enabled = gen.data();
    }
};

uint64_t lcg(uint64_t& s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return s >> 33;
}

void checkers_on_corruptions(const sevc_mesh& base) {
    double area = 0.0;
    expect(sevc_check_mesh(&base, &area) == 0, "valid mesh, host checker");
    expect(orc_check_mesh(&base, &area) == 0, "valid mesh, oracle checker");
    uint64_t seed = 12345;
    const int nv = base.n_vertices, ne = base.n_boundaries, ng = base.n_grains, S = base.max_vrt_per_grn;
    int caught = 0;
    for (int it = 0; it < 300; ++it) {
        Copy c(base);
        const int edits = 1 + static_cast<int>(lcg(seed) % 3);
        for (int e = 0; e < edits; ++e) {
            const int f = static_cast<int>(lcg(seed) % 8);
            auto pick = [&](uint64_t n) { return static_cast<size_t>(lcg(seed) % n); };
            auto id = [&](int hi) { return static_cast<int32_t>(static_cast<int64_t>(lcg(seed) % (hi + 5)) - 2); };
            switch (f) {
                case 0: c.vb[pick(c.vb.size())] = id(ne); break;
                case 1: c.vg[pick(c.vg.size())] = id(ng); break;
                case 2: c.bvx[pick(c.bvx.size())] = id(nv); break;
                case 3: c.gvx[pick(c.gvx.size())] = id(nv); break;
                case 4: c.glen[pick(c.glen.size())] = id(S); break;
                case 5: { size_t k = pick(c.ven.size()); c.ven[k] = 1 - c.ven[k]; } break;
                case 6: { size_t k = pick(c.ben.size()); c.ben[k] = 1 - c.ben[k]; } break;
                default: { size_t k = pick(c.gen.size()); c.gen[k] = 1 - c.gen[k]; } break;
            }
        }
        const int h = sevc_check_mesh(&c.m, &area), o = orc_check_mesh(&c.m, &area);
        expect(h == o, "host and oracle checkers agree on a corrupted mesh");
        caught += h > 0 ? 1 : 0;
    }
    expect(caught > 200, "most corruptions are caught");
}

void oracle_run(const sevc_options& o, const sevc_mesh& m, int steps, const char* what) {
    orc* s = orc_create(&o, &m);
    expect(s != nullptr, what);
    if (!s) return;
    sevc_step_stats st{};
    expect(orc_step(s, steps, &st) == 0, what);
    sevc_mesh out{};
    expect(sevc_mesh_alloc(&out, m.n_vertices, m.n_boundaries, m.n_grains, m.max_vrt_per_grn) == 0, what);
    out.domain_bound = m.domain_bound;
    expect(orc_get_mesh(s, &out) == 0, what);
    double area = 0.0;
    expect(sevc_check_mesh(&out, &area) == 0, what);
    expect(st.n_flips > 0, what);
    std::printf("%s: %d steps, %lld flips, %lld removals, %lld nucleations\n", what, st.steps_done,
                static_cast<long long>(st.n_flips), static_cast<long long>(st.n_removals),
                static_cast<long long>(st.n_nucleations));
    sevc_mesh_free(&out);
    orc_destroy(s);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s GOLDEN_DIR TMP_DIR\n", argv[0]);
        return 2;
    }
    const std::string golden = argv[1], tmp = argv[2];

    // the .ini parser on the example configuration
    sevc_options ini;
    expect(sevc_options_from_ini((golden + "/example.ini").c_str(), &ini) == 0, "example.ini parses");
    expect(sevc_options_from_ini((tmp + "/missing.ini").c_str(), &ini) != 0, "missing .ini is an error");

    // the 2,000-grain fixture through the product loader and the oracle's
    sevc_options o;
    sevc_options_default(&o);
    const std::string r = golden + "/ref2k/";
    std::snprintf(o.vertex_fname, sizeof o.vertex_fname, "%s", (r + "vertices.txt").c_str());
    std::snprintf(o.boundary_fname, sizeof o.boundary_fname, "%s", (r + "ridges.txt").c_str());
    std::snprintf(o.orientation_fname, sizeof o.orientation_fname, "%s", (r + "oris.txt").c_str());
    std::snprintf(o.SE_fname, sizeof o.SE_fname, "%s", (r + "SE/uniform.txt").c_str());
    o.n_grains = 2000;
    o.GB_eps = 0.0;
    o.SE_eps = 0.5;
    o.dt = 2e-4;
    o.do_nucleation = 1;
    o.nucleation_mode = SEVC_NUCL_RATE;
    o.nucleation_rate = 1e-3;
    o.nucleation_scaling = 0.9;
    o.adaptive_dt = 1;
    o.dt_safety = 0.5;
    o.flip_threshold = 1e-4;
    o.extra_capacity = 0.05;
    sevc_mesh m{}, mo{};
    expect(sevc_mesh_load_files(&o, &m) == 0, "product loader");
    expect(orc_load_files(&o, &mo) == 0, "oracle loader");
    expect(m.n_vertices == mo.n_vertices && m.n_boundaries == mo.n_boundaries && m.n_grains == mo.n_grains,
           "loaders agree on sizes");
    checkers_on_corruptions(m);

    // a snapshot round trip
    expect(sevc_export_snapshot(&m, &o, tmp.c_str(), 7, 0.25) == 0, "snapshot written");
    sevc_mesh back{};
    double t = 0.0;
    expect(sevc_load_snapshot((tmp + "/000007.txt").c_str(), m.max_vrt_per_grn, &back, &t) == 0, "snapshot read");
    double area = 0.0;
    expect(back.n_grains == m.n_grains && t == 0.25 && sevc_check_mesh(&back, &area) == 0, "snapshot round trip");
    sevc_mesh_free(&back);

    // the oracle's steps: flips, removals, two-triangle collapses, nucleation
    oracle_run(o, m, 300, "oracle on the 2,000-grain fixture");

    // the periodic Voronoi generator and a small tessellation's run to a few grains
    int32_t nv = 0, nr = 0;
    double* pos = nullptr;
    int32_t* rid = nullptr;
    expect(sevc_voronoi_generate(400, 9, &nv, &pos, &nr, &rid) == 0, "generator");
    std::vector<double> ori(400, 0.3), se(400, 0.2);
    sevc_options og = o;
    og.n_grains = 400;
    og.dt = 1e-3;
    og.leave_percent = 0.3;
    sevc_mesh mg{};
    expect(sevc_mesh_build(&og, nv, pos, nr, rid, 400, ori.data(), se.data(), &mg) == 0, "mesh from generator");
    checkers_on_corruptions(mg);
    oracle_run(og, mg, 400, "oracle on a generated tessellation");
    sevc_free(pos);
    sevc_free(rid);
    sevc_mesh_free(&mg);
    sevc_mesh_free(&m);
    orc_mesh_free(&mo);
    std::printf("%s\n", failures == 0 ? "ALL OK" : "FAILURES");
    return failures == 0 ? 0 : 1;
}
