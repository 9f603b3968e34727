// sevc_main.cpp -- command-line driver over the C ABI, in place of the
// reference's main (src/main.cu:21-424): `sevc -c config.ini [--graph]`.
//
// Reads the reference's .ini (options.h:69-206) and input files
// (io.h:66-264), steps the model on the GPU (main.cu:124-389) and writes
// snapshots in the export_data format (io.h:280-377) to <folder>/DATA/
// after step indices n with n % snap_gap == 0 and at the stop rule
// (main.cu:328), each after the check_data validation (utils.h:85-336).
// Unlike create_folders (io.h:21-52) an existing output folder is not
// deleted; files of the same names are overwritten.
#include <sys/stat.h>
#include <sys/types.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "../../include/sevc.h"

namespace {

int usage(const char* argv0) {
    std::fprintf(stderr,
                 "usage: %s -c config.ini [--graph]\n"
                 "  -c FILE   configuration in the format of config/example.ini\n"
                 "  --graph   replay each step as a HIP graph\n",
                 argv0);
    return 2;
}

int save(sevc_state* s, sevc_mesh* m, const sevc_options* o, const std::string& dir, int index, double t) {
    int rc = sevc_get_mesh(s, 0, m);
    if (rc != SEVC_OK) return rc;
    double area = 0.0;
    int bad = sevc_check_mesh(m, &area);
    if (bad != 0) {
        std::printf("Error while saving output file (%d violations). Exiting...\n", bad);
        return SEVC_E_STATE;
    }
    std::printf("Valid data. Saving...\n");
    return sevc_export_snapshot(m, o, dir.c_str(), index, t);
}

}  // namespace

int main(int argc, char** argv) {
    const char* ini = nullptr;
    bool graph = false;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "-c") == 0 && i + 1 < argc) ini = argv[++i];
        else if (std::strcmp(argv[i], "--graph") == 0) graph = true;
        else return usage(argv[0]);
    }
    if (!ini) return usage(argv[0]);

    sevc_options o;
    int rc = sevc_options_from_ini(ini, &o);
    if (rc != SEVC_OK) {
        std::fprintf(stderr, "%s: %s\n", ini, sevc_strerror(rc));
        return 1;
    }
    std::string folder = o.output_folder;
    if (folder.empty()) {
        char buf[64];
        std::time_t now = std::time(nullptr);
        std::strftime(buf, sizeof buf, "%Y%m%d_%H%M%S", std::localtime(&now));
        folder = buf;
    }
    ::mkdir(folder.c_str(), 0775);
    const std::string data = folder + "/DATA";
    ::mkdir(data.c_str(), 0775);

    sevc_mesh m;
    if ((rc = sevc_mesh_load_files(&o, &m)) != SEVC_OK) {
        std::fprintf(stderr, "loading input files: %s\n", sevc_strerror(rc));
        return 1;
    }
    sevc_state* s = nullptr;
    if ((rc = sevc_create(&o, &m, &s)) != SEVC_OK) {
        std::fprintf(stderr, "device state: %s\n", sevc_strerror(rc));
        sevc_mesh_free(&m);
        return 1;
    }
    sevc_mesh out;
    int32_t nv, ne, ng, S;
    sevc_member_sizes(s, 0, &nv, &ne, &ng, &S);
    sevc_mesh_alloc(&out, nv, ne, ng, S);

    int n0 = 0;
    for (int g = 0; g < m.n_grains; ++g) n0 += m.g_enabled[g] ? 1 : 0;
    const int min_grains = o.leave_percent < 0 ? -1 : static_cast<int>(o.leave_percent * n0);
    const int gap = o.snap_gap > 0 ? o.snap_gap : 0;
    int prog = 0, n_done = 0;
    sevc_step_stats st{};
    for (;;) {
        // next chunk ends at a step index n with n % gap == 0 (step n is the
        // (n+1)-th step), or runs 1000 steps when snapshots are off
        int chunk = 1000;
        if (gap > 0) chunk = (n_done == 0) ? 1 : gap - ((n_done - 1) % gap);
        rc = graph ? sevc_step_graph(s, chunk, &st) : sevc_step(s, chunk, &st);
        if (rc != SEVC_OK) {
            std::fprintf(stderr, "step: %s (%s)\n", sevc_strerror(rc), sevc_last_error(s));
            break;
        }
        n_done = st.n;
        std::printf("n = %d, t = %f\n", st.n - 1, st.t);
        std::printf("Number of grains: %d\n", st.n_grains);
        std::printf("Nucleated area %%: %.16f\n", st.nucleated_area_pct);
        // main.cu:328: after step n with n % snap_gap == 0, or when the grain
        // rule stops the run (the time rule ends the loop without a save)
        const bool snap = gap > 0 && st.steps_done > 0 && (st.n - 1) % gap == 0;
        const bool grain_stop = st.steps_done > 0 && min_grains > 0 && st.n_grains <= min_grains;
        if (snap || grain_stop) {
            std::printf("Saving at n=%d, t=%.16f, n_grains=%d, nucl_area=%.16f\n", st.n - 1, st.t, st.n_grains,
                        st.nucleated_area_pct);
            if ((rc = save(s, &out, &o, data, prog++, st.t)) != SEVC_OK) break;
        }
        if (st.done || st.steps_done == 0) break;
    }
    std::printf("\nPerformed %d iterations.\n", n_done);
    std::printf("Number of grains: %d\n", st.n_grains);
    sevc_destroy(s);
    sevc_mesh_free(&out);
    sevc_mesh_free(&m);
    return rc == SEVC_OK ? 0 : 1;
}
