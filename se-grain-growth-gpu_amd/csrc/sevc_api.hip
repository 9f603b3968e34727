// sevc_api.hip -- C ABI of the device state: creation (upload with global
// slot ids), the per-step launch sequence, statistics and download.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "sevc_state.h"
#include "sevc_check.h"
#include "../../include/sevc.h"

namespace sevc {
// kernels of sevc_kernels.hip
__global__ void k_begin(Dev d, int rescale_flipped);
__global__ void k_count_all(Dev d);
__global__ void k_member_reduce(Dev d);
__global__ void k_edge_len_energy(Dev d, int only_nucl);
__global__ void k_grain_area(Dev d, int only_nucl);
__global__ void k_vertex_energy(Dev d);
__global__ void k_vel_grain(Dev d, int tuning);
__global__ void k_vel_vertex(Dev d, int tuning);
template <int VT> __global__ void k_vel_vtile(Dev d, int tuning);
__global__ void k_vel_glane(Dev d, int tuning);
__global__ void k_vel_bigrain(Dev d, int tuning);
__global__ void k_velocity_generic(Dev d);
__global__ void k_velocity_fallback(Dev d);
__global__ void k_velocity_analytic(Dev d);
__global__ void k_grain_dadt(Dev d);
__global__ void k_translate(Dev d);
__global__ void k_text(Dev d);
__global__ void k_label(Dev d);
__global__ void k_poll_gather(Dev d);
__global__ void k_poll(Dev d);
__global__ void k_fix(Dev d, int stage);
__global__ void k_fix_clear(Dev d, int stage, int with_area);
__global__ void k_two_grain(Dev d);
__global__ void k_label3(Dev d);
__global__ void k_scan3(Dev d);
__global__ void k_t2_par(Dev d);
__global__ void k_t2_serial(Dev d);
__global__ void k_st2(Dev d);
__global__ void k_st2_fix_clear(Dev d, int nclear);
__global__ void k_translate_clear(Dev d, int nclear);
__global__ void k_flip_par(Dev d);
__global__ void k_flip_serial(Dev d);
__global__ void k_nucl_area_l1(Dev d);
__global__ void k_nucl_area_up(Dev d);
__global__ void k_begin_rate(Dev d, int rescale_flipped);
__global__ void k_nucl_manager(Dev d);
__global__ void k_nucl_insert(Dev d);
// kernels of sevc_adjoint.hip
__global__ void k_win_prepare(Dev d, double* ebase);
__global__ void k_win_area(Dev d, const double* X, double* A, int to_state);
__global__ void k_win_step(Dev d, const double* X, const double* A, double* Xn, int write_vel);
__global__ void k_win_end(Dev d, int steps);
__global__ void k_win_init(Dev d, double* lam, double* acc_mu, double* acc_s, double* acc_se);
__global__ void k_win_adjoint(Dev d, const double* X, const double* A, const double* lam, double* lam2,
                              double* acc_mu, double* acc_s, double* acc_se, const double* ebase);
__global__ void k_win_sum_l1(Dev d, const double* acc_mu, const double* acc_s, double* part);
__global__ void k_win_sum_up(Dev d, double* part, double* out);
__global__ void k_win_dse(Dev d, const double* acc_se, double* dse);
__global__ void k_hist_l1(Dev d, int nb, const double* ranges, double* hA, double* part);
__global__ void k_hist_member(Dev d, int nb, const double* part, const double* targets, double* hist, double* res,
                              double* loss, double* gcount);
__global__ void k_hist_weight(Dev d, int nb, const double* ranges, const double* hA, const double* res,
                              const double* gcount, double* wg);
__global__ void k_hist_seed(Dev d, const double* wg, double* seed);
}  // namespace sevc

using namespace sevc;

struct KTimer {
    std::string name;
    hipEvent_t a, b;
};

struct sevc_state {
    int B = 0;
    std::vector<sevc_options> opts;
    std::vector<int> nv, ne, ng, voff, eoff, goff;
    std::vector<int> Sm;                     // each member's max_vrt_per_grn (lists are stored at stride S)
    int V = 0, E = 0, G = 0, S = 0;
    Dev d{};
    std::vector<void*> allocs;
    hipStream_t stream = nullptr;
    std::string err;
    bool any_nucl = false, any_events = false, any_adaptive = false, any_analytic = false;
    bool any_mc = false, any_rate = false;   // nucleation modes present
    bool any_aniso = false;                  // a member with GB_eps != 0 (energies change with the topology)
    bool any_debug = false;                  // a member with debug_lvl >= 1 (checks inside the step)
    bool ben_dirty = true;                   // boundary energies not computed since creation / a new scale
    bool diag_next = false;
    bool any_stop = false;                   // a member with a stop rule (tmax, nmax, leave_percent)
    std::vector<MState> ms_host, ms_prev;
    std::vector<MState> ms_poll;             // members' state read while a call runs (all finished?)
    // graph replay: gexec holds kGraphSteps steps, gexec1 one step
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec1 = nullptr;
    hipGraph_t graph1 = nullptr;
    // kernel timing
    bool timing = false;
    bool timing_hot_only = false;
    int timing_every = 1;                // sevc_step times one step in timing_every
    long long timing_steps = 0;          // steps stepped since timing was enabled
    bool timing_skip = false;            // the step being launched is not sampled
    std::vector<KTimer> pending;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    std::map<std::string, std::pair<double, long long>> ktimes;
    // grains per tile of velocity pass 1 (sevc_autotune), table capacity
    int tile_grains = kTile;
    size_t tile_cap = 0;
    // velocity path: 0 two passes over grain tiles and vertices, 1 one pass
    // over tiles of vtile_size vertices, 2 two passes with one lane per
    // grain then per vertex (1 and 2 need a list stride S that is a
    // multiple of 4, at least 8)
    int vel_path = 0;
    int vtile_size = 64;
    size_t vtile_cap = 0;
    // device checkpoints of the whole state (sevc_checkpoint / sevc_restore)
    struct Ckpt {
        char* buf = nullptr;
        size_t bytes = 0;
        std::vector<MState> ms;
    };
    std::vector<Ckpt> ckpts;
    // topology-fixed window: checkpoints of x every K steps (ck), the
    // positions and areas of one segment (seg, segA; `holds` = its index),
    // and the reverse-pass arrays
    struct Window {
        int steps = 0, K = 0, nck = 0, holds = -1;
        bool ready = false;
        double *ck = nullptr, *seg = nullptr, *segA = nullptr;
        size_t ck_n = 0, seg_n = 0, segA_n = 0;
        double *lamA = nullptr, *lamB = nullptr, *acc_mu = nullptr, *acc_s = nullptr, *acc_se = nullptr;
        double *ebase = nullptr, *part = nullptr, *out = nullptr, *dse = nullptr;
        bool seeded = false;              // lamA holds dL/dx from sevc_area_histogram
        // area histogram loss
        double *hA = nullptr, *hwg = nullptr, *hpart = nullptr, *hrange = nullptr, *htarget = nullptr;
        double *hhist = nullptr, *hres = nullptr, *hloss = nullptr, *hcount = nullptr;
    } win;
};

namespace {

// Captured steps hold table sizes and parameters: dropped when they change.
void drop_graphs(sevc_state* s) {
    if (s->gexec) { (void)hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
    if (s->gexec1) { (void)hipGraphExecDestroy(s->gexec1); s->gexec1 = nullptr; }
    if (s->graph1) { (void)hipGraphDestroy(s->graph1); s->graph1 = nullptr; }
}

// A failed sevc_batch_create: the reason goes to sevc_last_error(NULL).
int create_fail(int code, const std::string& msg) {
    set_stateless_error(std::string("sevc_batch_create: ") + msg);
    return code;
}

int fail(sevc_state* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    return code;
}

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(s, SEVC_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(sevc_state* s, T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(s, SEVC_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return SEVC_OK;
}

hipEvent_t take_event(sevc_state* s) {
    if (s->pool_used == s->pool.size()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        s->pool.push_back(e);
    }
    return s->pool[s->pool_used++];
}

// Launch helper: optional per-kernel device timing on the state's stream.
template <typename K, typename... Args>
void launch(sevc_state* s, const char* name, K kernel, int grid, int block, Args... args) {
    if (grid <= 0) return;
    bool timed = s->timing && !s->timing_skip;
    if (timed && s->timing_hot_only)
        timed = std::strcmp(name, "vel_grain") == 0 || std::strcmp(name, "velocity") == 0 ||
                std::strcmp(name, "vel_fused") == 0 || std::strcmp(name, "vel_glane") == 0 ||
                std::strcmp(name, "vel_bigrain") == 0 ||
                std::strcmp(name, "translate") == 0 || std::strcmp(name, "text") == 0;
    if (timed) {
        KTimer t{name, take_event(s), take_event(s)};
        (void)hipEventRecord(t.a, s->stream);
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s->stream, args...);
        (void)hipEventRecord(t.b, s->stream);
        s->pending.push_back(t);
    } else {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s->stream, args...);
    }
}

void collect_timers(sevc_state* s) {
    for (auto& t : s->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        auto& e = s->ktimes[t.name];
        e.first += ms;
        e.second += 1;
    }
    s->pending.clear();
    s->pool_used = 0;
}

constexpr int kListGrid = 512;
constexpr int kGraphSteps = 8;           // steps per replay of the captured graph
int count_grid(const Dev& d) { return (d.ngb + d.neb + d.nvb + kCountWaves - 1) / kCountWaves; }
// one wave per flagged grain in the fix stages: enough workgroups that
// every grain of a busy step has its own
constexpr int kFixGrid = 2048;
// one wave per nucleation event
constexpr int kNuclInsGrid = 256;
// the flip stage flags four grains per flip: a step with thousands of
// flips (the Monte Carlo workload) flags tens of thousands
constexpr int kFlipFixGrid = 8192;
constexpr int kClearGrid = kFixGrid / 4;  // workgroups of k_translate_clear that clear (4 waves each; more cost the plain step)
constexpr int kTileMin = 16;             // smallest tile the autotuner tries

// Tile table of velocity pass 1: T consecutive grains of one member per
// tile.  Any T <= kTile gives the same results (each grain's terms are
// computed the same way whatever the tiling); it only changes the speed.
int set_tiles(sevc_state* s, int T) {
    std::vector<Blk> tt;
    for (int m = 0; m < s->B; ++m)
        for (int b = 0; b < s->ng[m]; b += T)
            tt.push_back(Blk{m, s->goff[m] + b, s->goff[m] + std::min(b + T, s->ng[m])});
    if (tt.size() > s->tile_cap || T < kTileMin || T > kTile) return SEVC_E_INVALID_ARG;
    if (hipMemcpy(const_cast<Blk*>(s->d.gtile), tt.data(), tt.size() * sizeof(Blk), hipMemcpyHostToDevice) !=
        hipSuccess)
        return SEVC_E_HIP;
    s->d.ntile = static_cast<int>(tt.size());
    s->d.tile_T = T;
    s->tile_grains = T;
    // a captured step holds the old table size
    drop_graphs(s);
    return SEVC_OK;
}

constexpr int kVTileMin = 32;            // smallest vertex tile of the one-pass velocity

bool vtile_ok(const sevc_state* s) { return s->S % 4 == 0 && s->S >= 8; }

// Enabled vertices in tiles of VT vertex slots whose polygons do not fit the
// one-pass kernel's staging (more than 7 VT list entries or 2 VT distinct
// grains; k_vel_vtile sends such tiles to the two-pass path), from host
// copies of the vertex grains (global ids), enable flags and list lengths.
// Meshes whose vertex ids follow space (the generator's) have none; a mesh
// read in arbitrary vertex order has most.
long long vtile_overflow(const sevc_state* s, int VT, const std::vector<int>& vg, const std::vector<uint8_t>& ven,
                         const std::vector<int>& glen) {
    long long bad = 0;
    std::vector<int> gs;
    for (int m = 0; m < s->B; ++m)
        for (int b = 0; b < s->nv[m]; b += VT) {
            gs.clear();
            int nen = 0;
            for (int i = s->voff[m] + b; i < s->voff[m] + std::min(b + VT, s->nv[m]); ++i) {
                if (!ven[i]) continue;
                ++nen;
                for (int j = 0; j < 3; ++j) gs.push_back(vg[3 * static_cast<size_t>(i) + j]);
            }
            std::sort(gs.begin(), gs.end());
            gs.erase(std::unique(gs.begin(), gs.end()), gs.end());
            long long pts = 0;
            for (int g : gs) pts += g >= 0 && g < s->G ? glen[g] : 0;
            if (pts > 7LL * VT || static_cast<long long>(gs.size()) > 2LL * VT) bad += nen;
        }
    return bad;
}

// The one-pass tile size whose staging overflows for at most 1 % of the
// enabled vertices (64 preferred), or 0 when neither fits (two-pass path).
int pick_vtile(const sevc_state* s, const std::vector<int>& vg, const std::vector<uint8_t>& ven,
               const std::vector<int>& glen) {
    if (!vtile_ok(s)) return 0;
    long long nen = 0;
    for (uint8_t e : ven) nen += e ? 1 : 0;
    for (int VT : {64, 32})
        if (100 * vtile_overflow(s, VT, vg, ven, glen) <= nen) return VT;
    return 0;
}

// Tile table of the one-pass velocity: VT consecutive vertex slots of one
// member per tile (VT = 32 or 64).  Results do not depend on the tiling.
int set_vtiles(sevc_state* s, int VT) {
    if (VT != 32 && VT != 64) return SEVC_E_INVALID_ARG;
    std::vector<Blk> tt;
    for (int m = 0; m < s->B; ++m)
        for (int b = 0; b < s->nv[m]; b += VT)
            tt.push_back(Blk{m, s->voff[m] + b, s->voff[m] + std::min(b + VT, s->nv[m])});
    if (tt.size() > s->vtile_cap) return SEVC_E_INVALID_ARG;
    if (hipMemcpy(const_cast<Blk*>(s->d.vtile), tt.data(), tt.size() * sizeof(Blk), hipMemcpyHostToDevice) !=
        hipSuccess)
        return SEVC_E_HIP;
    s->d.nvtile = static_cast<int>(tt.size());
    s->vtile_size = VT;
    drop_graphs(s);
    return SEVC_OK;
}

// The FD velocity of the active path (both are bit-identical).
constexpr int kBigGrid = 256;            // one wave per long grain (lane-per-grain path)

void launch_velocity(sevc_state* s, const Dev& d) {
    if (s->vel_path == 2) {
        launch(s, "vel_glane", k_vel_glane, d.ngb, kBlock, d, 0);
        launch(s, "vel_bigrain", k_vel_bigrain, kBigGrid, 64, d, 0);
        launch(s, "velocity", k_vel_vertex, d.nvb, kBlock, d, 0);
    } else if (s->vel_path == 1) {
        if (s->vtile_size == 32) launch(s, "vel_fused", k_vel_vtile<32>, d.nvtile, 128, d, 0);
        else launch(s, "vel_fused", k_vel_vtile<64>, d.nvtile, 256, d, 0);
    } else {
        launch(s, "vel_grain", k_vel_grain, d.ntile, 256, d, 0);
        launch(s, "velocity", k_vel_vertex, d.nvb, kBlock, d, 0);
    }
    launch(s, "velocity_fallback", k_velocity_fallback, 64, 256, d);
}

// The device check of check_data over the state's slots; gate = 1 keeps
// it to the active members with debug_lvl >= 1.
ChkView check_view(const Dev& d, int gate) {
    ChkView c{};
    c.vb = d.vb; c.vg = d.vg; c.ven = d.ven;
    c.bvx = d.bvx; c.benab = d.benab;
    c.gvx = d.gvx; c.glen = d.glen; c.genab = d.genab;
    c.S = d.S; c.V = d.V; c.E = d.E; c.G = d.G; c.B = d.B;
    c.vmem = d.vmem; c.emem = d.emem; c.gmem = d.gmem;
    c.prm = d.prm; c.ms = d.ms; c.gate = gate; c.bad = d.chk_bad;
    return c;
}

int check_grid(long long slots) { return static_cast<int>(std::min<long long>((slots + 255) / 256, 8192)); }

// A check inside the step (main.cu:185-196, 237-248, 269-280, 297-308):
// members with violations stop with SEVC_E_CHECK and the phase.
void launch_check(sevc_state* s, int phase) {
    const Dev& d = s->d;
    launch(s, "check", k_check_clear, 1, 256, d.chk_bad, s->B);
    launch(s, "check", k_check, check_grid((long long)d.V + d.E + d.G), 256, check_view(d, 1));
    launch(s, "check", k_check_stop, 1, 256, d.prm, d.ms, static_cast<const int*>(d.chk_bad), s->B, phase);
}

// One time step of every active member (main.cu:124-389).
void launch_step(sevc_state* s, bool diag) {
    Dev d = s->d;
    d.diag = diag ? 1 : 0;
    // Lengths come from k_text.  Energies change from step to step only with
    // GB_eps != 0; with GB_eps = 0 they are the scale, computed at the first
    // step after creation or after a new scale (for the members active then,
    // as the reference's per-step pass would).
    const bool all_energies = diag || s->any_aniso || s->ben_dirty;
    // the rate mode's candidate draw runs in the begin's launch
    if (s->any_nucl && s->any_rate)
        launch(s, "begin", k_begin_rate, 1 + d.ngb, kBlock, d, all_energies ? 0 : 1);
    else
        launch(s, "begin", k_begin, 1, 1024, d, all_energies ? 0 : 1);
    if (all_energies) launch(s, "edge_len_energy", k_edge_len_energy, d.neb, kBlock, d, 0);
    s->ben_dirty = false;
    if (diag || s->any_analytic) launch(s, "grain_area", k_grain_area, d.ngb, kBlock, d, 0);
    if (diag) launch(s, "vertex_energy", k_vertex_energy, d.nvb, kBlock, d);
    if (s->any_nucl) {
        launch(s, "nucl_manager", k_nucl_manager, s->B, kNuclThreads, d);
        launch(s, "nucl_insert", k_nucl_insert, kNuclInsGrid, 64, d);
        if (diag) launch(s, "vertex_energy", k_vertex_energy, d.nvb, kBlock, d);
    }
    if (s->any_debug) launch_check(s, CHK_NUCLEATION);
    if (diag) {
        launch(s, "velocity_generic", k_velocity_generic, d.nvb, kBlock, d);
    } else {
        launch_velocity(s, d);
    }
    if (s->any_analytic) launch(s, "velocity_analytic", k_velocity_analytic, d.nvb, kBlock, d);
    if (diag) launch(s, "grain_dadt", k_grain_dadt, d.ngb, kBlock, d);
    launch(s, "text", k_text, d.neb, kBlock, d);
    if (s->any_events) {
        launch(s, "label", k_label, d.neb, kBlock, d);
        launch(s, "poll_gather", k_poll_gather, kPollBig / 64, 64, d);
        launch(s, "poll", k_poll, 1, 1024, d);
        launch(s, "scan3", k_scan3, d.ngb, kBlock, d);
        launch(s, "two_grain", k_two_grain, 1, 64, d);
        if (s->any_debug) launch_check(s, CHK_TWO_GRAIN);
        launch(s, "label3", k_label3, d.ngb, kBlock, d);     // only after a two-triangle event
        launch(s, "t2_par", k_t2_par, kListGrid, 256, d);
        launch(s, "t2_serial", k_t2_serial, 1, 64, d);
        launch(s, "fix", k_fix, kFixGrid, 64, d, 1);
        if (s->any_debug) {
            // the checks read the cleared fix-up state
            launch(s, "fix_clear", k_fix_clear, kFixGrid, 64, d, 1, 1);
            launch_check(s, CHK_REMOVAL);
            launch(s, "st2", k_st2, kListGrid, 256, d);
        } else {
            launch(s, "st2_fix_clear", k_st2_fix_clear, kFixGrid + 4 * kListGrid, 64, d, kFixGrid);
        }
        launch(s, "flip_par", k_flip_par, kListGrid, 256, d);
        launch(s, "flip_serial", k_flip_serial, 1, 64, d);
        launch(s, "fix", k_fix, kFlipFixGrid, 64, d, 2);
        if (s->any_debug) {
            launch(s, "fix_clear", k_fix_clear, kFixGrid, 64, d, 2, 0);
            launch_check(s, CHK_FLIP);
        }
    }
    if (s->any_events && !s->any_debug)
        launch(s, "translate", k_translate_clear, kClearGrid + d.nvb, kBlock, d, kClearGrid);
    else
        launch(s, "translate", k_translate, d.nvb, kBlock, d);
    // the per-block slot counts are kept by the events themselves
    launch(s, "member_reduce", k_member_reduce, s->B, 1024, d);
}

void build_table(std::vector<Blk>& tbl, std::vector<int>& first, const std::vector<int>& off,
                 const std::vector<int>& n) {
    tbl.clear();
    first.assign(off.size() + 1, 0);
    for (size_t m = 0; m < off.size(); ++m) {
        first[m] = static_cast<int>(tbl.size());
        for (int b = 0; b < n[m]; b += kBlock)
            tbl.push_back(Blk{static_cast<int>(m), off[m] + b, off[m] + std::min(b + kBlock, n[m])});
    }
    first[off.size()] = static_cast<int>(tbl.size());
}

int fill_stats(sevc_state* s, int m, sevc_step_stats* st) {
    if (!st) return SEVC_OK;
    const MState& a = s->ms_host[static_cast<size_t>(m)];
    const MState& p = s->ms_prev[static_cast<size_t>(m)];
    std::memset(st, 0, sizeof(*st));
    st->steps_done = static_cast<int32_t>(a.steps - p.steps);
    st->n = a.n;
    st->t = a.t;
    st->dt_last = a.dt;
    st->n_grains = a.curr_grains;
    st->n_vertices = a.n_venabled;
    st->nucleated_area_pct = a.nucl_area_pct;
    st->done = a.done;
    st->status = a.status;
    st->n_flips = a.flips - p.flips;
    st->n_removals = a.removals - p.removals;
    st->n_two_grain = a.two - p.two;
    st->n_nucleations = a.nucl - p.nucl;
    st->n_vertex_steps = a.vsteps - p.vsteps;
    return SEVC_OK;
}

}  // namespace

extern "C" {

int sevc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* sevc_last_error(const sevc_state* s) { return s ? s->err.c_str() : stateless_error(); }

int32_t sevc_n_members(const sevc_state* s) { return s ? s->B : 0; }

int sevc_batch_create(const sevc_options* opts, const sevc_mesh* meshes, int32_t B, sevc_state** out) {
    if (!opts || !meshes || B <= 0 || !out) return create_fail(SEVC_E_INVALID_ARG, "null arguments or no members");
    *out = nullptr;
    // enumerated options and the domain: an unknown velocity model would
    // leave the velocities unwritten, an unknown nucleation mode would run
    // as another one
    for (int m = 0; m < B; ++m) {
        const sevc_options& o = opts[m];
        if (o.velocity_model != SEVC_VEL_FD && o.velocity_model != SEVC_VEL_ANALYTIC)
            return create_fail(SEVC_E_INVALID_ARG, "member " + std::to_string(m) + ": unknown velocity_model");
        if (o.nucleation_mode != SEVC_NUCL_MC && o.nucleation_mode != SEVC_NUCL_RATE)
            return create_fail(SEVC_E_INVALID_ARG, "member " + std::to_string(m) + ": unknown nucleation_mode");
        if (!(o.domain_bound > 0) || !(meshes[m].domain_bound > 0))
            return create_fail(SEVC_E_INVALID_ARG, "member " + std::to_string(m) + ": domain_bound not positive");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SEVC_E_NO_DEVICE;
    sevc_state* s = new sevc_state();
    s->B = B;
    s->opts.assign(opts, opts + B);
    s->S = 0;
    for (int m = 0; m < B; ++m) {
        const sevc_mesh& me = meshes[m];
        if (me.n_vertices <= 0 || me.n_boundaries <= 0 || me.n_grains <= 0 || me.max_vrt_per_grn < 3 ||
            !me.pos || !me.v_bnd || !me.v_grn || !me.v_enabled || !me.b_vtx || !me.b_enabled || !me.g_vtx ||
            !me.g_vlen || !me.g_enabled || !me.g_SE || !me.g_orientation) {
            delete s;
            return create_fail(SEVC_E_INVALID_ARG, "member " + std::to_string(m) + ": empty mesh or missing arrays");
        }
        // Kernels index through every stored id of an enabled entity, so
        // an inconsistent input is rejected here (utils.h:85-336 checks).
        double total_area = 0.0;
        if (sevc_check_mesh(&me, &total_area) != 0) {
            delete s;
            return SEVC_E_TOPOLOGY;
        }
        s->S = std::max(s->S, static_cast<int>(me.max_vrt_per_grn));
        s->Sm.push_back(me.max_vrt_per_grn);
        s->voff.push_back(s->V); s->eoff.push_back(s->E); s->goff.push_back(s->G);
        s->nv.push_back(me.n_vertices); s->ne.push_back(me.n_boundaries); s->ng.push_back(me.n_grains);
        s->V += me.n_vertices; s->E += me.n_boundaries; s->G += me.n_grains;
        if (opts[m].dt <= 0) { delete s; return create_fail(SEVC_E_INVALID_ARG, "member " + std::to_string(m) + ": dt <= 0"); }
        if (opts[m].events && opts[m].do_nucleation) {
            s->any_nucl = true;
            if (opts[m].nucleation_mode == SEVC_NUCL_MC) s->any_mc = true;
            else s->any_rate = true;
        }
        if (opts[m].events) s->any_events = true;
        if (opts[m].adaptive_dt) s->any_adaptive = true;
        if (opts[m].velocity_model == SEVC_VEL_ANALYTIC) s->any_analytic = true;
        if (opts[m].GB_eps != 0.0) s->any_aniso = true;
        if (opts[m].debug_lvl >= 1) s->any_debug = true;
    }
    if (s->S > 128) { delete s; return create_fail(SEVC_E_INVALID_ARG, "grain list stride above 128"); }
    const size_t V = static_cast<size_t>(s->V), E = static_cast<size_t>(s->E), G = static_cast<size_t>(s->G);
    const size_t S = static_cast<size_t>(s->S);
    Dev& d = s->d;
    d.S = s->S;
    d.B = B;
    d.V = s->V;
    d.E = s->E;
    d.G = s->G;
    d.debug = s->any_debug ? 1 : 0;
    d.poll_fast = kPollFast;
    d.poll_lds = kPollBig;
    d.free_rows = 0;
    int rc = SEVC_OK;
#define DA(p, n) do { if ((rc = dalloc(s, &(p), (n))) != SEVC_OK) { sevc_destroy(s); return rc; } } while (0)
    DA(d.pos, 2 * V); DA(d.vel, 2 * V); DA(d.vb, 3 * V); DA(d.vg, 3 * V); DA(d.ven, V);
    DA(d.venergy, V); DA(d.vparts, 4 * V);
    DA(d.mods, V);
    DA(d.nucl_r, V); DA(d.nucl_ori, V); DA(d.dE, V); DA(d.nfac, V); DA(d.vnuc, V);
    DA(d.bvx, 2 * E); DA(d.ben, E); DA(d.blen, E); DA(d.bt, E); DA(d.benab, E); DA(d.bcand, E); DA(d.conf, E);
    DA(d.pgat, 25 * kPollBig); DA(d.pgt, kPollBig); DA(d.blab, E); DA(d.l_flipped, E + 16); DA(d.n_flipped, 1);
    DA(d.vmem, V); DA(d.emem, E); DA(d.gmem, G);
    DA(d.chk_bad, static_cast<size_t>(B));
    DA(d.gvx, G * S); DA(d.glen, G); DA(d.garea, G); DA(d.gdadt, G); DA(d.gori, G); DA(d.gse, G);
    DA(d.genab, G); DA(d.gtype, G); DA(d.gcand, G); DA(d.gfix, G);
    DA(d.cnt, C_COUNT);
    DA(d.l_cand, E); DA(d.l_two, E); DA(d.l_t2, G); DA(d.l_t2c, G);
    DA(d.l_flipc, E);
    for (int k = 0; k < kFixStages; ++k) DA(d.l_fix[k], G);
    DA(d.l_dirty, G); DA(d.l_nucl, G); DA(d.l_conf, 4 * E + 16);
    DA(d.nscr, 12 * G + 16);
    DA(d.l_vfall, V + 16);
    DA(d.l_bigr, G + 16);
    DA(d.nrec, 4 * G + 16); DA(d.nal, G + 16); DA(d.l_nev, 8 * G + 16);

    // block tables
    std::vector<Blk> vt, et, gt;
    std::vector<int> vf, ef, gf;
    build_table(vt, vf, s->voff, s->nv);
    build_table(et, ef, s->eoff, s->ne);
    build_table(gt, gf, s->goff, s->ng);
    d.nvb = static_cast<int>(vt.size()); d.neb = static_cast<int>(et.size()); d.ngb = static_cast<int>(gt.size());
    Blk *dvt, *det, *dgt;
    DA(dvt, vt.size()); DA(det, et.size()); DA(dgt, gt.size());
    DA(d.vblk_first, vf.size()); DA(d.eblk_first, ef.size()); DA(d.gblk_first, gf.size());
    DA(d.bc_v, vt.size()); DA(d.bc_e, et.size()); DA(d.bc_g, gt.size());
    DA(d.partial, 2 * gt.size() + 2);
    d.vblk = dvt; d.eblk = det; d.gblk = dgt;
    {
        // room for the smallest tile size the autotuner tries
        size_t cap = 0;
        for (int m = 0; m < B; ++m) cap += (s->ng[m] + kTileMin - 1) / kTileMin;
        Blk* dtt;
        DA(dtt, cap);
        d.gtile = dtt;
        s->tile_cap = cap;
        if ((rc = set_tiles(s, kTile)) != SEVC_OK) {
            sevc_destroy(s);
            return create_fail(rc, "grain tile table");
        }
        size_t vcap = 0;
        for (int m = 0; m < B; ++m) vcap += (s->nv[m] + kVTileMin - 1) / kVTileMin;
        Blk* dvt2;
        DA(dvt2, vcap);
        d.vtile = dvt2;
        s->vtile_cap = vcap;
        if ((rc = set_vtiles(s, 64)) != SEVC_OK) {
            sevc_destroy(s);
            return create_fail(rc, "vertex tile table (64)");
        }
    }
    DA(d.vterm, 6 * V);
    DA(d.slist, static_cast<size_t>(d.ngb) + d.neb + d.nvb);
    Params* dprm;
    DA(dprm, static_cast<size_t>(B));
    DA(d.ms, static_cast<size_t>(B));
    d.prm = dprm;
#undef DA
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        sevc_destroy(s);
        return SEVC_E_HIP;
    }
#define UP(dst, src, n) \
    do { if (hipMemcpy((dst), (src), (n), hipMemcpyHostToDevice) != hipSuccess) { sevc_destroy(s); return SEVC_E_HIP; } } while (0)
    UP(dvt, vt.data(), vt.size() * sizeof(Blk));
    UP(det, et.data(), et.size() * sizeof(Blk));
    UP(dgt, gt.data(), gt.size() * sizeof(Blk));
    UP(d.vblk_first, vf.data(), vf.size() * sizeof(int));
    UP(d.eblk_first, ef.data(), ef.size() * sizeof(int));
    UP(d.gblk_first, gf.data(), gf.size() * sizeof(int));

    // host staging with global ids
    std::vector<double> pos(2 * V), vel(2 * V, 0.0), ben(E, 0.0), blen(E, 0.0), bt(E, SEVC_MAX_T_EXT);
    std::vector<double> garea(G, 0.0), gori(G, 0.0), gse(G, 0.0), venergy(V, 0.0), vparts(4 * V, 0.0), gdadt(G, 0.0);
    std::vector<int> vb(3 * V), vg(3 * V), bvx(2 * E), gvx(G * S, -1), glen(G, 0);
    std::vector<uint8_t> ven(V), benab(E), bcand(E, 0), genab(G), gtype(G, 0);
    std::vector<Params> prm(static_cast<size_t>(B));
    s->ms_host.assign(static_cast<size_t>(B), MState{});
    auto gid = [](int x, int off) { return x < 0 ? -1 : x + off; };
    for (int m = 0; m < B; ++m) {
        const sevc_mesh& me = meshes[m];
        const sevc_options& o = opts[m];
        const int vo = s->voff[m], eo = s->eoff[m], go = s->goff[m];
        const int Sm = me.max_vrt_per_grn;
        int ngr = 0, nve = 0;
        for (int i = 0; i < me.n_vertices; ++i) {
            const size_t I = static_cast<size_t>(vo + i);
            pos[2 * I] = me.pos[2 * i]; pos[2 * I + 1] = me.pos[2 * i + 1];
            if (me.vel) { vel[2 * I] = me.vel[2 * i]; vel[2 * I + 1] = me.vel[2 * i + 1]; }
            for (int k = 0; k < 3; ++k) {
                vb[3 * I + k] = gid(me.v_bnd[3 * i + k], eo);
                vg[3 * I + k] = gid(me.v_grn[3 * i + k], go);
            }
            ven[I] = me.v_enabled[i] ? 1 : 0;
            nve += ven[I];
            if (me.v_energy) venergy[I] = me.v_energy[i];
            if (me.v_vel_parts) for (int k = 0; k < 4; ++k) vparts[4 * I + k] = me.v_vel_parts[4 * i + k];
        }
        for (int b = 0; b < me.n_boundaries; ++b) {
            const size_t I = static_cast<size_t>(eo + b);
            bvx[2 * I] = gid(me.b_vtx[2 * b], vo); bvx[2 * I + 1] = gid(me.b_vtx[2 * b + 1], vo);
            benab[I] = me.b_enabled[b] ? 1 : 0;
            if (me.b_energy) ben[I] = me.b_energy[b];
            if (me.b_arclength) blen[I] = me.b_arclength[b];
            if (me.b_t_ext) bt[I] = me.b_t_ext[b];
            if (me.b_candidate) bcand[I] = me.b_candidate[b];
        }
        for (int g = 0; g < me.n_grains; ++g) {
            const size_t I = static_cast<size_t>(go + g);
            glen[I] = me.g_vlen[g];
            if (glen[I] > Sm || glen[I] < 0) {
                sevc_destroy(s);
                return create_fail(SEVC_E_INVALID_ARG, "member " + std::to_string(m) + ": grain list longer than its stride");
            }
            for (int k = 0; k < glen[I]; ++k) gvx[I * S + k] = gid(me.g_vtx[static_cast<size_t>(g) * Sm + k], vo);
            genab[I] = me.g_enabled[g] ? 1 : 0;
            ngr += genab[I];
            gori[I] = me.g_orientation[g];
            gse[I] = me.g_SE[g];
            if (me.g_area) garea[I] = me.g_area[g];
            if (me.g_type) gtype[I] = me.g_type[g];
            if (me.g_dAdt) gdadt[I] = me.g_dAdt[g];
        }
        Params& P = prm[static_cast<size_t>(m)];
        P.D = me.domain_bound;
        P.dt_max = o.dt; P.dt_min = o.dt_min; P.dt_safety = o.dt_safety; P.flip_thr = o.flip_threshold;
        P.mobility = o.mobility; P.gb_eps = o.GB_eps; P.gb_scaling = o.GB_scaling; P.mc_k = o.MC_k;
        P.nucl_rate = o.nucleation_rate; P.nucl_scaling = o.nucleation_scaling; P.tmax = o.tmax;
        P.seed = o.seed; P.nmax = o.nmax; P.adaptive = o.adaptive_dt; P.nucl_mode = o.nucleation_mode;
        P.do_nucl = o.do_nucleation; P.gap = o.nucleation_gap; P.events = o.events; P.vel_model = o.velocity_model;
        P.debug = o.debug_lvl;
        P.cap = me.max_vrt_per_grn;
        P.n_grains0 = ngr;
        P.gate_grains = o.nucleation_gate >= 0 ? o.nucleation_gate : 5 * ngr / 6;
        P.min_grains = o.leave_percent < 0 ? -1 : static_cast<int>(std::floor(o.leave_percent * ngr));
        if (P.tmax > 0 || P.nmax > 0 || P.min_grains > 0) s->any_stop = true;
        P.voff = vo; P.nv = me.n_vertices; P.eoff = eo; P.ne = me.n_boundaries; P.goff = go; P.ng = me.n_grains;
        MState& M = s->ms_host[static_cast<size_t>(m)];
        M.curr_grains = ngr;
        M.n_venabled = nve;
        M.dt = o.dt;
        M.tau_bits = 0x7FF0000000000000ull;
    }
    s->ms_prev = s->ms_host;
    {
        // the two-pass velocity (measured the faster at 1M grains); the
        // one-pass tiling that fits the vertex order is prepared for
        // sevc_autotune / sevc_set_velocity_config
        const int VT = pick_vtile(s, vg, ven, glen);
        if (VT > 0 && (rc = set_vtiles(s, VT)) != SEVC_OK) {
            sevc_destroy(s);
            return create_fail(rc, "vertex tile table");
        }
        s->vel_path = 0;
    }
    UP(d.pos, pos.data(), pos.size() * 8); UP(d.vel, vel.data(), vel.size() * 8);
    UP(d.vb, vb.data(), vb.size() * 4); UP(d.vg, vg.data(), vg.size() * 4); UP(d.ven, ven.data(), V);
    UP(d.venergy, venergy.data(), V * 8); UP(d.vparts, vparts.data(), 4 * V * 8);
    UP(d.bvx, bvx.data(), bvx.size() * 4); UP(d.ben, ben.data(), E * 8); UP(d.blen, blen.data(), E * 8);
    UP(d.bt, bt.data(), E * 8); UP(d.benab, benab.data(), E); UP(d.bcand, bcand.data(), E);
    UP(d.gvx, gvx.data(), gvx.size() * 4); UP(d.glen, glen.data(), G * 4); UP(d.garea, garea.data(), G * 8);
    UP(d.gdadt, gdadt.data(), G * 8); UP(d.gori, gori.data(), G * 8); UP(d.gse, gse.data(), G * 8);
    UP(d.genab, genab.data(), G); UP(d.gtype, gtype.data(), G);
    UP(dprm, prm.data(), prm.size() * sizeof(Params));
    {
        std::vector<int> vm(V), em(E), gm(G);
        for (int m = 0; m < B; ++m) {
            std::fill(vm.begin() + s->voff[m], vm.begin() + s->voff[m] + s->nv[m], m);
            std::fill(em.begin() + s->eoff[m], em.begin() + s->eoff[m] + s->ne[m], m);
            std::fill(gm.begin() + s->goff[m], gm.begin() + s->goff[m] + s->ng[m], m);
        }
        UP(d.vmem, vm.data(), V * 4); UP(d.emem, em.data(), E * 4); UP(d.gmem, gm.data(), G * 4);
    }
    UP(d.ms, s->ms_host.data(), s->ms_host.size() * sizeof(MState));
#undef UP
    // scratch initial values
    std::vector<int> minus1(std::max(4 * V, E), -1);
    std::vector<ModRec> mods_none(V);
    for (ModRec& r : mods_none) {
        for (int j = 0; j < 4; j++) { r.grn[j] = -1; r.vrt[j] = -1; r.prev[j] = -1; r.fix[j] = FIX_NONE; }
    }
    if (hipMemcpy(d.mods, mods_none.data(), V * sizeof(ModRec), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.conf, minus1.data(), E * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(d.gfix, 0, G * 4) != hipSuccess || hipMemset(d.gcand, 0, G) != hipSuccess ||
        hipMemset(d.vnuc, 0, V) != hipSuccess || hipMemset(d.cnt, 0, C_COUNT * 4) != hipSuccess ||
        hipMemset(d.n_flipped, 0, 4) != hipSuccess ||
        hipMemset(d.nucl_r, 0, V * 8) != hipSuccess || hipMemset(d.nucl_ori, 0, V * 8) != hipSuccess ||
        hipMemset(d.dE, 0, V * 8) != hipSuccess || hipMemset(d.nfac, 0, V * 8) != hipSuccess) {
        sevc_destroy(s);
        return SEVC_E_HIP;
    }
    // per-block (enabled, free) counts, which the first step's nucleation reads
    hipLaunchKernelGGL(k_count_all, dim3(count_grid(d)), dim3(64 * kCountWaves), 0, s->stream, d);
    if (hipStreamSynchronize(s->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
        sevc_destroy(s);
        return SEVC_E_HIP;
    }
    *out = s;
    return SEVC_OK;
}

int sevc_create(const sevc_options* o, const sevc_mesh* m, sevc_state** out) {
    return sevc_batch_create(o, m, 1, out);
}

void sevc_destroy(sevc_state* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    for (auto& c : s->ckpts) (void)hipFree(c.buf);
    for (void* p : s->allocs) (void)hipFree(p);
    for (auto e : s->pool) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int sevc_request_diagnostics(sevc_state* s) {
    if (!s) return SEVC_E_INVALID_ARG;
    s->diag_next = true;
    return SEVC_OK;
}

int sevc_enable_kernel_timing(sevc_state* s, int32_t on) {
    if (!s) return SEVC_E_INVALID_ARG;
    s->timing = on != 0;
    s->timing_hot_only = on == 2;
    s->timing_steps = 0;
    s->ktimes.clear();
    return SEVC_OK;
}

int sevc_set_timing_sample(sevc_state* s, int32_t every) {
    if (!s || every < 1) return SEVC_E_INVALID_ARG;
    s->timing_every = every;
    s->timing_steps = 0;
    return SEVC_OK;
}

int sevc_kernel_times(sevc_state* s, int32_t max_entries, char (*names)[48], double* ms, int64_t* launches) {
    if (!s) return SEVC_E_INVALID_ARG;
    int k = 0;
    for (auto& kv : s->ktimes) {
        if (k >= max_entries) break;
        std::snprintf(names[k], 48, "%s", kv.first.c_str());
        ms[k] = kv.second.first;
        launches[k] = kv.second.second;
        ++k;
    }
    return k;
}

static int finish_call(sevc_state* s) {
    Dev d = s->d;
    launch(s, "nucl_area_l1", k_nucl_area_l1, d.ngb, 64, d);
    launch(s, "nucl_area_up", k_nucl_area_up, s->B, 1024, d);
    s->ms_prev = s->ms_host;
    HIPCHK(hipMemcpyAsync(s->ms_host.data(), d.ms, s->ms_host.size() * sizeof(MState), hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->timing) collect_timers(s);
    // a member whose event met an inconsistent state or a full grain list
    // during this call stops alone; the call reports the first such member
    for (int m = 0; m < s->B; ++m) {
        const MState& a = s->ms_host[static_cast<size_t>(m)];
        const int was = s->ms_prev[static_cast<size_t>(m)].status;
        if (a.status != 0 && a.status != SEVC_E_CHECK && (was == 0 || s->B == 1))
            return fail(s, a.status, std::string("device event error: ") + sevc_strerror(a.status) +
                                         " (member " + std::to_string(m) + ", step " + std::to_string(a.n) + ")");
    }
    // a member whose check failed during this call (debug_lvl >= 1): the
    // reference's "Error during ..." lines (main.cu:193, 245, 277, 305)
    static const char* const kPhase[] = {"", "nucleation", "remove two grains configuration", "Grain Removal",
                                         "flipping"};
    for (int m = 0; m < s->B; ++m) {
        const MState& a = s->ms_host[static_cast<size_t>(m)];
        if (a.status == SEVC_E_CHECK && s->ms_prev[static_cast<size_t>(m)].status != SEVC_E_CHECK) {
            const int ph = a.chk_phase >= 1 && a.chk_phase <= 4 ? a.chk_phase : 0;
            return fail(s, SEVC_E_CHECK, std::string("Error during ") + kPhase[ph] + " (member " +
                                             std::to_string(m) + ", step " + std::to_string(a.n) + ")");
        }
    }
    return SEVC_OK;
}

// Every member finished (stop rule or error)?  Steps after that do nothing
// on the device, so a call stops launching them.  Read every stop_poll(s)
// steps: 64 when a member has a stop rule, else 512 (an error can still
// end every member).
static int stop_poll(const sevc_state* s) { return s->any_stop ? 64 : 512; }

static bool all_done(sevc_state* s) {
    if (s->ms_poll.size() != static_cast<size_t>(s->B)) s->ms_poll.resize(static_cast<size_t>(s->B));
    if (hipMemcpyAsync(s->ms_poll.data(), s->d.ms, sizeof(MState) * s->ms_poll.size(), hipMemcpyDeviceToHost,
                       s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return false;                            // the call's end reports the error
    for (const MState& M : s->ms_poll)
        if (!M.done) return false;
    return true;
}

int sevc_step(sevc_state* s, int32_t n_steps, sevc_step_stats* st) {
    if (!s || n_steps < 0) return SEVC_E_INVALID_ARG;
    if (n_steps > 0) s->win.ready = false;       // the topology may change
    const int poll = stop_poll(s);
    for (int k = 0; k < n_steps; ++k) {
        bool diag = s->diag_next;
        s->diag_next = false;
        s->timing_skip = s->timing && (s->timing_steps++ % s->timing_every) != 0;
        launch_step(s, diag);
        s->timing_skip = false;
        if (s->timing && s->pending.size() > 4000) {
            (void)hipStreamSynchronize(s->stream);
            collect_timers(s);
        }
        if ((k + 1) % poll == 0 && k + 1 < n_steps && all_done(s)) break;
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(s, SEVC_E_HIP, std::string("launch: ") + hipGetErrorString(le));
    int rc = finish_call(s);
    fill_stats(s, 0, st);
    return rc;
}

int sevc_step_graph(sevc_state* s, int32_t n_steps, sevc_step_stats* st) {
    if (!s || n_steps < 0) return SEVC_E_INVALID_ARG;
    if (n_steps > 0) s->win.ready = false;
    if (!s->gexec) {
        bool timing = s->timing;
        s->timing = false;
        // the graphs hold the steady step (kGraphSteps of them, and one); a
        // step that must recompute the energies runs outside them
        const bool dirty = s->ben_dirty;
        s->ben_dirty = false;
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kGraphSteps; ++k) launch_step(s, false);
        HIPCHK(hipStreamEndCapture(s->stream, &s->graph));
        HIPCHK(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
        HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        launch_step(s, false);
        HIPCHK(hipStreamEndCapture(s->stream, &s->graph1));
        HIPCHK(hipGraphInstantiate(&s->gexec1, s->graph1, nullptr, nullptr, 0));
        s->ben_dirty = dirty;
        s->timing = timing;
    }
    const int poll = stop_poll(s);
    for (int k = 0, next_poll = poll; k < n_steps;) {
        if (s->diag_next || s->ben_dirty) {
            const bool diag = s->diag_next;
            s->diag_next = false;
            launch_step(s, diag);
            k += 1;
        } else if (n_steps - k >= kGraphSteps) {
            HIPCHK(hipGraphLaunch(s->gexec, s->stream));
            k += kGraphSteps;
        } else {
            HIPCHK(hipGraphLaunch(s->gexec1, s->stream));
            k += 1;
        }
        if (k >= next_poll && k < n_steps) {
            next_poll = k + poll;
            if (all_done(s)) break;
        }
    }
    int rc = finish_call(s);
    fill_stats(s, 0, st);
    return rc;
}

int sevc_member_stats(sevc_state* s, int32_t m, sevc_step_stats* st) {
    if (!s || m < 0 || m >= s->B) return SEVC_E_INVALID_ARG;
    return fill_stats(s, m, st);
}

int sevc_member_sizes(const sevc_state* s, int32_t m, int32_t* nv, int32_t* ne, int32_t* ng, int32_t* S) {
    if (!s || m < 0 || m >= s->B) return SEVC_E_INVALID_ARG;
    if (nv) *nv = s->nv[m];
    if (ne) *ne = s->ne[m];
    if (ng) *ng = s->ng[m];
    if (S) *S = s->Sm[m];
    return SEVC_OK;
}

int sevc_get_mesh(sevc_state* s, int32_t m, sevc_mesh* out) {
    if (!s || !out || m < 0 || m >= s->B) return SEVC_E_INVALID_ARG;
    const int nv = s->nv[m], ne = s->ne[m], ng = s->ng[m], S = s->Sm[m];
    if (out->n_vertices != nv || out->n_boundaries != ne || out->n_grains != ng || out->max_vrt_per_grn != S)
        return fail(s, SEVC_E_INVALID_ARG, "output mesh sizes differ from the member's");
    HIPCHK(hipStreamSynchronize(s->stream));
    const Dev& d = s->d;
    const size_t vo = static_cast<size_t>(s->voff[m]), eo = static_cast<size_t>(s->eoff[m]),
                 go = static_cast<size_t>(s->goff[m]);
    auto down = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    };
    bool ok = true;
    out->domain_bound = s->opts[m].domain_bound;
    ok &= down(out->pos, d.pos + 2 * vo, 16 * static_cast<size_t>(nv));
    if (out->vel) ok &= down(out->vel, d.vel + 2 * vo, 16 * static_cast<size_t>(nv));
    ok &= down(out->v_bnd, d.vb + 3 * vo, 12 * static_cast<size_t>(nv));
    ok &= down(out->v_grn, d.vg + 3 * vo, 12 * static_cast<size_t>(nv));
    ok &= down(out->v_enabled, d.ven + vo, static_cast<size_t>(nv));
    if (out->v_energy) ok &= down(out->v_energy, d.venergy + vo, 8 * static_cast<size_t>(nv));
    if (out->v_vel_parts) ok &= down(out->v_vel_parts, d.vparts + 4 * vo, 32 * static_cast<size_t>(nv));
    ok &= down(out->b_vtx, d.bvx + 2 * eo, 8 * static_cast<size_t>(ne));
    if (out->b_energy) ok &= down(out->b_energy, d.ben + eo, 8 * static_cast<size_t>(ne));
    if (out->b_arclength) ok &= down(out->b_arclength, d.blen + eo, 8 * static_cast<size_t>(ne));
    if (out->b_t_ext) ok &= down(out->b_t_ext, d.bt + eo, 8 * static_cast<size_t>(ne));
    ok &= down(out->b_enabled, d.benab + eo, static_cast<size_t>(ne));
    if (out->b_candidate) ok &= down(out->b_candidate, d.bcand + eo, static_cast<size_t>(ne));
    // the member's lists at its own stride S (held at the batch's stride d.S;
    // no list is longer than S)
    ok &= hipMemcpy2D(out->g_vtx, 4 * static_cast<size_t>(S), d.gvx + go * d.S, 4 * static_cast<size_t>(d.S),
                      4 * static_cast<size_t>(S), static_cast<size_t>(ng), hipMemcpyDeviceToHost) == hipSuccess;
    ok &= down(out->g_vlen, d.glen + go, 4 * static_cast<size_t>(ng));
    if (out->g_area) ok &= down(out->g_area, d.garea + go, 8 * static_cast<size_t>(ng));
    if (out->g_dAdt) ok &= down(out->g_dAdt, d.gdadt + go, 8 * static_cast<size_t>(ng));
    ok &= down(out->g_orientation, d.gori + go, 8 * static_cast<size_t>(ng));
    ok &= down(out->g_SE, d.gse + go, 8 * static_cast<size_t>(ng));
    ok &= down(out->g_enabled, d.genab + go, static_cast<size_t>(ng));
    if (out->g_type) ok &= down(out->g_type, d.gtype + go, static_cast<size_t>(ng));
    if (!ok) return fail(s, SEVC_E_HIP, "device to host copy failed");
    auto loc = [](int32_t* a, size_t n, int off) {
        for (size_t i = 0; i < n; ++i) if (a[i] >= 0) a[i] -= off;
    };
    loc(out->v_bnd, 3 * static_cast<size_t>(nv), s->eoff[m]);
    loc(out->v_grn, 3 * static_cast<size_t>(nv), s->goff[m]);
    loc(out->b_vtx, 2 * static_cast<size_t>(ne), s->voff[m]);
    loc(out->g_vtx, static_cast<size_t>(ng) * S, s->voff[m]);
    return SEVC_OK;
}

int sevc_autotune(sevc_state* s, int32_t* tile_grains) {
    if (!s) return SEVC_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(s->stream));
    // candidates: the two-pass path with 16..64 grains per tile of its
    // first pass, and the one-pass path with 32 or 64 vertices per tile;
    // each timed over 3 launches of the whole velocity on the current state
    struct Cand { int path, tile; };
    std::vector<Cand> cands = {{0, 16}, {0, 24}, {0, 32}, {0, 48}, {0, 64}};
    if (vtile_ok(s)) {
        // one-pass tile sizes whose staging fits the current topology
        // (timing launches skip the two-pass fallback, so an overflowing
        // tiling would look faster than it is)
        std::vector<int> vg(3 * static_cast<size_t>(s->V)), glen(s->G);
        std::vector<uint8_t> ven(s->V);
        HIPCHK(hipMemcpy(vg.data(), s->d.vg, vg.size() * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(glen.data(), s->d.glen, glen.size() * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ven.data(), s->d.ven, ven.size(), hipMemcpyDeviceToHost));
        long long nen = 0;
        for (uint8_t e : ven) nen += e ? 1 : 0;
        for (int VT : {32, 64})
            if (100 * vtile_overflow(s, VT, vg, ven, glen) <= nen) cands.push_back({1, VT});
        cands.push_back({2, 0});
    }
    // the timed launches write velocities and grain areas: both are kept
    // and put back, so the state is left as it was
    double *keep_vel = nullptr, *keep_area = nullptr;
    const size_t nvel = 2 * static_cast<size_t>(s->V), narea = static_cast<size_t>(s->G);
    if (hipMalloc(&keep_vel, nvel * sizeof(double)) != hipSuccess ||
        hipMalloc(&keep_area, narea * sizeof(double)) != hipSuccess) {
        (void)hipFree(keep_vel);
        return fail(s, SEVC_E_NOMEM, "autotune: no room to keep velocities and areas");
    }
    HIPCHK(hipMemcpyAsync(keep_vel, s->d.vel, nvel * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(keep_area, s->d.garea, narea * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    Cand best{s->vel_path, s->vel_path == 1 ? s->vtile_size : (s->vel_path == 2 ? 0 : s->tile_grains)};
    float best_ms = 1e30f;
    int rc = SEVC_OK;
    for (const Cand& c : cands) {
        if ((rc = (c.path == 0 ? set_tiles(s, c.tile) : (c.path == 1 ? set_vtiles(s, c.tile) : SEVC_OK))) != SEVC_OK) break;
        const Dev d = s->d;
        auto once = [&]() {
            if (c.path == 1) {
                if (c.tile == 32) hipLaunchKernelGGL(k_vel_vtile<32>, dim3(d.nvtile), dim3(128), 0, s->stream, d, 1);
                else hipLaunchKernelGGL(k_vel_vtile<64>, dim3(d.nvtile), dim3(256), 0, s->stream, d, 1);
            } else if (c.path == 2) {
                (void)hipMemsetAsync(d.cnt + C_BIGR, 0, sizeof(int), s->stream);
                hipLaunchKernelGGL(k_vel_glane, dim3(d.ngb), dim3(kBlock), 0, s->stream, d, 1);
                hipLaunchKernelGGL(k_vel_bigrain, dim3(kBigGrid), dim3(64), 0, s->stream, d, 1);
                hipLaunchKernelGGL(k_vel_vertex, dim3(d.nvb), dim3(kBlock), 0, s->stream, d, 1);
            } else {
                hipLaunchKernelGGL(k_vel_grain, dim3(d.ntile), dim3(256), 0, s->stream, d, 1);
                hipLaunchKernelGGL(k_vel_vertex, dim3(d.nvb), dim3(kBlock), 0, s->stream, d, 1);
            }
        };
        once();   // warm-up
        (void)hipEventRecord(e0, s->stream);
        for (int r = 0; r < 3; ++r) once();
        (void)hipEventRecord(e1, s->stream);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = SEVC_E_HIP; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best_ms) { best_ms = ms; best = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the timing launches of the lane-per-grain pass listed long grains: the
    // next step starts from an empty list
    const bool back = hipMemsetAsync(s->d.cnt + C_BIGR, 0, sizeof(int), s->stream) == hipSuccess &&
                      hipMemcpyAsync(s->d.vel, keep_vel, nvel * sizeof(double), hipMemcpyDeviceToDevice,
                                     s->stream) == hipSuccess &&
                      hipMemcpyAsync(s->d.garea, keep_area, narea * sizeof(double), hipMemcpyDeviceToDevice,
                                     s->stream) == hipSuccess &&
                      hipStreamSynchronize(s->stream) == hipSuccess;
    (void)hipFree(keep_vel);
    (void)hipFree(keep_area);
    if (rc != SEVC_OK) return fail(s, rc, "autotune failed");
    if (!back) return fail(s, SEVC_E_HIP, "autotune: restoring velocities and areas failed");
    if ((rc = sevc_set_velocity_config(s, best.path, best.tile)) != SEVC_OK) return rc;
    if (tile_grains) *tile_grains = best.tile;
    return SEVC_OK;
}

int sevc_set_poll_limit(sevc_state* s, int32_t max_fast, int32_t max_lds) {
    if (!s || max_fast < 0 || max_fast > kPollFast || max_lds < 0 || max_lds > kPollBig) return SEVC_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->d.poll_fast = max_fast;
    s->d.poll_lds = max_lds;
    drop_graphs(s);
    return SEVC_OK;
}

int sevc_set_search_round(sevc_state* s, int32_t blocks) {
    if (!s || blocks < 0 || blocks > kNuclThreads) return SEVC_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->d.free_rows = blocks;
    drop_graphs(s);
    return SEVC_OK;
}

int sevc_velocity_config(const sevc_state* s, int32_t* path, int32_t* tile) {
    if (!s) return SEVC_E_INVALID_ARG;
    if (path) *path = s->vel_path;
    if (tile) *tile = s->vel_path == 1 ? s->vtile_size : (s->vel_path == 2 ? 0 : s->tile_grains);
    return SEVC_OK;
}

int sevc_set_velocity_config(sevc_state* s, int32_t path, int32_t tile) {
    if (!s) return SEVC_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(s->stream));
    int rc;
    if (path == 0) {
        if ((rc = set_tiles(s, tile)) != SEVC_OK) return fail(s, rc, "grain tile size outside 16..64");
    } else if (path == 1) {
        if (!vtile_ok(s)) return fail(s, SEVC_E_INVALID_ARG, "one-pass velocity needs a list stride that is a multiple of 4, at least 8");
        if ((rc = set_vtiles(s, tile)) != SEVC_OK) return fail(s, rc, "vertex tile size must be 32 or 64");
    } else if (path == 2) {
        if (!vtile_ok(s)) return fail(s, SEVC_E_INVALID_ARG, "lane-per-grain velocity needs a list stride that is a multiple of 4, at least 8");
    } else {
        return fail(s, SEVC_E_INVALID_ARG, "unknown velocity path");
    }
    s->vel_path = path;
    return SEVC_OK;
}

int sevc_get_positions(sevc_state* s, int32_t m, double* pos) {
    if (!s || !pos || m < 0 || m >= s->B) return SEVC_E_INVALID_ARG;
    HIPCHK(hipMemcpyAsync(pos, s->d.pos + 2 * static_cast<size_t>(s->voff[m]), 16 * static_cast<size_t>(s->nv[m]),
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return SEVC_OK;
}

int sevc_check_state(sevc_state* s, int32_t* bad) {
    if (!s || !bad) return SEVC_E_INVALID_ARG;
    const Dev& d = s->d;
    hipLaunchKernelGGL(k_check_clear, dim3(1), dim3(256), 0, s->stream, d.chk_bad, s->B);
    hipLaunchKernelGGL(k_check, dim3(check_grid((long long)d.V + d.E + d.G)), dim3(256), 0, s->stream,
                       check_view(d, 0));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(bad, d.chk_bad, sizeof(int32_t) * static_cast<size_t>(s->B), hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return SEVC_OK;
}

int sevc_check_mesh_device(const sevc_mesh* m, int32_t* bad) {
    if (!m || !bad) return SEVC_E_INVALID_ARG;
    const long long nv = m->n_vertices, ne = m->n_boundaries, ng = m->n_grains, S = m->max_vrt_per_grn;
    if (nv <= 0 || ne <= 0 || ng <= 0 || S <= 0 || !m->v_bnd || !m->v_grn || !m->v_enabled || !m->b_vtx ||
        !m->b_enabled || !m->g_vtx || !m->g_vlen || !m->g_enabled)
        return SEVC_E_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SEVC_E_NO_DEVICE;
    // one allocation: the arrays at 256-byte aligned offsets
    struct Part { const void* src; size_t bytes; size_t off; };
    Part parts[] = {{m->v_bnd, 12 * (size_t)nv, 0},   {m->v_grn, 12 * (size_t)nv, 0},
                    {m->v_enabled, (size_t)nv, 0},    {m->b_vtx, 8 * (size_t)ne, 0},
                    {m->b_enabled, (size_t)ne, 0},    {m->g_vtx, 4 * (size_t)(ng * S), 0},
                    {m->g_vlen, 4 * (size_t)ng, 0},   {m->g_enabled, (size_t)ng, 0},
                    {nullptr, sizeof(Params), 0},      {nullptr, sizeof(int), 0}};
    size_t total = 0;
    for (Part& p : parts) {
        p.off = total;
        total += (p.bytes + 255) / 256 * 256;
    }
    char* buf = nullptr;
    if (hipMalloc(&buf, total) != hipSuccess) return SEVC_E_NOMEM;
    Params P{};
    P.voff = 0; P.nv = (int)nv; P.eoff = 0; P.ne = (int)ne; P.goff = 0; P.ng = (int)ng;
    P.debug = 1;
    parts[8].src = &P;
    int rc = SEVC_OK;
    for (const Part& p : parts)
        if (p.src && hipMemcpy(buf + p.off, p.src, p.bytes, hipMemcpyHostToDevice) != hipSuccess) rc = SEVC_E_HIP;
    if (rc == SEVC_OK) {
        ChkView c{};
        c.vb = (const int*)(buf + parts[0].off); c.vg = (const int*)(buf + parts[1].off);
        c.ven = (const uint8_t*)(buf + parts[2].off);
        c.bvx = (const int*)(buf + parts[3].off); c.benab = (const uint8_t*)(buf + parts[4].off);
        c.gvx = (const int*)(buf + parts[5].off); c.glen = (const int*)(buf + parts[6].off);
        c.genab = (const uint8_t*)(buf + parts[7].off);
        c.S = (int)S; c.V = (int)nv; c.E = (int)ne; c.G = (int)ng; c.B = 1;
        c.prm = (const Params*)(buf + parts[8].off);
        c.gate = 0;
        c.bad = (int*)(buf + parts[9].off);
        hipLaunchKernelGGL(k_check_clear, dim3(1), dim3(256), 0, nullptr, c.bad, 1);
        hipLaunchKernelGGL(k_check, dim3(check_grid(nv + ne + ng)), dim3(256), 0, nullptr, c);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(bad, c.bad, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
            rc = SEVC_E_HIP;
    }
    (void)hipFree(buf);
    return rc;
}

extern "C++" {
namespace {
// Every device array a step reads and writes (the loop state of the members
// aside): what a checkpoint holds.  Parameters (Params) are not state.
std::vector<std::pair<void*, size_t>> state_arrays(const sevc_state* s) {
    const Dev& d = s->d;
    const size_t V = s->V, E = s->E, G = s->G, S = s->S;
    return {{d.pos, 16 * V}, {d.vel, 16 * V}, {d.vb, 12 * V}, {d.vg, 12 * V}, {d.ven, V},
            {d.venergy, 8 * V}, {d.vparts, 32 * V}, {d.mods, sizeof(ModRec) * V},
            {d.nucl_r, 8 * V}, {d.nucl_ori, 8 * V}, {d.dE, 8 * V}, {d.nfac, 8 * V}, {d.vnuc, V},
            {d.bvx, 8 * E}, {d.ben, 8 * E}, {d.blen, 8 * E}, {d.bt, 8 * E}, {d.benab, E}, {d.bcand, E},
            {d.conf, 4 * E}, {d.blab, E}, {d.l_flipped, 4 * (E + 16)}, {d.n_flipped, 4},
            {d.gvx, 4 * G * S}, {d.glen, 4 * G}, {d.garea, 8 * G}, {d.gdadt, 8 * G}, {d.gori, 8 * G},
            {d.gse, 8 * G}, {d.genab, G}, {d.gtype, G}, {d.gcand, G}, {d.gfix, 4 * G},
            {d.cnt, 4 * static_cast<size_t>(C_COUNT)}, {d.bc_v, 8 * static_cast<size_t>(d.nvb)},
            {d.bc_e, 8 * static_cast<size_t>(d.neb)}, {d.bc_g, 8 * static_cast<size_t>(d.ngb)},
            {d.ms, sizeof(MState) * static_cast<size_t>(s->B)}};
}
constexpr int kMaxCkpt = 64;
}  // namespace
}  // extern "C++"

int sevc_checkpoint(sevc_state* s, int32_t slot) {
    if (!s || slot < 0 || slot >= kMaxCkpt) return SEVC_E_INVALID_ARG;
    if (static_cast<int>(s->ckpts.size()) <= slot) s->ckpts.resize(slot + 1);
    auto arrays = state_arrays(s);
    size_t total = 0;
    for (auto& a : arrays) total += (a.second + 255) / 256 * 256;
    sevc_state::Ckpt& c = s->ckpts[slot];
    if (!c.buf) {
        HIPCHK(hipStreamSynchronize(s->stream));
        if (hipMalloc(&c.buf, total) != hipSuccess) return fail(s, SEVC_E_NOMEM, "no room for a checkpoint");
        c.bytes = total;
    }
    size_t off = 0;
    for (auto& a : arrays) {
        HIPCHK(hipMemcpyAsync(c.buf + off, a.first, a.second, hipMemcpyDeviceToDevice, s->stream));
        off += (a.second + 255) / 256 * 256;
    }
    HIPCHK(hipMemcpyAsync(s->ms_host.data(), s->d.ms, sizeof(MState) * s->ms_host.size(), hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    c.ms = s->ms_host;
    return SEVC_OK;
}

int sevc_restore(sevc_state* s, int32_t slot) {
    if (!s || slot < 0 || slot >= static_cast<int>(s->ckpts.size()) || !s->ckpts[slot].buf)
        return SEVC_E_INVALID_ARG;
    const sevc_state::Ckpt& c = s->ckpts[slot];
    size_t off = 0;
    for (auto& a : state_arrays(s)) {
        HIPCHK(hipMemcpyAsync(a.first, c.buf + off, a.second, hipMemcpyDeviceToDevice, s->stream));
        off += (a.second + 255) / 256 * 256;
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->ms_host = c.ms;
    s->ms_prev = c.ms;
    s->ben_dirty = true;
    s->win.ready = false;
    return SEVC_OK;
}

int sevc_set_member_mesh(sevc_state* s, int32_t m, const sevc_mesh* me) {
    if (!s || !me || m < 0 || m >= s->B) return SEVC_E_INVALID_ARG;
    const int nv = s->nv[m], ne = s->ne[m], ng = s->ng[m], Sm = s->Sm[m];
    if (me->n_vertices != nv || me->n_boundaries != ne || me->n_grains != ng || me->max_vrt_per_grn != Sm ||
        !me->pos || !me->v_bnd || !me->v_grn || !me->v_enabled || !me->b_vtx || !me->b_enabled || !me->g_vtx ||
        !me->g_vlen || !me->g_enabled || !me->g_SE || !me->g_orientation)
        return fail(s, SEVC_E_INVALID_ARG, "mesh sizes or arrays differ from the member's");
    double total_area = 0.0;
    if (sevc_check_mesh(me, &total_area) != 0) return fail(s, SEVC_E_TOPOLOGY, "mesh fails check_data");
    HIPCHK(hipStreamSynchronize(s->stream));
    const Dev& d = s->d;
    const int vo = s->voff[m], eo = s->eoff[m], go = s->goff[m];
    const size_t S = static_cast<size_t>(s->S);
    auto gid = [](int x, int off) { return x < 0 ? -1 : x + off; };
    std::vector<double> vel(2 * static_cast<size_t>(nv), 0.0), ben(ne), blen(ne, 0.0), bt(ne, SEVC_MAX_T_EXT),
        garea(ng, 0.0), gdadt(ng, 0.0);
    std::vector<int> vb(3 * static_cast<size_t>(nv)), vg(3 * static_cast<size_t>(nv)), bvx(2 * static_cast<size_t>(ne)),
        gvx(static_cast<size_t>(ng) * S, -1), glen(ng), conf(ne, -1);
    std::vector<uint8_t> ven(nv), benab(ne), bcand(ne, 0), genab(ng), gtype(ng, 0);
    int ngr = 0, nve = 0;
    for (int i = 0; i < nv; ++i) {
        if (me->vel) { vel[2 * i] = me->vel[2 * i]; vel[2 * i + 1] = me->vel[2 * i + 1]; }
        for (int k = 0; k < 3; ++k) {
            vb[3 * i + k] = gid(me->v_bnd[3 * i + k], eo);
            vg[3 * i + k] = gid(me->v_grn[3 * i + k], go);
        }
        ven[i] = me->v_enabled[i] ? 1 : 0;
        nve += ven[i];
    }
    for (int b = 0; b < ne; ++b) {
        bvx[2 * b] = gid(me->b_vtx[2 * b], vo);
        bvx[2 * b + 1] = gid(me->b_vtx[2 * b + 1], vo);
        benab[b] = me->b_enabled[b] ? 1 : 0;
        ben[b] = me->b_energy ? me->b_energy[b] : 0.0;
        if (me->b_arclength) blen[b] = me->b_arclength[b];
        if (me->b_t_ext) bt[b] = me->b_t_ext[b];
    }
    for (int g = 0; g < ng; ++g) {
        glen[g] = me->g_vlen[g];
        if (glen[g] > Sm || glen[g] < 0) return fail(s, SEVC_E_INVALID_ARG, "grain list longer than the stride");
        for (int k = 0; k < glen[g]; ++k) gvx[static_cast<size_t>(g) * S + k] = gid(me->g_vtx[static_cast<size_t>(g) * Sm + k], vo);
        genab[g] = me->g_enabled[g] ? 1 : 0;
        ngr += genab[g];
        if (me->g_area) garea[g] = me->g_area[g];
        if (me->g_type) gtype[g] = me->g_type[g];
    }
    std::vector<ModRec> mods(nv);
    for (ModRec& r : mods)
        for (int j = 0; j < 4; j++) { r.grn[j] = -1; r.vrt[j] = -1; r.prev[j] = -1; r.fix[j] = FIX_NONE; }
    bool ok = true;
    auto up = [&](void* dst, const void* src, size_t bytes) {
        ok = ok && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    up(d.pos + 2 * static_cast<size_t>(vo), me->pos, 16 * static_cast<size_t>(nv));
    up(d.vel + 2 * static_cast<size_t>(vo), vel.data(), 16 * static_cast<size_t>(nv));
    up(d.vb + 3 * static_cast<size_t>(vo), vb.data(), 12 * static_cast<size_t>(nv));
    up(d.vg + 3 * static_cast<size_t>(vo), vg.data(), 12 * static_cast<size_t>(nv));
    up(d.ven + vo, ven.data(), static_cast<size_t>(nv));
    up(d.mods + vo, mods.data(), sizeof(ModRec) * static_cast<size_t>(nv));
    up(d.bvx + 2 * static_cast<size_t>(eo), bvx.data(), 8 * static_cast<size_t>(ne));
    up(d.ben + eo, ben.data(), 8 * static_cast<size_t>(ne));
    up(d.blen + eo, blen.data(), 8 * static_cast<size_t>(ne));
    up(d.bt + eo, bt.data(), 8 * static_cast<size_t>(ne));
    up(d.benab + eo, benab.data(), static_cast<size_t>(ne));
    up(d.bcand + eo, bcand.data(), static_cast<size_t>(ne));
    up(d.conf + eo, conf.data(), 4 * static_cast<size_t>(ne));
    up(d.gvx + static_cast<size_t>(go) * S, gvx.data(), 4 * gvx.size());
    up(d.glen + go, glen.data(), 4 * static_cast<size_t>(ng));
    up(d.garea + go, garea.data(), 8 * static_cast<size_t>(ng));
    up(d.gori + go, me->g_orientation, 8 * static_cast<size_t>(ng));
    up(d.gse + go, me->g_SE, 8 * static_cast<size_t>(ng));
    up(d.genab + go, genab.data(), static_cast<size_t>(ng));
    up(d.gtype + go, gtype.data(), static_cast<size_t>(ng));
    ok = ok && hipMemset(d.gfix + go, 0, 4 * static_cast<size_t>(ng)) == hipSuccess &&
         hipMemset(d.gcand + go, 0, static_cast<size_t>(ng)) == hipSuccess &&
         hipMemset(d.vnuc + vo, 0, static_cast<size_t>(nv)) == hipSuccess;
    if (!ok) return fail(s, SEVC_E_HIP, "upload of the member's mesh failed");
    // loop state: counts of the new mesh; the step counter and time go on
    MState M;
    HIPCHK(hipMemcpy(&M, d.ms + m, sizeof(MState), hipMemcpyDeviceToHost));
    M.curr_grains = ngr;
    M.n_venabled = nve;
    HIPCHK(hipMemcpy(d.ms + m, &M, sizeof(MState), hipMemcpyHostToDevice));
    s->ms_host[static_cast<size_t>(m)].curr_grains = ngr;
    s->ms_host[static_cast<size_t>(m)].n_venabled = nve;
    hipLaunchKernelGGL(k_count_all, dim3(count_grid(d)), dim3(64 * kCountWaves), 0, s->stream, d);
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipGetLastError());
    s->ben_dirty = true;          // energies of the restored topology at the next step
    s->win.ready = false;
    drop_graphs(s);
    return SEVC_OK;
}

int sevc_set_member_params(sevc_state* s, int32_t m, const double* pos, const double* g_SE, double gb_scaling,
                           double mobility) {
    if (!s || m < 0 || m >= s->B) return SEVC_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(s->stream));
    const Dev& d = s->d;
    if (pos) HIPCHK(hipMemcpy(d.pos + 2 * static_cast<size_t>(s->voff[m]), pos, 16 * static_cast<size_t>(s->nv[m]),
                              hipMemcpyHostToDevice));
    if (g_SE) HIPCHK(hipMemcpy(d.gse + s->goff[m], g_SE, 8 * static_cast<size_t>(s->ng[m]), hipMemcpyHostToDevice));
    Params P;
    HIPCHK(hipMemcpy(&P, d.prm + m, sizeof(Params), hipMemcpyDeviceToHost));
    if (gb_scaling > 0) { P.gb_scaling = gb_scaling; s->opts[m].GB_scaling = gb_scaling; }
    if (mobility > 0) { P.mobility = mobility; s->opts[m].mobility = mobility; }
    HIPCHK(hipMemcpy(const_cast<Params*>(d.prm) + m, &P, sizeof(Params), hipMemcpyHostToDevice));
    if (gb_scaling > 0) s->ben_dirty = true;      // energies at the new scale on the next step
    drop_graphs(s);
    s->win.ready = false;
    return SEVC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// topology-fixed windows and their adjoint
// ---------------------------------------------------------------------

namespace {

// (Re)allocate a window buffer of at least n doubles.
int win_buffer(sevc_state* s, double** p, size_t* have, size_t n) {
    if (*p && *have >= n) return SEVC_OK;
    if (*p) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(*p);
        s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), static_cast<void*>(*p)));
        *p = nullptr;
    }
    *have = 0;
    int rc = dalloc(s, p, n);
    if (rc == SEVC_OK) *have = n;
    return rc;
}

// Reverse-pass arrays whose size depends only on the state.
int win_fixed_buffers(sevc_state* s) {
    auto& w = s->win;
    if (w.lamA) return SEVC_OK;
    const size_t V = s->V, n2 = 2 * V;
    int rc;
    if ((rc = dalloc(s, &w.lamA, n2)) || (rc = dalloc(s, &w.lamB, n2)) || (rc = dalloc(s, &w.acc_mu, V)) ||
        (rc = dalloc(s, &w.acc_s, V)) || (rc = dalloc(s, &w.acc_se, 3 * V)) ||
        (rc = dalloc(s, &w.ebase, static_cast<size_t>(s->E))) ||
        (rc = dalloc(s, &w.part, 4 * static_cast<size_t>(s->d.nvb))) ||
        (rc = dalloc(s, &w.out, 2 * static_cast<size_t>(s->B))) ||
        (rc = dalloc(s, &w.dse, static_cast<size_t>(s->G))))
        return rc;
    return SEVC_OK;
}

// Steps x_{c K} .. of segment c from the segment buffer's slot 0; areas of
// each step into segA.  Forward of the window (state outputs on) or its
// recomputation in the reverse pass (off).
void win_run_segment(sevc_state* s, int len, bool forward) {
    const Dev& d = s->d;
    auto& w = s->win;
    const size_t n2 = 2 * static_cast<size_t>(s->V);
    for (int i = 0; i < len; ++i) {
        const double* X = w.seg + i * n2;
        double* A = w.segA + static_cast<size_t>(i) * s->G;
        launch(s, "win_area", k_win_area, d.ngb, kBlock, d, X, A, forward ? 1 : 0);
        launch(s, "win_step", k_win_step, d.nvb, kBlock, d, X, static_cast<const double*>(A),
               w.seg + (i + 1) * n2, forward ? 1 : 0);
    }
}

}  // namespace

extern "C" {

int sevc_window_forward(sevc_state* s, int32_t steps, int32_t ckpt_every) {
    if (!s || steps < 0) return SEVC_E_INVALID_ARG;
    auto& w = s->win;
    const Dev& d = s->d;
    const int K = (ckpt_every <= 0 || ckpt_every > steps) ? std::max(steps, 1) : ckpt_every;
    const int nck = (steps + K - 1) / K;
    const size_t V = s->V, n2 = 2 * V;
    int rc;
    if ((rc = win_buffer(s, &w.ck, &w.ck_n, std::max<size_t>(1, nck) * n2)) != SEVC_OK ||
        (rc = win_buffer(s, &w.seg, &w.seg_n, (static_cast<size_t>(K) + 1) * n2)) != SEVC_OK ||
        (rc = win_buffer(s, &w.segA, &w.segA_n, static_cast<size_t>(K) * s->G)) != SEVC_OK)
        return rc;
    if ((rc = win_fixed_buffers(s)) != SEVC_OK) return rc;
    w.ready = false;
    w.seeded = false;
    launch(s, "win_prepare", k_win_prepare, d.neb, kBlock, d, w.ebase);
    HIPCHK(hipMemcpyAsync(w.seg, d.pos, n2 * 8, hipMemcpyDeviceToDevice, s->stream));
    int last = 0;
    for (int c = 0; c < nck; ++c) {
        const int len = std::min(K, steps - c * K);
        if (c > 0) HIPCHK(hipMemcpyAsync(w.seg, w.seg + last * n2, n2 * 8, hipMemcpyDeviceToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(w.ck + c * n2, w.seg, n2 * 8, hipMemcpyDeviceToDevice, s->stream));
        win_run_segment(s, len, true);
        last = len;
    }
    if (steps > 0) HIPCHK(hipMemcpyAsync(d.pos, w.seg + last * n2, n2 * 8, hipMemcpyDeviceToDevice, s->stream));
    launch(s, "win_end", k_win_end, (s->B + 63) / 64, 64, d, static_cast<int>(steps));
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(s, SEVC_E_HIP, std::string("launch: ") + hipGetErrorString(le));
    if ((rc = finish_call(s)) != SEVC_OK) return rc;
    w.steps = steps;
    w.K = K;
    w.nck = nck;
    w.holds = nck - 1;
    w.ready = true;
    return SEVC_OK;
}

int sevc_adjoint_window(sevc_state* s, const double* const* dL_dpos, sevc_grads* grads) {
    if (!s || !grads) return SEVC_E_INVALID_ARG;
    auto& w = s->win;
    if (!w.ready) return fail(s, SEVC_E_STATE, "no window: call sevc_window_forward first (after any step)");
    const Dev& d = s->d;
    const size_t n2 = 2 * static_cast<size_t>(s->V);
    if (dL_dpos) {
        for (int m = 0; m < s->B; ++m) {
            if (!dL_dpos[m]) return fail(s, SEVC_E_INVALID_ARG, "dL_dpos missing for a member");
            HIPCHK(hipMemcpyAsync(w.lamA + 2 * static_cast<size_t>(s->voff[m]), dL_dpos[m],
                                  16 * static_cast<size_t>(s->nv[m]), hipMemcpyHostToDevice, s->stream));
        }
    } else if (!w.seeded) {
        return fail(s, SEVC_E_STATE, "dL_dpos is NULL and no sevc_area_histogram seed is held");
    }
    w.seeded = false;
    launch(s, "win_init", k_win_init, d.nvb, kBlock, d, w.lamA, w.acc_mu, w.acc_s, w.acc_se);
    double *cur = w.lamA, *nxt = w.lamB;
    for (int c = w.nck - 1; c >= 0; --c) {
        const int len = std::min(w.K, w.steps - c * w.K);
        if (w.holds != c) {
            HIPCHK(hipMemcpyAsync(w.seg, w.ck + c * n2, n2 * 8, hipMemcpyDeviceToDevice, s->stream));
            win_run_segment(s, len, false);
            w.holds = c;
        }
        for (int i = len - 1; i >= 0; --i) {
            launch(s, "win_adjoint", k_win_adjoint, d.nvb, kBlock, d, static_cast<const double*>(w.seg + i * n2),
                   static_cast<const double*>(w.segA + static_cast<size_t>(i) * s->G),
                   static_cast<const double*>(cur), nxt, w.acc_mu, w.acc_s, w.acc_se,
                   static_cast<const double*>(w.ebase));
            std::swap(cur, nxt);
        }
    }
    launch(s, "win_sum_l1", k_win_sum_l1, (d.nvb + 255) / 256, 256, d, static_cast<const double*>(w.acc_mu),
           static_cast<const double*>(w.acc_s), w.part);
    if (s->B > 0)
        hipLaunchKernelGGL(k_win_sum_up, dim3(s->B, 2), dim3(256), 0, s->stream, d, w.part, w.out);
    launch(s, "win_dse", k_win_dse, d.ngb, kBlock, d, static_cast<const double*>(w.acc_se), w.dse);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(s, SEVC_E_HIP, std::string("launch: ") + hipGetErrorString(le));
    std::vector<double> out(2 * static_cast<size_t>(s->B));
    HIPCHK(hipMemcpyAsync(out.data(), w.out, out.size() * 8, hipMemcpyDeviceToHost, s->stream));
    for (int m = 0; m < s->B; ++m) {
        if (grads[m].d_pos)
            HIPCHK(hipMemcpyAsync(grads[m].d_pos, cur + 2 * static_cast<size_t>(s->voff[m]),
                                  16 * static_cast<size_t>(s->nv[m]), hipMemcpyDeviceToHost, s->stream));
        if (grads[m].d_SE)
            HIPCHK(hipMemcpyAsync(grads[m].d_SE, w.dse + s->goff[m], 8 * static_cast<size_t>(s->ng[m]),
                                  hipMemcpyDeviceToHost, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->timing) collect_timers(s);
    for (int m = 0; m < s->B; ++m) {
        grads[m].d_gb_scaling = out[2 * static_cast<size_t>(m)];
        grads[m].d_mobility = out[2 * static_cast<size_t>(m) + 1];
    }
    return SEVC_OK;
}

int sevc_area_histogram(sevc_state* s, int32_t n_bins, const double* ranges, const double* targets, double* hist,
                        double* loss, double* const* dL_dpos, int32_t seed_adjoint) {
    if (!s || !ranges || !targets || n_bins < 1 || n_bins > kMaxHistBins) return SEVC_E_INVALID_ARG;
    for (int m = 0; m < s->B; ++m)
        if (!(ranges[3 * m + 2] > 0) || !(ranges[3 * m + 1] > ranges[3 * m]))
            return fail(s, SEVC_E_INVALID_ARG, "histogram range needs a_max > a_min and width > 0");
    auto& w = s->win;
    const Dev& d = s->d;
    int rc;
    if ((rc = win_fixed_buffers(s)) != SEVC_OK) return rc;
    const size_t B = s->B, G = s->G, nb = static_cast<size_t>(n_bins);
    if (!w.hA) {
        if ((rc = dalloc(s, &w.hA, G)) || (rc = dalloc(s, &w.hwg, G)) ||
            (rc = dalloc(s, &w.hpart, static_cast<size_t>(d.ngb) * (kMaxHistBins + 1))) ||
            (rc = dalloc(s, &w.hrange, 3 * B)) || (rc = dalloc(s, &w.htarget, B * kMaxHistBins)) ||
            (rc = dalloc(s, &w.hhist, B * kMaxHistBins)) || (rc = dalloc(s, &w.hres, B * kMaxHistBins)) ||
            (rc = dalloc(s, &w.hloss, B)) || (rc = dalloc(s, &w.hcount, B)))
            return rc;
    }
    HIPCHK(hipMemcpyAsync(w.hrange, ranges, 3 * B * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(w.htarget, targets, B * nb * 8, hipMemcpyHostToDevice, s->stream));
    launch(s, "hist_l1", k_hist_l1, d.ngb, kBlock, d, n_bins, static_cast<const double*>(w.hrange), w.hA, w.hpart);
    launch(s, "hist_member", k_hist_member, s->B, kMaxHistBins + 64, d, n_bins,
           static_cast<const double*>(w.hpart), static_cast<const double*>(w.htarget), w.hhist, w.hres, w.hloss,
           w.hcount);
    const bool need_grad = seed_adjoint || dL_dpos;
    if (need_grad) {
        launch(s, "hist_weight", k_hist_weight, d.ngb, kBlock, d, n_bins, static_cast<const double*>(w.hrange),
               static_cast<const double*>(w.hA), static_cast<const double*>(w.hres),
               static_cast<const double*>(w.hcount), w.hwg);
        launch(s, "hist_seed", k_hist_seed, d.nvb, kBlock, d, static_cast<const double*>(w.hwg), w.lamA);
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(s, SEVC_E_HIP, std::string("launch: ") + hipGetErrorString(le));
    std::vector<double> hh(B * nb);
    // the kernels store member m's bins at m * n_bins
    if (hist) HIPCHK(hipMemcpyAsync(hh.data(), w.hhist, B * nb * 8, hipMemcpyDeviceToHost, s->stream));
    if (loss) HIPCHK(hipMemcpyAsync(loss, w.hloss, B * 8, hipMemcpyDeviceToHost, s->stream));
    if (dL_dpos)
        for (int m = 0; m < s->B; ++m)
            if (dL_dpos[m])
                HIPCHK(hipMemcpyAsync(dL_dpos[m], w.lamA + 2 * static_cast<size_t>(s->voff[m]),
                                      16 * static_cast<size_t>(s->nv[m]), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (hist) std::memcpy(hist, hh.data(), hh.size() * 8);
    if (s->timing) collect_timers(s);
    w.seeded = need_grad && seed_adjoint;
    return SEVC_OK;
}

}  // extern "C"

namespace sevc {
__global__ void k_selftest(int op, int n, const double* x, const double* y, double* out);
}

extern "C" int sevc_selftest_math(int32_t op, int32_t n, const double* x, const double* y, double* out) {
    if (n <= 0 || !x || !y || !out) return SEVC_E_INVALID_ARG;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    const size_t bytes = static_cast<size_t>(n) * sizeof(double);
    if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dy, bytes) != hipSuccess ||
        hipMalloc(&dout, bytes) != hipSuccess) {
        (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dout);
        return SEVC_E_NOMEM;
    }
    int rc = SEVC_OK;
    if (hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = SEVC_E_HIP;
    if (rc == SEVC_OK) {
        hipLaunchKernelGGL(sevc::k_selftest, dim3((n + 255) / 256), dim3(256), 0, 0, op, n, dx, dy, dout);
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = SEVC_E_HIP;
    }
    (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dout);
    return rc;
}
