// sevc_dev.h -- device helpers shared by the time step and the adjoint
// window kernels (entity access, set operations on the trivalent mesh,
// areas and the analytic force).  Include after sevc_state.h/sevc_geom.h.
#pragma once
#include <hip/hip_runtime.h>
#include "sevc_state.h"
#include "sevc_geom.h"
#include "../../include/sevc.h"
#include "../../include/sevc_math.h"

namespace sevc {

enum { T_COMMON = 0, T_NUCLEATED = 1 };                                  // geometry.h:19
constexpr unsigned kStreamMcSite = 1u, kStreamMcPick = 2u, kStreamRate = 3u;
constexpr int kMaxStride = 128;   // largest grain-list stride supported

// ---------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------

// Map the hardware workgroup id so that each of the 8 XCDs (workgroups are
// dealt round-robin over XCDs) processes one contiguous slice of the table:
// neighbouring slots then share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// An event of member m met an inconsistent state or a full grain list: the
// member records the first such code and stops (its later event entries of
// the step are skipped, as after a failed check); the other members of a
// batch go on.
__device__ __forceinline__ void set_error(const Dev& d, int m, int code) {
    MState& M = d.ms[m];
    if (atomicCAS(&M.status, 0, code) == 0) {
        M.active = 0;
        M.done = 1;
    }
}

__device__ __forceinline__ int other_end(const Dev& d, int b, int v) {
    return d.bvx[2 * b] == v ? d.bvx[2 * b + 1] : d.bvx[2 * b];
}

// geometry.h:837-848
__device__ __forceinline__ int grains_intersect(const Dev& d, int a, int b, int out[2]) {
    int k = 0;
    out[0] = out[1] = -1;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (d.vg[3 * a + i] == d.vg[3 * b + j]) {
                out[k++] = d.vg[3 * a + i];
                if (k == 2) return 2;
            }
    return k;
}

// geometry.h:886-898 (with grains_union, :863-875)
__device__ __forceinline__ void grains_symdiff(const Dev& d, int a, int b, int out[2]) {
    int gic[2], gun[4];
    out[0] = out[1] = -1;
    grains_intersect(d, a, b, gic);
    gun[3] = -1;
    for (int i = 0; i < 3; i++) gun[i] = d.vg[3 * a + i];
    for (int i = 0; i < 3; i++) {
        int g = d.vg[3 * b + i];
        if (g != gun[0] && g != gun[1] && g != gun[2]) { gun[3] = g; break; }
    }
    int k = 0;
    for (int i = 0; i < 4; i++)
        if (gun[i] != gic[0] && gun[i] != gic[1]) {
            out[k++] = gun[i];
            if (k == 2) return;
        }
}

__device__ __forceinline__ int* glist(const Dev& d, int g) { return d.gvx + (size_t)g * d.S; }

// Member owning a vertex / grain / edge slot (members own consecutive slot
// ranges; a batch keeps the owner of every slot, one load).
__device__ __forceinline__ int member_of_vertex(const Dev& d, int v) { return d.B == 1 ? 0 : d.vmem[v]; }
__device__ __forceinline__ int member_of_grain(const Dev& d, int g) { return d.B == 1 ? 0 : d.gmem[g]; }
__device__ __forceinline__ int member_of_edge(const Dev& d, int e) { return d.B == 1 ? 0 : d.emem[e]; }

// geometry.h:703-710
// The list is read eight entries at a time (independent loads), then
// searched in order: the first index holding v, or -1.
__device__ __forceinline__ int grain_contains_vertex(const Dev& d, int g, int v) {
    const int* L = glist(d, g);
    const int n = d.glen[g];
    for (int i0 = 0; i0 < n; i0 += 8) {
        int x[8];
#pragma unroll
        for (int q = 0; q < 8; q++) x[q] = i0 + q < n ? L[i0 + q] : -2;
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (x[q] == v) return i0 + q;
    }
    return -1;
}

// geometry.h:724-734, reading the list eight entries (and the next one) at a
// time: the first i (in order) with (L[i], L[(i+1) % n]) = (a, b) gives 1,
// = (b, a) gives 2; -1 if none.
__device__ __forceinline__ int grain_contains_consecutive(const Dev& d, int g, int a, int b) {
    const int* L = glist(d, g);
    const int n = d.glen[g];
    const int l0 = n > 0 ? L[0] : -2;
    for (int i0 = 0; i0 < n; i0 += 8) {
        int x[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int i = i0 + q;
            x[q] = i < n ? L[i] : (i == n ? l0 : -2);
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (i0 + q >= n) break;
            if (x[q] == a && x[q + 1] == b) return 1;
            if (x[q] == b && x[q + 1] == a) return 2;
        }
    }
    return -1;
}

// geometry.h:548-550, pow(c,3) as c*c*c
__device__ __forceinline__ double gb_energy_func(double d_alpha, double eps) {
    double c = sevc_cos(4 * d_alpha);
    return (1 + 0.5 * eps * (1 - c * c * c));
}

// geometry.h:559-564
__device__ __forceinline__ double boundary_compute_energy(const Dev& d, int b, double eps) {
    int gic[2];
    if (grains_intersect(d, d.bvx[2 * b], d.bvx[2 * b + 1], gic) < 2) return 1.0;
    return gb_energy_func(d.gori[gic[0]] - d.gori[gic[1]], eps);
}

// geometry.h:513-517 plus the flip-threshold extension
__device__ __forceinline__ bool is_candidate(double te, double len, double dt, double thr) {
    if (te >= 0 && te <= dt) return true;
    // a zero-length edge (just flipped, t_ext = no-motion value) is not shrinking
    if (thr > 0 && te >= 0 && len > 0 && len < thr) return true;
    return false;
}

__device__ __forceinline__ double member_dt(const Params& P, const MState& M) {
    if (!P.adaptive) return P.dt_max;
    unsigned long long bits = M.tau_bits;
    if (bits == 0x7FF0000000000000ull) return P.dt_max;
    double tau = __longlong_as_double((long long)bits);
    double dt = P.dt_safety * tau;
    if (dt < P.dt_min) dt = P.dt_min;
    if (dt > P.dt_max) dt = P.dt_max;
    return dt;
}

// compute_area of grain g (geometry.h:757-765 after :908-931) with the
// vertex positions X (2 doubles per vertex slot).
__device__ __forceinline__ double grain_area_at(const Dev& d, const double* X, int g, double D) {
    int n = d.glen[g];
    if (n <= 0) return 0.0;
    const int* L = glist(d, g);
    AreaAcc acc;
    int w = L[0];
    acc.start(X[2 * w], X[2 * w + 1]);
    for (int k = 1; k < n; k++) {
        w = L[k];
        acc.push(X[2 * w], X[2 * w + 1], D);
    }
    return acc.finish();
}

__device__ __forceinline__ double grain_area(const Dev& d, int g, double D) { return grain_area_at(d, d.pos, g, D); }

// geometry.h:369-376, as three 16-byte stores (grn, vrt, prev) and one
// 4-byte store (the four fix bytes) into the vertex's 64-byte record
__device__ __forceinline__ void reset_mods(const Dev& d, int v) {
    static_assert(offsetof(ModRec, grn) == 0 && offsetof(ModRec, vrt) == 16 && offsetof(ModRec, prev) == 32 &&
                      offsetof(ModRec, fix) == 48,
                  "record layout");
    int4* r = reinterpret_cast<int4*>(&d.mods[v]);
    r[0] = make_int4(-1, -1, -1, -1);
    r[1] = make_int4(-1, -1, -1, -1);
    r[2] = make_int4(-1, -1, -1, -1);
    *reinterpret_cast<unsigned*>(d.mods[v].fix) = 0x01010101u * (unsigned)FIX_NONE;
}

__device__ __forceinline__ void set_mod(const Dev& d, int v, int slot, int fix, int grn, int vrt) {
    d.mods[v].fix[slot] = (uint8_t)fix;
    d.mods[v].grn[slot] = grn;
    d.mods[v].vrt[slot] = vrt;
}

// A slot of a list whose length is *counter, for each active lane with
// `want` (-1 for the others), with one atomic per wave: lanes that append
// at the same point share it (the slots are consecutive in lane order).  The
// lists' order is unspecified, as with an atomic per lane; one atomic per
// wave keeps a step with thousands of events from queueing them on one
// address.
__device__ __forceinline__ int wave_slot(int* counter, bool want) {
    const unsigned long long m = __ballot(want);
    if (m == 0) return -1;
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, (int)__popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);
    return want ? base + (int)__popcll(m & ((1ull << lane) - 1ull)) : -1;
}

// Adds, for each active lane with `want`, one to the member counter
// `field` of member `mem`: one atomic per distinct member among the wave's
// lanes.
__device__ __forceinline__ void wave_count(const Dev& d, bool want, int mem, long long MState::*field) {
    unsigned long long m = __ballot(want);
    const int lane = (int)(threadIdx.x & 63);
    while (m) {
        const int leader = __ffsll((long long)m) - 1;
        const int k = __builtin_amdgcn_readlane(mem, leader);
        const unsigned long long same = __ballot(want && mem == k);
        if (lane == leader) atomicAdd(reinterpret_cast<unsigned long long*>(&(d.ms[k].*field)), (unsigned long long)__popcll(same));
        m &= ~same;
    }
}

// The per-table-block (enabled, free) slot counts follow every event that
// enables or disables a slot (k_count_all recounts them only after a host
// upload): one 64-bit atomic adds (+1, -1) or (-1, +1) to the pair.
__device__ __forceinline__ void slot_count_add(int2* bc, int blk, bool enable) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&bc[blk]),
              enable ? 0xFFFFFFFF00000001ull : 0x00000000FFFFFFFFull);
}

// Table block of a slot: members own consecutive runs of kBlock slots from
// their offsets.
__device__ __forceinline__ int vertex_block(const Dev& d, int v) {
    if (d.B == 1) return v / kBlock;
    const int m = d.vmem[v];
    return d.vblk_first[m] + (v - d.prm[m].voff) / kBlock;
}
__device__ __forceinline__ int edge_block(const Dev& d, int e) {
    if (d.B == 1) return e / kBlock;
    const int m = d.emem[e];
    return d.eblk_first[m] + (e - d.prm[m].eoff) / kBlock;
}
__device__ __forceinline__ int grain_block(const Dev& d, int g) {
    if (d.B == 1) return g / kBlock;
    const int m = d.gmem[g];
    return d.gblk_first[m] + (g - d.prm[m].goff) / kBlock;
}

// Set an enable flag, counting a change of state.  No other thread writes
// the same flag in the same launch (events that touch a slot are serial or
// disjoint), so the read needs no atomic.
__device__ __forceinline__ void set_vertex_enabled(const Dev& d, int v, bool on) {
    if ((d.ven[v] != 0) != on) {
        d.ven[v] = on ? 1 : 0;
        slot_count_add(d.bc_v, vertex_block(d, v), on);
    }
}
__device__ __forceinline__ void set_edge_enabled(const Dev& d, int e, bool on) {
    if ((d.benab[e] != 0) != on) {
        d.benab[e] = on ? 1 : 0;
        slot_count_add(d.bc_e, edge_block(d, e), on);
    }
}
__device__ __forceinline__ void set_grain_enabled(const Dev& d, int g, bool on) {
    if ((d.genab[g] != 0) != on) {
        d.genab[g] = on ? 1 : 0;
        slot_count_add(d.bc_g, grain_block(d, g), on);
    }
}

// Flag grain g for the fix stage `stage` (grain.fix = true) and list it once.
__device__ __forceinline__ void flag_grain(const Dev& d, int g, int stage) {
    const bool first = atomicExch(&d.gfix[g], 1) == 0;
    const int i = wave_slot(&d.cnt[C_FIX0 + stage], first);
    if (first) d.l_fix[stage][i] = g;
}

// Inclusive prefix sum of v over the 64 lanes of a wave (every lane active),
// by DPP lane moves: Hillis-Steele steps within each row of 16 lanes (row
// shifts by 1, 2, 4, 8; lanes shifted in from outside the row add 0), then
// the total of row 0 / row 2 (its lane 15, broadcast) into rows 1 / 3, and
// the total of rows 0-1 (lane 31, broadcast) into rows 2 and 3.  Rows that
// a broadcast does not write add 0.
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);    // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);    // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);    // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);    // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31 into rows 2, 3
    return v;
}

// Sum of v over the 64 lanes of a wave (every lane active), in every lane.
__device__ __forceinline__ int wave_sum(int v) {
    return __builtin_amdgcn_readlane(wave_inclusive_scan(v), 63);
}

// Minimum of v over the 64 lanes of a wave (every lane active), in every
// lane: the row-shift and broadcast pattern of wave_inclusive_scan with min
// for sum, on both 32-bit halves; lanes with no source keep the identity
// (all ones).
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#define SEVC_MIN_STEP(ctrl, rows)                                                                        \
    {                                                                                                    \
        const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(-1, (int)(unsigned)v, ctrl, rows, 0xf, false);  \
        const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(-1, (int)(unsigned)(v >> 32), ctrl, rows, 0xf, false); \
        const unsigned long long s = ((unsigned long long)hi << 32) | lo;                                \
        v = s < v ? s : v;                                                                               \
    }
    SEVC_MIN_STEP(0x111, 0xf)   // row_shr:1
    SEVC_MIN_STEP(0x112, 0xf)   // row_shr:2
    SEVC_MIN_STEP(0x114, 0xf)   // row_shr:4
    SEVC_MIN_STEP(0x118, 0xf)   // row_shr:8
    SEVC_MIN_STEP(0x142, 0xa)   // row_bcast:15 into rows 1, 3
    SEVC_MIN_STEP(0x143, 0xc)   // row_bcast:31 into rows 2, 3
#undef SEVC_MIN_STEP
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// Block lb of one of the three entity tables.  A single member's table is
// runs of kBlock consecutive slots from 0, so its bounds are computed: the
// block's own loads then need not wait for a table load.
__device__ __forceinline__ Blk table_blk(const Dev& d, const Blk* tbl, int lb) {
    if (d.B == 1 && (tbl == d.vblk || tbl == d.eblk || tbl == d.gblk)) {
        const int total = tbl == d.vblk ? d.V : (tbl == d.eblk ? d.E : d.G);
        const int b = lb * kBlock;
        return Blk{0, b, min(b + kBlock, total)};
    }
    return tbl[lb];
}

#define TABLE_BLOCK(tbl, nblk)                                   \
    int lb_ = xcd_remap((int)blockIdx.x, (int)gridDim.x);        \
    if (lb_ >= (nblk)) return;                                   \
    const Blk bd_ = table_blk(d, (tbl), lb_);                    \
    const Params& P = d.prm[bd_.m];                              \
    MState& M = d.ms[bd_.m];                                     \
    (void)P; (void)M;                                            \
    const int i = bd_.b + (int)threadIdx.x;

// Analytic -grad E at vertex v (extension; the adjoint differentiates
// exactly this): edges contribute -gamma_b wrap(p_v - p_nb)/L_b, grains
// -SE_g sign(A_g) dA_g/dp_v with dA/dx_k = (dy_prev + dy_next)/2 and
// dA/dy_k = -(dx_prev + dx_next)/2, d_prev = wrap(p_k - p_{k-1}),
// d_next = wrap(p_{k+1} - p_k).  Positions X, grain areas A.
__device__ __forceinline__ void analytic_force_at(const Dev& d, const double* X, const double* A, const Params& P,
                                                  int v, double* fx, double* fy) {
    const double D = P.D;
    double x = X[2 * v], y = X[2 * v + 1];
    double ex = 0.0, ey = 0.0, gx = 0.0, gy = 0.0;
    for (int j = 0; j < 3; j++) {
        int b = d.vb[3 * v + j];
        int nb = other_end(d, b, v);
        double dx = wrapd(x, X[2 * nb], D), dy = wrapd(y, X[2 * nb + 1], D);
        double L = mag(dx, dy);
        if (L > 0) {
            ex += d.ben[b] * (dx / L);
            ey += d.ben[b] * (dy / L);
        }
    }
    for (int j = 0; j < 3; j++) {
        int g = d.vg[3 * v + j];
        int n = d.glen[g];
        int k = grain_contains_vertex(d, g, v);
        if (k < 0 || n < 3) continue;
        const int* L = glist(d, g);
        int wp = L[(k + n - 1) % n], wn = L[(k + 1) % n];
        double dpx = wrapd(x, X[2 * wp], D), dpy = wrapd(y, X[2 * wp + 1], D);
        double dnx = wrapd(X[2 * wn], x, D), dny = wrapd(X[2 * wn + 1], y, D);
        double sg = A[g] < 0 ? -1.0 : 1.0;
        double c = d.gse[g] * sg * 0.5;
        gx += c * (dpy + dny);
        gy += c * (-(dpx + dnx));
    }
    *fx = -(ex + gx);
    *fy = -(ey + gy);
}

}  // namespace sevc
