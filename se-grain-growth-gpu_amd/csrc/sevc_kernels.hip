// sevc_kernels.hip -- the vertex-model time step on CDNA4 (gfx950), FP64.
//
// One launch sequence per step covers every member of a batch.  All control
// decisions (who is active, dt, whether an event list is empty) are taken on
// the device, so a step never waits for the host and can be captured in a
// HIP graph.  Kernel order follows main.cu:124-389.
//
// Parity rules (the oracle follows the same):
//  * events that touch disjoint vertex sets commute, so independent T1 and
//    T2 events run in parallel; events whose vertices collide run afterwards
//    on one lane in entity-index order, which is the serial order of the
//    reference's <<<1,1>>> kernels (main.cu:234,259);
//  * deferred grain-list edits (topological.h:1209-1227) touch only the
//    edited grain's own list, so they run one lane per flagged grain.
#include <hip/hip_runtime.h>

#include <climits>

#include "sevc_state.h"
#include "sevc_geom.h"
#include "../../include/sevc.h"
#include "../../include/sevc_math.h"
#include "sevc_dev.h"

namespace sevc {

// ---------------------------------------------------------------------
// step begin / end
// ---------------------------------------------------------------------

// Whether a member runs the coming step: the loop condition of main.cu:120
// (time rule) and the grain rule, from the loop state before the step.
__device__ __forceinline__ bool member_runs(const Params& P, const MState& M) {
    if (M.done) return false;
    const bool tok = (P.tmax <= 0 && P.nmax <= 0) || (M.t <= P.tmax && P.tmax > 0) || (M.n <= P.nmax && P.nmax > 0);
    const bool gok = (P.min_grains <= 0) || (M.curr_grains >= P.min_grains);
    return tok && gok;
}

// Whether a rate-mode member nucleates in the coming step (every gap-th).
__device__ __forceinline__ bool rate_nucleates(const Params& P, const MState& M) {
    if (!(P.events && P.do_nucl) || P.nucl_mode != SEVC_NUCL_RATE) return false;
    const int gap = P.gap > 0 ? P.gap : 1;
    return (M.n % gap) == 0;
}

// Loop condition of main.cu:120 (time rule), grain rule, nucleation gate
// main.cu:137-141, for every member, by one workgroup.  The per-step
// counters were cleared at the end of the previous step (k_member_reduce)
// or at set-up.
__device__ __forceinline__ void begin_members(const Dev& d, int rescale_flipped) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int m = tid; m < d.B; m += nt) {
        const Params& P = d.prm[m];
        MState& M = d.ms[m];
        M.tau_bits = 0x7FF0000000000000ull;
        M.nucl_now = 0;
        if (M.done) { M.active = 0; continue; }
        if (!member_runs(P, M)) { M.active = 0; M.done = 1; continue; }
        M.active = 1;
        M.vsteps += M.n_venabled;
        if (P.events && P.do_nucl) {
            int gap = P.gap > 0 ? P.gap : 1;
            if (P.nucl_mode == SEVC_NUCL_MC) {
                if (M.curr_grains <= P.gate_grains) M.allow_nucl = 1;   // main.cu:137-139
                M.nucl_now = ((M.n % gap) == 0 && M.allow_nucl) ? 1 : 0;
            } else {
                M.nucl_now = rate_nucleates(P, M) ? 1 : 0;
            }
        }
    }
    // flip (topological.h:286) leaves the flipped edge's energy without the
    // GB_scaling factor until compute_boundary_energies (calculus.h:104-114)
    // runs at the next step's start.  Steps that skip that pass (GB_eps = 0,
    // where every other energy is the scale itself) recompute only the
    // edges flipped since, for the members active in this step (steps that
    // run k_edge_len_energy recompute every energy there).
    __syncthreads();
    const int nf = rescale_flipped ? min(*d.n_flipped, d.E) : 0;
    for (int k = tid; k < nf; k += nt) {
        const int b = d.l_flipped[k];
        const int m = member_of_edge(d, b);
        const Params& P = d.prm[m];
        if (!d.ms[m].active || !d.benab[b]) continue;
        const double e = (P.gb_eps == 0.0) ? 1.0 : boundary_compute_energy(d, b, P.gb_eps);
        d.ben[b] = e * P.gb_scaling;
    }
    __syncthreads();
    if (tid == 0) *d.n_flipped = 0;
}

__global__ void __launch_bounds__(1024) k_begin(Dev d, int rescale_flipped) {
    begin_members(d, rescale_flipped);
}

// Per table block, for grains, edges and vertices in one launch: the
// numbers of enabled and of free slots, after the host uploads a state.
// The steps keep them from then on (each event that enables or disables a
// slot adds to its block's pair: set_*_enabled).  The enabled counts feed
// the grain / vertex totals of main.cu:317 (one partial per table block,
// summed per member by k_member_reduce); the pairs also serve the
// nucleation (free-slot search, alive-vertex draw), which reads them before
// the step's insertions.  One wave
// per table block: each lane reads four flags 64 bytes apart (four
// coalesced 64-byte loads per wave) and the wave counts them by ballots.
__global__ void __launch_bounds__(64 * kCountWaves) k_count_all(Dev d) {
    const int lane = threadIdx.x & 63;
    int lb = (int)blockIdx.x * kCountWaves + (int)(threadIdx.x >> 6);
    if (lb >= d.ngb + d.neb + d.nvb) return;
    const Blk* tbl;
    const uint8_t* en;
    int2* out;
    if (lb < d.ngb) {
        tbl = d.gblk; en = d.genab; out = d.bc_g;
    } else if (lb < d.ngb + d.neb) {
        lb -= d.ngb; tbl = d.eblk; en = d.benab; out = d.bc_e;
    } else {
        lb -= d.ngb + d.neb; tbl = d.vblk; en = d.ven; out = d.bc_v;
    }
    const Blk bd = table_blk(d, tbl, lb);
    uint8_t f[kBlock / 64];
#pragma unroll
    for (int j = 0; j < kBlock / 64; j++) {
        const int i = bd.b + 64 * j + lane;
        f[j] = i < bd.e ? en[i] : 0;
    }
    int r = 0;
#pragma unroll
    for (int j = 0; j < kBlock / 64; j++) r += (int)__popcll(__ballot(f[j] != 0));
    if (lane == 0) out[lb] = make_int2(r, (bd.e - bd.b) - r);
}

// One workgroup per member: sums the per-block partials of the enabled
// grain and vertex counts.
__global__ void __launch_bounds__(1024) k_member_reduce(Dev d) {
    __shared__ int sa[16], sb[16];
    const int m = blockIdx.x;
    // the step's work lists are done with: clear their counters for the next step
    if (m == 0 && (int)threadIdx.x < C_COUNT) d.cnt[threadIdx.x] = 0;
    if (m >= d.B) return;
    MState& M = d.ms[m];
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    // each lane reads its share of the grain and vertex partials, up to 16
    // of each at a time (independent loads, issued with the member state;
    // a single member's ranges are the whole tables), then the waves reduce
    // (wave_sum) and thread 0 sums the wave results: integer sums (at most
    // the slot counts), so any order gives the same value
    constexpr int kBatch = 16;
    const int glo = d.B == 1 ? 0 : d.gblk_first[m], ghi = d.B == 1 ? d.ngb : d.gblk_first[m + 1];
    const int vlo = d.B == 1 ? 0 : d.vblk_first[m], vhi = d.B == 1 ? d.nvb : d.vblk_first[m + 1];
    const int ngb = ghi - glo, nvb = vhi - vlo, nmax = max(ngb, nvb);
    int a = 0, b2 = 0;
    for (int k0 = tid; k0 < nmax; k0 += kBatch * nt) {
        int tg[kBatch], tv[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; q++) {
            const int k = k0 + q * nt;
            tg[q] = k < ngb ? d.bc_g[glo + k].x : 0;
            tv[q] = k < nvb ? d.bc_v[vlo + k].x : 0;
        }
#pragma unroll
        for (int q = 0; q < kBatch; q++) { a += tg[q]; b2 += tv[q]; }
    }
    if (!M.active) return;
    {
        a = wave_sum(a);
        b2 = wave_sum(b2);
        if (lane == 0) { sa[wave] = a; sb[wave] = b2; }
        __syncthreads();
        if (tid == 0) {
            int ra = 0, rb = 0;
            for (int w = 0; w < nw; w++) { ra += sa[w]; rb += sb[w]; }
            M.curr_grains = (int)ra;
            M.n_venabled = (int)rb;
            // main.cu:388-393: step counter, time and stop rules
            const Params& P = d.prm[m];
            M.dt = member_dt(P, M);
            M.n += 1;
            M.steps += 1;
            if (P.adaptive) M.t += M.dt;
            else M.t = M.n * P.dt_max;
            if (M.status != 0) M.done = 1;
            if (M.curr_grains <= P.min_grains && P.min_grains > 0) M.done = 1;
        }
    }
}


// ---------------------------------------------------------------------
// calculus.h: lengths, energies, areas, velocities, translation
// ---------------------------------------------------------------------

// compute_boundary_arclengths (calculus.h:83-93) fused with
// compute_boundary_energies (calculus.h:104-114).
__global__ void k_edge_len_energy(Dev d, int only_nucl) {
    TABLE_BLOCK(d.eblk, d.neb)
    if (!M.active || (only_nucl && !M.nucl_now)) return;
    if (i >= bd_.e || !d.benab[i]) return;
    int a = d.bvx[2 * i], b = d.bvx[2 * i + 1];
    d.blen[i] = arclen(d.pos[2 * a], d.pos[2 * a + 1], d.pos[2 * b], d.pos[2 * b + 1], P.D);
    double e = (P.gb_eps == 0.0) ? 1.0 : boundary_compute_energy(d, i, P.gb_eps);
    d.ben[i] = e * P.gb_scaling;
}

// The stored length of edge b seen from its end v: |wrap(p_v - p_nb)|,
// bit-identical to compute_arclength from either end (wrap_dist is odd and
// vector2_mag sign-symmetric).
__device__ __forceinline__ double edge_len_from(const Dev& d, int b, int v, double D) {
    const int nb = other_end(d, b, v);
    return arclen(d.pos[2 * v], d.pos[2 * v + 1], d.pos[2 * nb], d.pos[2 * nb + 1], D);
}

// compute_grain_areas, calculus.h:224-234
__global__ void k_grain_area(Dev d, int only_nucl) {
    TABLE_BLOCK(d.gblk, d.ngb)
    if (!M.active || (only_nucl && !M.nucl_now)) return;
    if (i >= bd_.e || !d.genab[i]) return;
    d.garea[i] = grain_area(d, i, P.D);
}

// compute_vertex_energies, calculus.h:123-133 (diagnostic)
__global__ void k_vertex_energy(Dev d) {
    TABLE_BLOCK(d.vblk, d.nvb)
    if (!M.active || i >= bd_.e || !d.ven[i]) return;
    double e = 0.0;
    for (int j = 0; j < 3; j++) {
        int b = d.vb[3 * i + j], g = d.vg[3 * i + j];
        e += (d.blen[b] * d.ben[b] * 0.5);
        e += (d.gse[g] * d.garea[g] / d.glen[g]);
    }
    d.venergy[i] = e;
}

// energy_diff, calculus.h:19-73: energy change for a grad_eps shift of one
// coordinate of vertex v, from 3 perturbed lengths and 3 perturbed areas.
__device__ __forceinline__ double energy_diff(const Dev& d, const Params& P, int comp, int v,
                                              double* bnd_out, double* grn_out) {
    const double D = P.D;
    double nx = d.pos[2 * v], ny = d.pos[2 * v + 1];
    if (comp == 0) nx += SEVC_GRAD_EPS;
    else ny += SEVC_GRAD_EPS;
    nx = dom_mod(nx, D);
    ny = dom_mod(ny, D);
    double arcl[3], arcl0[3];
    for (int j = 0; j < 3; j++) {
        int b = d.vb[3 * v + j];
        int nb = other_end(d, b, v);
        arcl[j] = arclen(nx, ny, d.pos[2 * nb], d.pos[2 * nb + 1], D);
        arcl0[j] = arclen(d.pos[2 * v], d.pos[2 * v + 1], d.pos[2 * nb], d.pos[2 * nb + 1], D);   // stored length
    }
    double areas[3];
    for (int j = 0; j < 3; j++) {
        int g = d.vg[3 * v + j];
        int n = d.glen[g];
        const int* L = glist(d, g);
        double a = 0.0;
        if (n > 0) {
            AreaAcc acc;
            int w = L[0];
            double x = d.pos[2 * w], y = d.pos[2 * w + 1];
            if (w == v) { x = nx; y = ny; }
            acc.start(x, y);
            for (int k = 1; k < n; k++) {
                w = L[k];
                x = d.pos[2 * w]; y = d.pos[2 * w + 1];
                if (w == v) { x = nx; y = ny; }
                acc.push(x, y, D);
            }
            a = acc.finish();
        }
        areas[j] = a;
    }
    double bnd_term = 0.0, grn_term = 0.0;
    for (int j = 0; j < 3; j++) {
        int b = d.vb[3 * v + j], g = d.vg[3 * v + j];
        bnd_term += (d.ben[b] * (arcl[j] - arcl0[j]));
        grn_term += (d.gse[g] * (fabs(areas[j]) - fabs(d.garea[g])));
    }
    *bnd_out = bnd_term;
    *grn_out = grn_term;
    return bnd_term + grn_term;
}

// Analytic -grad E at the current positions and grain areas.
__device__ __forceinline__ void analytic_force(const Dev& d, const Params& P, int v, double* fx, double* fy) {
    analytic_force_at(d, d.pos, d.garea, P, v, fx, fy);
}

// Velocity, pass 1 of 2 (stored-energy part of energy_diff,
// calculus.h:39-61, and compute_grain_areas, calculus.h:224-234): one
// workgroup per tile of consecutive grains.  The tile's vertex lists and
// the positions they name are staged in LDS with one wide pass over
// memory; each grain's area is computed once (and stored, as
// compute_grain_areas does); then every (grain, list slot) pair computes
// the two shifted areas of its vertex and writes SE*(|A'|-|A|) for x and
// y into the vertex's term slot of that grain (slot j where
// vertex.grains[j] == grain).
//
// Both shifted areas come from one pass over the polygon (the reference's
// compute_area on the x-shifted and on the y-shifted copy,
// geometry.h:757-765 after :908-931).  With the shifted point (px, y) the Y
// track of the unwrap is the unperturbed one and the X track differs from
// it only in the deltas that enter and leave position k (the vertex's
// slot); likewise for (x, qy).  Those four deltas are computed first from
// the two list neighbours and every other delta is shared.  Each track is
// still the reference's sequential sum U_i = U_{i-1} + delta_i, and each
// area the reference's left-to-right shoelace sum, so the results are
// bit-identical.  Pass 2 sends vertices with dom_mod(x) != x or
// dom_mod(y) != y to the two-pass path, which this form does not cover.
//
// The (grain, slot) pairs of the shifted areas are dealt to the lanes in
// order of their loop length after slot k, n - k (then grain), so that the
// lanes of a wave run that loop, the longer one, for the same number of
// iterations: pair (g, k) goes to position start(n - k) + (rank of g among
// the tile's grains with at least n - k entries).
constexpr int kTilePts = 448;   // with the per-slot and per-grain arrays: 8 workgroups of 256 per CU
static_assert(kTilePts >= kMaxStride, "a tile must hold the longest grain list");
__global__ void __launch_bounds__(256) k_vel_grain(Dev d, int tuning) {
    __shared__ double2 Ps[kTilePts];
    __shared__ double2 Ds[kTilePts];        // wrap(P_k - P_{k-1}) of list slot k > 0
    __shared__ int Ls[kTilePts];
    __shared__ unsigned char pgl[kTilePts];
    __shared__ unsigned char vmask[kTilePts]; // bit j: vg[j] of the slot's vertex is the slot's grain
    __shared__ unsigned short ord[kTilePts];  // the list entry of each term lane
    static_assert(kTilePts <= 65535, "entries fit 16 bits");
    __shared__ int off[kTile + 1];
    __shared__ double area_s[kTile], se_s[kTile];
    __shared__ unsigned char en_s[kTile];
    __shared__ int gend_s;
    const int lb = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (lb >= d.ntile) return;
    // a single member's tiles are consecutive runs of tile_T grains: their
    // bounds need no table load, so the tile's first loads (below) issue at
    // once, together with the member's state
    Blk bd;
    if (d.B == 1) {
        bd.m = 0;
        bd.b = lb * d.tile_T;
        bd.e = min(bd.b + d.tile_T, d.G);
    } else {
        bd = d.gtile[lb];
    }
    const int tid = threadIdx.x, nt = blockDim.x;
    // the first kRowEntries list entries of the tile's grains (one 16-byte
    // load per thread), staged in the delta array until the first deltas
    // overwrite them, and the first sub-tile's enable flags, lengths and
    // stored energies: loaded before the member test (in bounds whatever
    // the member's state); a tile split for its length reads its later
    // parts itself
    constexpr int kRowEntries = 16;
    static_assert(kRowEntries * kTile * 4 <= (int)sizeof(Ds), "rows fit the delta array");
    int* rowbuf = reinterpret_cast<int*>(Ds);
    const bool rows = (d.S % 4 == 0) && d.S >= kRowEntries;
    int4 row0 = make_int4(0, 0, 0, 0);
    const int grow = bd.b + tid / (kRowEntries / 4);
    const bool has_row = rows && tid < kTile * (kRowEntries / 4) && grow < bd.e;
    if (has_row) row0 = reinterpret_cast<const int4*>(glist(d, grow))[tid % (kRowEntries / 4)];
    uint8_t en0 = 0;
    int len0 = 0;
    double se0 = 0.0;
    if (tid < kTile && bd.b + tid < bd.e) {
        en0 = d.genab[bd.b + tid];
        len0 = d.glen[bd.b + tid];
        se0 = d.gse[bd.b + tid];
    }
    const Params& P = d.prm[bd.m];
    const MState& M = d.ms[bd.m];
    // tuning: timed by sevc_autotune on any member, without writing areas
    // (the terms it writes are scratch of the next step's pass 2)
    if ((!M.active && !tuning) || P.vel_model != SEVC_VEL_FD) return;
    const double D = P.D;
    if (has_row) reinterpret_cast<int4*>(rowbuf)[tid] = row0;
    for (int gs = bd.b; gs < bd.e;) {
        const bool from_rows = rows && gs == bd.b;
        if (tid < kTile) {
            // wave 0: inclusive scan of list lengths; the tile takes the
            // longest prefix of grains whose lists fit in kTilePts points
            const int g = gs + tid;
            const bool in = g < bd.e;
            const bool first = gs == bd.b;
            const uint8_t en8 = first ? en0 : (in ? d.genab[g] : 0);
            const int len = first ? len0 : (in ? d.glen[g] : 0);
            const bool en = en8 != 0;
            const int n = en ? len : 0;
            // the values the later phases need from global memory, loaded here
            en_s[tid] = en ? 1 : 0;
            se_s[tid] = first ? se0 : (in ? d.gse[g] : 0.0);
            static_assert(kTile == 64, "one wave scans a sub-tile's lengths");
            const int inc = wave_inclusive_scan(n);
            const unsigned long long over = __ballot(g >= bd.e || inc > kTilePts);
            const int cnt = over ? __ffsll((long long)over) - 1 : kTile;
            const int nm = tid < cnt ? n : 0;   // entries of this sub-tile's grain
            if (tid < cnt) {
                off[tid] = inc - n;
                for (int q = inc - n; q < inc; q++) pgl[q] = (unsigned char)tid;   // the grain of each point
            }
            if (tid == cnt - 1) off[cnt] = inc;
            if (tid == 0) {
                gend_s = gs + cnt;
                if (cnt == 0) off[0] = 0;
            }
            // term lanes by loop length r = n - k: every grain with n >= r
            // has one slot with that length (k = n - r)
            const unsigned long long below = (1ull << tid) - 1ull;
            int base = 0;
            for (int r = 1;; r++) {
                const unsigned long long has = __ballot(nm >= r);
                if (!has) break;
                if (nm >= r) ord[base + (int)__popcll(has & below)] = (unsigned short)(inc - r);
                base += (int)__popcll(has);
            }
        }
        __syncthreads();
        const int gend = gend_s, ngr = gend - gs, npts = off[ngr];
        for (int p = tid; p < npts; p += nt) {
            const int g = pgl[p], e = p - off[g];
            const int w = (from_rows && e < kRowEntries) ? rowbuf[g * kRowEntries + e] : glist(d, gs + g)[e];
            Ls[p] = w;
            Ps[p] = make_double2(d.pos[2 * w], d.pos[2 * w + 1]);
            unsigned char mk = 0;
            for (int j = 0; j < 3; j++) mk |= (d.vg[3 * w + j] == gs + g) ? (unsigned char)(1 << j) : 0;
            vmask[p] = mk;
        }
        __syncthreads();
        for (int p = tid; p < npts; p += nt) {
            // the unwrap deltas of the unperturbed polygon, shared by its area
            // and by every shifted area of the grain (the same wrap_dist calls
            // the sequential unwrap makes)
            const int b0 = off[pgl[p]];
            Ds[p] = (p > b0) ? make_double2(wrapd(Ps[p].x, Ps[p - 1].x, D), wrapd(Ps[p].y, Ps[p - 1].y, D))
                             : make_double2(0.0, 0.0);
        }
        __syncthreads();
        for (int g = tid; g < ngr; g += nt) {
            const int b0 = off[g], n = off[g + 1] - b0;
            // compute_area from the deltas: U_k = U_{k-1} + D_k, left to right
            double ux = 0, uy = 0, a = 0;
            for (int k = 1; k < n; k++) {
                const double2 du = Ds[b0 + k];
                const double nx = ux + du.x, ny = uy + du.y;
                a += ux * ny - uy * nx;
                ux = nx;
                uy = ny;
            }
            const double z = 0;
            a += ux * z - uy * z;
            a *= 0.5;
            if (n <= 0) a = 0.0;
            const int G = gs + g;
            if (en_s[g] && !tuning) d.garea[G] = a;
            area_s[g] = a;
        }
        __syncthreads();
        for (int q = tid; q < npts; q += nt) {
            const int p = ord[q];
            const int g = pgl[p], b0 = off[g], n = off[g + 1] - b0, k = p - b0;
            const int v = Ls[p];
            const double x = Ps[p].x, y = Ps[p].y;
            const double px = dom_mod(x + SEVC_GRAD_EPS, D), qy = dom_mod(y + SEVC_GRAD_EPS, D);
            // the unshifted track up to U_{k-1}, with the sum a of
            // U_i x U_{i+1} for i <= k - 2 (both shifted polygons share it)
            double ux = 0, uy = 0, a = 0;
            for (int i = 1; i < k; i++) {
                const double2 di = Ds[b0 + i];
                const double nx = ux + di.x, ny = uy + di.y;
                a += ux * ny - uy * nx;
                ux = nx;
                uy = ny;
            }
            // index k: the x-shifted polygon's point (xX, Y_k), the
            // y-shifted one's (X_k, yY); index k + 1: deltas from the shifted
            // point; later the shared deltas
            double sx = a, sy = a, xX = 0, yY = 0;
            if (k > 0) {
                const double ex = wrapd(px, Ps[p - 1].x, D), ey = wrapd(qy, Ps[p - 1].y, D);
                const double2 dk = Ds[p];
                const double nx = ux + dk.x, ny = uy + dk.y;
                xX = ux + ex;
                yY = uy + ey;
                sx += ux * ny - uy * xX;
                sy += ux * yY - uy * nx;
                ux = nx;
                uy = ny;
            }
            if (k + 1 < n) {
                const double ex1 = wrapd(Ps[p + 1].x, px, D), ey1 = wrapd(Ps[p + 1].y, qy, D);
                const double2 d1 = Ds[p + 1];
                const double nx = ux + d1.x, ny = uy + d1.y;
                const double nxX = xX + ex1, nyY = yY + ey1;
                sx += xX * ny - uy * nxX;
                sy += ux * nyY - yY * nx;
                ux = nx; uy = ny; xX = nxX; yY = nyY;
                for (int i = k + 2; i < n; i++) {
                    const double2 di = Ds[b0 + i];
                    const double nx2 = ux + di.x, ny2 = uy + di.y;
                    const double nxX2 = xX + di.x, nyY2 = yY + di.y;
                    sx += xX * ny2 - uy * nxX2;
                    sy += ux * nyY2 - yY * nx2;
                    ux = nx2; uy = ny2; xX = nxX2; yY = nyY2;
                }
            }
            const double z = 0;
            sx += xX * z - uy * z;
            sy += ux * z - yY * z;
            const double ax = sx * 0.5, ay = sy * 0.5;
            const double a0 = fabs(area_s[g]), se = se_s[g];
            const double tx = se * (fabs(ax) - a0), ty = se * (fabs(ay) - a0);
            const unsigned mk = vmask[p];
            for (int j = 0; j < 3; j++)
                if ((mk >> j) & 1u) {
                    d.vterm[6 * (size_t)v + 2 * j] = tx;
                    d.vterm[6 * (size_t)v + 2 * j + 1] = ty;
                }
        }
        __syncthreads();
        gs = gend;
    }
}

// Velocity pass 1 with one lane per grain (an alternative to the tiles of
// k_vel_grain; both feed k_vel_vertex): the lane loads the grain's length,
// SE and first kGL list entries together, then its vertices' positions and
// grain triples; forms the unwrap deltas, the grain's area
// (compute_grain_areas, calculus.h:224-234) and, for every slot k, the two
// shifted areas of energy_diff (calculus.h:39-61) from the unshifted
// track's prefix at k - 1 (the shifted polygons' tracks equal it up to
// there); writes SE*(|A'| - |A|) for x and y into the vertex's term slot of
// this grain.  No LDS, no barrier: every array index is a compile-time
// constant, so the polygon lives in registers.  Grains longer than kGL go
// to a list for k_vel_bigrain.
constexpr int kGL = 8;

// One shifted pair (x, y) of areas of a polygon in registers for slot k:
// prefix (ux, uy) = U_{k-1}, a_pre = the sum of U_i x U_{i+1} for i <= k-2.
template <int K>
__device__ __forceinline__ void slot_areas(const double2 (&Pq)[kGL], const double (&dx)[kGL],
                                           const double (&dy)[kGL], int n, double D, double ux, double uy,
                                           double a_pre, double* ax_out, double* ay_out) {
    const double px = dom_mod(Pq[K].x + SEVC_GRAD_EPS, D), qy = dom_mod(Pq[K].y + SEVC_GRAD_EPS, D);
    double ax = a_pre, ay = a_pre, xX, yY;
    if (K == 0) {
        xX = 0.0;
        yY = 0.0;
    } else {
        const double ex = wrapd(px, Pq[K - 1].x, D), ey = wrapd(qy, Pq[K - 1].y, D);
        const double nx = ux + dx[K], ny = uy + dy[K];
        xX = ux + ex;
        yY = uy + ey;
        ax += ux * ny - uy * xX;
        ay += ux * yY - uy * nx;
        ux = nx;
        uy = ny;
    }
    if (K + 1 < kGL && K + 1 < n) {
        const double ex1 = wrapd(Pq[K + 1].x, px, D), ey1 = wrapd(Pq[K + 1].y, qy, D);
        const double nx = ux + dx[K + 1], ny = uy + dy[K + 1];
        const double nxX = xX + ex1, nyY = yY + ey1;
        ax += xX * ny - uy * nxX;
        ay += ux * nyY - yY * nx;
        ux = nx; uy = ny; xX = nxX; yY = nyY;
#pragma unroll
        for (int i = K + 2; i < kGL; i++) {
            if (i < n) {
                const double nx2 = ux + dx[i], ny2 = uy + dy[i];
                const double nxX2 = xX + dx[i], nyY2 = yY + dy[i];
                ax += xX * ny2 - uy * nxX2;
                ay += ux * nyY2 - yY * nx2;
                ux = nx2; uy = ny2; xX = nxX2; yY = nyY2;
            }
        }
    }
    const double z = 0;
    ax += xX * z - uy * z;
    ay += ux * z - yY * z;
    *ax_out = ax * 0.5;
    *ay_out = ay * 0.5;
}

template <int K>
__device__ __forceinline__ void slot_terms(const Dev& d, int g, const int (&L)[kGL], const int (&jq)[kGL],
                                           const double2 (&Pq)[kGL], const double (&dx)[kGL],
                                           const double (&dy)[kGL], int n, double D, double a0, double se,
                                           double& ux, double& uy, double& a_pre) {
    if (K >= n) return;
    double ax, ay;
    slot_areas<K>(Pq, dx, dy, n, D, ux, uy, a_pre, &ax, &ay);
    const double tx = se * (fabs(ax) - a0), ty = se * (fabs(ay) - a0);
    const int v = L[K], j = jq[K];
    if (j >= 0) reinterpret_cast<double2*>(d.vterm)[3 * (size_t)v + j] = make_double2(tx, ty);
    // advance the unshifted prefix to (U_K, S_{K-1})
    if (K > 0) {
        const double nx = ux + dx[K], ny = uy + dy[K];
        a_pre += ux * ny - uy * nx;
        ux = nx;
        uy = ny;
    }
}

__global__ void __launch_bounds__(256) k_vel_glane(Dev d, int tuning) {
    TABLE_BLOCK(d.gblk, d.ngb)
    const bool in = i < bd_.e;
    uint8_t en = 0;
    int n = 0;
    double se = 0.0;
    int4 r0 = make_int4(-1, -1, -1, -1), r1 = r0;
    if (in) {
        en = d.genab[i];
        n = d.glen[i];
        se = d.gse[i];
        const int4* row = reinterpret_cast<const int4*>(glist(d, i));
        r0 = row[0];
        r1 = row[1];
    }
    if ((!M.active && !tuning) || P.vel_model != SEVC_VEL_FD || !en) return;
    if (n > kGL) {
        const int s = atomicAdd(&d.cnt[C_BIGR], 1);
        if (s < d.G) d.l_bigr[s] = i;
        return;
    }
    const double D = P.D;
    const int L[kGL] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    const double2* X = reinterpret_cast<const double2*>(d.pos);
    double2 Pq[kGL];
    int jq[kGL];
#pragma unroll
    for (int q = 0; q < kGL; q++) Pq[q] = q < n ? X[L[q]] : make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < kGL; q++) {
        int j = -1;
        if (q < n) {
            const int v = L[q];
            const int a = d.vg[3 * v], b = d.vg[3 * v + 1], c = d.vg[3 * v + 2];
            j = (a == i) ? 0 : ((b == i) ? 1 : ((c == i) ? 2 : -1));
        }
        jq[q] = j;
    }
    double dx[kGL], dy[kGL];
    dx[0] = 0.0;
    dy[0] = 0.0;
#pragma unroll
    for (int q = 1; q < kGL; q++) {
        dx[q] = q < n ? wrapd(Pq[q].x, Pq[q - 1].x, D) : 0.0;
        dy[q] = q < n ? wrapd(Pq[q].y, Pq[q - 1].y, D) : 0.0;
    }
    // compute_area (geometry.h:757-765 after :908-931): U_q = U_{q-1} + D_q
    double ux = 0.0, uy = 0.0, a = 0.0;
#pragma unroll
    for (int q = 1; q < kGL; q++) {
        if (q < n) {
            const double nx = ux + dx[q], ny = uy + dy[q];
            a += ux * ny - uy * nx;
            ux = nx;
            uy = ny;
        }
    }
    const double z = 0;
    a += ux * z - uy * z;
    a *= 0.5;
    if (n <= 0) a = 0.0;
    if (!tuning) d.garea[i] = a;
    const double a0 = fabs(a);
    double pux = 0.0, puy = 0.0, pre = 0.0;
    slot_terms<0>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
    slot_terms<1>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
    slot_terms<2>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
    slot_terms<3>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
    slot_terms<4>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
    slot_terms<5>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
    slot_terms<6>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
    slot_terms<7>(d, i, L, jq, Pq, dx, dy, n, D, a0, se, pux, puy, pre);
}

// Velocity pass 1 for the grains k_vel_glane listed (more than kGL
// vertices): one wave per grain.  The list, positions and the vertices'
// grain slots are staged in LDS, lane 0 forms the area and the prefix of
// the unshifted track, then lane k computes slot k's two shifted areas from
// the prefix at k - 1 (as slot_areas does) and writes its terms.
__global__ void __launch_bounds__(64) k_vel_bigrain(Dev d, int tuning) {
    __shared__ double2 Ps[kMaxStride], Ds[kMaxStride], Us[kMaxStride];
    __shared__ double Ss[kMaxStride];
    __shared__ int Ls[kMaxStride];
    __shared__ double a_s;
    const int n_items = min(d.cnt[C_BIGR], d.G);
    const int lane = threadIdx.x;
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int g = d.l_bigr[item];
        const Params& P = d.prm[member_of_grain(d, g)];
        const double D = P.D;
        const int n = min(d.glen[g], kMaxStride);
        const double se = d.gse[g];
        const int* Lg = glist(d, g);
        for (int k = lane; k < n; k += 64) {
            Ls[k] = Lg[k];
            Ps[k] = reinterpret_cast<const double2*>(d.pos)[Ls[k]];
        }
        __syncthreads();
        for (int k = lane; k < n; k += 64)
            Ds[k] = k > 0 ? make_double2(wrapd(Ps[k].x, Ps[k - 1].x, D), wrapd(Ps[k].y, Ps[k - 1].y, D))
                          : make_double2(0.0, 0.0);
        __syncthreads();
        if (lane == 0) {
            // Us[k] = U_k, Ss[k] = t_0 + ... + t_k (left to right)
            double ux = 0.0, uy = 0.0, a = 0.0;
            Us[0] = make_double2(0.0, 0.0);
            for (int k = 1; k < n; k++) {
                const double nx = ux + Ds[k].x, ny = uy + Ds[k].y;
                a += ux * ny - uy * nx;
                Ss[k - 1] = a;
                ux = nx;
                uy = ny;
                Us[k] = make_double2(ux, uy);
            }
            const double z = 0;
            a += ux * z - uy * z;
            a *= 0.5;
            a_s = a;
            if (!tuning) d.garea[g] = a;
        }
        __syncthreads();
        const double a0 = fabs(a_s);
        for (int k = lane; k < n; k += 64) {
            const double px = dom_mod(Ps[k].x + SEVC_GRAD_EPS, D), qy = dom_mod(Ps[k].y + SEVC_GRAD_EPS, D);
            double ux = 0.0, uy = 0.0, ax = 0.0, ay = 0.0, xX = 0.0, yY = 0.0;
            if (k > 0) {
                ux = Us[k - 1].x;
                uy = Us[k - 1].y;
                ax = ay = (k >= 2) ? Ss[k - 2] : 0.0;
                const double ex = wrapd(px, Ps[k - 1].x, D), ey = wrapd(qy, Ps[k - 1].y, D);
                const double nx = ux + Ds[k].x, ny = uy + Ds[k].y;
                xX = ux + ex;
                yY = uy + ey;
                ax += ux * ny - uy * xX;
                ay += ux * yY - uy * nx;
                ux = nx;
                uy = ny;
            }
            if (k + 1 < n) {
                const double ex1 = wrapd(Ps[k + 1].x, px, D), ey1 = wrapd(Ps[k + 1].y, qy, D);
                const double nx = ux + Ds[k + 1].x, ny = uy + Ds[k + 1].y;
                const double nxX = xX + ex1, nyY = yY + ey1;
                ax += xX * ny - uy * nxX;
                ay += ux * nyY - yY * nx;
                ux = nx; uy = ny; xX = nxX; yY = nyY;
                for (int i = k + 2; i < n; i++) {
                    const double2 di = Ds[i];
                    const double nx2 = ux + di.x, ny2 = uy + di.y;
                    const double nxX2 = xX + di.x, nyY2 = yY + di.y;
                    ax += xX * ny2 - uy * nxX2;
                    ay += ux * nyY2 - yY * nx2;
                    ux = nx2; uy = ny2; xX = nxX2; yY = nyY2;
                }
            }
            const double z = 0;
            ax += xX * z - uy * z;
            ay += ux * z - yY * z;
            ax *= 0.5;
            ay *= 0.5;
            const int v = Ls[k];
            const int a = d.vg[3 * v], b = d.vg[3 * v + 1], c = d.vg[3 * v + 2];
            const int j = (a == g) ? 0 : ((b == g) ? 1 : ((c == g) ? 2 : -1));
            if (j >= 0)
                reinterpret_cast<double2*>(d.vterm)[3 * (size_t)v + j] =
                    make_double2(se * (fabs(ax) - a0), se * (fabs(ay) - a0));
        }
        __syncthreads();
    }
}

// A vertex for the two-pass velocity (k_velocity_fallback).  Each vertex is
// listed at most once per step (the list holds V entries; the counter is
// cleared by k_begin); timing launches of the autotuner list nothing.
__device__ __forceinline__ void push_vfall(const Dev& d, int v) {
    const int s = atomicAdd(&d.cnt[C_VFALL], 1);
    if (s < d.V) d.l_vfall[s] = v;
}

// Velocity, pass 2 of 2: the line-tension part of energy_diff
// (calculus.h:28-37,57-60) -- six shifted lengths -- plus the three
// stored-energy terms of pass 1 summed in vertex.grains order, then
// v = -dE/grad_eps * mobility (calculus.h:149-150).  Vertices whose
// unshifted coordinate dom_mod would move (a coordinate equal to D) are
// listed for the two-pass path.
__global__ void __launch_bounds__(256) k_vel_vertex(Dev d, int tuning) {
    TABLE_BLOCK(d.vblk, d.nvb)
    const int v = i;
    const bool in = i < bd_.e;
    const uint8_t en = in ? d.ven[i] : 0;
    // the vertex's data (position, edges, their ends, the neighbours'
    // positions, the stored-energy terms) is loaded together with the
    // member's state
    const double2* X = reinterpret_cast<const double2*>(d.pos);
    double2 p = make_double2(0.0, 0.0), q[3], t[3];
    int b[3] = {0, 0, 0};
    if (en) {
        p = X[v];
        int2 ends[3];
#pragma unroll
        for (int j = 0; j < 3; j++) b[j] = d.vb[3 * v + j];
#pragma unroll
        for (int j = 0; j < 3; j++) ends[j] = reinterpret_cast<const int2*>(d.bvx)[b[j]];
#pragma unroll
        for (int j = 0; j < 3; j++) q[j] = X[ends[j].x == v ? ends[j].y : ends[j].x];
#pragma unroll
        for (int j = 0; j < 3; j++) t[j] = reinterpret_cast<const double2*>(d.vterm)[3 * (size_t)v + j];
    }
    if ((!M.active && !tuning) || P.vel_model != SEVC_VEL_FD || !en) return;
    const double D = P.D;
    const double x = p.x, y = p.y;
    const double px = dom_mod(x + SEVC_GRAD_EPS, D), pyy = dom_mod(y, D);
    const double qx = dom_mod(x, D), qy = dom_mod(y + SEVC_GRAD_EPS, D);
    if (!((pyy == y) && (qx == x))) {
        if (!tuning) push_vfall(d, v);
        return;
    }
    const bool uni = (P.gb_eps == 0.0);
    double bx = 0.0, by = 0.0, gx = 0.0, gy = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const double nbx = q[j].x, nby = q[j].y;
        const double lx = arclen(px, pyy, nbx, nby, D);
        const double ly = arclen(qx, qy, nbx, nby, D);
        // the stored arclength of b (compute_boundary_arclengths at the step's
        // positions, refreshed for edges moved by nucleation): wrap_dist is odd
        // and vector2_mag sign-symmetric, so |wrap(p_v - p_nb)| has its bits
        const double L0 = arclen(x, y, nbx, nby, D);
        const double gam = uni ? P.gb_scaling : d.ben[b[j]];
        bx += (gam * (lx - L0));
        by += (gam * (ly - L0));
    }
    gx += t[0].x; gy += t[0].y;
    gx += t[1].x; gy += t[1].y;
    gx += t[2].x; gy += t[2].y;
    d.vel[2 * v] = (-(bx + gx) / SEVC_GRAD_EPS) * P.mobility;
    d.vel[2 * v + 1] = (-(by + gy) / SEVC_GRAD_EPS) * P.mobility;
}

// The reference's two energy_diff passes for one vertex (also stores the
// velocity parts exported by export_data, calculus.h:62-69).
__device__ void velocity_generic(const Dev& d, const Params& P, int v, bool parts) {
    double bx, gx, by, gy;
    double ex = energy_diff(d, P, 0, v, &bx, &gx);
    double ey = energy_diff(d, P, 1, v, &by, &gy);
    d.vel[2 * v] = (-ex / SEVC_GRAD_EPS) * P.mobility;
    d.vel[2 * v + 1] = (-ey / SEVC_GRAD_EPS) * P.mobility;
    if (parts) {
        d.vparts[4 * v + 0] = -bx / SEVC_GRAD_EPS;
        d.vparts[4 * v + 1] = -gx / SEVC_GRAD_EPS;
        d.vparts[4 * v + 2] = -by / SEVC_GRAD_EPS;
        d.vparts[4 * v + 3] = -gy / SEVC_GRAD_EPS;
    }
}

// Every FD vertex through the two-pass path (diagnostic steps).
__global__ void k_velocity_generic(Dev d) {
    TABLE_BLOCK(d.vblk, d.nvb)
    if (!M.active || P.vel_model != SEVC_VEL_FD || i >= bd_.e || !d.ven[i]) return;
    velocity_generic(d, P, i, d.diag != 0);
}

__global__ void k_velocity_fallback(Dev d) {
    const int n = min(d.cnt[C_VFALL], d.V);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int v = d.l_vfall[k];
        velocity_generic(d, d.prm[member_of_vertex(d, v)], v, false);
    }
}

// Velocity in one pass over tiles of consecutive vertex slots (energy_diff,
// calculus.h:19-73, for both components, compute_vertex_velocities,
// :143-154, and compute_grain_areas, :224-234).  A workgroup takes VT
// vertices of one member:
//  1. their positions, grains and boundaries (coalesced), and the distinct
//     grains they touch, gathered in an LDS hash set;
//  2. those grains' lengths, stored energies and list rows, then the
//     positions of every listed vertex (one gather each), staged in LDS;
//     meanwhile each vertex lane loads its boundaries' other ends and their
//     positions;
//  3. the unwrap deltas wrap(P_k - P_{k-1}) of every staged polygon;
//  4. one lane per (vertex, grain slot j): the grain's area and the two
//     areas with the vertex shifted by grad_eps in x and in y, in one pass
//     over the polygon's deltas (see below), giving the stored-energy terms
//     SE*(|A'| - |A|); the lane of the grain's first list entry writes the
//     grain's area (exactly one tile holds that vertex);
//  5. one lane per vertex: the six shifted lengths, the terms summed in
//     vertex.grains order, v = -dE/grad_eps * mobility.
// A grain touched by several tiles is staged by each (its areas are
// recomputed, identically); every (vertex, grain) term is computed once.
// No intermediate leaves the workgroup.
//
// The shifted areas are the reference's compute_area (geometry.h:757-765
// after the sequential unwrap of :908-931) on the polygon with point k
// replaced.  Unwrapped points U_i = U_{i-1} + D_i (U_0 = 0) and the
// left-to-right sum of U_i x U_{i+1} are carried for the unshifted polygon;
// up to U_{k-1} the shifted polygons' tracks are that track, at i = k and
// k + 1 their deltas are the ones to and from the shifted point, later they
// add the shared deltas D_i.  Every sum is the reference's, term by term,
// so the results are bit-identical.  Vertices whose unshifted coordinate
// dom_mod would move go to the two-pass path (k_velocity_fallback), as do
// all vertices of a tile whose polygons exceed the staging room.
template <int VT>
struct VTileCfg {
    static constexpr int kPts = 7 * VT;          // staged list entries (a 1M-grain tessellation needs <= 6 VT)
    static constexpr int kGr = 2 * VT;           // distinct grains (about 0.8 VT at 64)
    static constexpr int kHash = 4 * VT;         // hash slots (power of two; > 3 VT keys, so probing ends)
    static constexpr int kThreads = 4 * VT;      // one lane per (vertex, slot) needs 3 VT
    static constexpr int kRow = 8;               // list entries loaded with the grain
};

template <int VT>
__global__ void __launch_bounds__(4 * VT) k_vel_vtile(Dev d, int tuning) {
    using C = VTileCfg<VT>;
    __shared__ double2 Ps[C::kPts];          // raw positions of the staged entries
    __shared__ double2 Ds[C::kPts];          // unwrap deltas; before that the grains' list rows
    __shared__ int Ls[C::kPts];              // vertex of each staged entry
    __shared__ unsigned char pg[C::kPts];    // its grain (index into ug)
    __shared__ int hkey[C::kHash];           // hash set of grain ids (-1: empty)
    __shared__ unsigned char hval[C::kHash]; // index of the grain in ug
    __shared__ int ug[C::kGr];               // distinct grains
    __shared__ int uoff[C::kGr + 1];         // their first staged entry
    __shared__ int ulen[C::kGr];             // their list lengths
    __shared__ double use_[C::kGr];          // their SE
    __shared__ int vslot[3 * VT];            // hash slot of vertex.grains[j]
    __shared__ int vid[VT];                  // vertex of each tile lane (-1: not computed here)
    __shared__ double2 vterm_s[3 * VT];      // (x, y) stored-energy terms of (vertex, j)
    __shared__ int nu_s, nover_s;
    static_assert(C::kGr <= 255, "grain index fits a byte");
    static_assert((C::kHash & (C::kHash - 1)) == 0, "hash size is a power of two");
    int* rowbuf = reinterpret_cast<int*>(Ds);   // kGr rows of kRow entries
    static_assert(sizeof(Ds) >= sizeof(int) * C::kGr * C::kRow, "row staging fits the delta array");

    const int lb = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (lb >= d.nvtile) return;
    const Blk bd = d.vtile[lb];
    const Params& P = d.prm[bd.m];
    const MState& M = d.ms[bd.m];
    // tuning: timed by sevc_autotune on any member (velocities and areas
    // are recomputed by every step before they are read)
    if ((!M.active && !tuning) || P.vel_model != SEVC_VEL_FD) return;
    const double D = P.D;
    const int tid = threadIdx.x;
    const int nt = C::kThreads;

    // 1. the tile's vertices (lane t < VT) and the hash set of their grains
    const int v = bd.b + tid;
    const bool vlane = tid < VT;
    bool ven = false;
    double2 pv = make_double2(0.0, 0.0);
    int vgj[3] = {-1, -1, -1}, vbj[3] = {-1, -1, -1};
    if (vlane && v < bd.e) {
        // one round: the flag and the slot's data (valid addresses for any slot)
        ven = d.ven[v] != 0;
        pv = reinterpret_cast<const double2*>(d.pos)[v];
#pragma unroll
        for (int j = 0; j < 3; j++) { vgj[j] = d.vg[3 * v + j]; vbj[j] = d.vb[3 * v + j]; }
    }
    for (int h = tid; h < C::kHash; h += nt) hkey[h] = -1;
    if (tid == 0) { nu_s = 0; nover_s = 0; }
    __syncthreads();
    // the other ends of the vertex's boundaries: loads issued now, used last
    int2 ends[3] = {make_int2(0, 0), make_int2(0, 0), make_int2(0, 0)};
    if (ven) {
#pragma unroll
        for (int j = 0; j < 3; j++) ends[j] = reinterpret_cast<const int2*>(d.bvx)[vbj[j]];
    }
    if (vlane) vid[tid] = ven ? v : -1;
    if (ven) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int g = vgj[j];
            unsigned h = ((unsigned)g * 2654435761u) >> 7;
            for (;;) {
                h &= C::kHash - 1;
                const int old = atomicCAS(&hkey[h], -1, g);
                if (old == -1) {
                    const int u = atomicAdd(&nu_s, 1);
                    if (u < C::kGr) {
                        hval[h] = (unsigned char)u;
                        ug[u] = g;
                    }
                    break;
                }
                if (old == g) break;
                h++;
            }
            vslot[3 * tid + j] = (int)h;
        }
    }
    __syncthreads();
    const int nu = min(nu_s, C::kGr);

    // 2. the grains' lengths, SE and first list entries; the neighbours'
    // positions (loads issued now, used last)
    double2 q[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    if (ven) {
        const double2* X = reinterpret_cast<const double2*>(d.pos);
#pragma unroll
        for (int j = 0; j < 3; j++) q[j] = X[ends[j].x == v ? ends[j].y : ends[j].x];
    }
    for (int u = tid; u < nu; u += nt) {
        const int g = ug[u];
        ulen[u] = d.glen[g];
        use_[u] = d.gse[g];
    }
    for (int c = tid; c < nu * (C::kRow / 4); c += nt) {
        const int u = c / (C::kRow / 4), part = c % (C::kRow / 4);
        const int4 r = reinterpret_cast<const int4*>(glist(d, ug[u]))[part];
        reinterpret_cast<int4*>(rowbuf)[c] = r;
    }
    __syncthreads();
    // inclusive scan of the lengths by wave 0 (nu <= kGr, a few per lane)
    if (tid < 64) {
        constexpr int kPer = (C::kGr + 63) / 64;
        int loc[kPer], sum = 0;
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const int u = tid * kPer + e;
            loc[e] = u < nu ? ulen[u] : 0;
            sum += loc[e];
        }
        const int inc = wave_inclusive_scan(sum);
        int run = inc - sum;
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const int u = tid * kPer + e;
            if (u < nu) { uoff[u] = run; run += loc[e]; }
        }
        if (tid == 63) {
            uoff[nu] = inc;
            if (inc > C::kPts || nu_s > C::kGr) nover_s = 1;
        }
    }
    __syncthreads();
    const int npts = uoff[nu];
    if (nover_s) {
        // polygons too long to stage: the tile's vertices take the two-pass
        // path, which reads the stored areas; the areas this tile would have
        // written (grains whose first entry is one of its vertices) are
        // written here
        if (ven) {
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (glist(d, vgj[j])[0] == v) d.garea[vgj[j]] = grain_area(d, vgj[j], D);
            if (!tuning) push_vfall(d, v);
        }
        return;
    }
    // each grain's entries: its vertices and grain index (entries past the
    // first kRow are read from the list itself)
    for (int u = tid; u < nu; u += nt) {
        const int b0 = uoff[u], n = uoff[u + 1] - b0;
        const int* row = rowbuf + u * C::kRow;
        int tmp[C::kRow];
#pragma unroll
        for (int i = 0; i < C::kRow; i++) tmp[i] = row[i];
        for (int i = 0; i < n; i++) pg[b0 + i] = (unsigned char)u;
#pragma unroll
        for (int i = 0; i < C::kRow; i++)
            if (i < n) Ls[b0 + i] = tmp[i];
        if (n > C::kRow) {
            const int* L = glist(d, ug[u]);
            for (int i = C::kRow; i < n; i++) Ls[b0 + i] = L[i];
        }
    }
    __syncthreads();
    {
        const double2* X = reinterpret_cast<const double2*>(d.pos);
        for (int p = tid; p < npts; p += nt) Ps[p] = X[Ls[p]];
    }
    __syncthreads();

    // 3. unwrap deltas (the row staging in Ds is no longer needed)
    for (int p = tid; p < npts; p += nt) {
        const int b0 = uoff[pg[p]];
        Ds[p] = (p > b0) ? make_double2(wrapd(Ps[p].x, Ps[p - 1].x, D), wrapd(Ps[p].y, Ps[p - 1].y, D))
                         : make_double2(0.0, 0.0);
    }
    __syncthreads();

    // 4. one lane per (vertex, slot j): lane t = j * VT + vertex
    if (tid < 3 * VT) {
        const int j = tid / VT, vi = tid % VT;
        const int w = vid[vi];
        double2 res = make_double2(0.0, 0.0);
        if (w >= 0) {
            const int u = hval[vslot[3 * vi + j]];
            const int b0 = uoff[u], n = uoff[u + 1] - b0;
            int k = -1;
            for (int i = 0; i < n; i++)
                if (Ls[b0 + i] == w) { k = i; break; }
            if (k >= 0) {
                const double2 pw = Ps[b0 + k];
                const double px = dom_mod(pw.x + SEVC_GRAD_EPS, D), qy = dom_mod(pw.y + SEVC_GRAD_EPS, D);
                // unshifted track up to U_{k-1}: sum a of U_i x U_{i+1}, i <= k-2
                double ux = 0.0, uy = 0.0, a = 0.0;
                for (int i = 1; i < k; i++) {
                    const double2 di = Ds[b0 + i];
                    const double nx = ux + di.x, ny = uy + di.y;
                    a += ux * ny - uy * nx;
                    ux = nx;
                    uy = ny;
                }
                // tracks at index k: the x-shifted polygon's point (xX, Y_k),
                // the y-shifted one's (X_k, yY)
                double ax = a, ay = a, xX, yY;
                if (k == 0) {
                    xX = 0.0;
                    yY = 0.0;
                } else {
                    const double2 pm = Ps[b0 + k - 1];
                    const double2 dk = Ds[b0 + k];
                    const double ex = wrapd(px, pm.x, D), ey = wrapd(qy, pm.y, D);
                    const double nx = ux + dk.x, ny = uy + dk.y;
                    xX = ux + ex;
                    yY = uy + ey;
                    a += ux * ny - uy * nx;
                    ax += ux * ny - uy * xX;
                    ay += ux * yY - uy * nx;
                    ux = nx;
                    uy = ny;
                }
                if (k + 1 < n) {
                    const double2 pn = Ps[b0 + k + 1];
                    const double2 d1 = Ds[b0 + k + 1];
                    const double ex1 = wrapd(pn.x, px, D), ey1 = wrapd(pn.y, qy, D);
                    const double nx = ux + d1.x, ny = uy + d1.y;
                    const double nxX = xX + ex1, nyY = yY + ey1;
                    a += ux * ny - uy * nx;
                    ax += xX * ny - uy * nxX;
                    ay += ux * nyY - yY * nx;
                    ux = nx; uy = ny; xX = nxX; yY = nyY;
                    for (int i = k + 2; i < n; i++) {
                        const double2 di = Ds[b0 + i];
                        const double nx2 = ux + di.x, ny2 = uy + di.y;
                        const double nxX2 = xX + di.x, nyY2 = yY + di.y;
                        a += ux * ny2 - uy * nx2;
                        ax += xX * ny2 - uy * nxX2;
                        ay += ux * nyY2 - yY * nx2;
                        ux = nx2; uy = ny2; xX = nxX2; yY = nyY2;
                    }
                }
                const double z = 0;
                a += ux * z - uy * z;
                ax += xX * z - uy * z;
                ay += ux * z - yY * z;
                a *= 0.5;
                ax *= 0.5;
                ay *= 0.5;
                if (n <= 0) a = 0.0;
                const double a0 = fabs(a), se = use_[u];
                res = make_double2(se * (fabs(ax) - a0), se * (fabs(ay) - a0));
                if (k == 0) d.garea[ug[u]] = a;       // compute_grain_areas: one writer per grain
            }
        }
        vterm_s[tid] = res;
    }
    __syncthreads();

    // 5. one lane per vertex: line tension and the sum (calculus.h:57-61,149-150)
    if (!ven) return;
    const double x = pv.x, y = pv.y;
    const double px = dom_mod(x + SEVC_GRAD_EPS, D), pyy = dom_mod(y, D);
    const double qx = dom_mod(x, D), qy = dom_mod(y + SEVC_GRAD_EPS, D);
    if (!((pyy == y) && (qx == x))) {
        if (!tuning) push_vfall(d, v);
        return;
    }
    const bool uni = (P.gb_eps == 0.0);
    double bx = 0.0, by = 0.0, gx = 0.0, gy = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        // with pyy == y and qx == x the shifted lengths share the unshifted
        // length's other wrapped difference
        const double dx0 = wrapd(x, q[j].x, D), dy0 = wrapd(y, q[j].y, D);
        const double lx = mag(wrapd(px, q[j].x, D), dy0);
        const double ly = mag(dx0, wrapd(qy, q[j].y, D));
        const double L0 = mag(dx0, dy0);
        const double gam = uni ? P.gb_scaling : d.ben[vbj[j]];
        bx += (gam * (lx - L0));
        by += (gam * (ly - L0));
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const double2 t = vterm_s[j * VT + tid];
        gx += t.x;
        gy += t.y;
    }
    reinterpret_cast<double2*>(d.vel)[v] =
        make_double2((-(bx + gx) / SEVC_GRAD_EPS) * P.mobility, (-(by + gy) / SEVC_GRAD_EPS) * P.mobility);
}

template __global__ void k_vel_vtile<32>(Dev d, int tuning);
template __global__ void k_vel_vtile<64>(Dev d, int tuning);

// Analytic -grad E velocities (extension).
__global__ void k_velocity_analytic(Dev d) {
    TABLE_BLOCK(d.vblk, d.nvb)
    if (!M.active || P.vel_model != SEVC_VEL_ANALYTIC || i >= bd_.e || !d.ven[i]) return;
    double fx, fy;
    analytic_force(d, P, i, &fx, &fy);
    d.vel[2 * i] = fx * P.mobility;
    d.vel[2 * i + 1] = fy * P.mobility;
}

// compute_grain_dAdts, calculus.h:243-253 (diagnostic)
__global__ void k_grain_dadt(Dev d) {
    TABLE_BLOCK(d.gblk, d.ngb)
    if (!M.active || i >= bd_.e || !d.genab[i]) return;
    int n = d.glen[i];
    double r = 0.0;
    if (n > 0) {
        const int* L = glist(d, i);
        DadtAcc acc;
        int w = L[0];
        acc.start(d.pos[2 * w], d.pos[2 * w + 1], d.vel[2 * w], d.vel[2 * w + 1]);
        for (int k = 1; k < n; k++) {
            w = L[k];
            acc.push(d.pos[2 * w], d.pos[2 * w + 1], d.vel[2 * w], d.vel[2 * w + 1], P.D);
        }
        r = acc.finish();
    }
    d.gdadt[i] = r;
}

// update_vertices_positions, calculus.h:391-404
__device__ __forceinline__ void translate_block(const Dev& d, int lb) {
    const Blk bd_ = table_blk(d, d.vblk, lb);
    const Params& P = d.prm[bd_.m];
    const MState& M = d.ms[bd_.m];
    const int i = bd_.b + (int)threadIdx.x;
    // the vertex's data is loaded together with the member's state
    const bool in = i < bd_.e;
    const double2 p = in ? reinterpret_cast<const double2*>(d.pos)[i] : make_double2(0.0, 0.0);
    const double2 v = in ? reinterpret_cast<const double2*>(d.vel)[i] : make_double2(0.0, 0.0);
    const uint8_t en = in ? d.ven[i] : 0;
    if (!M.active || !en) return;
    const double dt = member_dt(P, M);
    double x = p.x, y = p.y;
    x += v.x * dt;
    y += v.y * dt;
    reinterpret_cast<double2*>(d.pos)[i] = make_double2(dom_mod(x, P.D), dom_mod(y, P.D));
}

__global__ void k_translate(Dev d) {
    const int lb = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (lb < d.nvb) translate_block(d, lb);
}

// k_translate and the clearing after the flip stage's fix-up (k_fix_clear,
// stage 2, no areas) in one launch: the first `nclear` workgroups (in
// dispatch order, so that their chains of loads start first) clear, eight
// listed grains per wave; the others translate table blocks [0, nvb).
// Disjoint data: positions, and fix flags and vertex edit records.
__global__ void k_translate_clear(Dev d, int nclear) {
    const int lb = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (lb >= nclear) {
        translate_block(d, lb - nclear);
        return;
    }
    // eight listed grains per wave, eight lanes each
    const int n = d.cnt[C_FIX0 + 2];
    const int lane = (int)threadIdx.x & 63, wpb = (int)(blockDim.x >> 6);
    const int sub = lane >> 3, k = lane & 7;
    for (int base = 8 * (lb * wpb + (int)(threadIdx.x >> 6)); base < n; base += 8 * nclear * wpb) {
        const int item = base + sub;
        if (item >= n) continue;
        const int g = d.l_fix[2][item];
        if (k == 0) d.gfix[g] = 0;
        if (!d.genab[g]) continue;
        const int* L = glist(d, g);
        const int m = d.glen[g];
        for (int i = k; i < m; i += 8) reset_mods(d, L[i]);
    }
}

// ---------------------------------------------------------------------
// topological.h: extinction times, labels, votes
// ---------------------------------------------------------------------

// compute_t_ext, topological.h:303-334, fused with the adaptive-dt minimum:
// the smallest positive t_ext of the member's enabled edges not shorter than
// the flip threshold, reduced per block and then into the member's value by
// one atomic minimum per block (skipped when the block's is not smaller).
// The minimum does not depend on the order, so the result equals a serial
// scan.
__global__ void k_text(Dev d) {
    __shared__ unsigned long long wmin[kBlock / 64];
    const int lb_ = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    if (lb_ >= d.neb) return;
    const Blk bd_ = table_blk(d, d.eblk, lb_);
    const Params& P = d.prm[bd_.m];
    MState& M = d.ms[bd_.m];
    const int i = bd_.b + (int)threadIdx.x;
    unsigned long long v = 0x7FF0000000000000ull;
    // the edge's data is loaded together with the member's state
    const bool in = i < bd_.e;
    const uint8_t en = in ? d.benab[i] : 0;
    const int2 ab = in ? reinterpret_cast<const int2*>(d.bvx)[i] : make_int2(0, 0);
    double2 pa = make_double2(0.0, 0.0), pb = pa, va = pa, vb = pa;
    if (en) {
        pa = reinterpret_cast<const double2*>(d.pos)[ab.x];
        pb = reinterpret_cast<const double2*>(d.pos)[ab.y];
        va = reinterpret_cast<const double2*>(d.vel)[ab.x];
        vb = reinterpret_cast<const double2*>(d.vel)[ab.y];
    }
    if (M.active && en) {
        const double D = P.D;
        d.bcand[i] = 0;
        // adjust_origin_for_points on (ini, end): U0 = 0, U1 = 0 + wrap(end - ini)
        double x1 = 0 + wrapd(pb.x, pa.x, D);
        double y1 = 0 + wrapd(pb.y, pa.y, D);
        double dx = x1 - 0, dy = y1 - 0;
        double norm = mag(dx, dy);
        // compute_boundary_arclengths (calculus.h:83-93): the same |wrap(end - ini)|
        d.blen[i] = norm;
        double tx, ty;
        if (norm == 0) { tx = 0; ty = 0; }
        else {
            double ang = sevc_atan2(dy, dx);
            sevc_sincos(ang, &ty, &tx);
        }
        double px = vb.x - va.x;
        double py = vb.y - va.y;
        double aux = tx * px + ty * py;
        double te = (aux == 0) ? 100.0 : -norm / aux;
        d.bt[i] = te;
        if (P.adaptive && te > 0 && !(norm < P.flip_thr)) v = (unsigned long long)__double_as_longlong(te);
    }
    v = wave_min_u64(v);
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = wmin[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = wmin[w] < r ? wmin[w] : r;
        // the member's minimum (k_begin set it to +inf): the smallest value
        // whatever the order, so the result equals a serial scan; a block
        // whose minimum is not below the value already there skips the atomic
        if (r < M.tau_bits) atomicMin(&M.tau_bits, r);
    }
}

// label_boundaries, topological.h:343-353 (every slot, quirk A-6), and the
// compaction of enabled candidates.
__global__ void k_label(Dev d) {
    __shared__ int wcnt[kBlock / 64], bbase;
    TABLE_BLOCK(d.eblk, d.neb)
    // the edge's data is loaded together with the member's state
    const bool in = i < bd_.e;
    const double te = in ? d.bt[i] : 0.0, len = in ? d.blen[i] : 0.0;
    const uint8_t en = in ? d.benab[i] : 0;
    if (!M.active || !P.events) return;             // the same for the whole workgroup
    const double dt = member_dt(P, M);
    const bool c = in && is_candidate(te, len, dt, P.flip_thr);
    if (in) {
        d.bcand[i] = c ? 1 : 0;
        d.blab[i] = (uint8_t)((c && en ? 1 : 0) | (en ? 2 : 0));  // for k_scan3: one byte per edge (bit 2: k_poll)
    }
    // the workgroup's candidates take consecutive list positions with one
    // atomic (a step with many candidates would otherwise queue one per wave
    // on the counter)
    const bool cand = c && en;
    const unsigned long long bal = __ballot(cand);
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    if (lane == 0) wcnt[wave] = (int)__popcll(bal);
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) {
        const int x = wcnt[w];
        before += w < wave ? x : 0;
        total += x;
    }
    if (threadIdx.x == 0 && total > 0) bbase = atomicAdd(&d.cnt[C_CAND], total);
    __syncthreads();
    if (cand) {
        const int k = bbase + before + (int)__popcll(bal & ((1ull << lane) - 1ull));
        d.l_cand[k] = i;
        d.conf[i] = k;          // list position, for k_poll_gather
    }
}

// The boundaries polling and the conflict check look at, for each listed
// candidate c at position k (nc <= kPollFast), mapped to list positions
// through conf (k_label wrote conf[c] = k; -1 elsewhere), as columns of
// d.pgat (column f at f * kPollFast):
//   0..5    the boundaries of c's ends a0 (slots 0-2) and a1 (3-5), vb order;
//   6..23   for slot j = 0..5 and r = 0..2 (column 6 + 3j + r): boundary r
//           of the far end of the slot's edge x, -1 when x is c, x is
//           disabled, or that boundary is x itself;
//   24      bits 0-1: slot 0 of end a0 / a1 is c itself; bits 2-7: slot j
//           holds an enabled edge other than c.
// d.pgt[k] = t_ext of c.  Many workgroups issue these gathers, so that the
// one-workgroup polling kernel reads contiguous columns only.
__global__ void k_poll_gather(Dev d) {
    const int nc = d.cnt[C_CAND];
    if (nc > d.poll_fast && nc > d.poll_lds) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        const int c = d.l_cand[k];
        const int a[2] = {d.bvx[2 * c], d.bvx[2 * c + 1]};
        d.pgt[k] = d.bt[c];
        int eb[6];
#pragma unroll
        for (int j = 0; j < 6; j++) eb[j] = d.vb[3 * a[j / 3] + j % 3];
        int bits = (eb[0] == c ? 1 : 0) | (eb[3] == c ? 2 : 0);
        int en[6], x0[6], x1[6];
#pragma unroll
        for (int j = 0; j < 6; j++) {
            d.pgat[j * kPollBig + k] = d.conf[eb[j]];
            en[j] = eb[j] != c && d.benab[eb[j]];
            x0[j] = d.bvx[2 * eb[j]];
            x1[j] = d.bvx[2 * eb[j] + 1];
        }
        int yw[6][3];
#pragma unroll
        for (int j = 0; j < 6; j++) {
            if (en[j]) bits |= 4 << j;
            const int W = en[j] ? (x0[j] == a[j / 3] ? x1[j] : x0[j]) : 0;
#pragma unroll
            for (int r = 0; r < 3; r++) yw[j][r] = en[j] ? d.vb[3 * W + r] : 0;
        }
#pragma unroll
        for (int j = 0; j < 6; j++)
#pragma unroll
            for (int r = 0; r < 3; r++)
                d.pgat[(6 + 3 * j + r) * kPollBig + k] = (en[j] && yw[j][r] != eb[j]) ? d.conf[yw[j][r]] : -1;
        d.pgat[24 * kPollBig + k] = bits;
    }
}

// polling_system (topological.h:549-578) as Jacobi rounds to a fixpoint,
// then check_conflictive_boundaries (:83-121) and
// unlabel_conflictive_boundaries (:129-155), over the compacted candidate
// list.  Only vertices that touch a candidate can vote for one, so the
// rounds are exact on the list.  One workgroup, in one of three ways by the
// list's length: up to 2,048 candidates with their mapped neighbours in
// registers; up to 16,384 with flags and t_ext in LDS and the mapped
// neighbours re-read from k_poll_gather's columns; longer lists by the
// general rounds over global memory.
__global__ void __launch_bounds__(1024) k_poll(Dev d) {
    __shared__ int changed;
    // LDS of the first two ways: [cf int, cid int, ct double] x kPollFast,
    // or [ct double, cf byte] x kPollBig
    __shared__ __align__(16) unsigned char pbuf[9 * kPollBig];
    static_assert(16 * kPollFast <= 9 * kPollBig, "the register way's arrays fit the buffer");
    const int nc = d.cnt[C_CAND];
    if (nc == 0) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    auto vote = [&](int v) {
        double min_t = 99999999;
        int arg = 0;
        for (int k = 0; k < 3; k++) {
            int b = d.vb[3 * v + k];
            if (d.bt[b] <= min_t && d.bcand[b]) { min_t = d.bt[b]; arg = k; }
        }
        return d.vb[3 * v + arg];
    };
    if (nc <= d.poll_fast && 2 * nt == kPollFast) {
        // Up to two candidates per lane (kPollFast = 2 x 1024).  Their list
        // positions, t_ext and flags sit in LDS; k_poll_gather has mapped
        // every boundary the rounds and the conflict check look at to its
        // list position (-1: not a listed candidate, whose flag cannot be
        // set since the ends of an enabled edge touch enabled edges only
        // and k_label lists every enabled candidate).  The rounds and the
        // check then read LDS only.
        constexpr int kPer = kPollFast / 1024;
        int* cf = reinterpret_cast<int*>(pbuf);
        int* cid = cf + kPollFast;
        double* ct = reinterpret_cast<double*>(pbuf + 8 * kPollFast);
        int ep[kPer][6], bits[kPer];
        bool two[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int k = tid + u * nt;
            const bool on = k < nc;
            if (on) {
                cid[k] = d.l_cand[k];
                ct[k] = d.pgt[k];
                cf[k] = 1;
            }
#pragma unroll
            for (int j = 0; j < 6; j++) ep[u][j] = on ? d.pgat[j * kPollBig + k] : -1;
            bits[u] = on ? d.pgat[24 * kPollBig + k] : 0;
        }
        __syncthreads();
        // vertices_vote (topological.h:362-386) at end e of the candidate at
        // position k: the end's flagged candidate with the smallest t_ext,
        // later slots winning ties (<=), or slot 0 if none; true when the
        // vote goes to the candidate itself
        auto votes_self = [&](int u, int e, int k) {
            double min_t = 99999999;
            int arg = -1;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const int p = ep[u][3 * e + r];
                if (p >= 0 && cf[p] && ct[p] <= min_t) { min_t = ct[p]; arg = p; }
            }
            return arg >= 0 ? arg == k : ((bits[u] >> e) & 1) != 0;
        };
        for (;;) {
            if (tid == 0) changed = 0;
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                const int k = tid + u * nt;
                two[u] = k < nc && cf[k] && votes_self(u, 0, k) && votes_self(u, 1, k);
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kPer; u++)
                if (two[u])
#pragma unroll
                    for (int j = 0; j < 6; j++) {
                        const int p = ep[u][j];
                        if (p >= 0 && p != tid + u * nt && cf[p]) { cf[p] = 0; changed = 1; }
                    }
            __syncthreads();
            int ch = changed;
            __syncthreads();
            if (!ch) break;
        }
        // check_conflictive_boundaries and unlabel_conflictive_boundaries
        // on the final flags, as one decision per flagged candidate c: c is
        // dropped when one of its enabled neighbours x without a flag has a
        // flagged candidate at each end (the first in slot order, bi at the
        // end shared with c, be at the far end) and c is not the winner of
        // the two (smaller t_ext, lower index on a tie; symmetric in bi,
        // be).  The reference writes that winner into x and lets c read it
        // back; here c forms it itself.
        auto flagp = [&](int p) { return p >= 0 && cf[p] != 0; };
        bool keep[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int k = tid + u * nt;
            keep[u] = k < nc && cf[k] != 0;
            int pw[6][3];
#pragma unroll
            for (int j = 0; j < 6; j++)
#pragma unroll
                for (int r = 0; r < 3; r++) pw[j][r] = keep[u] ? d.pgat[(6 + 3 * j + r) * kPollBig + k] : -1;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                if (!keep[u] || !((bits[u] >> (2 + j)) & 1) || flagp(ep[u][j])) continue;
                int pbi = -1, pbe = -1;
                const int e3 = 3 * (j / 3);
#pragma unroll
                for (int r = 2; r >= 0; r--) {           // first in slot order
                    if (flagp(ep[u][e3 + r])) pbi = ep[u][e3 + r];
                    if (flagp(pw[j][r])) pbe = pw[j][r];
                }
                if (pbi < 0 || pbe < 0) continue;
                const double ti = ct[pbi], tb = ct[pbe];
                const int bi = cid[pbi], be = cid[pbe];
                int w;
                if (ti == tb) w = bi < be ? bi : be;
                else w = ti < tb ? bi : be;
                if (w != cid[k]) keep[u] = false;
            }
        }
        __syncthreads();                 // every flag read
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int k = tid + u * nt;
            if (k < nc) {
                const int c = cid[k];
                d.bcand[c] = keep[u] ? 1 : 0;
                d.blab[c] = (uint8_t)(keep[u] ? 7 : 3);     // + the label after polling (bit 2)
                d.conf[c] = -1;
            }
        }
        return;
    } else if (nc <= d.poll_lds && nt * 32 >= kPollBig) {
        // The same rounds and check with the flags (bytes) and t_ext of up
        // to kPollBig candidates in LDS; each candidate's mapped neighbours
        // and bits are read again from k_poll_gather's columns (coalesced,
        // L2-resident) where needed, and the per-candidate results of a
        // lane are bits of one word (at most 16 candidates per lane).
        double* ct = reinterpret_cast<double*>(pbuf);
        uint8_t* cf = pbuf + 8 * kPollBig;
        const int* G = d.pgat;
        for (int k = tid; k < nc; k += nt) {
            ct[k] = d.pgt[k];
            cf[k] = 1;
        }
        __syncthreads();
        // vertices_vote at end e of the candidate at position k (as in the
        // register way above)
        auto votes_self = [&](int e, int k, int bk) {
            double min_t = 99999999;
            int arg = -1;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const int p = G[(3 * e + r) * kPollBig + k];
                if (p >= 0 && cf[p] && ct[p] <= min_t) { min_t = ct[p]; arg = p; }
            }
            return arg >= 0 ? arg == k : ((bk >> e) & 1) != 0;
        };
        for (;;) {
            if (tid == 0) changed = 0;
            unsigned two = 0;
            // four of the lane's candidates at a time: their mapped
            // neighbours (columns 0-5) and bits read together, then the
            // votes from LDS
            constexpr int kChunk = 4;
            for (int u0 = 0; tid + u0 * nt < nc; u0 += kChunk) {
                int bk[kChunk], pp[kChunk][6];
#pragma unroll
                for (int q = 0; q < kChunk; q++) {
                    const int k = tid + (u0 + q) * nt;
                    const bool on = k < nc;
                    bk[q] = on ? G[24 * kPollBig + k] : 0;
#pragma unroll
                    for (int c = 0; c < 6; c++) pp[q][c] = on ? G[c * kPollBig + k] : -1;
                }
#pragma unroll
                for (int q = 0; q < kChunk; q++) {
                    const int k = tid + (u0 + q) * nt;
                    if (k >= nc || !cf[k]) continue;
                    bool both = true;
#pragma unroll
                    for (int e = 0; e < 2; e++) {
                        double min_t = 99999999;
                        int arg = -1;
#pragma unroll
                        for (int r = 0; r < 3; r++) {
                            const int p = pp[q][3 * e + r];
                            if (p >= 0 && cf[p] && ct[p] <= min_t) { min_t = ct[p]; arg = p; }
                        }
                        both = both && (arg >= 0 ? arg == k : ((bk[q] >> e) & 1) != 0);
                    }
                    if (both) two |= 1u << (u0 + q);
                }
            }
            __syncthreads();
            for (int u = 0, k = tid; k < nc; u++, k += nt)
                if ((two >> u) & 1u)
#pragma unroll
                    for (int j = 0; j < 6; j++) {
                        const int p = G[j * kPollBig + k];
                        if (p >= 0 && p != k && cf[p]) { cf[p] = 0; changed = 1; }
                    }
            __syncthreads();
            const int ch = changed;
            __syncthreads();
            if (!ch) break;
        }
        auto flagp = [&](int p) { return p >= 0 && cf[p] != 0; };
        unsigned keep = 0;
        for (int u = 0, k = tid; k < nc; u++, k += nt) {
            bool kp = cf[k] != 0;
            if (kp) {
                // every column of the candidate in one round of loads
                const int bk = G[24 * kPollBig + k];
                int ep[6], fw[18];
#pragma unroll
                for (int j = 0; j < 6; j++) ep[j] = G[j * kPollBig + k];
#pragma unroll
                for (int x = 0; x < 18; x++) fw[x] = G[(6 + x) * kPollBig + k];
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    if (!kp || !((bk >> (2 + j)) & 1) || flagp(ep[j])) continue;
                    int pbi = -1, pbe = -1;
                    const int e3 = 3 * (j / 3);
#pragma unroll
                    for (int r = 2; r >= 0; r--) {           // first in slot order
                        if (flagp(ep[e3 + r])) pbi = ep[e3 + r];
                        const int pw = fw[3 * j + r];
                        if (flagp(pw)) pbe = pw;
                    }
                    if (pbi < 0 || pbe < 0) continue;
                    const double ti = ct[pbi], tb = ct[pbe];
                    const int bi = d.l_cand[pbi], be = d.l_cand[pbe];
                    int w;
                    if (ti == tb) w = bi < be ? bi : be;
                    else w = ti < tb ? bi : be;
                    if (w != d.l_cand[k]) kp = false;
                }
            }
            if (kp) keep |= 1u << u;
        }
        __syncthreads();                 // every flag read
        for (int u = 0, k = tid; k < nc; u++, k += nt) {
            const int c = d.l_cand[k];
            const bool kp = (keep >> u) & 1u;
            d.bcand[c] = kp ? 1 : 0;
            d.blab[c] = (uint8_t)(kp ? 7 : 3);
            d.conf[c] = -1;
        }
        return;
    } else for (;;) {
        if (tid == 0) changed = 0;
        // phase A: votes on the flags at the start of the round
        for (int k = tid; k < nc; k += nt) {
            int c = d.l_cand[k];
            int two = 0;
            if (d.bcand[c]) {
                int nvote = (vote(d.bvx[2 * c]) == c) + (vote(d.bvx[2 * c + 1]) == c);
                two = (nvote == 2);
            }
            d.conf[c] = two;     // scratch: 2-vote flag of this round
        }
        __syncthreads();
        // phase B: a 2-vote edge un-labels its four neighbours
        for (int k = tid; k < nc; k += nt) {
            int c = d.l_cand[k];
            if (d.conf[c] == 1) {
                for (int e = 0; e < 2; e++) {
                    int v = d.bvx[2 * c + e];
                    for (int q = 0; q < 3; q++) {
                        int x = d.vb[3 * v + q];
                        if (x != c && d.bcand[x]) { d.bcand[x] = 0; changed = 1; }
                    }
                }
            }
        }
        __syncthreads();
        int ch = changed;
        __syncthreads();
        if (!ch) break;
    }
    for (int k = tid; k < nc; k += nt) d.conf[d.l_cand[k]] = -1;
    __syncthreads();
    // check_conflictive_boundaries on the non-candidate neighbours
    for (int k = tid; k < nc; k += nt) {
        int c = d.l_cand[k];
        if (!d.bcand[c]) continue;
        for (int e = 0; e < 2; e++) {
            int v = d.bvx[2 * c + e];
            for (int q = 0; q < 3; q++) {
                int x = d.vb[3 * v + q];
                if (x == c || !d.benab[x] || d.bcand[x]) continue;
                int ini = d.bvx[2 * x], end = d.bvx[2 * x + 1];
                int bi = -1, be = -1;
                for (int r = 0; r < 3; r++) {
                    int y = d.vb[3 * ini + r];
                    if (y != x && d.bcand[y]) { bi = y; break; }
                }
                for (int r = 0; r < 3; r++) {
                    int y = d.vb[3 * end + r];
                    if (y != x && d.bcand[y]) { be = y; break; }
                }
                if (bi >= 0 && be >= 0) {
                    int w;
                    if (d.bt[bi] == d.bt[be]) w = bi < be ? bi : be;
                    else w = d.bt[bi] < d.bt[be] ? bi : be;
                    d.conf[x] = w;
                    int slot = atomicAdd(&d.cnt[C_CONF], 1);
                    d.l_conf[slot] = x;
                }
            }
        }
    }
    __syncthreads();
    // unlabel_conflictive_boundaries: reads conf only, writes its own flag
    for (int k = tid; k < nc; k += nt) {
        int c = d.l_cand[k];
        if (!d.bcand[c]) continue;
        bool drop = false;
        for (int e = 0; e < 2 && !drop; e++) {
            int v = d.bvx[2 * c + e];
            for (int q = 0; q < 3; q++) {
                int x = d.vb[3 * v + q];
                if (x != c && d.conf[x] != -1 && d.conf[x] != c) { drop = true; break; }
            }
        }
        if (drop) d.bcand[c] = 0;
    }
    __syncthreads();
    const int ncf = d.cnt[C_CONF];
    for (int k = tid; k < ncf; k += nt) d.conf[d.l_conf[k]] = -1;
    // the label after polling of every listed candidate, bit 2 of its byte
    for (int k = tid; k < nc; k += nt) {
        const int c = d.l_cand[k];
        d.blab[c] = (uint8_t)(d.bcand[c] ? 7 : 3);
    }
}

// ---------------------------------------------------------------------
// deferred grain-list edits (topological.h:1035-1163)
// ---------------------------------------------------------------------

// grain_add_vertex(grn, vrt, vrt_prev), geometry.h:629-645
__device__ __forceinline__ void list_add_after(const Dev& d, int g, int v, int prev) {
    int* L = glist(d, g);
    int n = d.glen[g], pos = -1;
    for (int i = 0; i < n; i++) if (L[i] == prev) { pos = i + 1; break; }
    if (pos < 0) { set_error(d, member_of_grain(d, g), SEVC_E_STATE); return; }
    if (n >= d.prm[member_of_grain(d, g)].cap) { set_error(d, member_of_grain(d, g), SEVC_E_GRAIN_FULL); return; }
    for (int i = n; i > pos; i--) L[i] = L[i - 1];
    L[pos] = v;
    d.glen[g] = n + 1;
}

// grain_remove_vertex, geometry.h:653-666
__device__ __forceinline__ void list_remove(const Dev& d, int g, int v) {
    int* L = glist(d, g);
    int n = d.glen[g];
    for (int i = 0; i < n; i++)
        if (L[i] == v) {
            for (int j = i; j < n - 1; j++) L[j] = L[j + 1];
            d.glen[g] = n - 1;
            L[n - 1] = -1;
            return;
        }
}

// grain_replace_vertex, geometry.h:686-694
__device__ __forceinline__ void list_replace(const Dev& d, int g, int old_v, int new_v) {
    int* L = glist(d, g);
    int n = d.glen[g];
    for (int i = 0; i < n; i++) if (L[i] == old_v) { L[i] = new_v; return; }
}

// perform_ADD_ops, perform_REMOVE_ops, perform_REPLACE_ops for one grain.
__device__ void fix_one_grain(const Dev& d, int g) {
    if (!d.genab[g]) return;
    int* L = glist(d, g);
    for (int i = 0; i < d.glen[g]; i++) {                       // :1040-1067
        int v = L[i];
        for (int j = 0; j < 4; j++)
            if (d.mods[v].fix[j] == FIX_ADD && d.mods[v].grn[j] == g)
                list_add_after(d, g, d.mods[v].vrt[j], d.mods[v].prev[j]);
    }
    {                                                           // :1086-1106
        // The reference gathers the REMOVE commands of the list's vertices
        // and deletes each named vertex once; deleting a set of distinct
        // entries leaves the same list in any order, so mark the doomed
        // positions first and compact once.
        const int n = d.glen[g];
        uint64_t doomed[kMaxStride / 64] = {0, 0};
        for (int i = 0; i < n; i++) {
            int v = L[i];
            for (int j = 0; j < 4; j++) {
                if (d.mods[v].fix[j] != FIX_REMOVE || d.mods[v].grn[j] != g) continue;
                int w = d.mods[v].vrt[j];
                for (int q = 0; q < n; q++)
                    if (L[q] == w) { doomed[q >> 6] |= 1ull << (q & 63); break; }
            }
        }
        int w = 0;
        for (int i = 0; i < n; i++)
            if (!((doomed[i >> 6] >> (i & 63)) & 1ull)) L[w++] = L[i];
        for (int i = w; i < n; i++) L[i] = -1;
        d.glen[g] = w;
    }
    for (int i = 0; i < d.glen[g]; i++) {                       // :1124-1136
        int v = L[i];
        for (int j = 0; j < 4; j++)
            if (d.mods[v].fix[j] == FIX_REPLACE && d.mods[v].grn[j] == g)
                list_replace(d, g, d.mods[v].prev[j], d.mods[v].vrt[j]);
    }
}

// fix_grain_lists for the flagged grains of a stage, one wave (workgroup of
// 64) per grain.  The wave stages the grain's list and the four deferred-edit
// records of every listed vertex in LDS with one round of independent loads;
// lane 0 then replays perform_ADD_ops, perform_REMOVE_ops and
// perform_REPLACE_ops (topological.h:1035-1136) on the staged copy in exactly
// the order fix_one_grain uses on global memory (an entry inserted or
// replaced during the replay reads its vertex's records from global memory,
// as fix_one_grain does), and the wave writes the list back, including the
// -1 padding fix_one_grain leaves after the removals.  Records do not change
// during this kernel, and each grain edits only its own list.
__global__ void __launch_bounds__(64) k_fix(Dev d, int stage) {
    __shared__ int Ls[kMaxStride];
    __shared__ short src[kMaxStride];            // index of the staged records, -1: read globally
    __shared__ unsigned char mf[4 * kMaxStride];
    __shared__ int mg[4 * kMaxStride], mv[4 * kMaxStride], mp[4 * kMaxStride];
    __shared__ unsigned short kinds[kMaxStride];   // per staged entry: 4 bits each for ADD, REMOVE, REPLACE records of this grain
    __shared__ int n_add_s, n_fin_s;
    const int n_items = d.cnt[C_FIX0 + stage];
    const int lane = threadIdx.x;
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const int g = d.l_fix[stage][item];
        // the grain's flag, length and whole list row in one round of loads
        int* L = glist(d, g);
        const uint8_t en = d.genab[g];
        const int n0 = d.glen[g];
        int row[kMaxStride / 64];
#pragma unroll
        for (int q = 0; q < kMaxStride / 64; q++) row[q] = lane + 64 * q < d.S ? L[lane + 64 * q] : -1;
        if (!en) continue;
        const int cap = d.prm[member_of_grain(d, g)].cap;
        for (int i = lane; i < n0; i += 64) {
            const int v = row[i >> 6];
            Ls[i] = v;
            src[i] = (short)i;
            unsigned kd = 0;
            for (int j = 0; j < 4; j++) {
                const int f = d.mods[v].fix[j], gg = d.mods[v].grn[j];
                mf[4 * i + j] = (unsigned char)f;
                mg[4 * i + j] = gg;
                mv[4 * i + j] = d.mods[v].vrt[j];
                mp[4 * i + j] = d.mods[v].prev[j];
                if (gg == g) {
                    if (f == FIX_ADD) kd |= 1u << j;
                    else if (f == FIX_REMOVE) kd |= 1u << (4 + j);
                    else if (f == FIX_REPLACE) kd |= 1u << (8 + j);
                }
            }
            kinds[i] = (unsigned short)kd;
        }
        __syncthreads();
        if (lane == 0) {
            auto rec = [&](int es, int v, int j, int& f, int& gg, int& vv, int& pp) {
                if (es >= 0) {
                    const int q = 4 * es + j;
                    f = mf[q]; gg = mg[q]; vv = mv[q]; pp = mp[q];
                } else {
                    f = d.mods[v].fix[j]; gg = d.mods[v].grn[j]; vv = d.mods[v].vrt[j]; pp = d.mods[v].prev[j];
                }
            };
            int n = n0;
            // a staged entry's kind bits name the records that can apply; an
            // entry inserted or replaced during the replay checks all four
            for (int i = 0; i < n; i++) {                            // ADD, :1040-1067
                const int es = src[i], v = Ls[i];
                const unsigned km = es >= 0 ? (kinds[es] & 0xFu) : 0xFu;
                for (int j = 0; j < 4; j++) {
                    if (!((km >> j) & 1u)) continue;
                    int f, gg, vv, pp;
                    rec(es, v, j, f, gg, vv, pp);
                    if (f != FIX_ADD || gg != g) continue;
                    int pos = -1;                                    // grain_add_vertex, geometry.h:629-645
                    for (int q = 0; q < n; q++) if (Ls[q] == pp) { pos = q + 1; break; }
                    if (pos < 0) { set_error(d, member_of_grain(d, g), SEVC_E_STATE); continue; }
                    if (n >= cap) { set_error(d, member_of_grain(d, g), SEVC_E_GRAIN_FULL); continue; }
                    for (int q = n; q > pos; q--) { Ls[q] = Ls[q - 1]; src[q] = src[q - 1]; }
                    Ls[pos] = vv;
                    src[pos] = -1;
                    n++;
                }
            }
            const int n_add = n;
            uint64_t doomed[kMaxStride / 64] = {0, 0};               // REMOVE, :1086-1106
            for (int i = 0; i < n; i++) {
                const int es = src[i], v = Ls[i];
                const unsigned km = es >= 0 ? ((kinds[es] >> 4) & 0xFu) : 0xFu;
                for (int j = 0; j < 4; j++) {
                    if (!((km >> j) & 1u)) continue;
                    int f, gg, vv, pp;
                    rec(es, v, j, f, gg, vv, pp);
                    if (f != FIX_REMOVE || gg != g) continue;
                    for (int q = 0; q < n; q++)
                        if (Ls[q] == vv) { doomed[q >> 6] |= 1ull << (q & 63); break; }
                }
            }
            int w = 0;
            for (int i = 0; i < n; i++)
                if (!((doomed[i >> 6] >> (i & 63)) & 1ull)) { Ls[w] = Ls[i]; src[w] = src[i]; w++; }
            n = w;
            for (int i = 0; i < n; i++) {                            // REPLACE, :1124-1136
                const int es = src[i], v = Ls[i];
                const unsigned km = es >= 0 ? ((kinds[es] >> 8) & 0xFu) : 0xFu;
                for (int j = 0; j < 4; j++) {
                    if (!((km >> j) & 1u)) continue;
                    int f, gg, vv, pp;
                    rec(es, v, j, f, gg, vv, pp);
                    if (f != FIX_REPLACE || gg != g) continue;
                    for (int q = 0; q < n; q++)                      // grain_replace_vertex(pp -> vv)
                        if (Ls[q] == pp) { Ls[q] = vv; src[q] = -1; break; }
                }
            }
            n_add_s = n_add;
            n_fin_s = n;
        }
        __syncthreads();
        const int n_add = n_add_s, n_fin = n_fin_s;
        for (int i = lane; i < n_add; i += 64) L[i] = i < n_fin ? Ls[i] : -1;
        if (lane == 0) d.glen[g] = n_fin;
        __syncthreads();
    }
}

// clear_grains_fix_state (topological.h:1150-1163) for the flagged grains;
// with with_area also the area update of main.cu:266 for them and for the
// grains edited by the two-triangle stage (its dirty list).  One wave
// (workgroup of 64) per grain: the lanes reset the listed vertices' records
// and load the list's points in parallel, each forming the wrapped delta to
// its predecessor; lane 0 then accumulates the shoelace sum over the deltas
// in list order, the same arithmetic as grain_area.
// Items of k_fix_clear's work for one 64-lane workgroup: workgroup `blk` of
// `nblk` takes items blk, blk + nblk, ...
__device__ __forceinline__ void fix_clear_items(const Dev& d, int stage, int with_area, int blk, int nblk) {
    __shared__ double sdx[kMaxStride], sdy[kMaxStride];
    const int n = d.cnt[C_FIX0 + stage];
    const int nd = with_area ? min(d.cnt[C_DIRTY], d.G) : 0;
    const int lane = threadIdx.x;
    for (int item = blk; item < n + nd; item += nblk) {
        int g;
        if (item < n) {
            g = d.l_fix[stage][item];
            if (lane == 0) d.gfix[g] = 0;
            if (!d.genab[g]) continue;
            const int* L = glist(d, g);
            const int m = d.glen[g];
            for (int i = lane; i < m; i += 64) reset_mods(d, L[i]);
            if (!with_area) continue;
        } else {
            g = d.l_dirty[item - n];
            if (!d.genab[g]) continue;
        }
        const int m = d.glen[g];
        const int* L = glist(d, g);
        const double D = d.prm[member_of_grain(d, g)].D;
        for (int i = 1 + lane; i < m; i += 64) {
            const int w = L[i], wp = L[i - 1];
            sdx[i] = wrapd(d.pos[2 * w], d.pos[2 * wp], D);
            sdy[i] = wrapd(d.pos[2 * w + 1], d.pos[2 * wp + 1], D);
        }
        __syncthreads();
        if (lane == 0) {
            double a = 0.0;
            if (m > 0) {
                AreaAcc acc;
                acc.start(0.0, 0.0);
                for (int i = 1; i < m; i++) acc.push_delta(sdx[i], sdy[i]);
                a = acc.finish();
            }
            d.garea[g] = a;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(64) k_fix_clear(Dev d, int stage, int with_area) {
    fix_clear_items(d, stage, with_area, (int)blockIdx.x, (int)gridDim.x);
}

// ---------------------------------------------------------------------
// two adjacent 3-sided grains (topological.h:22-73, 840-985)
// ---------------------------------------------------------------------

__device__ bool two_qualifies(const Dev& d, int e) {
    if (!d.benab[e]) return false;
    int gic[2];
    if (grains_intersect(d, d.bvx[2 * e], d.bvx[2 * e + 1], gic) < 2) return false;
    return d.glen[gic[0]] == 3 && d.glen[gic[1]] == 3;
}

// remove_two_grains, topological.h:840-955.  Returns CE (the edge whose
// endpoint moved) or -1.
__device__ int remove_two_grains(const Dev& d, int bnd, const int gnic[2]) {
    int A = d.bvx[2 * bnd], B = d.bvx[2 * bnd + 1];
    int AC = -1, BD = -1, CE = -1, DF = -1;
    for (int i = 0; i < 3; i++) if (d.vb[3 * A + i] == bnd) { AC = d.vb[3 * A + (i + 1) % 3]; break; }
    for (int j = 0; j < 3; j++) if (d.vb[3 * B + j] == bnd) { BD = d.vb[3 * B + (j + 1) % 3]; break; }
    if (AC < 0 || BD < 0) { set_error(d, member_of_edge(d, bnd), SEVC_E_STATE); return -1; }
    int C = other_end(d, AC, A);
    int Dv = other_end(d, BD, B);
    for (int k = 0; k < 3; k++) if (d.vb[3 * C + k] == AC) { CE = d.vb[3 * C + (k + 2) % 3]; break; }
    for (int l = 0; l < 3; l++) if (d.vb[3 * Dv + l] == BD) { DF = d.vb[3 * Dv + (l + 2) % 3]; break; }
    if (CE < 0 || DF < 0) { set_error(d, member_of_edge(d, bnd), SEVC_E_STATE); return -1; }
    int F = other_end(d, DF, Dv);
    if (d.bvx[2 * CE] == C) d.bvx[2 * CE] = F;
    else d.bvx[2 * CE + 1] = F;
    for (int p = 0; p < 3; p++) if (d.vb[3 * F + p] == DF) { d.vb[3 * F + p] = CE; break; }
    int G3, G4;
    if (grain_contains_vertex(d, gnic[0], A) >= 0) { G3 = gnic[0]; G4 = gnic[1]; }
    else { G3 = gnic[1]; G4 = gnic[0]; }
    set_mod(d, A, 0, FIX_REMOVE, G3, A); flag_grain(d, G3, 0);
    set_mod(d, A, 1, FIX_REMOVE, G3, C);
    set_mod(d, A, 2, FIX_REMOVE, G3, Dv);
    set_mod(d, B, 0, FIX_REMOVE, G4, B); flag_grain(d, G4, 0);
    set_mod(d, B, 1, FIX_REMOVE, G4, C);
    set_mod(d, B, 2, FIX_REMOVE, G4, Dv);
    for (int i = 0; i < 3; i++) {
        set_edge_enabled(d, d.vb[3 * A + i], false);
        set_edge_enabled(d, d.vb[3 * B + i], false);
    }
    set_edge_enabled(d, DF, false);
    set_edge_enabled(d, bnd, false);
    set_vertex_enabled(d, A, false); set_vertex_enabled(d, Dv, false);
    set_vertex_enabled(d, C, false); set_vertex_enabled(d, B, false);
    return CE;
}

__device__ void sort_ints(int* a, int n) {
    for (int i = 1; i < n; i++) {
        int x = a[i], j = i - 1;
        while (j >= 0 && a[j] > x) { a[j + 1] = a[j]; j--; }
        a[j + 1] = x;
    }
}


// The loop of main.cu:227-250 on one lane: serial passes in edge order over
// the qualifying edges (re-checked when reached, as the <<<1,1>>> scan of
// topological.h:965-985 sees them), each followed by the list fix-up of the
// flagged grains.  The next pass's qualifying edges can only be edges of
// grains whose list was just edited.
__global__ void k_two_grain(Dev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int n = d.cnt[C_TWO];
    if (n == 0) return;
    int* lst = d.l_two;
    for (int it = 0; it < 20 && n > 0; it++) {
        sort_ints(lst, n);
        for (int k = 0; k < n; k++) {
            int e = lst[k];
            if (!two_qualifies(d, e)) continue;
            if (!d.ms[member_of_edge(d, e)].active) continue;
            int gic[2], gnic[2];
            grains_intersect(d, d.bvx[2 * e], d.bvx[2 * e + 1], gic);
            grains_symdiff(d, d.bvx[2 * e], d.bvx[2 * e + 1], gnic);
            if (gnic[0] < 0 || gnic[1] < 0) { set_error(d, member_of_edge(d, e), SEVC_E_STATE); continue; }
            int CE = remove_two_grains(d, e, gnic);
            set_grain_enabled(d, gic[0], false);
            set_grain_enabled(d, gic[1], false);
            d.ms[member_of_edge(d, e)].two += 1;
            // the T2 labels k_scan3 took before this stage are redone
            d.cnt[C_RELABEL] = 1;
            d.cnt[C_T2] = 0;
            // an edge whose endpoint moved may qualify later in this pass
            if (CE > e && two_qualifies(d, CE)) {
                bool have = false;
                for (int q = k + 1; q < n; q++) if (lst[q] == CE) have = true;
                if (!have) {
                    int p = n++;
                    while (p > k + 1 && lst[p - 1] > CE) { lst[p] = lst[p - 1]; p--; }
                    lst[p] = CE;
                }
            }
        }
        // fix_grain_lists for the flagged grains
        const int nf = d.cnt[C_FIX0];
        for (int q = 0; q < nf; q++) fix_one_grain(d, d.l_fix[0][q]);
        int nn = 0;
        for (int q = 0; q < nf; q++) {
            int g = d.l_fix[0][q];
            d.gfix[g] = 0;
            int s = atomicAdd(&d.cnt[C_DIRTY], 1);
            if (s < d.G) d.l_dirty[s] = g;
            else set_error(d, member_of_grain(d, g), SEVC_E_STATE);
            if (!d.genab[g]) continue;
            const int* L = glist(d, g);
            for (int i = 0; i < d.glen[g]; i++) reset_mods(d, L[i]);
        }
        // candidates of the next pass: the edited grains' own edges
        for (int q = 0; q < nf; q++) {
            int g = d.l_fix[0][q];
            if (!d.genab[g]) continue;
            const int* L = glist(d, g);
            int len = d.glen[g];
            for (int k = 0; k < len; k++) {
                int a = L[k], c = L[(k + 1) % len];
                for (int r = 0; r < 3; r++) {
                    int e = d.vb[3 * a + r];
                    if (other_end(d, e, a) != c) continue;
                    if (!two_qualifies(d, e)) continue;
                    bool have = false;
                    for (int z = 0; z < nn; z++) if (lst[z] == e) have = true;
                    if (!have) lst[nn++] = e;
                }
            }
        }
        d.cnt[C_FIX0] = 0;
        n = nn;
    }
}

// Two-triangle candidates and T2 labels in one pass over the grains,
// sharing the loads of a 3-sided grain's list, its vertices' edges and
// grains and those edges' ends and flags.  Two-triangle candidates
// (topological.h:22-73): edges whose two grains are both 3-sided, found
// from the 3-sided grains; for edge (a, c) of 3-sided grain g the two grains
// are vg(a) ∩ vg(c), the set grains_intersect forms from the edge's ends,
// and the edge is listed once, by the lower of the two grains.  The T2
// labels (as k_label3) are those before the two-triangle stage; k_label3
// redoes them if that stage changes anything.
__global__ void k_scan3(Dev d) {
    __shared__ int q_s[kBlock];       // the block's enabled 3-sided grains
    __shared__ int nq;
    TABLE_BLOCK(d.gblk, d.ngb)
    // the grain's data is loaded together with the member's state
    const bool in = i < bd_.e;
    const uint8_t en = in ? d.genab[i] : 0;
    const int len = in ? d.glen[i] : 0;
    if (threadIdx.x == 0) nq = 0;
    __syncthreads();
    const bool run = M.active && P.events;          // the same for the whole workgroup
    if (run && en) {
        if (len == 3) q_s[atomicAdd(&nq, 1)] = i;
        else d.gcand[i] = 0;
    }
    __syncthreads();
    // the 3-sided grains go to the first lanes, so that one wave carries
    // their chains of loads and the others finish at once
    if (!run || (int)threadIdx.x >= nq) return;
    const int g = q_s[threadIdx.x];
    uint8_t cand = 0;
    {
        const int* L = glist(d, g);
        const int v0 = L[0], v1 = L[1], v2 = L[2];
        const int vv[3] = {v0, v1, v2};
        int bb[9], gv[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            bb[e] = d.vb[3 * vv[e / 3] + e % 3];
            gv[e] = d.vg[3 * vv[e / 3] + e % 3];
        }
        // the edges' ends and the byte k_label and k_poll left: bit 0 the
        // label before polling (is_candidate at this step's dt) of an enabled
        // edge, bit 1 enabled, bit 2 the label after polling (an enabled
        // edge's current label is set only if it was listed and kept); two
        // loads per edge
        int ini[9], end[9];
        uint8_t bc[9], lab[9], be[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const int b = bb[e];
            ini[e] = d.bvx[2 * b];
            end[e] = d.bvx[2 * b + 1];
            lab[e] = d.blab[b];
        }
#pragma unroll
        for (int e = 0; e < 9; e++) {
            be[e] = (lab[e] >> 1) & 1;
            bc[e] = (lab[e] >> 2) & 1;
        }
        // label_3sided_grains (as k_label3)
        auto in_g = [&](int x) { return x == v0 || x == v1 || x == v2; };
        bool hit = false;
        int ans = 0;
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const bool inn = in_g(ini[e]) && in_g(end[e]);
            hit = hit || (inn && bc[e] && be[e]);
            ans += (inn && (lab[e] & 1)) ? 1 : 0;
        }
        ans /= 2;
        if (hit || ans == 3) cand = 1;
        // two-triangle candidates: edge (a, c) of the grain with both of its
        // grains 3-sided
        for (int k = 0; k < 3; k++) {
            const int kc = (k + 1) % 3;
            const int c = vv[kc];
            for (int q = 0; q < 3; q++) {
                const int e = 3 * k + q;
                const int oe = (ini[e] == vv[k]) ? end[e] : ini[e];
                if (oe != c || !be[e]) continue;
                int gic[2] = {-1, -1}, cnt = 0;
                for (int x = 0; x < 3 && cnt < 2; x++)
                    for (int y = 0; y < 3; y++)
                        if (gv[3 * k + x] == gv[3 * kc + y]) { gic[cnt++] = gv[3 * k + x]; break; }
                if (cnt < 2) continue;
                if (d.glen[gic[0]] == 3 && d.glen[gic[1]] == 3) {
                    const int lo = gic[0] < gic[1] ? gic[0] : gic[1];
                    if (lo == g) {
                        const int s = atomicAdd(&d.cnt[C_TWO], 1);     // rare: two adjacent 3-sided grains
                        d.l_two[s] = bb[e];
                    }
                }
            }
        }
    }
    d.gcand[g] = cand;
    const int s = wave_slot(&d.cnt[C_T2], cand != 0);
    if (cand) d.l_t2[s] = g;
}

// ---------------------------------------------------------------------
// T2: 3-sided grain removal (topological.h:614-668, 687-822, 994-1004)
// ---------------------------------------------------------------------

// clear_grains_state + label_3sided_grains; candidates are compacted.
// k_scan3 takes these labels before the two-triangle stage; this kernel
// redoes them when that stage removed a pair of grains (a step without such
// an event leaves every datum the labels read unchanged).
__global__ void k_label3(Dev d) {
    if (d.cnt[C_RELABEL] == 0) return;
    TABLE_BLOCK(d.gblk, d.ngb)
    // the grain's data is loaded together with the member's state
    const bool in = i < bd_.e;
    const uint8_t en = in ? d.genab[i] : 0;
    const int len = in ? d.glen[i] : 0;
    if (!M.active || !P.events || !en) return;
    const int g = i;
    uint8_t cand = 0;
    if (len == 3) {
        // every datum of the 9 (vertex, edge) visits is loaded up front (three
        // rounds of independent loads), then the visits run in the order of
        // label_3sided_grains; "in g" is membership in the 3-entry list
        // (grain_contains_vertex)
        const double dt = member_dt(P, M);
        const int* L = glist(d, g);
        const int v0 = L[0], v1 = L[1], v2 = L[2];
        const int vv[3] = {v0, v1, v2};
        int bb[9];
#pragma unroll
        for (int e = 0; e < 9; e++) bb[e] = d.vb[3 * vv[e / 3] + e % 3];
        int ini[9], end[9];
        uint8_t bc[9], be[9];
        double te[9], bl[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const int b = bb[e];
            ini[e] = d.bvx[2 * b];
            end[e] = d.bvx[2 * b + 1];
            bc[e] = d.bcand[b];
            be[e] = d.benab[b];
            te[e] = d.bt[b];
            bl[e] = d.blen[b];
        }
        auto in_g = [&](int x) { return x == v0 || x == v1 || x == v2; };
        // label_3sided_grains stops counting a vertex's edges at the first
        // labelled one, but then the grain is a candidate whatever the count,
        // so counting every visit gives the same label
        bool hit = false;
        int ans = 0;
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const bool inn = in_g(ini[e]) && in_g(end[e]);
            hit = hit || (inn && bc[e] && be[e]);
            ans += (inn && is_candidate(te[e], bl[e], dt, P.flip_thr) && be[e]) ? 1 : 0;
        }
        ans /= 2;
        if (hit || ans == 3) cand = 1;
    }
    d.gcand[g] = cand;
    const int s = wave_slot(&d.cnt[C_T2], cand != 0);
    if (cand) d.l_t2[s] = g;
}

// remove_grain, topological.h:687-822
// grains_intersect on vertex-grain triples already in registers.
__device__ __forceinline__ int rg_intersect(const int* ga, const int* gb, int out[2]) {
    int k = 0;
    out[0] = out[1] = -1;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (ga[i] == gb[j]) {
                out[k++] = ga[i];
                if (k == 2) return 2;
            }
    return k;
}

// remove_3sided_grain.  Everything the removal reads is loaded first (the
// grain's vertices; their edges, grains and positions; those edges' ends),
// then the reference's sequence of tests and writes runs on local copies
// in the same order -- a later test sees an earlier write exactly as it
// would in memory -- and the writes go out.
__device__ void remove_grain(const Dev& d, int g, double D) {
    const int* Lg = glist(d, g);
    const int A = Lg[0], B = Lg[1], C = Lg[2];
    int vA[3], vB[3], vC[3], gA[3], gB[3], gC[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        vA[i] = d.vb[3 * A + i]; vB[i] = d.vb[3 * B + i]; vC[i] = d.vb[3 * C + i];
        gA[i] = d.vg[3 * A + i]; gB[i] = d.vg[3 * B + i]; gC[i] = d.vg[3 * C + i];
    }
    const double xp0 = d.pos[2 * A], xp1 = d.pos[2 * B], xp2 = d.pos[2 * C];
    const double yp0 = d.pos[2 * A + 1], yp1 = d.pos[2 * B + 1], yp2 = d.pos[2 * C + 1];
    int xA[3][2], xB[3][2], xC[3][2];              // ends of the vertices' edges
#pragma unroll
    for (int i = 0; i < 3; i++) {
        xA[i][0] = d.bvx[2 * vA[i]]; xA[i][1] = d.bvx[2 * vA[i] + 1];
        xB[i][0] = d.bvx[2 * vB[i]]; xB[i][1] = d.bvx[2 * vB[i] + 1];
        xC[i][0] = d.bvx[2 * vC[i]]; xC[i][1] = d.bvx[2 * vC[i] + 1];
    }
    double nAB = arclen(xp0, yp0, xp1, yp1, D);
    double nBC = arclen(xp1, yp1, xp2, yp2, D);
    double nCA = arclen(xp2, yp2, xp0, yp0, D);
    double denom = nAB + nBC + nCA;
    // adjust_origin_for_points on (A, B, C)
    double ux1 = 0 + wrapd(xp1, xp0, D), uy1 = 0 + wrapd(yp1, yp0, D);
    double ux2 = ux1 + wrapd(xp2, xp1, D), uy2 = uy1 + wrapd(yp2, yp1, D);
    double ux0 = 0, uy0 = 0;
    double ax = xp0 + (nBC * ux0 + nCA * ux1 + nAB * ux2) / denom;
    double ay = yp0 + (nBC * uy0 + nCA * uy1 + nAB * uy2) / denom;
    d.pos[2 * A] = dom_mod(ax, D);
    d.pos[2 * A + 1] = dom_mod(ay, D);
    int gic[2], G1, G2, G3;
    rg_intersect(gB, gC, gic); G3 = (gic[0] == g) ? gic[1] : gic[0];
    rg_intersect(gA, gB, gic); G2 = (gic[0] == g) ? gic[1] : gic[0];
    rg_intersect(gA, gC, gic); G1 = (gic[0] == g) ? gic[1] : gic[0];
    if (G1 < 0 || G2 < 0 || G3 < 0) { set_error(d, member_of_grain(d, g), SEVC_E_STATE); return; }
    d.vel[2 * A] = 0.0;
    d.vel[2 * A + 1] = 0.0;
    // The searches below take the first match in index order, written as
    // unrolled loops (constant indices keep the small arrays in registers).
    bool found = false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int ei = xC[i][0], ee = xC[i][1];
        if (!found && ((ei == B && ee == C) || (ee == B && ei == C))) { set_edge_enabled(d, vC[i], false); found = true; }
    }
    int BEb = -1, CF = -1, be0 = -1, be1 = -1, cf0 = -1, cf1 = -1;
#pragma unroll
    for (int i = 2; i >= 0; i--) {
        const int ei = xB[i][0], ee = xB[i][1];
        if (!(ei == A || ee == A) && !(ei == C || ee == C)) { BEb = vB[i]; be0 = ei; be1 = ee; }
    }
#pragma unroll
    for (int i = 2; i >= 0; i--) {
        const int ei = xC[i][0], ee = xC[i][1];
        if (!(ei == A || ee == A) && !(ei == B || ee == B)) { CF = vC[i]; cf0 = ei; cf1 = ee; }
    }
    if (BEb < 0 || CF < 0) { set_error(d, member_of_grain(d, g), SEVC_E_STATE); return; }
    // the ends of BEb and CF after the rewiring (CF's as seen after BEb's,
    // should they be the same edge).  Only the end that changes is written:
    // the other end of BEb or CF can be another removal's vertex, which that
    // removal rewires in the same launch.
    if (be0 == B) { be0 = A; d.bvx[2 * BEb] = A; }
    else { be1 = A; d.bvx[2 * BEb + 1] = A; }
    if (CF == BEb) { cf0 = be0; cf1 = be1; }
    if (cf0 == C) { cf0 = A; d.bvx[2 * CF] = A; }
    else { cf1 = A; d.bvx[2 * CF + 1] = A; }
    if (CF == BEb) { be0 = cf0; be1 = cf1; }
    // current ends of edge e (one of A's), after the two rewirings
    auto ends = [&](int e, int i, int& e0, int& e1) {
        if (e == CF) { e0 = cf0; e1 = cf1; }
        else if (e == BEb) { e0 = be0; e1 = be1; }
        else { e0 = xA[i][0]; e1 = xA[i][1]; }
    };
    int curA[3] = {vA[0], vA[1], vA[2]};
    found = false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int e0, e1;
        ends(curA[i], i, e0, e1);
        if (!found && (e0 == B || e1 == B)) {
            set_edge_enabled(d, curA[i], false); curA[i] = BEb; d.vb[3 * A + i] = BEb; found = true;
        }
    }
    found = false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int e0, e1;
        if (curA[i] == BEb) { e0 = be0; e1 = be1; }
        else ends(curA[i], i, e0, e1);
        if (!found && (e0 == C || e1 == C)) { set_edge_enabled(d, curA[i], false); d.vb[3 * A + i] = CF; found = true; }
    }
    set_vertex_enabled(d, B, false);
    set_vertex_enabled(d, C, false);
    set_grain_enabled(d, g, false);
    set_mod(d, B, 0, FIX_REMOVE, G2, B); flag_grain(d, G2, 1);
    set_mod(d, C, 0, FIX_REMOVE, G1, C); flag_grain(d, G1, 1);
    set_mod(d, B, 1, FIX_REMOVE, G3, B); flag_grain(d, G3, 1);
    set_mod(d, C, 1, FIX_REPLACE, G3, A);
    d.mods[C].prev[1] = C;
    for (int i = 0; i < 3; i++) if (gA[i] == g) d.vg[3 * A + i] = G3;
}


// T2 candidates in parallel: a candidate with no other enabled candidate
// among the grains of its vertices is removed at once; the others go to the
// serial list (k_t2_serial, index order).  A removed grain never has a
// candidate neighbour (it would have seen that neighbour, which is enabled
// until its own check passes), and a removal writes only data that no
// non-adjacent candidate's check reads, so the outcome does not depend on
// the schedule.
__global__ void k_t2_par(Dev d) {
    const int n = d.cnt[C_T2];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int g = d.l_t2[k];
        if (!d.ms[member_of_grain(d, g)].active) continue;   // stopped by a failed check or an error
        // the grains of the three vertices and their flags, in three rounds
        // of independent loads
        const int* L = glist(d, g);
        int vv[3], h[9];
        uint8_t he[9], hc[9];
#pragma unroll
        for (int q = 0; q < 3; q++) vv[q] = L[q];
#pragma unroll
        for (int e = 0; e < 9; e++) h[e] = d.vg[3 * vv[e / 3] + e % 3];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const bool other = h[e] != g && h[e] >= 0;
            he[e] = other ? d.genab[h[e]] : 0;
            hc[e] = other ? d.gcand[h[e]] : 0;
        }
        bool clash = false;
#pragma unroll
        for (int e = 0; e < 9; e++) clash = clash || (he[e] && hc[e]);
        const int s = wave_slot(&d.cnt[C_T2C], clash);
        if (clash) {
            d.l_t2c[s] = g;
        } else {
            int m = member_of_grain(d, g);
            remove_grain(d, g, d.prm[m].D);
        }
        wave_count(d, !clash, member_of_grain(d, g), &MState::removals);
    }
}

__global__ void k_t2_serial(Dev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int n = d.cnt[C_T2C];
    if (n == 0) return;
    sort_ints(d.l_t2c, n);
    for (int k = 0; k < n; k++) {
        int g = d.l_t2c[k];
        if (!(d.genab[g] && d.gcand[g])) continue;
        int m = member_of_grain(d, g);
        if (!d.ms[m].active) continue;
        remove_grain(d, g, d.prm[m].D);
        d.ms[m].removals += 1;
    }
}

// unlabel_boundaries_st2, topological.h:588-606
__device__ __forceinline__ void st2_items(const Dev& d, int k0, int stride) {
    const int n = d.cnt[C_CAND];
    for (int k = k0; k < n; k += stride) {
        int b = d.l_cand[k];
        if (!(d.benab[b] && d.bcand[b])) continue;
        if (!d.ms[member_of_edge(d, b)].active) continue;
        int gic[2];
        if (grains_intersect(d, d.bvx[2 * b], d.bvx[2 * b + 1], gic) < 2) { d.bcand[b] = 0; continue; }
        if (d.glen[gic[0]] == 3 || d.glen[gic[1]] == 3) d.bcand[b] = 0;
    }
}

__global__ void k_st2(Dev d) {
    st2_items(d, (int)(blockIdx.x * blockDim.x + threadIdx.x), (int)(gridDim.x * blockDim.x));
}

// k_st2 and the clearing after the T2 stage's fix-up (k_fix_clear, stage 1
// with areas) in one launch of 64-lane workgroups: the first `nclear` take
// the clearing, the others the candidate list.  They touch disjoint data:
// st2 reads the edges' flags and ends, the ends' grains and the grains'
// lengths and writes edge labels; the clearing reads lists and positions
// and writes fix flags, vertex edit records and areas.
__global__ void __launch_bounds__(64) k_st2_fix_clear(Dev d, int nclear) {
    const int b = (int)blockIdx.x;
    if (b < nclear) {
        fix_clear_items(d, 1, 1, b, nclear);
        return;
    }
    st2_items(d, (b - nclear) * 64 + (int)threadIdx.x, ((int)gridDim.x - nclear) * 64);
}

// ---------------------------------------------------------------------
// T1: flips (topological.h:174-291, 1016-1026)
// ---------------------------------------------------------------------

// What a flip reads first about its edge (A, B): the ends' edges, grains
// and positions.  The parallel flips load it together with their conflict
// check.
struct FlipIn {
    int A, B;
    int vA[3], vB[3], gA[3], gB[3];
    double pax, pay, pbx, pby;
};

__device__ __forceinline__ FlipIn flip_in(const Dev& d, int A, int B) {
    FlipIn f;
    f.A = A;
    f.B = B;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        f.vA[k] = d.vb[3 * A + k]; f.vB[k] = d.vb[3 * B + k];
        f.gA[k] = d.vg[3 * A + k]; f.gB[k] = d.vg[3 * B + k];
    }
    const double2 pa = reinterpret_cast<const double2*>(d.pos)[A], pb = reinterpret_cast<const double2*>(d.pos)[B];
    f.pax = pa.x; f.pay = pa.y; f.pbx = pb.x; f.pby = pb.y;
    return f;
}

// flip, topological.h:174-291.  Every datum the flip reads is loaded first
// (the edge's ends; their edges, grains and positions, given in `in`; the
// ends of the rewired edges; the lists of the four grains around; their
// orientations), then the reference's sequence of tests and writes runs on
// local copies in the same order and the writes go out.
__device__ __forceinline__ void flip(const Dev& d, int bnd, const Params& P, const FlipIn& in) {
    const double D = P.D;
    const int A = in.A, B = in.B;
    int vA[3], vB[3], gA[3], gB[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        vA[k] = in.vA[k]; vB[k] = in.vB[k];
        gA[k] = in.gA[k]; gB[k] = in.gB[k];
    }
    const double pax = in.pax, pay = in.pay, pbx = in.pbx, pby = in.pby;
    int gic[2], gnic[2];
    rg_intersect(gA, gB, gic);
    {
        // grains_symdiff (geometry.h:886-898) on the triples
        int gun[4];
        gnic[0] = gnic[1] = -1;
        gun[3] = -1;
        for (int k = 0; k < 3; k++) gun[k] = gA[k];
        for (int k = 0; k < 3; k++) {
            const int g = gB[k];
            if (g != gun[0] && g != gun[1] && g != gun[2]) { gun[3] = g; break; }
        }
        int c = 0;
        for (int k = 0; k < 4; k++)
            if (gun[k] != gic[0] && gun[k] != gic[1]) {
                gnic[c++] = gun[k];
                if (c == 2) break;
            }
    }
    int i = -1, j = -1, AC = -1, BEb = -1;
    for (int k = 0; k < 3; k++) if (vA[k] == bnd) { i = (k + 1) % 3; AC = vA[i]; break; }
    for (int k = 0; k < 3; k++) if (vB[k] == bnd) { j = (k + 1) % 3; BEb = vB[j]; break; }
    if (i < 0 || j < 0 || gic[1] < 0 || gnic[1] < 0) { set_error(d, member_of_edge(d, bnd), SEVC_E_STATE); return; }
    // reads that do not depend on the writes below, issued together
    const int ac0 = d.bvx[2 * AC];
    const int be0 = d.bvx[2 * BEb];
    const int c2[2] = {grain_contains_consecutive(d, gic[0], A, B), grain_contains_consecutive(d, gic[1], A, B)};
    const int c4[2] = {grain_contains_consecutive(d, gic[0], B, A), grain_contains_consecutive(d, gic[1], B, A)};
    const int h3[2] = {grain_contains_vertex(d, gnic[0], A), grain_contains_vertex(d, gnic[1], A)};
    const int h1[2] = {grain_contains_vertex(d, gnic[0], B), grain_contains_vertex(d, gnic[1], B)};
    const double oi0 = d.gori[gic[0]], oi1 = d.gori[gic[1]], on0 = d.gori[gnic[0]], on1 = d.gori[gnic[1]];
    d.vb[3 * A + i] = BEb; vA[i] = BEb;
    d.vb[3 * B + j] = AC; vB[j] = AC;
    if (ac0 == A) d.bvx[2 * AC] = B; else d.bvx[2 * AC + 1] = B;
    // BEb's ends as they are now (AC and BEb are distinct edges of a valid
    // mesh; if they coincide, the write above is seen)
    int nb0 = be0;
    if (BEb == AC) nb0 = (ac0 == A) ? B : ac0;
    if (nb0 == B) d.bvx[2 * BEb] = A; else d.bvx[2 * BEb + 1] = A;
    int g1, g2, g3, g4;
    for (g2 = 0; g2 < 2; g2++)
        if (c2[g2] == 1) {
            set_mod(d, A, 0, FIX_REMOVE, gic[g2], A);
            flag_grain(d, gic[g2], 2);
            break;
        }
    for (g4 = 0; g4 < 2; g4++)
        if (c4[g4] == 1) {
            set_mod(d, B, 0, FIX_REMOVE, gic[g4], B);
            flag_grain(d, gic[g4], 2);
            break;
        }
    for (g3 = 0; g3 < 2; g3++)
        if (h3[g3] >= 0) {
            set_mod(d, A, 1, FIX_ADD, gnic[g3], B);
            d.mods[A].prev[1] = A;
            flag_grain(d, gnic[g3], 2);
            break;
        }
    for (g1 = 0; g1 < 2; g1++)
        if (h1[g1] >= 0) {
            set_mod(d, B, 1, FIX_ADD, gnic[g1], A);
            d.mods[B].prev[1] = B;
            flag_grain(d, gnic[g1], 2);
            break;
        }
    if (g1 > 1 || g2 > 1 || g3 > 1 || g4 > 1) { set_error(d, member_of_edge(d, bnd), SEVC_E_STATE); return; }
    for (int k = 0; k < 3; k++) {
        if (gA[k] == gic[g2]) { gA[k] = gnic[g1]; d.vg[3 * A + k] = gnic[g1]; }
        if (gB[k] == gic[g4]) { gB[k] = gnic[g3]; d.vg[3 * B + k] = gnic[g3]; }
    }
    // vector2_mean, geometry.h:311-317
    double dx = wrapd(pbx, pax, D);
    double dy = wrapd(pby, pay, D);
    double cx = pax + dx / 2, cy = pay + dy / 2;
    cx = dom_mod(cx, D);
    cy = dom_mod(cy, D);
    d.pos[2 * A] = cx; d.pos[2 * A + 1] = cy;
    d.pos[2 * B] = cx; d.pos[2 * B + 1] = cy;
    d.vb[3 * A + 1] = vA[0]; d.vb[3 * A] = vA[1];
    d.vb[3 * B + 1] = vB[0]; d.vb[3 * B] = vB[1];
    // boundary_compute_energy on the rewired grains (no GB_scaling, topological.h:286)
    double e = 1.0;
    {
        int gn[2];
        if (rg_intersect(gA, gB, gn) == 2) {
            auto ori = [&](int g) {
                return g == gic[0] ? oi0 : g == gic[1] ? oi1 : g == gnic[0] ? on0 : g == gnic[1] ? on1 : d.gori[g];
            };
            e = gb_energy_func(ori(gn[0]) - ori(gn[1]), P.gb_eps);
        }
    }
    d.ben[bnd] = e;
    {
        // rescaled at the next step's start (k_begin) when no full pass runs
        const int k = atomicAdd(d.n_flipped, 1);
        if (k < d.E) d.l_flipped[k] = bnd;
    }
    d.vel[2 * A] = 0; d.vel[2 * A + 1] = 0;
    d.vel[2 * B] = 0; d.vel[2 * B + 1] = 0;
}

// Flips in parallel: a candidate with no other enabled candidate at either
// end flips at once, the others go to the serial list (k_flip_serial, index
// order).  A flip changes the boundary lists of its own two vertices and the
// ends of two edges; a candidate reads those only if it shares a vertex, and
// then both are in the serial list, so the outcome does not depend on the
// schedule.
__global__ void k_flip_par(Dev d) {
    const int n = d.cnt[C_CAND];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int c = d.l_cand[k];
        // the candidate's flags and ends; its ends' edges, grains and
        // positions (what the flip reads first); those edges' flags: three
        // rounds of independent loads
        const uint8_t ce = d.benab[c], cc = d.bcand[c];
        const int2 ab = reinterpret_cast<const int2*>(d.bvx)[c];
        if (!(ce && cc)) continue;
        if (!d.ms[member_of_edge(d, c)].active) continue;
        const FlipIn fi = flip_in(d, ab.x, ab.y);
        int x[6];
        uint8_t xe[6], xc[6];
#pragma unroll
        for (int q = 0; q < 6; q++) x[q] = q < 3 ? fi.vA[q] : fi.vB[q - 3];
#pragma unroll
        for (int q = 0; q < 6; q++) {
            xe[q] = x[q] != c ? d.benab[x[q]] : 0;
            xc[q] = x[q] != c ? d.bcand[x[q]] : 0;
        }
        bool clash = false;
#pragma unroll
        for (int q = 0; q < 6; q++) clash = clash || (xe[q] && xc[q]);
        const int s = wave_slot(&d.cnt[C_FLIPC], clash);
        if (clash) {
            d.l_flipc[s] = c;
        } else {
            int m = member_of_edge(d, c);
            flip(d, c, d.prm[m], fi);
        }
        wave_count(d, !clash, member_of_edge(d, c), &MState::flips);
    }
}

__global__ void k_flip_serial(Dev d) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int n = d.cnt[C_FLIPC];
    if (n == 0) return;
    sort_ints(d.l_flipc, n);
    for (int k = 0; k < n; k++) {
        int c = d.l_flipc[k];
        if (!(d.benab[c] && d.bcand[c])) continue;
        int m = member_of_edge(d, c);
        if (!d.ms[m].active) continue;
        flip(d, c, d.prm[m], flip_in(d, d.bvx[2 * c], d.bvx[2 * c + 1]));
        d.ms[m].flips += 1;
    }
}

// ---------------------------------------------------------------------
// deterministic sums (nucleated area, main.cu:321)
// ---------------------------------------------------------------------

// Level 1: one wave per 256-grain table block.  The lanes load the block's
// flags and areas (four coalesced rounds) into LDS, then lane 0 sums the
// nucleated areas left to right (the chunking of the oracle's sum).
__global__ void __launch_bounds__(64) k_nucl_area_l1(Dev d) {
    __shared__ double a_s[kBlock];
    __shared__ uint8_t ok_s[kBlock];
    const int lb = blockIdx.x;
    if (lb >= d.ngb) return;
    const Blk bd = table_blk(d, d.gblk, lb);
    const int lane = threadIdx.x;
#pragma unroll
    for (int q = 0; q < kBlock / 64; q++) {
        const int g = bd.b + 64 * q + lane;
        const bool ok = g < bd.e && d.genab[g] && d.gtype[g] == T_NUCLEATED;
        ok_s[64 * q + lane] = ok ? 1 : 0;
        a_s[64 * q + lane] = ok ? d.garea[g] : 0.0;
    }
    __syncthreads();
    if (lane == 0) {
        double s = 0.0;
        const int n = bd.e - bd.b;
        for (int i = 0; i < n; i++)
            if (ok_s[i]) s += a_s[i];
        d.partial[lb] = s;
    }
}

// Upper levels for each member: chunks of 256 partials left to right,
// repeated until one value remains.  One workgroup per member; a level of
// up to kAreaLds partials is first copied to LDS with coalesced loads and
// each lane then sums its chunk there; levels alternate between the two
// halves of `partial` so no lane overwrites a value another lane still
// reads.
constexpr int kAreaLds = 8192;
__global__ void __launch_bounds__(1024) k_nucl_area_up(Dev d) {
    __shared__ double v_s[kAreaLds];
    int m = blockIdx.x;
    if (m >= d.B) return;
    int first = d.gblk_first[m], cnt = d.gblk_first[m + 1] - first;
    double* src = d.partial + first;
    double* dst = d.partial + d.ngb + first;
    while (cnt > 1) {
        const int nn = (cnt + 255) / 256;
        const bool lds = cnt <= kAreaLds;                // the same for the whole workgroup
        if (lds) {
            for (int i0 = threadIdx.x; i0 < cnt; i0 += 4 * blockDim.x) {   // 4 loads per lane and round
                double t[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = i0 + q * (int)blockDim.x;
                    t[q] = i < cnt ? src[i] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int i = i0 + q * (int)blockDim.x;
                    if (i < cnt) v_s[i] = t[q];
                }
            }
            __syncthreads();
        }
        const double* from = lds ? v_s : src;
        for (int c = threadIdx.x; c < nn; c += blockDim.x) {
            double s = 0.0;
            int e = (c + 1) * 256 < cnt ? (c + 1) * 256 : cnt;
            for (int q0 = c * 256; q0 < e; q0 += 8) {    // 8 values read ahead, added in order
                double t[8];
#pragma unroll
                for (int q = 0; q < 8; q++) t[q] = q0 + q < e ? from[q0 + q] : 0.0;
#pragma unroll
                for (int q = 0; q < 8; q++)
                    if (q0 + q < e) s += t[q];
            }
            dst[c] = s;
        }
        __syncthreads();
        double* t = src; src = dst; dst = t;
        cnt = nn;
    }
    if (threadIdx.x == 0) {
        const Params& P = d.prm[m];
        double a = cnt == 1 ? src[0] : 0.0;
        d.ms[m].nucl_area_pct = 100 * a / (P.D * P.D);
    }
}


// ---------------------------------------------------------------------
// nucleation (nucleation2.cu)
// ---------------------------------------------------------------------

struct NuclGeom {
    double Apos[2], Bpos[2], Cpos[2];
    double Dpos[2], Epos[2], Fpos[2];   // positions of D, E, F as read
    int D, E, F, AD, AE, AF, g1, g2, g3;
    double dE;
};

// generate_grain_for_nucleation, nucleation2.cu:116-212
__device__ __forceinline__ int generate_grain(const Dev& d, const Params& P, int A, double eps, double scaling, int do_scale,
                              NuclGeom* r) {
    const double Dm = P.D;
    r->AD = d.vb[3 * A + 0]; r->AE = d.vb[3 * A + 1]; r->AF = d.vb[3 * A + 2];
    r->D = other_end(d, r->AD, A);
    r->E = other_end(d, r->AE, A);
    r->F = other_end(d, r->AF, A);
    int gAD[2], gAE[2], gAF[2];
    grains_intersect(d, A, r->D, gAD);
    grains_intersect(d, A, r->E, gAE);
    grains_intersect(d, A, r->F, gAF);
    if (gAD[0] == gAE[0] || gAD[0] == gAE[1]) { r->g2 = gAD[0]; r->g1 = gAD[1]; }
    else if (gAD[1] == gAE[0] || gAD[1] == gAE[1]) { r->g2 = gAD[1]; r->g1 = gAD[0]; }
    else return -1;
    if (gAE[0] == r->g2) r->g3 = gAE[1];
    else if (gAE[1] == r->g2) r->g3 = gAE[0];
    else return -1;
    if (r->g1 < 0 || r->g2 < 0 || r->g3 < 0) return -1;
    const double ax = d.pos[2 * A], ay = d.pos[2 * A + 1], rr = d.nucl_r[A];
    const int nb[3] = {r->D, r->E, r->F};
    double* out[3] = {r->Apos, r->Bpos, r->Cpos};
    r->Dpos[0] = d.pos[2 * nb[0]]; r->Dpos[1] = d.pos[2 * nb[0] + 1];
    r->Epos[0] = d.pos[2 * nb[1]]; r->Epos[1] = d.pos[2 * nb[1] + 1];
    r->Fpos[0] = d.pos[2 * nb[2]]; r->Fpos[1] = d.pos[2 * nb[2] + 1];
    for (int k = 0; k < 3; k++) {
        double tx = wrapd(d.pos[2 * nb[k]], ax, Dm);
        double ty = wrapd(d.pos[2 * nb[k] + 1], ay, Dm);
        double w = rr / mag(tx, ty);
        if (do_scale) w *= 0.99;
        double px = w * tx + ax, py = w * ty + ay;
        out[k][0] = dom_mod(px, Dm);
        out[k][1] = dom_mod(py, Dm);
    }
    double LAB = arclen(r->Apos[0], r->Apos[1], r->Bpos[0], r->Bpos[1], Dm);
    double LBC = arclen(r->Bpos[0], r->Bpos[1], r->Cpos[0], r->Cpos[1], Dm);
    double LCA = arclen(r->Cpos[0], r->Cpos[1], r->Apos[0], r->Apos[1], Dm);
    AreaAcc acc;
    acc.start(ax, ay); acc.push(r->Cpos[0], r->Cpos[1], Dm); acc.push(r->Apos[0], r->Apos[1], Dm);
    double A1 = acc.finish();
    acc.start(ax, ay); acc.push(r->Apos[0], r->Apos[1], Dm); acc.push(r->Bpos[0], r->Bpos[1], Dm);
    double A2 = acc.finish();
    acc.start(ax, ay); acc.push(r->Bpos[0], r->Bpos[1], Dm); acc.push(r->Cpos[0], r->Cpos[1], Dm);
    double A3 = acc.finish();
    double E0 = 0.0, E1;
    for (int i = 0; i < 3; i++) {
        if (do_scale) E0 += 0.99 * rr * d.ben[d.vb[3 * A + i]];
        else E0 += rr * d.ben[d.vb[3 * A + i]];
    }
    E0 += A1 * d.gse[r->g1] + A2 * d.gse[r->g2] + A3 * d.gse[r->g3];
    const double ori = d.nucl_ori[A];
    E1 = gb_energy_func(ori - d.gori[r->g1], eps) * scaling * LCA;
    E1 += gb_energy_func(ori - d.gori[r->g2], eps) * scaling * LAB;
    E1 += gb_energy_func(ori - d.gori[r->g3], eps) * scaling * LBC;
    r->dE = E1 - E0;
    return 0;
}

// compute_nucleation_factor, nucleation2.cu:239-282, with the random pair
// Philox(seed; step, local vertex id, stream 1).
// compute_nucleation_factor (nucleation2.cu) for one alive vertex v of a
// Monte Carlo member: the nucleus radius and orientation from the vertex's
// own Philox pair, dE and the factor dE - dE'.  The manager needs it only
// for the vertex it picks, so it runs there, for that vertex alone.
__device__ void nucl_factor_at(const Dev& d, const Params& P, const MState& M, int v) {
    d.dE[v] = 0;
    d.nfac[v] = 0;
    double ALs[3];
    for (int q = 0; q < 3; q++) ALs[q] = edge_len_from(d, d.vb[3 * v + q], v, P.D);
    double minAL = ALs[0];
    for (int q = 0; q < 3; q++) if (minAL > ALs[q]) minAL = ALs[q];
    double u0, u1;
    sevc_uniform2(P.seed, (uint32_t)M.n, (uint32_t)(v - P.voff), kStreamMcSite, 0u, &u0, &u1);
    d.nucl_r[v] = 0.8 * minAL * u0 + 0.1;
    d.nucl_ori[v] = 2 * SEVC_PI * u1;
    NuclGeom g0, g1;
    double de = 0.0, de_ = 0.0;
    if (generate_grain(d, P, v, P.gb_eps, P.gb_scaling, 0, &g0) == 0) de = g0.dE;
    d.dE[v] = de;
    if (generate_grain(d, P, v, P.gb_eps, P.gb_scaling, 1, &g1) == 0) de_ = g1.dE;
    d.nfac[v] = de - de_;
}

// Rate mode: grain g of a nucleating member is a candidate with
// probability nucleation_rate * SE_g (Philox(seed; step, local grain, 3, 0),
// first uniform).  The second uniform of the same draw picks the site, the
// list entry floor(u1 * n) of the grain; the candidate's record holds the
// site (-1 when its shortest boundary has zero length: no candidate), the
// site's three grains and that shortest length.
// One grain table block of the rate mode's candidate draw.  It runs in the
// launch of the step's begin (k_begin_rate), so it decides whether its
// member runs and nucleates from the loop state itself, by the same rules
// (member_runs, rate_nucleates) that the begin workgroup writes into it.
__device__ __forceinline__ void rate_cand_block(const Dev& d, int lb) {
    const Blk bd_ = table_blk(d, d.gblk, lb);
    const Params& P = d.prm[bd_.m];
    const MState& M = d.ms[bd_.m];
    const int i = bd_.b + (int)threadIdx.x;
    // the grain's data is loaded together with the member's state
    const bool in = i < bd_.e;
    const uint8_t en = in ? d.genab[i] : 0;
    const int len = in ? d.glen[i] : 0;
    const double se = in ? d.gse[i] : 0.0;
    if (!member_runs(P, M) || !rate_nucleates(P, M)) return;
    if (!en || len < 1) return;
    double p = P.nucl_rate * se;
    double u0, u1;
    sevc_uniform2(P.seed, (uint32_t)M.n, (uint32_t)(i - P.goff), kStreamRate, 0u, &u0, &u1);
    if (!(u0 < p)) return;
    int k = (int)(u1 * (double)len);
    if (k >= len) k = len - 1;
    const int A = glist(d, i)[k];
    int vb[3], vg[3];
#pragma unroll
    for (int q = 0; q < 3; q++) { vb[q] = d.vb[3 * A + q]; vg[q] = d.vg[3 * A + q]; }
    double al[3];
#pragma unroll
    for (int q = 0; q < 3; q++) al[q] = edge_len_from(d, vb[q], A, P.D);
    double minAL = al[0];
    for (int q = 0; q < 3; q++) if (minAL > al[q]) minAL = al[q];
    const int s = wave_slot(&d.cnt[C_NUCL], true);
    d.l_nucl[s] = i;
    d.nrec[4 * s] = (minAL > 0) ? A : -1;
#pragma unroll
    for (int q = 0; q < 3; q++) d.nrec[4 * s + 1 + q] = vg[q];
    d.nal[s] = minAL;
}

// The step's begin (workgroup 0, dispatched first) and the rate mode's
// candidate draw (one workgroup per grain table block) in one launch.
__global__ void __launch_bounds__(kBlock) k_begin_rate(Dev d, int rescale_flipped) {
    if (blockIdx.x == 0) {
        begin_members(d, rescale_flipped);
        return;
    }
    const int lb = xcd_remap((int)blockIdx.x - 1, (int)gridDim.x - 1);
    if (lb < d.ngb) rate_cand_block(d, lb);
}

// Exclusive prefix sum of one int per thread over the workgroup: an
// in-wave scan, the wave totals scanned by wave 0, two barriers.
// Returns the prefix; *total receives the sum.  All threads must call it;
// blockDim.x is a multiple of 64 (at most 1024) and `sh` holds >= 65 ints.
__device__ int wg_exclusive_scan(int x, int* sh, int* total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = (int)(blockDim.x >> 6);
    const int inc = wave_inclusive_scan(x);
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        const int v = lane < nw ? sh[lane] : 0;
        const int vi = wave_inclusive_scan(v);
        if (lane < nw) sh[lane] = vi - v;
        if (lane == nw - 1) sh[64] = vi;
    }
    __syncthreads();
    const int r = sh[wave] + inc - x;
    *total = sh[64];
    __syncthreads();
    return r;
}

__device__ __forceinline__ int block_matches(const int2* bc, int b, int want) {
    const int2 c = bc[b];
    return want ? c.x : c.y;
}

// Split of a member's table blocks into one run of consecutive blocks per
// lane, with the number of wanted slots in each run and its exclusive
// prefix over the workgroup.  The searches below walk only the runs that
// hold the ranks they need.
struct RunScan {
    int lo, hi;     // this lane's blocks [lo, hi)
    int cnt, pre;   // wanted slots in the run, and before it
    int total;      // wanted slots of the member
};

__device__ RunScan run_scan(int first, int nblk, const int2* bc, int want, int* sh) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int C = (nblk + nt - 1) / nt;
    RunScan r;
    r.lo = first + min(nblk, tid * C);
    r.hi = first + min(nblk, (tid + 1) * C);
    int c = 0;
    for (int b = r.lo; b < r.hi; b += 8) {            // 8 independent loads per round
        int t[8];
#pragma unroll
        for (int q = 0; q < 8; q++) t[q] = b + q < r.hi ? block_matches(bc, b + q, want) : 0;
#pragma unroll
        for (int q = 0; q < 8; q++) c += t[q];
    }
    r.cnt = c;
    r.pre = wg_exclusive_scan(c, sh, &r.total);
    return r;
}

// One kind of slot of a member to collect: its blocks [first, first+nblk)
// of the block table, which cover slots [off, off+n) kBlock at a time, the
// per-block (enabled, free) counts, the enable flags, how many free slots
// are wanted and where they go.
struct SlotKind {
    int first, nblk, off, n;
    const int2* bc;
    const uint8_t* en;
    int R;
    int* out;
};

// Rows per kind in one round of the free-slot search (grains, boundaries,
// vertices: a trivalent mesh has 3 boundaries and 2 vertices per grain, so
// a round covers about the same grains of every kind).
constexpr int kFreeRun = 12;
constexpr int kFreeRunOf[3] = {4, 12, 8};
constexpr int kFreeRowBase[3] = {0, 4, 16};   // first row of each kind in the LDS tables (24 rows)

// Shape of a round: rows of kind t, and blocks per row.  By default a row
// is one block per thread and a kind has kFreeRunOf rows; lim = 1..1024
// makes a round one row of lim blocks (several rounds on small states).
__device__ __forceinline__ int free_rows_of(int t, int lim) {
    return lim > 0 ? 1 : (t == 0 ? kFreeRunOf[0] : (t == 1 ? kFreeRunOf[1] : kFreeRunOf[2]));
}
__device__ __forceinline__ int free_row_len(int lim) {
    const int nt = blockDim.x;
    return lim > 0 && lim < nt ? lim : nt;
}

// Round r of a kind's free-slot search covers its blocks [r C W, (r+1) C W)
// (C = free_rows_of(kind, lim) <= kRun rows of W = free_row_len(lim)
// blocks): thread tid < W reads the free count of block r C W + q W + tid
// into y[q] (0 past the kind's blocks, for q >= C and for tid >= W),
// coalesced over the workgroup.
template <int kRun>
__device__ __forceinline__ void read_free_rows(const SlotKind& K, int r, int C, int W, int (&y)[kFreeRun]) {
    const int tid = threadIdx.x;
    const int b0 = r * C * W + tid;
#pragma unroll
    for (int q = 0; q < kFreeRun; q++) {
        const int b = b0 + q * W;
        y[q] = q < kRun && q < C && tid < W && b < K.nblk ? K.bc[K.first + b].y : 0;
    }
}

__device__ __forceinline__ void read_free_rounds(const SlotKind (&K)[3], int r, const bool (&need)[3], int lim,
                                                 int (&y)[3][kFreeRun]) {
    const int W = free_row_len(lim);
    read_free_rows<kFreeRunOf[0]>(K[0], need[0] ? r : 0, need[0] ? free_rows_of(0, lim) : 0, W, y[0]);
    read_free_rows<kFreeRunOf[1]>(K[1], need[1] ? r : 0, need[1] ? free_rows_of(1, lim) : 0, W, y[1]);
    read_free_rows<kFreeRunOf[2]>(K[2], need[2] ? r : 0, need[2] ? free_rows_of(2, lim) : 0, W, y[2]);
}

// The first R free slots (enable flag 0) of each of three kinds, in index
// order, cooperatively in one workgroup of at least 3 waves.  In rounds
// (one while a kind has at most kFreeRunOf nt blocks) every lane holds one
// free count per row of each kind (read_free_rows; y holds round 0 on
// entry, read by the caller ahead of time).  The waves scan each row
// (wave_inclusive_scan), waves 0-2 scan the (row, wave) totals of one kind each in
// block order, and every lane then knows the first rank of each of its
// blocks: it lists (kind, block, first rank, free count) for those that
// hold a wanted rank in `bl` (LDS when the caller's bound on the entries
// fits there, else global scratch).  The rounds stop once each kind has R
// ranks.  The waves then sweep the listed blocks, kSweep at a time, and
// write each free slot at its rank.  got[t] = min(R_t, free slots of kind
// t).  `sh` holds >= 24 * 16 + 3 ints; `bl` holds nbl >= R_0 + R_1 + R_2
// entries (a listed block holds a wanted rank, so no more are listed).
__device__ __forceinline__ void collect_free3(const SlotKind K0, const SlotKind K1, const SlotKind K2,
                                              int (&y)[3][kFreeRun], int lim, int* got, int* sh, int4* bl, int nbl) {
    const SlotKind K[3] = {K0, K1, K2};    // by value and indexed by constants only: kept in registers
    __shared__ int nl;
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    constexpr int kTot = 24 * 16;           // sh[kTot + t]: the round's free slots of kind t
    const int W = free_row_len(lim);        // blocks per row
    if (tid == 0) nl = 0;                   // seen by all after the first barrier below
    int off[3];                             // free slots in the blocks of earlier rounds
    bool need[3];
#pragma unroll
    for (int t = 0; t < 3; t++) { off[t] = 0; need[t] = K[t].R > 0 && K[t].nblk > 0; }
    for (int r = 0;; r++) {                 // the same for every lane
        if (r > 0) read_free_rounds(K, r, need, lim, y);
        // inclusive scan of every row within the wave, kept with the count
        // as (scan << 16) | count (a wave's sum is at most 64 * 256); wave
        // totals to LDS
        int pk[3][kFreeRun];
#pragma unroll
        for (int t = 0; t < 3; t++)
#pragma unroll
            for (int q = 0; q < kFreeRunOf[t]; q++) {
                const int c = y[t][q];              // 0 for kinds done and rows past free_rows_of
                const int v = wave_inclusive_scan(c);
                pk[t][q] = (v << 16) | c;
                if (lane == 63) sh[(kFreeRowBase[t] + q) * 16 + wave] = v;
            }
        __syncthreads();
        // wave t: exclusive scan of kind t's (row, wave) totals in block order
        if (wave < 3) {
            const int t = wave;
            const int base = (t == 0 ? kFreeRowBase[0] : (t == 1 ? kFreeRowBase[1] : kFreeRowBase[2])) * 16;
            const int rows = free_rows_of(t, lim);
            int carry = 0;
            for (int c0 = 0; c0 < rows * nw; c0 += 64) {
                const int k = c0 + lane;
                const int v = k < rows * nw ? sh[base + (k / nw) * 16 + k % nw] : 0;
                const int vi = wave_inclusive_scan(v);
                if (k < rows * nw) sh[base + (k / nw) * 16 + k % nw] = carry + vi - v;
                carry += __builtin_amdgcn_readlane(vi, 63);
            }
            if (lane == 0) sh[kTot + t] = carry;
        }
        __syncthreads();
        int tot[3];
#pragma unroll
        for (int t = 0; t < 3; t++) {
            tot[t] = sh[kTot + t];
            if (!need[t]) continue;
#pragma unroll
            for (int q = 0; q < kFreeRunOf[t]; q++) {
                const int c = pk[t][q] & 0xFFFF;
                const int rank = off[t] + sh[(kFreeRowBase[t] + q) * 16 + wave] + (pk[t][q] >> 16) - c;
                if (c > 0 && rank < K[t].R) {
                    const int e = atomicAdd(&nl, 1);
                    if (e < nbl) bl[e] = make_int4(t, K[t].first + (r * free_rows_of(t, lim) + q) * W + tid, rank, c);
                }
            }
        }
        __syncthreads();                    // sh is read above and written by the next round
        bool more = false;
#pragma unroll
        for (int t = 0; t < 3; t++) {
            if (need[t]) off[t] += tot[t];
            need[t] = need[t] && off[t] < K[t].R && (r + 1) * free_rows_of(t, lim) * W < K[t].nblk;
            more = more || need[t];
        }
        if (!more) break;
    }
    const int n = min(nl, nbl);
    const unsigned long long below = (1ull << lane) - 1ull;
    constexpr int kSweep = 8;
    for (int q0 = wave; q0 < n; q0 += nw * kSweep) {
        int4 e[kSweep];
#pragma unroll
        for (int j = 0; j < kSweep; j++) {
            const int q = q0 + nw * j;
            e[j] = q < n ? bl[q] : make_int4(0, 0, 0, -1);
        }
        bool hit[kSweep][kBlock / 64];
        int blo[kSweep];
#pragma unroll
        for (int j = 0; j < kSweep; j++) {
            // the kind's fields by selection (no indexing of K by a run-time t)
            const int t = e[j].x;
            const int first = t == 0 ? K[0].first : (t == 1 ? K[1].first : K[2].first);
            const int off = t == 0 ? K[0].off : (t == 1 ? K[1].off : K[2].off);
            const int nn = t == 0 ? K[0].n : (t == 1 ? K[1].n : K[2].n);
            const uint8_t* en = t == 0 ? K[0].en : (t == 1 ? K[1].en : K[2].en);
            // padding entries (w < 0) get the empty range [0, 0), so no lane loads
            const bool listed = e[j].w >= 0;
            blo[j] = listed ? off + kBlock * (e[j].y - first) : 0;
            const int bhi = listed ? min(blo[j] + kBlock, off + nn) : 0;
#pragma unroll
            for (int k = 0; k < kBlock / 64; k++) {
                const int idx = blo[j] + 64 * k + lane;
                hit[j][k] = idx < bhi && en[idx] == 0;
            }
        }
#pragma unroll
        for (int j = 0; j < kSweep; j++) {
            const int t = e[j].x;
            int* out = t == 0 ? K[0].out : (t == 1 ? K[1].out : K[2].out);
            const int R = t == 0 ? K[0].R : (t == 1 ? K[1].R : K[2].R);
            int rank = e[j].z;
#pragma unroll
            for (int k = 0; k < kBlock / 64; k++) {
                const unsigned long long hb = __ballot(hit[j][k]);
                if (hit[j][k]) {
                    const int r = rank + (int)__popcll(hb & below);
                    if (r < R) out[r] = blo[j] + 64 * k + lane;
                }
                rank += (int)__popcll(hb);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 3; t++) got[t] = off[t] < K[t].R ? off[t] : K[t].R;
}

// The k-th (0-based) slot of a member whose enable flag equals `want`,
// given the member's run scan; -1 if there is none.  The lane whose run
// holds rank k finds the block, wave 0 sweeps that block.
__device__ int find_kth(const Blk* tbl, const RunScan& rs, const int2* bc, const uint8_t* en, int want, int k) {
    __shared__ int fb, fr, found;
    const int tid = threadIdx.x;
    if (tid == 0) { fb = -1; fr = 0; found = -1; }
    __syncthreads();
    if (k >= rs.pre && k < rs.pre + rs.cnt) {
        int rank = rs.pre;
        bool found_b = false;
        for (int b0 = rs.lo; b0 < rs.hi && !found_b; b0 += 8) {
            int t[8];
#pragma unroll
            for (int q = 0; q < 8; q++) t[q] = (b0 + q < rs.hi) ? block_matches(bc, b0 + q, want) : 0;
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (!found_b) {
                    if (k < rank + t[q]) { fb = b0 + q; fr = k - rank; found_b = true; }
                    else rank += t[q];
                }
        }
    }
    __syncthreads();
    const int bb = fb;
    if (bb >= 0 && tid < 64) {
        const int hi = tbl[bb].e, lane = tid;
        const unsigned long long below = (1ull << lane) - 1ull;
        int rank = 0;
        for (int i = tbl[bb].b; i < hi && rank <= fr; i += 64) {
            const int idx = i + lane;
            const bool hit = idx < hi && ((en[idx] != 0) == (want != 0));
            const unsigned long long hb = __ballot(hit);
            if (hit && rank + (int)__popcll(hb & below) == fr) found = idx;
            rank += (int)__popcll(hb);
        }
    }
    __syncthreads();
    const int f = found;
    __syncthreads();
    return f;
}

// One workgroup per member: picks the nucleation events of this step and
// assigns their free slots; k_nucl_insert applies them.
//  MC mode (nucleation2.cu:292-443, 33-72, 474): one uniformly chosen alive
//  vertex, Metropolis acceptance, first free grain / 3 boundaries / 2
//  vertices.
//  RATE mode: candidates in grain order; a candidate is dropped when an
//  earlier candidate's site shares a grain with it; kept events take the
//  free slots in rank order.  Kept events are grain-disjoint, so they can
//  be applied in parallel.
// Scratch: 12 ints per grain slot of the member, at 12*goff in d.nscr:
//  [0,ng) candidates, [ng,4ng) their site grains, [4ng,5ng) kept sites,
//  [5ng,6ng) free grains, [6ng,9ng) free boundaries, [9ng,11ng) free
//  vertices (held in LDS instead while 6 K <= kSlotLds for K events),
//  [11ng,12ng) candidate sites.
constexpr int kNuclLds = 1024;       // rate-mode candidates held in LDS
constexpr int kNuclHash = 4096;      // grain table of the clash test (> 3 kNuclLds keys: never full)
constexpr int kHashEmpty = INT_MIN;
constexpr int kSlotLds = 1024;       // listed blocks of the free-slot search held in LDS

__device__ __forceinline__ int nucl_hash(int g) { return (int)(((unsigned)g * 2654435761u) >> 20); }

__global__ void __launch_bounds__(kNuclThreads) k_nucl_manager(Dev d) {
    __shared__ int sh[kNuclThreads + 1];
    __shared__ int nkeep, ncand, ev0;
    __shared__ int c_s[kNuclLds], site_s[kNuclLds], sg_s[3 * kNuclLds], k_s[kNuclLds], q_s[kNuclLds];
    __shared__ double al_s[kNuclLds];         // shortest edge at each candidate's site
    __shared__ int hk[kNuclHash], hv[kNuclHash];
    __shared__ int4 bl_s[kSlotLds];
    __shared__ int fo_s[kSlotLds];            // the free slots found (grains, boundaries, vertices)
    const int m = blockIdx.x;
    if (m >= d.B) return;
    const Params& P = d.prm[m];
    MState& M = d.ms[m];
    // a single member's block ranges are the whole tables
    const int gb0 = d.B == 1 ? 0 : d.gblk_first[m], gb1 = d.B == 1 ? d.ngb : d.gblk_first[m + 1];
    const int eb0 = d.B == 1 ? 0 : d.eblk_first[m], eb1 = d.B == 1 ? d.neb : d.eblk_first[m + 1];
    const int vb0 = d.B == 1 ? 0 : d.vblk_first[m], vb1 = d.B == 1 ? d.nvb : d.vblk_first[m + 1];
    if (!M.active || !M.nucl_now) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ng = P.ng;
    int* base = d.nscr + 12 * (size_t)P.goff;
    int* cand = base;
    int* sgr = base + ng;
    int* kept = base + 4 * (size_t)ng;
    int* site = base + 11 * (size_t)ng;
    // Free slots of each kind, taken in rank order before any insertion of
    // this step: the first round of their search reads its free counts here,
    // so that the loads overlap the selection below.
    SlotKind kinds[3] = {{gb0, gb1 - gb0, P.goff, P.ng, d.bc_g, d.genab, 0, base + 5 * (size_t)ng},
                         {eb0, eb1 - eb0, P.eoff, P.ne, d.bc_e, d.benab, 0, base + 6 * (size_t)ng},
                         {vb0, vb1 - vb0, P.voff, P.nv, d.bc_v, d.ven, 0, base + 9 * (size_t)ng}};
    int fy[3][kFreeRun];
    const bool all3[3] = {true, true, true};
    read_free_rounds(kinds, 0, all3, d.free_rows, fy);
    if (tid == 0) { nkeep = 0; ncand = 0; }
    __syncthreads();
    if (P.nucl_mode == SEVC_NUCL_MC) {
        const int first = d.vblk_first[m], nb = d.vblk_first[m + 1] - first;
        const RunScan rs = run_scan(first, nb, d.bc_v, 1, sh);
        const int n_alive = rs.total;
        if (n_alive == 0) return;
        double u0, u1;
        sevc_uniform2(P.seed, (uint32_t)M.n, 0u, kStreamMcPick, 0u, &u0, &u1);
        int k = (int)(u0 * (double)n_alive);
        if (k >= n_alive) k = n_alive - 1;
        const int A = find_kth(d.vblk, rs, d.bc_v, d.ven, 1, k);
        if (tid == 0 && A >= 0) {
            nucl_factor_at(d, P, M, A);
            d.vnuc[A] = 0;
            if (d.dE[A] < 0) d.vnuc[A] = 1;
            else if (d.nfac[A] < 0) {
                double thr = sevc_exp(-P.mc_k * d.dE[A]);
                if (u1 < thr) d.vnuc[A] = 1;
            }
            if (d.vnuc[A]) { kept[0] = A; nkeep = 1; }
        }
    } else {
        // this member's candidates (grain, list index), then grain order
        const int total = d.cnt[C_NUCL];
        for (int q = tid; q < total; q += nt) {
            const int g = d.l_nucl[q];
            if (g >= P.goff && g < P.goff + ng) {
                const int e = atomicAdd(&ncand, 1);
                cand[e] = g;
                if (e < kNuclLds) { k_s[e] = g; q_s[e] = q; }
            }
        }
        __syncthreads();
        const int nc = ncand;
        if (nc <= kNuclLds) {
            // In LDS.  Rank sort (distinct grain ids); the sorted list
            // indices wait in site_s until the records replace them.
            for (int e = tid; e < nc; e += nt) {
                const int g = k_s[e];
                int r = 0;
                for (int f = 0; f < nc; f++) r += k_s[f] < g ? 1 : 0;
                c_s[r] = g;
                site_s[r] = q_s[e];
            }
            for (int i = tid; i < kNuclHash; i += nt) { hk[i] = kHashEmpty; hv[i] = INT_MAX; }
            __syncthreads();
            for (int e = tid; e < nc; e += nt) {
                const int q = site_s[e];
                site_s[e] = d.nrec[4 * q];
#pragma unroll
                for (int a = 0; a < 3; a++) sg_s[3 * e + a] = d.nrec[4 * q + 1 + a];
                al_s[e] = d.nal[q];
            }
            __syncthreads();
            // every grain at a valid site -> the first candidate (in grain
            // order) whose site touches it
            for (int e = tid; e < nc; e += nt) {
                if (site_s[e] < 0) continue;
                for (int a = 0; a < 3; a++) {
                    const int key = sg_s[3 * e + a];
                    for (int h = nucl_hash(key);; h = (h + 1) & (kNuclHash - 1)) {
                        const int prev = atomicCAS(&hk[h], kHashEmpty, key);
                        if (prev == kHashEmpty || prev == key) { atomicMin(&hv[h], e); break; }
                    }
                }
            }
            __syncthreads();
            // kept iff no earlier candidate touches one of its grains
            for (int e = tid; e < nc; e += nt) {
                bool clash = site_s[e] < 0;
                for (int a = 0; a < 3 && !clash; a++) {
                    const int key = sg_s[3 * e + a];
                    for (int h = nucl_hash(key);; h = (h + 1) & (kNuclHash - 1))
                        if (hk[h] == key) { clash = hv[h] < e; break; }
                }
                k_s[e] = clash ? 0 : 1;
            }
            __syncthreads();
            // stable compaction of the kept candidates, one workgroup scan
            // per nt candidates (values are read before the scan, written
            // after)
            int off = 0;
            for (int b0 = 0; b0 < nc; b0 += nt) {
                const int e = b0 + tid;
                const int flag = e < nc ? k_s[e] : 0;
                const int g = flag ? c_s[e] : 0, A = flag ? site_s[e] : 0;
                const double la = flag ? al_s[e] : 0.0;
                int tot;
                const int pre = wg_exclusive_scan(flag, sh, &tot);
                if (flag) {
                    c_s[off + pre] = g;
                    site_s[off + pre] = A;
                    al_s[off + pre] = la;
                }
                off += tot;
                __syncthreads();
            }
            if (tid == 0) nkeep = off;
            __syncthreads();
            for (int r = tid; r < nkeep; r += nt) {
                const int g = c_s[r], A = site_s[r];
                double u2, u3;
                sevc_uniform2(P.seed, (uint32_t)M.n, (uint32_t)(g - P.goff), kStreamRate, 1u, &u2, &u3);
                d.nucl_r[A] = P.nucl_scaling * al_s[r] * u2;
                d.nucl_ori[A] = 2 * SEVC_PI * u3;
                kept[r] = A;
            }
        } else {
            // Beyond kNuclLds: the same selection on the global scratch,
            // with the clash test over all earlier candidates.
            int* uns = kept;                      // unsorted copy, later the keep flags
            for (int e = tid; e < nc; e += nt) uns[e] = cand[e];
            __syncthreads();
            for (int e = tid; e < nc; e += nt) {
                const int g = uns[e];
                int r = 0;
                for (int f = 0; f < nc; f++) r += uns[f] < g ? 1 : 0;
                cand[r] = g;
            }
            __syncthreads();
            for (int e = tid; e < nc; e += nt) {
                const int g = cand[e];
                double u0, u1;
                sevc_uniform2(P.seed, (uint32_t)M.n, (uint32_t)(g - P.goff), kStreamRate, 0u, &u0, &u1);
                const int len = d.glen[g];
                int k = (int)(u1 * (double)len);
                if (k >= len) k = len - 1;
                const int A = glist(d, g)[k];
                double al[3];
                for (int q = 0; q < 3; q++) al[q] = edge_len_from(d, d.vb[3 * A + q], A, P.D);
                double minAL = al[0];
                for (int q = 0; q < 3; q++) if (minAL > al[q]) minAL = al[q];
                site[e] = (minAL > 0) ? A : -1;    // a site on a zero-length boundary is not a candidate
                for (int a = 0; a < 3; a++) sgr[3 * e + a] = d.vg[3 * A + a];
            }
            __syncthreads();
            for (int e = tid; e < nc; e += nt) {
                bool clash = site[e] < 0;
                for (int f = 0; f < e && !clash; f++) {
                    if (site[f] < 0) continue;
                    for (int a = 0; a < 3 && !clash; a++)
                        for (int b = 0; b < 3; b++)
                            if (sgr[3 * e + a] == sgr[3 * f + b]) { clash = true; break; }
                }
                uns[e] = clash ? 0 : 1;
            }
            __syncthreads();
            int off = 0;
            for (int b0 = 0; b0 < nc; b0 += nt) {
                const int e = b0 + tid;
                const int flag = e < nc ? uns[e] : 0;
                const int g = flag ? cand[e] : 0, A = flag ? site[e] : 0;
                int tot;
                const int pre = wg_exclusive_scan(flag, sh, &tot);
                if (flag) {
                    cand[off + pre] = g;
                    site[off + pre] = A;
                }
                off += tot;
                __syncthreads();
            }
            if (tid == 0) nkeep = off;
            __syncthreads();
            for (int r = tid; r < nkeep; r += nt) {
                const int g = cand[r], A = site[r];
                double u2, u3;
                sevc_uniform2(P.seed, (uint32_t)M.n, (uint32_t)(g - P.goff), kStreamRate, 1u, &u2, &u3);
                double al[3];
                for (int q = 0; q < 3; q++) al[q] = edge_len_from(d, d.vb[3 * A + q], A, P.D);
                double minAL = al[0];
                for (int q = 0; q < 3; q++) if (minAL > al[q]) minAL = al[q];
                d.nucl_r[A] = P.nucl_scaling * minAL * u2;
                d.nucl_ori[A] = 2 * SEVC_PI * u3;
                kept[r] = A;
            }
        }
    }
    __syncthreads();
    const int K = nkeep;
    if (K == 0) return;
    kinds[0].R = K;
    kinds[1].R = 3 * K;
    kinds[2].R = 2 * K;
    const bool in_lds = 6 * K <= kSlotLds;
    if (in_lds) {
        kinds[0].out = fo_s;
        kinds[1].out = fo_s + K;
        kinds[2].out = fo_s + 4 * K;
    }
    int got[3];
    collect_free3(kinds[0], kinds[1], kinds[2], fy, d.free_rows, got, sh, in_lds ? bl_s : d.slist + gb0 + eb0 + vb0,
                  in_lds ? kSlotLds : (gb1 - gb0) + (eb1 - eb0) + (vb1 - vb0));
    const int* fg = kinds[0].out;
    const int* fe = kinds[1].out;
    const int* fv = kinds[2].out;
    // event r needs grain slot r, boundary slots 3r..3r+2, vertex slots 2r, 2r+1
    const int n_ev = min(K, min(got[0], min(got[1] / 3, got[2] / 2)));
    if (tid == 0) ev0 = n_ev > 0 ? atomicAdd(&d.cnt[C_NEV], n_ev) : 0;
    __syncthreads();
    for (int r = tid; r < n_ev; r += nt) {
        int* ev = d.l_nev + 8 * (size_t)(ev0 + r);
        ev[0] = m;
        ev[1] = kept[r];
        ev[2] = fv[2 * r];
        ev[3] = fv[2 * r + 1];
        ev[4] = fe[3 * r];
        ev[5] = fe[3 * r + 1];
        ev[6] = fe[3 * r + 2];
        ev[7] = fg[r];
    }
}

// nucleate (nucleation2.cu:474-676, after its checks) for the events the
// managers listed, one wave (workgroup of 64) per event, followed by the
// recomputation of main.cu:171-177 restricted to what an event changes:
// the lengths and energies (calculus.h:83-114) of its six boundaries and the
// areas (:224-234) of its four grains.  Events are grain-disjoint: no event
// reads a vertex, boundary or grain list another one writes, so they run
// concurrently, and each event's refresh reads only its own writes.
// The lists of the three grains at the site are staged in LDS with one
// round of loads; lane 0 applies the edits of grain_add_vertex and
// grain_remove_vertex (geometry.h:629-666) to the staged copies in the
// reference's order (a grain named twice keeps one copy, so the edits
// compose), and the wave writes them back.
__global__ void __launch_bounds__(64) k_nucl_insert(Dev d) {
    __shared__ int Ls[3][kMaxStride];
    __shared__ int n_s[3];
    __shared__ double sdx[4][kMaxStride], sdy[4][kMaxStride];
    const int lane = threadIdx.x;
    const int n_ev = min(d.cnt[C_NEV], d.G);
    for (int item = blockIdx.x; item < n_ev; item += gridDim.x) {
        const int* ev = d.l_nev + 8 * (size_t)item;
        const int m = ev[0], A = ev[1], B = ev[2], C = ev[3], AB = ev[4], BC = ev[5], CA = ev[6], g = ev[7];
        const Params& P = d.prm[m];
        NuclGeom r;
        // nucleation2.cu:526 passes (scaling, eps) swapped; only dE is affected
        if (generate_grain(d, P, A, P.gb_scaling, P.gb_eps, 1, &r) != 0) continue;
        const int gl[3] = {r.g1, r.g2, r.g3};
        int sl[3];
        sl[0] = 0;
        sl[1] = gl[1] == gl[0] ? 0 : 1;
        sl[2] = gl[2] == gl[0] ? 0 : (gl[2] == gl[1] ? sl[1] : 2);
        int n_old[3];
#pragma unroll
        for (int t = 0; t < 3; t++) n_old[t] = sl[t] == t ? d.glen[gl[t]] : 0;
#pragma unroll
        for (int t = 0; t < 3; t++) {
            if (sl[t] != t) continue;
            const int* L = glist(d, gl[t]);
            for (int i = lane; i < n_old[t]; i += 64) Ls[t][i] = L[i];
        }
        // what the serial edits below read, loaded with the lists (nothing
        // writes these entries before they are read)
        int vgA[3];
#pragma unroll
        for (int k = 0; k < 3; k++) vgA[k] = d.vg[3 * A + k];
        const int ae0 = d.bvx[2 * r.AE], ae1 = d.bvx[2 * r.AE + 1], af0 = d.bvx[2 * r.AF];
        if (lane == 0) { n_s[0] = n_old[0]; n_s[1] = n_old[1]; n_s[2] = n_old[2]; }
        __syncthreads();
        if (lane == 0) {
            const int S = P.cap;
            auto ladd = [&](int t, int v, int prev) {               // grain_add_vertex
                const int n = n_s[t];
                int pos = -1;
                for (int i = 0; i < n; i++) if (Ls[t][i] == prev) { pos = i + 1; break; }
                if (pos < 0) { set_error(d, m, SEVC_E_STATE); return; }
                if (n >= S) { set_error(d, m, SEVC_E_GRAIN_FULL); return; }
                for (int q = n; q > pos; q--) Ls[t][q] = Ls[t][q - 1];
                Ls[t][pos] = v;
                n_s[t] = n + 1;
            };
            auto lremove = [&](int t, int v) {                      // grain_remove_vertex
                const int n = n_s[t];
                for (int i = 0; i < n; i++)
                    if (Ls[t][i] == v) {
                        for (int j = i; j < n - 1; j++) Ls[t][j] = Ls[t][j + 1];
                        n_s[t] = n - 1;
                        return;
                    }
            };
            // init_vertex (geometry.h:346-361) for B and C
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int v = q == 0 ? B : C;
                d.pos[2 * v] = q == 0 ? r.Bpos[0] : r.Cpos[0];
                d.pos[2 * v + 1] = q == 0 ? r.Bpos[1] : r.Cpos[1];
                d.vel[2 * v] = 0.0; d.vel[2 * v + 1] = 0.0;
                d.venergy[v] = 0.0;
                set_vertex_enabled(d, v, true);
                d.vnuc[v] = 0;
                reset_mods(d, v);
            }
            d.pos[2 * A] = r.Apos[0]; d.pos[2 * A + 1] = r.Apos[1];
            // init_boundary (geometry.h:491-504) for AB, BC, CA and the
            // boundary slots they and AD, AE, AF take at A, B, C: each
            // vertex's table was cleared, so add_boundary fills it in order
            // (vb[A] = AB, CA, AD; vb[B] = AB, BC; vb[C] = BC, CA), after
            // which the reference swaps the first two slots of B and C and
            // puts AE, AF third
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int b = q == 0 ? AB : (q == 1 ? BC : CA);
                d.bvx[2 * b] = q == 0 ? A : (q == 1 ? B : C);
                d.bvx[2 * b + 1] = q == 0 ? B : (q == 1 ? C : A);
                d.blen[b] = 0;
                d.bt[b] = SEVC_MAX_T_EXT;
                set_edge_enabled(d, b, true);
                d.bcand[b] = 0;
                d.conf[b] = -1;
            }
            d.vb[3 * A] = AB; d.vb[3 * A + 1] = CA; d.vb[3 * A + 2] = r.AD;
            d.vb[3 * B] = BC; d.vb[3 * B + 1] = AB; d.vb[3 * B + 2] = r.AE;
            d.vb[3 * C] = CA; d.vb[3 * C + 1] = BC; d.vb[3 * C + 2] = r.AF;
            // nucleate_grain (geometry.h:596-606); its list is written below
            d.glen[g] = 3;
            d.garea[g] = 0;
            d.gse[g] = 0;
            set_grain_enabled(d, g, true);
            d.gcand[g] = 0;
            d.gfix[g] = 0;
            d.gtype[g] = T_NUCLEATED;
            d.gori[g] = d.nucl_ori[A];
            for (int k = 0; k < 3; k++) if (vgA[k] == r.g3) { d.vg[3 * A + k] = g; break; }
            d.vg[3 * B] = g; d.vg[3 * B + 1] = r.g2; d.vg[3 * B + 2] = r.g3;
            d.vg[3 * C] = g; d.vg[3 * C + 1] = r.g3; d.vg[3 * C + 2] = r.g1;
            ladd(sl[0], C, A);
            if (ae0 == A) ladd(sl[1], B, ae1);
            else if (ae1 == A) ladd(sl[1], B, ae0);
            ladd(sl[2], C, A);
            ladd(sl[2], B, C);
            lremove(sl[2], A);
            if (ae0 == A) d.bvx[2 * r.AE] = B;
            else if (ae1 == A) d.bvx[2 * r.AE + 1] = B;
            if (af0 == A) d.bvx[2 * r.AF] = C;
            else d.bvx[2 * r.AF + 1] = C;
            atomicAdd((unsigned long long*)&d.ms[m].nucl, 1ull);
        }
        __syncthreads();
        // write the edited lists back (-1 after the end, as the reference's
        // removal leaves) and the new grain's list
#pragma unroll
        for (int t = 0; t < 3; t++) {
            if (sl[t] != t) continue;
            const int n_new = n_s[t];
            int* L = glist(d, gl[t]);
            for (int i = lane; i < max(n_old[t], n_new); i += 64) L[i] = i < n_new ? Ls[t][i] : -1;
            if (lane == 0) d.glen[gl[t]] = n_new;
        }
        {
            int* Lg = glist(d, g);
            for (int i = lane; i < d.S; i += 64) Lg[i] = i == 0 ? A : (i == 1 ? B : (i == 2 ? C : -1));
        }
        __syncthreads();
        // refresh: the six boundaries (lanes 0-5), then the areas of g1, g2,
        // g3 and g from the wrapped deltas of their lists, summed in list
        // order by lanes 0-3 as grain_area does
        const double Dm = P.D;
        if (lane < 6) {
            // the six edges and their ends' positions after the event: AB, BC,
            // CA between the new vertices, AD at A, AE rewired to B, AF to C
            // (arclen gives the same bits in either end order)
            const int b = lane == 0 ? AB : lane == 1 ? BC : lane == 2 ? CA : lane == 3 ? r.AD : lane == 4 ? r.AE : r.AF;
            // by selection (no pointer into r, which would put it in scratch)
            const bool a0 = lane == 0 || lane == 3, b0 = lane == 1 || lane == 4;
            const double x0 = a0 ? r.Apos[0] : (b0 ? r.Bpos[0] : r.Cpos[0]);
            const double y0 = a0 ? r.Apos[1] : (b0 ? r.Bpos[1] : r.Cpos[1]);
            const double x1 = lane == 0 ? r.Bpos[0] : lane == 1 ? r.Cpos[0] : lane == 2 ? r.Apos[0]
                            : lane == 3 ? r.Dpos[0] : lane == 4 ? r.Epos[0] : r.Fpos[0];
            const double y1 = lane == 0 ? r.Bpos[1] : lane == 1 ? r.Cpos[1] : lane == 2 ? r.Apos[1]
                            : lane == 3 ? r.Dpos[1] : lane == 4 ? r.Epos[1] : r.Fpos[1];
            d.blen[b] = arclen(x0, y0, x1, y1, Dm);
            const double e = (P.gb_eps == 0.0) ? 1.0 : boundary_compute_energy(d, b, P.gb_eps);
            d.ben[b] = e * P.gb_scaling;
        }
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int n = n_s[sl[t]];
            for (int i = 1 + lane; i < n; i += 64) {
                const int w = Ls[sl[t]][i], wp = Ls[sl[t]][i - 1];
                sdx[t][i] = wrapd(d.pos[2 * w], d.pos[2 * wp], Dm);
                sdy[t][i] = wrapd(d.pos[2 * w + 1], d.pos[2 * wp + 1], Dm);
            }
        }
        if (lane < 2) {                                  // g: A, B, C
            const int w = lane == 0 ? B : C, wp = lane == 0 ? A : B;
            sdx[3][1 + lane] = wrapd(d.pos[2 * w], d.pos[2 * wp], Dm);
            sdy[3][1 + lane] = wrapd(d.pos[2 * w + 1], d.pos[2 * wp + 1], Dm);
        }
        __syncthreads();
        if (lane < 4) {
            const int t = lane;
            const int gg = t == 0 ? gl[0] : (t == 1 ? gl[1] : (t == 2 ? gl[2] : g));
            const int st = t == 0 ? sl[0] : (t == 1 ? sl[1] : sl[2]);
            const int n = t < 3 ? n_s[st] : 3;
            if (d.genab[gg]) {
                double a = 0.0;
                if (n > 0) {
                    AreaAcc acc;
                    acc.start(0.0, 0.0);
                    for (int i = 1; i < n; i++) acc.push_delta(sdx[t][i], sdy[t][i]);
                    a = acc.finish();
                }
                d.garea[gg] = a;
            }
        }
        __syncthreads();
    }
}

}  // namespace sevc

namespace sevc {
__global__ void k_selftest(int op, int n, const double* x, const double* y, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (op == 10 || op == 11) {
        // wave scan / sum of integer inputs (every lane of the wave active)
        const int v = i < n ? (int)x[i] : 0;
        const int r = op == 10 ? wave_inclusive_scan(v) : wave_sum(v);
        if (i < n) out[i] = (double)r;
        return;
    }
    if (op == 14) {
        // wave minimum of the inputs' bit patterns (lanes past n: all ones)
        const unsigned long long v = i < n ? (unsigned long long)__double_as_longlong(x[i]) : ~0ull;
        const unsigned long long r = wave_min_u64(v);
        if (i < n) out[i] = __longlong_as_double((long long)r);
        return;
    }
    if (i >= n) return;
    double a = x[i], b = y[i], r = 0.0;
    switch (op) {
        case 0: r = sevc_atan2(a, b); break;
        case 1: r = sevc_sin(a); break;
        case 2: r = sevc_cos(a); break;
        case 3: r = sevc_exp(a); break;
        case 4: r = a / b; break;
        case 5: r = sqrt(a); break;
        case 6: r = mag(a, b); break;
        case 7: r = wrapd(a, b, 1.0); break;
        case 9: r = sqrt_1_2(a); break;
        case 12: { double c; sevc_sincos(a, &r, &c); break; }
        case 13: { double sn; sevc_sincos(a, &sn, &r); break; }
        case 8: {
            uint32_t c[4] = {(uint32_t)a, (uint32_t)b, 0u, 0u}, k[2] = {1u, 2u}, o[4];
            sevc_philox4x32_10(c, k, o);
            r = (double)o[0];
            break;
        }
        default: r = 0.0;
    }
    out[i] = r;
}
}  // namespace sevc
