// sevc_state.h -- device state of a (batched) vertex-model simulation.
//
// All members of a batch live in one set of structure-of-arrays buffers;
// topology indices are global slot ids (member offset added at upload), so
// every per-entity kernel covers all members in one launch.  Per-member
// parameters and loop state are small device arrays read once per
// workgroup through the block tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sevc {

constexpr int kBlock = 256;          // elements per table block (one workgroup)
constexpr int kFixStages = 3;        // 0: two-grain, 1: T2 removal, 2: flips
constexpr int kTile = 64;            // grains per velocity tile (pass 1)
constexpr int kMaxHistBins = 64;     // bins of the device area histogram
constexpr int kNuclThreads = 1024;   // workgroup of the nucleation manager (one per member)
constexpr int kPollFast = 2048;      // candidate lists up to this length poll with registers and LDS (two per lane)
constexpr int kPollBig = 16384;      // lists up to this length poll with flags and t_ext in LDS
constexpr int kCountWaves = 4;       // table blocks per workgroup of k_count_all (one wave each)

// Per-member run parameters (from sevc_options).
struct Params {
    double D, dt_max, dt_min, dt_safety, flip_thr, mobility;
    double gb_eps, gb_scaling, mc_k, nucl_rate, nucl_scaling, tmax;
    unsigned long long seed;
    int nmax, adaptive, nucl_mode, do_nucl, gap, events, vel_model, min_grains, n_grains0;
    int gate_grains;                 // MC nucleation opens once enabled grains <= this (5 n_grains0 / 6)
    int debug;                       // debug_lvl: >= 1 runs the checks of check_data inside the step
    int cap;                         // max_vrt_per_grn of this member (lists are stored at the batch's stride S >= cap)
    int voff, nv, eoff, ne, goff, ng;
};

// Per-member loop state (main.cu:30-113), updated on the device.
struct MState {
    int n, active, done, curr_grains, allow_nucl, nucl_now, status, n_venabled;
    int chk_phase;                   // phase of a failed check (debug_lvl >= 1), else 0
    double t, dt;
    unsigned long long tau_bits;     // min positive t_ext (as ordered bits)
    long long flips, removals, two, nucl, vsteps, steps;
    double nucl_area_pct;
};

// A workgroup's range of one member's entities: [b, e).
struct Blk { int m, b, e; };

enum Counter {
    C_CAND = 0, C_TWO, C_T2, C_T2C, C_FLIPC,
    C_FIX0, C_FIX1, C_FIX2, C_DIRTY, C_NUCL, C_CONF, C_NEV, C_VFALL, C_RELABEL, C_BIGR, C_COUNT
};

enum { FIX_ADD = 0, FIX_REMOVE = 1, FIX_REPLACE = 2, FIX_NONE = 3 };   // geometry.h:17

// A vertex's four deferred grain-list edits (mod_t, geometry.h:26-31), in
// one 64-byte record so that reading or clearing them touches one line.
struct alignas(64) ModRec {
    int grn[4], vrt[4], prev[4];
    uint8_t fix[4];
};
static_assert(sizeof(ModRec) == 64, "one line per vertex");

struct Dev {
    // vertices
    double* pos; double* vel; int* vb; int* vg; uint8_t* ven;
    double* venergy; double* vparts;
    ModRec* mods;                        // deferred grain-list edits, one 64-B record per vertex
    double* nucl_r; double* nucl_ori; double* dE; double* nfac; uint8_t* vnuc;
    // boundaries
    int* bvx; double* ben; double* blen; double* bt; uint8_t* benab; uint8_t* bcand; int* conf;
    int* pgat; double* pgt;          // polling gathers: 25 columns of kPollBig, t_ext
    uint8_t* blab;                  // per edge, from k_label: bit 0 candidate and enabled, bit 1 enabled
    int* l_flipped; int* n_flipped;  // edges flipped since the last step start (energies to rescale)
    int* vmem; int* emem; int* gmem;  // owning member of each vertex / edge / grain slot (batches)
    // grains
    int* gvx; int* glen; double* garea; double* gdadt; double* gori; double* gse;
    uint8_t* genab; uint8_t* gtype; uint8_t* gcand; int* gfix;
    int S;
    int V, E, G;                          // total slots over members
    // members
    const Params* prm; MState* ms; int B;
    const Blk* vblk; const Blk* eblk; const Blk* gblk;
    int nvb, neb, ngb;
    const Blk* gtile; int ntile;         // tiles of tile_T consecutive grains of one member (velocity pass 1)
    int tile_T;
    const Blk* vtile; int nvtile;        // vertex tiles of the one-pass velocity
    double* vterm;                       // 6 per vertex: stored-energy terms (x,y) per grain slot
    int4* slist;                         // free-slot search: (kind, block, rank) lists, ngb+neb+nvb entries
    int* vblk_first; int* eblk_first; int* gblk_first;   // first table block of each member (B+1)
    // work lists and scratch
    int* cnt;
    int* l_cand; int* l_two; int* l_t2; int* l_t2c; int* l_flipc;
    int* l_fix[kFixStages]; int* l_dirty; int* l_nucl; int* l_conf; int* l_vfall;
    int* l_bigr;                          // grains longer than the lane-per-grain velocity pass holds
    int* nrec; double* nal;               // rate-mode candidate sites: (site or -1, its 3 grains), shortest edge
    int* l_nev;                           // nucleation events of the step: (member, A, B, C, AB, BC, CA, g)
    int2* bc_v; int2* bc_e; int2* bc_g;   // per table block: (enabled, free) slot counts
    double* partial;                      // per grain block partial sums (2 levels)
    int* nscr;                            // nucleation scratch, 12 ints per grain slot
    int diag;                             // compute export diagnostics this step
    int debug;                            // a member runs the checks of check_data (debug_lvl >= 1)
    int* chk_bad;                         // violations per member found by the last check
    int poll_fast;                        // longest candidate list polled from registers (<= kPollFast)
    int poll_lds;                         // longest candidate list polled from LDS (<= kPollBig)
    int free_rows;                        // blocks per kind of one free-slot search round (0: the default)
};

}  // namespace sevc
