// sevc_check.h -- the device check of check_data (utils.h:85-336): what
// its kernels read.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "sevc_state.h"

namespace sevc {

// Topology arrays (global slot ids), each member's slot ranges (Params
// voff/nv, eoff/ne, goff/ng) and where the per-member counts go.
struct ChkView {
    const int* vb; const int* vg; const uint8_t* ven;
    const int* bvx; const uint8_t* benab;
    const int* gvx; const int* glen; const uint8_t* genab;
    int S, V, E, G, B;
    const int* vmem; const int* emem; const int* gmem;   // owner of each slot (read when B > 1)
    const Params* prm;
    const MState* ms;    // gate != 0: only members with debug >= 1 that are active
    int gate;
    int* bad;            // violations per member (B ints)
};

// Phases of the checks inside a step (main.cu:185, :237, :269, :297).
enum { CHK_NUCLEATION = 1, CHK_TWO_GRAIN = 2, CHK_REMOVAL = 3, CHK_FLIP = 4 };

__global__ void k_check_clear(int* bad, int B);
__global__ void k_check(ChkView c);
__global__ void k_check_stop(const Params* prm, MState* ms, const int* bad, int B, int phase);

// Message of the last failed call made without a state on this thread.
const char* stateless_error();
void set_stateless_error(const std::string& msg);

}  // namespace sevc
