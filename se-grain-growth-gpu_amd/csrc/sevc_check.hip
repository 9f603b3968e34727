// sevc_check.hip -- check_data (utils.h:85-336) on the device: the
// conditions that make the reference report invalid data, counted per
// member with one lane per edge, vertex and grain slot.  The same kernels
// serve the checks a step runs for members with debug_lvl >= 1 (after
// nucleation, the two-triangle stage, grain removal and flips,
// main.cu:185-196,237-248,269-280,297-308) and the checks of a live state
// or of a host mesh copied to the device.
//
// The reference holds pointers, which cannot leave the arrays; here an id
// outside its member's slot range counts once for the entity holding it
// and the checks that would follow it are skipped, and a grain list is read
// up to min(vlen, S) entries, so no lane reads outside the arrays whatever
// the input holds.  The host checker (sevc_check_mesh) counts the same way.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "sevc_state.h"
#include "sevc_check.h"
#include "../../include/sevc.h"

namespace sevc {

namespace {

struct Ranges {
    int v0, v1, e0, e1, g0, g1;
    __device__ bool inV(int x) const { return x >= v0 && x < v1; }
    __device__ bool inE(int x) const { return x >= e0 && x < e1; }
    __device__ bool inG(int x) const { return x >= g0 && x < g1; }
};

// Enabled edge b (utils.h:91-158).  Returns its violations.
__device__ int check_edge(const ChkView& c, const Ranges& r, int b) {
    const int a = c.bvx[2 * b], e = c.bvx[2 * b + 1];
    if (!r.inV(a) || !r.inV(e)) return 1;
    int bad = 0;
    if (a == e) bad++;                                              // :94-98
    if (!c.ven[a] || !c.ven[e]) bad++;                              // :99-105
    int va[3], ve[3], ga[3], ge[3];
    for (int j = 0; j < 3; j++) {
        va[j] = c.vb[3 * a + j]; ve[j] = c.vb[3 * e + j];
        ga[j] = c.vg[3 * a + j]; ge[j] = c.vg[3 * e + j];
    }
    if (va[0] != b && va[1] != b && va[2] != b) bad++;              // :107-118
    if (ve[0] != b && ve[1] != b && ve[2] != b) bad++;              // :119-130
    bool dis = false;                                               // :132-138
    for (int j = 0; j < 3; j++) if (!r.inE(ve[j]) || !c.benab[ve[j]]) dis = true;
    if (dis) bad++;
    dis = false;                                                    // :139-145
    for (int j = 0; j < 3; j++) if (!r.inE(va[j]) || !c.benab[va[j]]) dis = true;
    if (dis) bad++;
    // :147-153, grains_intersect (geometry.h:837-848); fewer than two
    // shared grains counts too (the reference would follow a null pointer)
    int gic[2] = {-1, -1}, k = 0;
    for (int i = 0; i < 3 && k < 2; i++)
        for (int j = 0; j < 3 && k < 2; j++)
            if (ga[i] == ge[j]) gic[k++] = ga[i];
    if (k < 2 || gic[0] == gic[1]) bad++;
    return bad;
}

// Enabled vertex v (utils.h:161-263).
__device__ int check_vertex(const ChkView& c, const Ranges& r, int v) {
    int vb[3], vg[3];
    for (int j = 0; j < 3; j++) { vb[j] = c.vb[3 * v + j]; vg[j] = c.vg[3 * v + j]; }
    for (int j = 0; j < 3; j++)
        if (!r.inE(vb[j]) || !r.inG(vg[j])) return 1;
    int bad = 0;
    if (vb[0] == vb[1] || vb[1] == vb[2] || vb[2] == vb[0]) bad++;  // :164-169
    if (vg[0] == vg[1] || vg[1] == vg[2] || vg[2] == vg[0]) bad++;  // :182-187
    int bi[3], be[3];
    for (int j = 0; j < 3; j++) { bi[j] = c.bvx[2 * vb[j]]; be[j] = c.bvx[2 * vb[j] + 1]; }
    for (int j = 0; j < 3; j++) if (!c.benab[vb[j]]) bad++;        // :189-194
    for (int j = 0; j < 3; j++) if (!c.genab[vg[j]]) bad++;        // :196-201
    for (int j = 0; j < 3; j++) if (bi[j] == be[j]) bad++;         // :203-211
    int nb[3];                                                      // :213-225
    for (int j = 0; j < 3; j++) nb[j] = bi[j] == v ? be[j] : bi[j];
    if (nb[0] == nb[1] || nb[1] == nb[2] || nb[2] == nb[0]) bad++;
    for (int j = 0; j < 3; j++) if (bi[j] != v && be[j] != v) bad++;   // :227-234
    for (int j = 0; j < 3; j++) {                                   // :236-253
        const int g = vg[j];
        const int n = min(c.glen[g], c.S);
        const int* L = c.gvx + (size_t)g * c.S;
        bool found = false;
        for (int k = 0; k < n; k++) found |= (L[k] == v);
        if (!found) bad++;
    }
    return bad;
}

// Enabled grain g (utils.h:266-320; the area sum and the warnings are not
// counted).
__device__ int check_grain(const ChkView& c, const Ranges& r, int g) {
    const int* L = c.gvx + (size_t)g * c.S;
    const int vlen = c.glen[g];
    const int n = min(vlen, c.S);
    int bad = 0;
    for (int j = 0; j < n; j++) {                                   // :280-285
        const int v = L[j];
        if (!r.inV(v) || !c.ven[v]) bad++;
    }
    for (int j = 0; j < n; j++) {                                   // :287-301
        const int v = L[j];
        if (!r.inV(v)) continue;
        if (c.vg[3 * v] != g && c.vg[3 * v + 1] != g && c.vg[3 * v + 2] != g) { bad++; break; }
    }
    int cnt = 0;                                                    // :303-316
    for (int j = 0; j < c.S; j++) cnt += L[j] >= 0 ? 1 : 0;
    if (cnt != vlen) bad++;
    return bad;
}

}  // namespace

__global__ void k_check_clear(int* bad, int B) {
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < B; m += gridDim.x * blockDim.x) bad[m] = 0;
}

// One lane per slot: edges [0, E), then vertices, then grains.
__global__ void __launch_bounds__(256) k_check(ChkView c) {
    const long long total = (long long)c.E + c.V + c.G;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        int kind, i, m;
        if (t < c.E) {
            kind = 0; i = (int)t;
            if (!c.benab[i]) continue;
            m = c.B == 1 ? 0 : c.emem[i];
        } else if (t < (long long)c.E + c.V) {
            kind = 1; i = (int)(t - c.E);
            if (!c.ven[i]) continue;
            m = c.B == 1 ? 0 : c.vmem[i];
        } else {
            kind = 2; i = (int)(t - c.E - c.V);
            if (!c.genab[i]) continue;
            m = c.B == 1 ? 0 : c.gmem[i];
        }
        const Params& P = c.prm[m];
        if (c.gate && !(P.debug >= 1 && c.ms[m].active)) continue;
        const Ranges r{P.voff, P.voff + P.nv, P.eoff, P.eoff + P.ne, P.goff, P.goff + P.ng};
        const int bad = kind == 0 ? check_edge(c, r, i) : (kind == 1 ? check_vertex(c, r, i) : check_grain(c, r, i));
        if (bad > 0) atomicAdd(&c.bad[m], bad);
    }
}

// After a check inside a step: a member with debug_lvl >= 1 and any
// violation stops here, as the reference leaves its loop (main.cu:192-195
// and the like): status SEVC_E_CHECK, the phase recorded, inactive for
// the rest of the step.
__global__ void k_check_stop(const Params* prm, MState* ms, const int* bad, int B, int phase) {
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < B; m += gridDim.x * blockDim.x) {
        MState& M = ms[m];
        if (prm[m].debug < 1 || !M.active || bad[m] == 0) continue;
        M.status = SEVC_E_CHECK;
        M.chk_phase = phase;
        M.active = 0;
        M.done = 1;
    }
}


// ---------------------------------------------------------------------
// check_data of meshes already in device memory, allocating nothing
// ---------------------------------------------------------------------

// The scratch of sevc_check_meshes_device_noalloc: one parameter record
// (the slot ranges k_check reads) and one violation counter per mesh, for
// at most kNoAllocMeshes meshes per call.  Each GPU holds its own copy for
// the life of the process; calls that overlap on one GPU share it.
constexpr int kNoAllocMeshes = 64;
constexpr size_t kNoAllocPrmBytes = (sizeof(Params) * kNoAllocMeshes + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = kNoAllocPrmBytes + sizeof(int) * kNoAllocMeshes;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Slot counts of the meshes of one call, passed by value.
struct NoAllocSizes {
    int n;
    int nv[kNoAllocMeshes], ne[kNoAllocMeshes], ng[kNoAllocMeshes];
};

// Fill the parameter records and clear the counters (a kernel, not a
// memset or a copy, so that the call can be captured into a graph).
__global__ void k_noalloc_setup(unsigned char* scr, NoAllocSizes z) {
    Params* prm = reinterpret_cast<Params*>(scr);
    int* cnt = reinterpret_cast<int*>(scr + kNoAllocPrmBytes);
    const int m = threadIdx.x;
    if (m >= z.n) return;
    Params P{};
    P.voff = 0; P.nv = z.nv[m]; P.eoff = 0; P.ne = z.ne[m]; P.goff = 0; P.ng = z.ng[m];
    P.debug = 1;
    prm[m] = P;
    cnt[m] = 0;
}

__global__ void k_noalloc_out(const unsigned char* scr, int n, int* bad) {
    const int* cnt = reinterpret_cast<const int*>(scr + kNoAllocPrmBytes);
    const int m = threadIdx.x;
    if (m < n) bad[m] = cnt[m];
}

namespace {
thread_local std::string g_stateless_err;
std::mutex g_scratch_mu;
void* g_scratch_ptr[256] = {};        // device address of `scratch` per device id

int stateless_fail(int code, const std::string& msg) {
    g_stateless_err = msg;
    return code;
}
}  // namespace

const char* stateless_error() { return g_stateless_err.c_str(); }
void set_stateless_error(const std::string& msg) { g_stateless_err = msg; }

}  // namespace sevc

using namespace sevc;

extern "C" int sevc_check_meshes_device_noalloc(const sevc_mesh* meshes, int32_t n_meshes, int32_t* bad,
                                                void* stream) {
    if (!meshes || !bad) return stateless_fail(SEVC_E_INVALID_ARG, "null meshes or output");
    if (n_meshes <= 0 || n_meshes > kNoAllocMeshes)
        return stateless_fail(SEVC_E_INVALID_ARG, "n_meshes = " + std::to_string(n_meshes) +
                                                      " outside [1, " + std::to_string(kNoAllocMeshes) +
                                                      "], the meshes one call's static scratch holds");
    NoAllocSizes z{};
    z.n = n_meshes;
    for (int m = 0; m < n_meshes; ++m) {
        const sevc_mesh& h = meshes[m];
        if (h.n_vertices <= 0 || h.n_boundaries <= 0 || h.n_grains <= 0 || h.max_vrt_per_grn <= 0 || !h.v_bnd ||
            !h.v_grn || !h.v_enabled || !h.b_vtx || !h.b_enabled || !h.g_vtx || !h.g_vlen || !h.g_enabled)
            return stateless_fail(SEVC_E_INVALID_ARG, "mesh " + std::to_string(m) + ": empty or missing arrays");
        z.nv[m] = h.n_vertices; z.ne[m] = h.n_boundaries; z.ng[m] = h.n_grains;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 256)
        return stateless_fail(SEVC_E_NO_DEVICE, "no current HIP device");
    unsigned char* scr = nullptr;
    {
        // the scratch's device address on this device, looked up once (the
        // symbol's host-side name is no device address)
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return stateless_fail(SEVC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                                      hipGetErrorString(e));
            g_scratch_ptr[dev] = p;
        }
        scr = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(kNoAllocMeshes), 0, st, scr, z);
    const Params* prm = reinterpret_cast<const Params*>(scr);
    int* cnt = reinterpret_cast<int*>(scr + kNoAllocPrmBytes);
    for (int m = 0; m < n_meshes; ++m) {
        const sevc_mesh& h = meshes[m];
        ChkView c{};
        c.vb = h.v_bnd; c.vg = h.v_grn; c.ven = h.v_enabled;
        c.bvx = h.b_vtx; c.benab = h.b_enabled;
        c.gvx = h.g_vtx; c.glen = h.g_vlen; c.genab = h.g_enabled;
        c.S = h.max_vrt_per_grn; c.V = h.n_vertices; c.E = h.n_boundaries; c.G = h.n_grains; c.B = 1;
        c.prm = prm + m;
        c.gate = 0;
        c.bad = cnt + m;
        const long long slots = (long long)h.n_vertices + h.n_boundaries + h.n_grains;
        const int grid = static_cast<int>(slots / 256 + 1 < 8192 ? slots / 256 + 1 : 8192);
        hipLaunchKernelGGL(k_check, dim3(grid), dim3(256), 0, st, c);
    }
    hipLaunchKernelGGL(k_noalloc_out, dim3(1), dim3(kNoAllocMeshes), 0, st, static_cast<const unsigned char*>(scr),
                       static_cast<int>(n_meshes), bad);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return stateless_fail(SEVC_E_HIP, std::string("launch: ") + hipGetErrorString(le));
    return SEVC_OK;
}
