// sevc_adjoint.hip -- topology-fixed windows and their adjoint (extension).
//
// Within a window the topology is held fixed (no T1/T2/nucleation), dt is
// the fixed opts.dt and the velocity is the analytic model
//   x_{n+1} = dom_mod(x_n + dt * mu * F(x_n)),   F = -grad E,
//   E = sum_b gamma_b L_b + sum_g SE_g |A_g|,    gamma_b = e_b * GB_scaling.
// The reverse pass runs lambda_n = J_n^T lambda_{n+1} with
// J_n = I - dt mu H(x_n) (H the Hessian of E; dom_mod has derivative 1), and
// accumulates the parameter terms lambda_{n+1} . dx_{n+1}/dtheta.
//
// Every kernel takes the positions it works on explicitly (a checkpoint or
// a slot of the segment buffer), one lane per vertex / grain / edge, and
// evaluates the same operations in the same order as the oracle's
// orc_window_forward / orc_adjoint_window, so results are bit-identical.
#include <hip/hip_runtime.h>
#include "sevc_dev.h"

namespace sevc {

// Boundary energies for the window (compute_boundary_energies,
// calculus.h:104-114) and their derivative e_b = d gamma_b / d GB_scaling.
__global__ void k_win_prepare(Dev d, double* ebase) {
    TABLE_BLOCK(d.eblk, d.neb)
    if (i >= bd_.e || !d.benab[i]) return;
    const double e = (P.gb_eps == 0.0) ? 1.0 : boundary_compute_energy(d, i, P.gb_eps);
    d.ben[i] = e * P.gb_scaling;
    ebase[i] = e;
}

// Grain areas at positions X (compute_grain_areas, calculus.h:224-234);
// also into the state's area array when to_state is set.
__global__ void k_win_area(Dev d, const double* X, double* A, int to_state) {
    TABLE_BLOCK(d.gblk, d.ngb)
    if (i >= bd_.e || !d.genab[i]) return;
    const double a = grain_area_at(d, X, i, P.D);
    A[i] = a;
    if (to_state) d.garea[i] = a;
}

// One window step x_n -> x_{n+1}: analytic velocity (calculus.h:143-154
// with the analytic model) and update_vertices_positions
// (calculus.h:391-404), written to a separate array so that neighbours
// still read x_n.  Disabled slots are copied.
__global__ void k_win_step(Dev d, const double* X, const double* A, double* Xn, int write_vel) {
    TABLE_BLOCK(d.vblk, d.nvb)
    if (i >= bd_.e) return;
    if (!d.ven[i]) {
        Xn[2 * i] = X[2 * i];
        Xn[2 * i + 1] = X[2 * i + 1];
        return;
    }
    double fx, fy;
    analytic_force_at(d, X, A, P, i, &fx, &fy);
    const double vx = fx * P.mobility, vy = fy * P.mobility;
    if (write_vel) {
        d.vel[2 * i] = vx;
        d.vel[2 * i + 1] = vy;
    }
    double x = X[2 * i], y = X[2 * i + 1];
    x += vx * P.dt_max;
    y += vy * P.dt_max;
    Xn[2 * i] = dom_mod(x, P.D);
    Xn[2 * i + 1] = dom_mod(y, P.D);
}

// Window end: step counter and time as k_end advances them per step.
__global__ void k_win_end(Dev d, int steps) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= d.B) return;
    const Params& P = d.prm[m];
    MState& M = d.ms[m];
    for (int k = 0; k < steps; k++) {
        M.n += 1;
        if (P.adaptive) M.t += P.dt_max;
        else M.t = M.n * P.dt_max;
    }
}

// lambda <- 0 on disabled slots; accumulators <- 0.
__global__ void k_win_init(Dev d, double* lam, double* acc_mu, double* acc_s, double* acc_se) {
    TABLE_BLOCK(d.vblk, d.nvb)
    if (i >= bd_.e) return;
    if (!d.ven[i]) {
        lam[2 * i] = 0.0;
        lam[2 * i + 1] = 0.0;
    }
    acc_mu[i] = 0.0;
    acc_s[i] = 0.0;
    for (int j = 0; j < 3; j++) acc_se[3 * i + j] = 0.0;
}

// One reverse step at vertex v, with lambda = dL/dx_{n+1} and the state
// x_n (positions X, grain areas A):
//   lambda_n  = lambda - dt mu (H lambda)_v
//   acc_mu   += dt (lambda . F_v)
//   acc_s    += dt mu (lambda . dF_v/dGB_scaling),  dF_v/ds = -sum_b e_b u_b
//   acc_se_j += dt mu (lambda . dF_v/dSE_g),        dF_v/dSE_g = -sign(A_g) dA_g/dp_v
// An edge b to w (d = wrap(p_v - p_w), L = |d| > 0, u = d/L) adds
// (gamma_b / L)(I - u u^T)(lambda_v - lambda_w) to (H lambda)_v; a grain g
// with list neighbours prev, next adds
// SE_g sign(A_g)/2 (lambda_next.y - lambda_prev.y, lambda_prev.x - lambda_next.x).
__global__ void k_win_adjoint(Dev d, const double* X, const double* A, const double* lam, double* lam2,
                              double* acc_mu, double* acc_s, double* acc_se, const double* ebase) {
    TABLE_BLOCK(d.vblk, d.nvb)
    if (i >= bd_.e) return;
    const int v = i;
    if (!d.ven[v]) {
        lam2[2 * v] = 0.0;
        lam2[2 * v + 1] = 0.0;
        return;
    }
    const double D = P.D, mu = P.mobility, dt = P.dt_max, dtm = dt * mu;
    const double x = X[2 * v], y = X[2 * v + 1];
    const double lx = lam[2 * v], ly = lam[2 * v + 1];
    double ex = 0.0, ey = 0.0, gx = 0.0, gy = 0.0;
    double hex = 0.0, hey = 0.0, hgx = 0.0, hgy = 0.0, sx = 0.0, sy = 0.0;
    for (int j = 0; j < 3; j++) {
        const int b = d.vb[3 * v + j];
        const int w = other_end(d, b, v);
        const double dx = wrapd(x, X[2 * w], D), dy = wrapd(y, X[2 * w + 1], D);
        const double L = mag(dx, dy);
        if (L > 0) {
            const double ux = dx / L, uy = dy / L;
            const double gam = d.ben[b];
            ex += gam * ux;
            ey += gam * uy;
            const double wx = lx - lam[2 * w], wy = ly - lam[2 * w + 1];
            const double uw = ux * wx + uy * wy;
            const double c = gam / L;
            hex += c * (wx - ux * uw);
            hey += c * (wy - uy * uw);
            const double eb = ebase[b];
            sx += eb * ux;
            sy += eb * uy;
        }
    }
    for (int j = 0; j < 3; j++) {
        const int g = d.vg[3 * v + j];
        const int n = d.glen[g];
        const int k = grain_contains_vertex(d, g, v);
        if (k < 0 || n < 3) continue;
        const int* Lg = glist(d, g);
        const int wp = Lg[(k + n - 1) % n], wn = Lg[(k + 1) % n];
        const double dpx = wrapd(x, X[2 * wp], D), dpy = wrapd(y, X[2 * wp + 1], D);
        const double dnx = wrapd(X[2 * wn], x, D), dny = wrapd(X[2 * wn + 1], y, D);
        const double sg = A[g] < 0 ? -1.0 : 1.0;
        const double c = d.gse[g] * sg * 0.5;
        gx += c * (dpy + dny);
        gy += c * (-(dpx + dnx));
        hgx += c * (lam[2 * wn + 1] - lam[2 * wp + 1]);
        hgy += c * (lam[2 * wp] - lam[2 * wn]);
        const double q = sg * 0.5;
        const double tse = -(lx * (q * (dpy + dny)) + ly * (q * (-(dpx + dnx))));
        acc_se[3 * v + j] += dtm * tse;
    }
    const double fx = -(ex + gx), fy = -(ey + gy);
    const double hx = hex + hgx, hy = hey + hgy;
    acc_mu[v] += dt * (lx * fx + ly * fy);
    acc_s[v] += dtm * (-(lx * sx + ly * sy));
    lam2[2 * v] = lx - dtm * hx;
    lam2[2 * v + 1] = ly - dtm * hy;
}

// Level 1 of the deterministic parameter sums: one lane per 256-vertex
// table block sums its enabled slots left to right (the chunks of the
// oracle's orc_sum_chunked).  part[lb] for mu, part[2*nvb + lb] for s.
__global__ void k_win_sum_l1(Dev d, const double* acc_mu, const double* acc_s, double* part) {
    const int lb = blockIdx.x * blockDim.x + threadIdx.x;
    if (lb >= d.nvb) return;
    const Blk bd = d.vblk[lb];
    double a = 0.0, b = 0.0;
    for (int v = bd.b; v < bd.e; v++)
        if (d.ven[v]) {
            a += acc_mu[v];
            b += acc_s[v];
        }
    part[lb] = a;
    part[2 * d.nvb + lb] = b;
}

// Upper levels per member (chunks of 256 partials left to right until one
// remains), one workgroup per member and quantity (blockIdx.y); levels
// alternate between two halves so no lane overwrites a value still read.
// out[2m] = dL/dGB_scaling, out[2m+1] = dL/dmobility.
__global__ void k_win_sum_up(Dev d, double* part, double* out) {
    const int m = blockIdx.x, which = blockIdx.y;      // 0: mu, 1: s
    if (m >= d.B) return;
    const int first = d.vblk_first[m];
    int cnt = d.vblk_first[m + 1] - first;
    double* src = part + 2 * (size_t)d.nvb * which + first;
    double* dst = src + d.nvb;
    while (cnt > 1) {
        const int nn = (cnt + 255) / 256;
        for (int c = threadIdx.x; c < nn; c += blockDim.x) {
            double s = 0.0;
            const int e = (c + 1) * 256 < cnt ? (c + 1) * 256 : cnt;
            for (int q = c * 256; q < e; q++) s += src[q];
            dst[c] = s;
        }
        __syncthreads();
        double* t = src;
        src = dst;
        dst = t;
        cnt = nn;
    }
    if (threadIdx.x == 0) out[2 * m + (which == 0 ? 1 : 0)] = cnt == 1 ? src[0] : 0.0;
}

// dL/dSE_g: the grain's vertices' terms summed in list order.
__global__ void k_win_dse(Dev d, const double* acc_se, double* dse) {
    TABLE_BLOCK(d.gblk, d.ngb)
    if (i >= bd_.e) return;
    double r = 0.0;
    if (d.genab[i]) {
        const int* L = glist(d, i);
        for (int k = 0; k < d.glen[i]; k++) {
            const int v = L[k];
            for (int j = 0; j < 3; j++)
                if (d.vg[3 * v + j] == i) {
                    r += acc_se[3 * v + j];
                    break;
                }
        }
    }
    dse[i] = r;
}

}  // namespace sevc

namespace sevc {

// ---------------------------------------------------------------------
// smooth area histogram loss (calibration), on the device
// ---------------------------------------------------------------------
//
// Member m: c_k = a_min + (k + 1/2)(a_max - a_min)/nb, width w,
//   h_k = (1/G) sum_g exp(-z_gk^2 / 2),  z_gk = (|A_g| - c_k)/w,
//   L   = sum_k (h_k - t_k)^2,
// over its enabled grains with at least 3 vertices (G of them).  Sums run in
// a fixed order (grain index within 256-grain blocks, then blocks), so the
// result does not depend on scheduling.

__device__ __forceinline__ double hist_centre(const double* rg, int nb, int k) {
    return rg[0] + (k + 0.5) * (rg[1] - rg[0]) / nb;
}

// Level 1: one workgroup per grain table block.  Stores the signed areas;
// part[lb*(nb+1) + k] = sum over the block's grains of exp(-z^2/2) for bin
// k, part[lb*(nb+1) + nb] = number of grains in the histogram.
__global__ void __launch_bounds__(256) k_hist_l1(Dev d, int nb, const double* ranges, double* hA, double* part) {
    __shared__ double a_s[kBlock];
    __shared__ unsigned char ok_s[kBlock];
    TABLE_BLOCK(d.gblk, d.ngb)
    const int t = threadIdx.x;
    const bool ok = i < bd_.e && d.genab[i] && d.glen[i] >= 3;
    double a = 0.0;
    if (ok) {
        a = grain_area_at(d, d.pos, i, P.D);
        hA[i] = a;
    } else if (i < bd_.e) {
        hA[i] = 0.0;
    }
    a_s[t] = fabs(a);
    ok_s[t] = ok ? 1 : 0;
    __syncthreads();
    const int n = bd_.e - bd_.b;
    const double* rg = ranges + 3 * bd_.m;
    if (t < nb) {
        const double c = hist_centre(rg, nb, t), w = rg[2];
        double s = 0.0;
        for (int q = 0; q < n; q++)
            if (ok_s[q]) {
                const double z = (a_s[q] - c) / w;
                s += sevc_exp(-0.5 * (z * z));
            }
        part[(size_t)lb_ * (nb + 1) + t] = s;
    } else if (t == nb) {
        int cnt = 0;
        for (int q = 0; q < n; q++) cnt += ok_s[q];
        part[(size_t)lb_ * (nb + 1) + nb] = cnt;
    }
}

// Level 2: one workgroup per member.  hist[m*nb+k] = h_k, res[m*nb+k] =
// h_k - t_k, loss[m] = sum_k (h_k - t_k)^2, gcount[m] = G.
__global__ void k_hist_member(Dev d, int nb, const double* part, const double* targets, double* hist, double* res,
                              double* loss, double* gcount) {
    __shared__ double r_s[kMaxHistBins];
    __shared__ double g_s;
    const int m = blockIdx.x;
    if (m >= d.B) return;
    const int first = d.gblk_first[m], cnt = d.gblk_first[m + 1] - first;
    const int t = threadIdx.x;
    if (t == nb) {
        double g = 0.0;
        for (int b = first; b < first + cnt; b++) g += part[(size_t)b * (nb + 1) + nb];
        g_s = g;
    }
    __syncthreads();
    if (t < nb) {
        double s = 0.0;
        for (int b = first; b < first + cnt; b++) s += part[(size_t)b * (nb + 1) + t];
        const double G = g_s > 0 ? g_s : 1.0;
        const double h = s / G;
        const double r = h - targets[(size_t)m * nb + t];
        hist[(size_t)m * nb + t] = h;
        res[(size_t)m * nb + t] = r;
        r_s[t] = r;
    }
    __syncthreads();
    if (t == 0) {
        double L = 0.0;
        for (int k = 0; k < nb; k++) L += r_s[k] * r_s[k];
        loss[m] = L;
        gcount[m] = g_s;
    }
}

// Per grain: wg = dL/dA_g = sign(A_g) sum_k 2 r_k exp(-z^2/2) (-z/w) / G.
__global__ void k_hist_weight(Dev d, int nb, const double* ranges, const double* hA, const double* res,
                              const double* gcount, double* wg) {
    TABLE_BLOCK(d.gblk, d.ngb)
    if (i >= bd_.e) return;
    if (!(d.genab[i] && d.glen[i] >= 3)) {
        wg[i] = 0.0;
        return;
    }
    const double* rg = ranges + 3 * bd_.m;
    const double* r = res + (size_t)bd_.m * nb;
    const double a = fabs(hA[i]), w = rg[2];
    const double G = gcount[bd_.m] > 0 ? gcount[bd_.m] : 1.0;
    double s = 0.0;
    for (int k = 0; k < nb; k++) {
        const double z = (a - hist_centre(rg, nb, k)) / w;
        s += 2.0 * r[k] * sevc_exp(-0.5 * (z * z)) * (-z / w);
    }
    s = s / G;
    wg[i] = hA[i] < 0 ? -s : s;
}

// Per vertex: dL/dp_v = sum over its grains of wg * dA_g/dp_v with
// dA/dx_v = (dy_prev + dy_next)/2, dA/dy_v = -(dx_prev + dx_next)/2.
__global__ void k_hist_seed(Dev d, const double* wg, double* seed) {
    TABLE_BLOCK(d.vblk, d.nvb)
    if (i >= bd_.e) return;
    const int v = i;
    double gx = 0.0, gy = 0.0;
    if (d.ven[v]) {
        const double D = P.D, x = d.pos[2 * v], y = d.pos[2 * v + 1];
        for (int j = 0; j < 3; j++) {
            const int g = d.vg[3 * v + j];
            const int n = d.glen[g];
            const int k = grain_contains_vertex(d, g, v);
            if (k < 0 || n < 3 || !d.genab[g]) continue;
            const int* L = glist(d, g);
            const int wp = L[(k + n - 1) % n], wn = L[(k + 1) % n];
            const double dpx = wrapd(x, d.pos[2 * wp], D), dpy = wrapd(y, d.pos[2 * wp + 1], D);
            const double dnx = wrapd(d.pos[2 * wn], x, D), dny = wrapd(d.pos[2 * wn + 1], y, D);
            const double c = wg[g] * 0.5;
            gx += c * (dpy + dny);
            gy += c * (-(dpx + dnx));
        }
    }
    seed[2 * v] = gx;
    seed[2 * v + 1] = gy;
}

}  // namespace sevc
