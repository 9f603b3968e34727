// sevc_geom.h -- device-side periodic geometry of the vertex model.
//
// Same formulas and operation order as geometry.h of the reference; every
// translation unit that includes this must be compiled with
// -ffp-contract=off so that no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sevc {

// dom_mod, geometry.h:133-137.  Written with selects so that the compiler
// emits v_cndmask rather than divergent branches; the candidate values are
// the reference's a + D and a - D.
__device__ __forceinline__ double dom_mod(double a, double D) {
    const double lo = a + D, hi = a - D;
    const double r = (a >= D) ? hi : a;
    return (a < 0) ? lo : r;
}

// wrap_dist, geometry.h:147-152 (select form, same candidate values).
__device__ __forceinline__ double wrapd(double dest, double orig, double D) {
    const double d = dest - orig;
    const double h = D / 2.0;
    const double a = d - D, b = D + d;
    const double r = (d > +h) ? a : d;
    return (d < -h) ? b : r;
}

// Correctly rounded square root of s in [1, 2] (or NaN): the sequence the
// compiler emits for sqrt(double) -- reciprocal square root, one
// Newton-Raphson refinement of (g, h) ~ (sqrt(s), 1/(2 sqrt(s))), then two
// corrections by the exact residual s - g*g -- without its scaling of
// arguments below 2^-767 and its pass-through of 0 and infinity, neither of
// which can occur in this range.  Bit-identical to sqrt(s) there.
__device__ __forceinline__ double sqrt_1_2(double s) {
    const double y0 = __builtin_amdgcn_rsq(s);
    double g = s * y0, h = y0 * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, s);
    g = fma(d, h, g);
    d = fma(-g, g, s);
    g = fma(d, h, g);
    return g;
}

// vector2_mag, geometry.h:217-223: p*sqrt(1+(q/p)^2), p = max |.|, q = min |.|
// (1 + (q/p)^2 lies in [1, 2]; p = 0 gives NaN there, replaced by 0).
__device__ __forceinline__ double mag(double x, double y) {
    const double ax = fabs(x), ay = fabs(y);
    const double p = ax > ay ? ax : ay;
    const double q = ax < ay ? ax : ay;
    const double r = q / p;
    const double m = p * sqrt_1_2(1.0 + r * r);
    return (p == 0) ? 0.0 : m;
}

// compute_arclength, geometry.h:328-330: |wrap(a - b)|
__device__ __forceinline__ double arclen(double ax, double ay, double bx, double by, double D) {
    return mag(wrapd(ax, bx, D), wrapd(ay, by, D));
}

// Streaming form of compute_area (geometry.h:757-765 after the in-place
// unwrap of geometry.h:908-931).  The unwrapped point U_k is
// U_{k-1} + wrap(P_k - P_{k-1}) with U_0 = 0, and the shoelace term
// U_{k-1} x U_k is accumulated as soon as U_k is known, which reproduces
// the reference's left-to-right sum term by term; the closing term
// U_{n-1} x U_0 is added last.
struct AreaAcc {
    double px, py;   // previous raw point
    double ux, uy;   // previous unwrapped point
    double area;
    __device__ __forceinline__ void start(double x0, double y0) {
        px = x0; py = y0; ux = 0; uy = 0; area = 0;
    }
    __device__ __forceinline__ void push(double cx, double cy, double D) {
        double dx = wrapd(cx, px, D);
        double dy = wrapd(cy, py, D);
        px = cx; py = cy;
        push_delta(dx, dy);
    }
    // the same step from a precomputed wrapped delta (raw point not tracked)
    __device__ __forceinline__ void push_delta(double dx, double dy) {
        double nx = ux + dx, ny = uy + dy;
        area += ux * ny - uy * nx;
        ux = nx; uy = ny;
    }
    __device__ __forceinline__ double finish() {
        const double zx = 0, zy = 0;
        area += ux * zy - uy * zx;
        area *= 0.5;
        return area;
    }
};

// compute_dAdt (geometry.h:779-787) in the same streaming form: terms
// (V_k x U_{k+1}) + (U_k x V_{k+1}) in the reference's sign convention.
struct DadtAcc {
    double px, py, ux, uy, vx, vy, d;
    double u0x, u0y, v0x, v0y;
    __device__ __forceinline__ void start(double x0, double y0, double vx0, double vy0) {
        px = x0; py = y0; ux = 0; uy = 0; vx = vx0; vy = vy0; d = 0;
        u0x = 0; u0y = 0; v0x = vx0; v0y = vy0;
    }
    __device__ __forceinline__ void push(double cx, double cy, double cvx, double cvy, double D) {
        double dx = wrapd(cx, px, D), dy = wrapd(cy, py, D);
        px = cx; py = cy;
        double nx = ux + dx, ny = uy + dy;
        d += (vx * ny + ux * cvy);
        d -= (vy * nx + uy * cvx);
        ux = nx; uy = ny; vx = cvx; vy = cvy;
    }
    __device__ __forceinline__ double finish() {
        d += (vx * u0y + ux * v0y);
        d -= (vy * u0x + uy * v0x);
        return d;
    }
};

}  // namespace sevc
