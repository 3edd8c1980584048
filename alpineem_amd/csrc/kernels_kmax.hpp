// kernels_kmax.hpp -- device kernels of the kmax (constant-k-vector) build.
//
// The reference's kmax build (main_fdtd/fdtd_solver.f90 under
// use_kmax_version) makes every field complex and the periodic walls
// Bloch-periodic.  Every update of the ordinary build has real
// coefficients, so a complex field's real and imaginary parts evolve by the
// ordinary update each; a kmax simulation here is two ordinary simulations
// (the parts) that meet only in three places, which are these kernels:
//  - the Bloch copies of the periodic walls (a complex multiply by
//    exp(-+i k (N-2) d), f90:5387-5428 for H, 6685-6736 for E);
//  - the modal plane source that replaces TF/SF (Inc of f90:8860-8885,
//    applied at f90:4558-4625 on the H plane for TE, 5857-5948 on the E plane
//    for TM): its complex value goes to the two parts;
//  - nothing else: the TE/TM plane sums (f90:7105-7190) and the port lines
//    are sums and updates with real weights, run on each part as they are.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace aemk {

// Bloch copies of one periodic axis for both parts.  (c, s): the phase
// factor exp(i th) the copy multiplies by -- for H (last cell from the
// first) th = -k (N-2) d, for E (first cell from the last) th = +k (N-2) d.
// The complex product is (a c - b s, a s + b c) as the reference forms it.
template <bool E>
__global__ void k_kmax_pbc(Grid g, int axis, float c, float s, float* ar, float* ai, float* br, float* bi) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int v = blockIdx.y * blockDim.y + threadIdx.y + 1;
    long long dst, src;
    if (axis == 0) {          // E: Ez, Ey (1,j,k) <- (NX-1,j,k);  H: Hz, Hy (NX-1,j,k) <- (1,j,k)
        if (u > g.NY - 1 || v > g.NZ - 1) return;
        dst = off(g, E ? 1 : g.NX - 1, u, v);
        src = off(g, E ? g.NX - 1 : 1, u, v);
    } else if (axis == 1) {   // E: Ez, Ex;  H: Hz, Hx
        if (u > g.NX - 1 || v > g.NZ - 1) return;
        dst = off(g, u, E ? 1 : g.NY - 1, v);
        src = off(g, u, E ? g.NY - 1 : 1, v);
    } else {                  // E: Ey, Ex;  H: Hy, Hx
        if (u > g.NX - 1 || v > g.NY - 1) return;
        dst = off(g, u, v, E ? 1 : g.NZ - 1);
        src = off(g, u, v, E ? g.NZ - 1 : 1);
    }
    const float a0 = ar[src], b0 = ai[src], a1 = br[src], b1 = bi[src];
    ar[dst] = a0 * c - b0 * s;
    ai[dst] = a0 * s + b0 * c;
    br[dst] = a1 * c - b1 * s;
    bi[dst] = a1 * s + b1 * c;
}

// The modal plane source of one step (Inc, f90:8860-8885):
//   Inc = w g(t) exp(i (th(t) - (k1 d1 + k2 d2)))
// with g(t), th(t) the per-step Gaussian factor and carrier phase (tables
// from the setup) and (k1 d1 + k2 d2) the Bloch phase of the component's own
// position.  Each part of the field gets its part of
//   F += (dt / (mu0 or eps0 * d_a)) Inc * k_exception.
// The cosine and sine of the phase are the FP32 values correctly rounded
// from FP64, which equal the host libm's (the reference's) except where that
// libm itself is not correctly rounded.
struct KmaxSrc {
    Grid g;
    int plane;                 // kpl_H or kpl_E
    int lo[3], hi[3];          // H-plane ranges of the transverse axes (the TF/SF bounds: 1..N-1 on periodic axes)
    float w[3];                // WE (TE: H plane) or WH (TM: E plane)
    float kexc[3];
    float kv[3];
    float d[3];
    float c0;                  // dt / (mu0 d_a) or dt / (eps0 d_a)
    const float* gss;          // [T] Gaussian factor of this plane
    const float* ph;           // [T] carrier phase of this plane
    float *xr, *yr, *zr;       // the real part's field components
    float *xi, *yi, *zi;       // the imaginary part's
    const int* step;
};

__device__ __forceinline__ void kmax_add(const KmaxSrc& a, float w, float gs, float tht, float kd, float exc,
                                         float* Fr, float* Fi, long long o) {
    const float th = tht - kd;
    double sd, cd;
    sincos((double)th, &sd, &cd);
    const float amp = w * gs;
    const float ir = amp * (float)cd, ii = amp * (float)sd;
    Fr[o] = Fr[o] + a.c0 * ir * exc;
    Fi[o] = Fi[o] + a.c0 * ii * exc;
}

// Source on the H plane (TE, time n-1, x = 0).  Grid over (u, v), the two
// transverse axes of the source axis A in the reference's (outer, inner)
// roles; both components at every cell.
template <int A>
__global__ void k_kmax_src_h(KmaxSrc a) {
    const int n = *a.step;
    const float gs = a.gss[n - 1], tht = a.ph[n - 1];
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y * blockDim.y + threadIdx.y;
    const Grid& g = a.g;
    if (A == 0) {            // f90:4566-4583: j = ylow..yhigh, k = zlow..zhigh
        const int j = a.lo[1] + u, k = a.lo[2] + v;
        if (j > a.hi[1] || k > a.hi[2]) return;
        const long long o = off(g, a.plane, j, k);
        kmax_add(a, a.w[1], gs, tht, a.kv[1] * (((float)j + 0.5f) * a.d[1]) + a.kv[2] * ((float)k * a.d[2]),
                 a.kexc[2], a.zr, a.zi, o);
        kmax_add(a, a.w[2], gs, tht, a.kv[1] * ((float)j * a.d[1]) + a.kv[2] * (((float)k + 0.5f) * a.d[2]),
                 a.kexc[1], a.yr, a.yi, o);
    } else if (A == 1) {     // f90:4587-4604: i = xlow..xhigh, k = zlow..zhigh
        const int i = a.lo[0] + u, k = a.lo[2] + v;
        if (i > a.hi[0] || k > a.hi[2]) return;
        const long long o = off(g, i, a.plane, k);
        kmax_add(a, a.w[0], gs, tht, a.kv[0] * (((float)i + 0.5f) * a.d[0]) + a.kv[2] * ((float)k * a.d[2]),
                 a.kexc[2], a.zr, a.zi, o);
        kmax_add(a, a.w[2], gs, tht, a.kv[0] * ((float)i * a.d[0]) + a.kv[2] * (((float)k + 0.5f) * a.d[2]),
                 a.kexc[0], a.xr, a.xi, o);
    } else {                 // f90:4608-4622: i = xlow..xhigh, j = ylow..yhigh
        const int i = a.lo[0] + u, j = a.lo[1] + v;
        if (i > a.hi[0] || j > a.hi[1]) return;
        const long long o = off(g, i, j, a.plane);
        kmax_add(a, a.w[1], gs, tht, a.kv[1] * (((float)j + 0.5f) * a.d[1]) + a.kv[0] * ((float)i * a.d[0]),
                 a.kexc[0], a.xr, a.xi, o);
        kmax_add(a, a.w[0], gs, tht, a.kv[1] * ((float)j * a.d[1]) + a.kv[0] * (((float)i + 0.5f) * a.d[0]),
                 a.kexc[1], a.yr, a.yi, o);
    }
}

// Source on the E plane (TM, time n-1/2, x = -1/2), f90:5861-5945: each
// component has its own range (index 2.. on the component's transverse
// edge axis, 1.. on the other); grid over (u, v) = indices from 1.
template <int A>
__global__ void k_kmax_src_e(KmaxSrc a) {
    const int n = *a.step;
    const float gs = a.gss[n - 1], tht = a.ph[n - 1];
    const int u = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int v = 1 + blockIdx.y * blockDim.y + threadIdx.y;
    const Grid& g = a.g;
    if (A == 0) {            // Ez: k = 1..NZ-1, j = 2..NY-1;  Ey: k = 2..NZ-1, j = 1..NY-1
        const int j = u, k = v;
        if (j > g.NY - 1 || k > g.NZ - 1) return;
        const long long o = off(g, a.plane, j, k);
        if (j >= 2)
            kmax_add(a, a.w[1], gs, tht, a.kv[1] * ((float)j * a.d[1]) + a.kv[2] * (((float)k + 0.5f) * a.d[2]),
                     a.kexc[2], a.zr, a.zi, o);
        if (k >= 2)
            kmax_add(a, a.w[2], gs, tht, a.kv[1] * (((float)j + 0.5f) * a.d[1]) + a.kv[2] * ((float)k * a.d[2]),
                     a.kexc[1], a.yr, a.yi, o);
    } else if (A == 1) {     // Ez: k = 1..NZ-1, i = 2..NX-1;  Ex: k = 2..NZ-1, i = 1..NX-1
        const int i = u, k = v;
        if (i > g.NX - 1 || k > g.NZ - 1) return;
        const long long o = off(g, i, a.plane, k);
        if (i >= 2)
            kmax_add(a, a.w[0], gs, tht, a.kv[0] * ((float)i * a.d[0]) + a.kv[2] * (((float)k + 0.5f) * a.d[2]),
                     a.kexc[2], a.zr, a.zi, o);
        if (k >= 2)
            kmax_add(a, a.w[2], gs, tht, a.kv[0] * (((float)i + 0.5f) * a.d[0]) + a.kv[2] * ((float)k * a.d[2]),
                     a.kexc[0], a.xr, a.xi, o);
    } else {                 // Ex: j = 2..NY-1, i = 1..NX-1;  Ey: j = 1..NY-1, i = 2..NX-1
        const int i = u, j = v;
        if (i > g.NX - 1 || j > g.NY - 1) return;
        const long long o = off(g, i, j, a.plane);
        if (j >= 2)
            kmax_add(a, a.w[1], gs, tht, a.kv[1] * ((float)j * a.d[1]) + a.kv[0] * (((float)i + 0.5f) * a.d[0]),
                     a.kexc[0], a.xr, a.xi, o);
        if (i >= 2)
            kmax_add(a, a.w[0], gs, tht, a.kv[1] * (((float)j + 0.5f) * a.d[1]) + a.kv[0] * ((float)i * a.d[0]),
                     a.kexc[1], a.yr, a.yi, o);
    }
}

// The modal source from a table of per-step increments built on the host
// with the host libm's sincosf -- the function the reference's CEXP calls
// (about 1% of its results are not the correctly rounded value, so a device
// evaluation would differ in the last bit there).  Cells of two groups (the
// two field components the plane receives); the table holds, per step,
// ncell real then ncell imaginary increments c0 Inc k_exception.
struct KmaxTab {
    const long long* off;      // [ncell] field offsets: group 0 first
    const float* inc;          // [T][2][ncell]
    int n0, ncell;
    float *r0, *i0, *r1, *i1;  // the two groups' fields, real and imaginary part
    const int* step;
};

__global__ void k_kmax_src_tab(KmaxTab t) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= t.ncell) return;
    const int n = *t.step;
    const float* row = t.inc + (size_t)(n - 1) * 2 * t.ncell;
    const long long o = t.off[c];
    float* Fr = c < t.n0 ? t.r0 : t.r1;
    float* Fi = c < t.n0 ? t.i0 : t.i1;
    Fr[o] = Fr[o] + row[c];
    Fi[o] = Fi[o] + row[t.ncell + c];
}

}  // namespace aemk
