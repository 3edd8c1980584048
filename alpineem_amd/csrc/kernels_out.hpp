// kernels_out.hpp -- field exports and the numeric watchdog of libaem.
//
// k_video     per-step cell-centred 2-D slice of all six components
//             (fdtd_solver.f90:6833-6990), kept on the device for the whole
//             run and written as <name>_video_*.bin after the loop
//             (f90:8788-8807);
// k_full      cell-centred copy of the whole grid for the per-step
//             <name>_full_*.bin records (f90:6992-7003);
// k_watch     sets a flag when any field value is NaN or infinite.  The
//             reference has no such check: an unstable deck (Courant factor
//             below its 0.98 limit, f90:1312) runs on to NaN silently.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace aemk {

struct Video {
    int slice, loc;      // slice axis (0 x, 1 y, 2 z) and its 1-based cell index
    int v1, v2, T;       // slice extents and the run's step count
    float* V[6];         // [T][v2][v1] per component: Ex, Ey, Ez, Hx, Hy, Hz
};

// One cell of the slice per thread.  E is the four-edge average at the
// step, H the two-face sum over 4 (f90:6856-6858); H is also written one
// frame ahead, which the next step overwrites (f90:6863-6873).  The step
// counter holds the current (1-based) step.
__global__ void k_video(Grid g, Video vd, const float* __restrict__ Ex, const float* __restrict__ Ey,
                        const float* __restrict__ Ez, const float* __restrict__ Hx, const float* __restrict__ Hy,
                        const float* __restrict__ Hz, const int* __restrict__ step) {
    const int a = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
    const int b = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
    if (a > vd.v1 || b > vd.v2) return;
    const int n = *step;
    int i, j, k;
    if (vd.slice == 0) { i = vd.loc; j = a; k = b; }
    else if (vd.slice == 1) { j = vd.loc; i = a; k = b; }
    else { k = vd.loc; i = a; j = b; }
    const long long o = off(g, i, j, k), px = g.PX, pl = g.plane;
    const long long ps = (long long)vd.v1 * vd.v2;
    const long long w = (long long)(a - 1) + (long long)vd.v1 * (b - 1);
    const long long cur = (long long)(n - 1) * ps + w;
    vd.V[0][cur] = (Ex[o] + Ex[o + px] + Ex[o + pl] + Ex[o + px + pl]) / 4.0f;
    vd.V[1][cur] = (Ey[o] + Ey[o + 1] + Ey[o + pl] + Ey[o + 1 + pl]) / 4.0f;
    vd.V[2][cur] = (Ez[o] + Ez[o + 1] + Ez[o + px] + Ez[o + 1 + px]) / 4.0f;
    const float hx = (Hx[o] + Hx[o + 1]) / 4.0f;
    const float hy = (Hy[o] + Hy[o + px]) / 4.0f;
    const float hz = (Hz[o] + Hz[o + pl]) / 4.0f;
    vd.V[3][cur] = hx;
    vd.V[4][cur] = hy;
    vd.V[5][cur] = hz;
    if (n < vd.T) {
        vd.V[3][cur + ps] = hx;
        vd.V[4][cur + ps] = hy;
        vd.V[5][cur + ps] = hz;
    }
}

// Cell-centred fields of cells [1,nx]x[1,ny]x[1,nz] (the user grid without
// the periodic ghost layer) into out[c][(i-1) + nx((j-1) + ny(k-1))]: E the
// four-edge average, H the two-face average (f90:6995-7018).
__global__ void k_full(Grid g, int nx, int ny, int nz, float* const* __restrict__ out, const float* __restrict__ Ex,
                       const float* __restrict__ Ey, const float* __restrict__ Ez, const float* __restrict__ Hx,
                       const float* __restrict__ Hy, const float* __restrict__ Hz) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
    const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
    const int k = (int)blockIdx.z + 1;
    if (i > nx || j > ny || k > nz) return;
    const long long o = off(g, i, j, k), px = g.PX, pl = g.plane;
    const long long w = (long long)(i - 1) + (long long)nx * ((long long)(j - 1) + (long long)ny * (k - 1));
    out[0][w] = (Ex[o] + Ex[o + px] + Ex[o + pl] + Ex[o + px + pl]) / 4.0f;
    out[1][w] = (Ey[o] + Ey[o + 1] + Ey[o + pl] + Ey[o + 1 + pl]) / 4.0f;
    out[2][w] = (Ez[o] + Ez[o + 1] + Ez[o + px] + Ez[o + 1 + px]) / 4.0f;
    out[3][w] = (Hx[o] + Hx[o + 1]) / 2.0f;
    out[4][w] = (Hy[o] + Hy[o + px]) / 2.0f;
    out[5][w] = (Hz[o] + Hz[o + pl]) / 2.0f;
}

struct WatchSet {
    const float* f[6];
};

// *flag = 1 if any of the six arrays (n4 float4 each) holds a NaN or an
// infinity.  Grid-stride, one float4 per thread and iteration; a flag that
// is already set stays set.
__global__ void __launch_bounds__(256) k_watch(WatchSet ws, long long n4, int* __restrict__ flag) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    bool bad = false;
    for (int c = 0; c < 6; ++c) {
        const float4* p = reinterpret_cast<const float4*>(ws.f[c]);
        for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
            const float4 v = p[q];
            bad |= !(__builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z) &&
                     __builtin_isfinite(v.w));
        }
    }
    if (bad) *flag = 1;   // every writer stores the same value
}

}  // namespace aemk
