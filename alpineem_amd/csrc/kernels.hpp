// kernels.hpp -- HIP kernels of the FDTD leapfrog step for gfx950 (MI355X).
//
// One step of the reference loop (main_fdtd/fdtd_solver.f90:4440-8299) is a
// fixed sequence of launches on one stream:
//   H: bulk -> sheets -> TF/SF (y,z,x faces) -> ports -> CPML (x,y,z slabs) -> PBC
//   E: bulk -> sheets -> Drude -> TF/SF (y,z,x) -> ports -> CPML (x,y,z) -> PBC -> E_old
//   outputs: plane sums, NTFF currents + retarded-time binning, phase centre,
//            step counter
// Decks with nothing between the bulk H and bulk E updates of a core cell
// (CPML, ports, materials; no sheets, TF/SF, PBC, Drude) run the bulk H and
// E of the core box as one single-pass launch (k_step_sp) between two buffer
// sets, the y/z CPML slabs as ping-pong bulk launches after it.
// Every FP32 expression keeps the reference's left-to-right order and the
// file is compiled with -ffp-contract=off, so the bulk, CPML, sheet, Drude
// and port arithmetic is bit-identical to the serial reference.
//
// Layout: every 3-D array uses off(i,j,k) = (i-1) + PX*((j-1) + NY*(k-1)),
// i,j,k 1-based (x fastest); plane = PX*NY.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace aemk {

struct Grid {
    int NX, NY, NZ, PX;
    long long plane;
};

__device__ __forceinline__ long long off(const Grid& g, int i, int j, int k) {
    return (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * (long long)(k - 1));
}

// Division of a field quantity by a real constant d.  The ordinary build
// divides reals: x / d.  The kmax build's fields are complex, and Fortran
// divides a complex by a real as complex by complex; its flang build does
// that with the C99 routine __divsc3, which for a divisor (d, 0) computes,
// with e = logb(d) and d' = d 2^-e, each part as ((x d') / (d' d')) 2^-e.
// That is not x / d to the last bit, so a kmax simulation (cplx = 1) uses it.
struct RDiv {
    float d;       // the divisor
    float ds;      // d 2^-e
    float dd;      // ds * ds, rounded
    int e;         // logb(d)
    int cplx;
};

inline RDiv make_rdiv(float d, bool cplx) {
    RDiv r{d, 0.0f, 0.0f, 0, cplx ? 1 : 0};
    if (cplx) {
        r.e = ilogbf(d);
        r.ds = scalbnf(d, -r.e);
        r.dd = r.ds * r.ds;
    }
    return r;
}

__device__ __forceinline__ float rdiv(float x, const RDiv& r) {
    return r.cplx ? ldexpf((x * r.ds) / r.dd, -r.e) : x / r.d;
}

// ---------------------------------------------------------------------------
// CPML state (H: f90:5227-5375, E: f90:6522-6670).  psi of face f / slot s is
// stored restricted to its slab, with the slab coordinate t replacing the
// normal index:  x faces [k][j][t],  y faces [k][t][i],  z faces [t][j][i].
// ---------------------------------------------------------------------------
struct Cpml {
    int n[6];                  // npml per face
    const float* b[6];         // bh (H) or be (E)
    const float* c[6];         // ch or ce
    float* psi[6][2];          // per face, two components
    float* pso[6][2];          // where a ping-pong bulk launch stores the new psi (x faces)
};

// Slab membership along one axis for H (layers 1..n1-1 and N+1-n2..N-1,
// local index q = i or N-i) and for E (layers 2..n1 and N+1-n2..N-1, q = i or
// N+1-i).  Returns the face (2a or 2a+1) or -1; q is the 1-based coefficient
// index, t the 0-based slab storage index.
__device__ __forceinline__ int slab_h(int i, int N, int n1, int n2, int& q) {
    if (i <= n1 - 1) { q = i; return 0; }
    if (i >= N + 1 - n2) { q = N - i; return 1; }
    return -1;
}
__device__ __forceinline__ int slab_e(int i, int N, int n1, int n2, int& q, int& t) {
    if (i >= 2 && i <= n1) { q = i; t = i - 2; return 0; }
    if (i >= N + 1 - n2 && i <= N - 1) { q = N + 1 - i; t = q - 2; return 1; }
    return -1;
}

// ---------------------------------------------------------------------------
// Bulk launches cover sets of boxes of cells.  With CPML fusion the domain is
// split into a core box (every x, y and z outside the y/z slabs; only the
// x-slab test is compiled in, which costs no occupancy) and one launch over
// the y-slab and z-slab boxes with every slab test compiled in
// (C = kCx|kCy|kCz).  A launch is a 1-D grid; each box owns a contiguous
// range of work-groups, ordered x fastest, then y, then z chunks.
// ---------------------------------------------------------------------------
enum : int { kCx = 1, kCy = 2, kCz = 4 };

// Bulk work-group tile: 64 x 4 columns (one wave64 per 64-cell x row, two
// 128-byte lines), each thread marching Box::kc planes.  Measured at 400^3
// on MI355X against 32x8, 128x2, 64x8, 128x4 and 128x8 tiles with 4-16
// planes and an XCD-grouped block order: 64x4 and 64x8 with 4 planes were
// fastest (the x-halo line is shared by twice as many cells as with 32-wide
// rows); grouping blocks by XCD was slower in every case.
constexpr int kTX = 64, kTY = 4;

struct Box {
    int i0, j0, k0;   // first cell (1-based)
    int ni, nj, nk;   // extents
    int kc;           // planes marched by one thread
};

constexpr int kMaxBoxes = 4;

struct BoxSet {
    Box b[kMaxBoxes];
    int first[kMaxBoxes + 1];   // first work-group of each box; first[n] = total
    int n;
};

// Box of this work-group and its (x, y, z-chunk) block coordinates.  The
// selection uses constant indices only, so the box table stays in kernel
// arguments (scalar registers) instead of a private copy.
__device__ __forceinline__ Box box_of(const BoxSet& s, int& bx, int& by, int& bz) {
    const int g = (int)blockIdx.x;
    Box b = s.b[0];
    int f = s.first[0];
#pragma unroll
    for (int q = 1; q < kMaxBoxes; ++q)
        if (q < s.n && g >= s.first[q]) {
            b = s.b[q];
            f = s.first[q];
        }
    const int nbx = (b.ni + kTX - 1) / kTX, nby = (b.nj + kTY - 1) / kTY;
    const int l = g - f;
    bx = l % nbx;
    by = (l / nbx) % nby;
    bz = l / (nbx * nby);
    return b;
}

// Non-temporal access for the cell's own field values, which are read and
// written exactly once per launch (the neighbour reads stay cached).
template <class T>
__device__ __forceinline__ T ld_once(const T* p) { return __builtin_nontemporal_load(p); }
template <class T>
__device__ __forceinline__ void st_once(T v, T* p) { __builtin_nontemporal_store(v, p); }

// The same by byte offset from an array base.
template <class T, class OffT>
__device__ __forceinline__ T ldo(const T* p, OffT ob) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + ob);
}
template <class OffT>
__device__ __forceinline__ float ldo_once(const float* p, OffT ob) {
    return __builtin_nontemporal_load(reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + ob));
}
template <class OffT>
__device__ __forceinline__ void sto_once(float v, float* p, OffT ob) {
    __builtin_nontemporal_store(v, reinterpret_cast<float*>(reinterpret_cast<char*>(p) + ob));
}

// ---------------------------------------------------------------------------
// Bulk H (f90:4442-4484).  One thread per (i,j) column marching b.kc planes;
// Ex(k+1), Ey(k+1) are carried in registers to the next plane.  Slab cells
// of the box then add their CPML corrections in the reference's per-cell
// order (x, then y, then z slab), reusing the loaded neighbours.
// ---------------------------------------------------------------------------
template <int C, bool PP, typename OffT>
__device__ __forceinline__ void h_bulk_body(const Grid& g, const Box& b, int bx, int by, int bz,
                                            float* __restrict__ Hx, float* __restrict__ Hy, float* __restrict__ Hz,
                                            float* __restrict__ Hxo, float* __restrict__ Hyo,
                                            float* __restrict__ Hzo, const float* __restrict__ Ex,
                                            const float* __restrict__ Ey, const float* __restrict__ Ez,
                                            const float* __restrict__ dhx, const float* __restrict__ dhy,
                                            const float* __restrict__ dhz, float DB, const Cpml& cp, float dx,
                                            float dy, float dz) {
    const int i = b.i0 + bx * kTX + (int)threadIdx.x;
    const int j = b.j0 + by * kTY + (int)threadIdx.y;
    if (i >= b.i0 + b.ni || j >= b.j0 + b.nj) return;
    const int k0 = b.k0 + bz * b.kc;
    const int k1 = min(k0 + b.kc, b.k0 + b.nk) - 1;
    const float dx_i = dhx[i - 1], dy_j = dhy[j - 1];
    int qx = 0, qy = 0;
    const int fx = (C & kCx) ? slab_h(i, g.NX, cp.n[0], cp.n[1], qx) : -1;
    const int fy = (C & kCy) ? slab_h(j, g.NY, cp.n[2], cp.n[3], qy) : -1;
    // byte offsets from the array bases (OffT: 32 bits where the arrays stay
    // under 4 GiB, so one VGPR addresses every array)
    const OffT PB = (OffT)g.PX * 4u, ZB = (OffT)g.plane * 4u;
    OffT o = (OffT)off(g, i, j, k0) * 4u;
    float ey_k = ldo(Ey, o), ex_k = ldo(Ex, o);
    for (int k = k0; k <= k1; ++k) {
        const float ey_k1 = ldo(Ey, o + ZB);
        const float ex_k1 = ldo(Ex, o + ZB);
        const float ez = ldo(Ez, o);
        const float ez_xp = ldo(Ez, o + 4u);
        const float ez_yp = ldo(Ez, o + PB);
        const float ey_xp = ldo(Ey, o + 4u);
        const float ex_yp = ldo(Ex, o + PB);
        const float dz_k = dhz[k - 1];
        float hx = ldo_once(Hx, o) + DB * ((ez - ez_yp) * dy_j + (ey_k1 - ey_k) * dz_k);
        float hy = ldo_once(Hy, o) + DB * ((ez_xp - ez) * dx_i + (ex_k - ex_k1) * dz_k);
        float hz = ldo_once(Hz, o) + DB * ((ey_k - ey_xp) * dx_i + (ex_yp - ex_k) * dy_j);
        if (C) {
            if ((C & kCx) && fx >= 0) {                      // psi_Hyx, psi_Hzx
                const int f = fx;
                const int ntf = cp.n[f] - 1;
                const long long po = (long long)(qx - 1) + (long long)ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
                const float cb = cp.b[f][qx - 1], cc = cp.c[f][qx - 1];
                float p0 = cb * cp.psi[f][0][po] + cc * (ez_xp - ez) / dx;
                (PP ? cp.pso[f][0] : cp.psi[f][0])[po] = p0;
                hy = hy + DB * p0;
                float p1 = cb * cp.psi[f][1][po] + cc * (ey_k - ey_xp) / dx;
                (PP ? cp.pso[f][1] : cp.psi[f][1])[po] = p1;
                hz = hz + DB * p1;
            }
            if ((C & kCy) && fy >= 0) {                      // psi_Hxy, psi_Hzy
                const int f = 2 + fy;
                const int ntf = cp.n[f] - 1;
                const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)(qy - 1) + (long long)ntf * (k - 1));
                const float cb = cp.b[f][qy - 1], cc = cp.c[f][qy - 1];
                float p0 = cb * cp.psi[f][0][po] + cc * (ez - ez_yp) / dy;
                (PP ? cp.pso[f][0] : cp.psi[f][0])[po] = p0;
                hx = hx + DB * p0;
                float p1 = cb * cp.psi[f][1][po] + cc * (ex_yp - ex_k) / dy;
                (PP ? cp.pso[f][1] : cp.psi[f][1])[po] = p1;
                hz = hz + DB * p1;
            }
            int qz = 0;
            const int fz = (C & kCz) ? slab_h(k, g.NZ, cp.n[4], cp.n[5], qz) : -1;
            if ((C & kCz) && fz >= 0) {                      // psi_Hxz, psi_Hyz
                const int f = 4 + fz;
                const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * (qz - 1));
                const float cb = cp.b[f][qz - 1], cc = cp.c[f][qz - 1];
                float p0 = cb * cp.psi[f][0][po] + cc * (ey_k1 - ey_k) / dz;
                (PP ? cp.pso[f][0] : cp.psi[f][0])[po] = p0;
                hx = hx + DB * p0;
                float p1 = cb * cp.psi[f][1][po] + cc * (ex_k - ex_k1) / dz;
                (PP ? cp.pso[f][1] : cp.psi[f][1])[po] = p1;
                hy = hy + DB * p1;
            }
        }
        sto_once(hx, PP ? Hxo : Hx, o);
        sto_once(hy, PP ? Hyo : Hy, o);
        sto_once(hz, PP ? Hzo : Hz, o);
        ey_k = ey_k1;
        ex_k = ex_k1;
        o += ZB;
    }
}

// ---------------------------------------------------------------------------
// Bulk E (f90:5667-5710) with the per-cell coefficient tuple index.
// Component ranges: Ex i>=1,j>=2,k>=2; Ey i>=2,j>=1,k>=2; Ez i>=2,j>=2,k>=1.
// Slab cells of the box append their CPML corrections as in k_h_bulk.
// ---------------------------------------------------------------------------
template <typename IdxT, int C, bool PP, typename OffT>
__device__ __forceinline__ void e_bulk_body(const Grid& g, const Box& b, int bx, int by, int bz,
                                            float* __restrict__ Ex, float* __restrict__ Ey, float* __restrict__ Ez,
                                            float* __restrict__ Exo, float* __restrict__ Eyo,
                                            float* __restrict__ Ezo, const float* __restrict__ Hx, const float* __restrict__ Hy,
                                            const float* __restrict__ Hz, const IdxT* __restrict__ idx,
                                            const float* __restrict__ tup, const float* __restrict__ dex,
                                            const float* __restrict__ dey, const float* __restrict__ dez,
                                            const Cpml& cp, float dx, float dy, float dz) {
    const int i = b.i0 + bx * kTX + (int)threadIdx.x;
    const int j = b.j0 + by * kTY + (int)threadIdx.y;
    if (i >= b.i0 + b.ni || j >= b.j0 + b.nj) return;
    const int k0 = b.k0 + bz * b.kc;
    const int k1 = min(k0 + b.kc, b.k0 + b.nk) - 1;
    const float dx_i = dex[i - 1], dy_j = dey[j - 1];
    const bool ix = i >= 2, jy = j >= 2;
    int qx = 0, tx = 0, qy = 0, ty = 0;
    const int fx = (C & kCx) ? slab_e(i, g.NX, cp.n[0], cp.n[1], qx, tx) : -1;
    const int fy = (C & kCy) ? slab_e(j, g.NY, cp.n[2], cp.n[3], qy, ty) : -1;
    const OffT PB = (OffT)g.PX * 4u, ZB = (OffT)g.plane * 4u;
    OffT o = (OffT)off(g, i, j, k0) * 4u;   // byte offset, as in h_bulk_body
    OffT oi = (OffT)off(g, i, j, k0) * (OffT)sizeof(IdxT);
    const OffT ZI = (OffT)g.plane * (OffT)sizeof(IdxT);
    float hx_km = 0.f, hy_km = 0.f;
    if (k0 >= 2) {
        hx_km = ldo(Hx, o - ZB);
        hy_km = ldo(Hy, o - ZB);
    }
    // i-1 / j-1 neighbours through clamped offsets: at i = 1 (j = 1) the load
    // reads the cell itself, and every component that would use it (the
    // outer PEC tangential E) is left unwritten, so all loads of an
    // iteration issue together without branches.
    const OffT dxm = ix ? 4u : 0u, dym = jy ? PB : (OffT)0;
    int t_next = (int)ldo(idx, oi);   // coefficient index, loaded one plane ahead
    for (int k = k0; k <= k1; ++k) {
        const float hx = ldo(Hx, o), hy = ldo(Hy, o), hz = ldo(Hz, o);
        const int t = t_next * 6;
        t_next = (int)ldo(idx, oi + (k < k1 ? ZI : (OffT)0));
        const float dz_k = dez[k - 1];
        const bool ex_on = jy && k >= 2, ey_on = ix && k >= 2, ez_on = ix && jy;
        const float hz_ym = ldo(Hz, o - dym);
        const float hz_xm = ldo(Hz, o - dxm);
        const float hy_xm = ldo(Hy, o - dxm);
        const float hx_ym = ldo(Hx, o - dym);
        const float ex0 = ldo_once(Ex, o), ey0 = ldo_once(Ey, o), ez0 = ldo_once(Ez, o);
        const float gax = tup[t + 0], gbx = tup[t + 1], gay = tup[t + 2], gby = tup[t + 3], gaz = tup[t + 4],
                    gbz = tup[t + 5];
        float ex = gax * ex0 + gbx * ((hz - hz_ym) * dy_j + (hy_km - hy) * dz_k);
        float ey = gay * ey0 + gby * ((hz_xm - hz) * dx_i + (hx - hx_km) * dz_k);
        float ez = gaz * ez0 + gbz * ((hy - hy_xm) * dx_i + (hx_ym - hx) * dy_j);
        if (C) {
            if ((C & kCx) && fx >= 0) {                      // psi_Eyx, psi_Ezx
                const int f = fx;
                const int ntf = cp.n[f] - 1;
                const long long po = (long long)tx + (long long)ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
                const float cb = cp.b[f][qx - 1], cc = cp.c[f][qx - 1];
                if (k >= 2) {
                    float p = cb * cp.psi[f][0][po] + cc * (hz_xm - hz) / dx;
                    (PP ? cp.pso[f][0] : cp.psi[f][0])[po] = p;
                    ey = ey + gby * p;
                }
                if (jy) {
                    float p = cb * cp.psi[f][1][po] + cc * (hy - hy_xm) / dx;
                    (PP ? cp.pso[f][1] : cp.psi[f][1])[po] = p;
                    ez = ez + gbz * p;
                }
            }
            if ((C & kCy) && fy >= 0) {                      // psi_Exy, psi_Ezy
                const int f = 2 + fy;
                const int ntf = cp.n[f] - 1;
                const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)ty + (long long)ntf * (k - 1));
                const float cb = cp.b[f][qy - 1], cc = cp.c[f][qy - 1];
                if (k >= 2) {
                    float p = cb * cp.psi[f][0][po] + cc * (hz - hz_ym) / dy;
                    (PP ? cp.pso[f][0] : cp.psi[f][0])[po] = p;
                    ex = ex + gbx * p;
                }
                if (ix) {
                    float p = cb * cp.psi[f][1][po] + cc * (hx_ym - hx) / dy;
                    (PP ? cp.pso[f][1] : cp.psi[f][1])[po] = p;
                    ez = ez + gbz * p;
                }
            }
            int qz = 0, tz = 0;
            const int fz = (C & kCz) ? slab_e(k, g.NZ, cp.n[4], cp.n[5], qz, tz) : -1;
            if ((C & kCz) && fz >= 0) {                      // psi_Exz, psi_Eyz
                const int f = 4 + fz;
                const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * tz);
                const float cb = cp.b[f][qz - 1], cc = cp.c[f][qz - 1];
                if (jy) {
                    float p = cb * cp.psi[f][0][po] + cc * (hy_km - hy) / dz;
                    (PP ? cp.pso[f][0] : cp.psi[f][0])[po] = p;
                    ex = ex + gbx * p;
                }
                if (ix) {
                    float p = cb * cp.psi[f][1][po] + cc * (hx - hx_km) / dz;
                    (PP ? cp.pso[f][1] : cp.psi[f][1])[po] = p;
                    ey = ey + gby * p;
                }
            }
        }
        // components outside the reference's update ranges get their own
        // value back; storing every component keeps the loads unconditional
        sto_once(ex_on ? ex : ex0, PP ? Exo : Ex, o);
        sto_once(ey_on ? ey : ey0, PP ? Eyo : Ey, o);
        sto_once(ez_on ? ez : ez0, PP ? Ezo : Ez, o);
        hx_km = hx;
        hy_km = hy;
        o += ZB;
        oi += ZI;
    }
}

// ---------------------------------------------------------------------------
// Bulk kernels.  The single-simulation form takes its arguments by value and
// a box set; the batched form runs many independent simulations in one
// launch, each work-group finding its (member, box) in a table of entries
// sorted by first work-group.  Both call the same body, so a member of a
// batch computes exactly what it computes alone.
// ---------------------------------------------------------------------------
struct HArgs {
    Grid g;
    float *Hx, *Hy, *Hz;
    float *Hxo, *Hyo, *Hzo;   // new H of a ping-pong launch (unused in place)
    const float *Ex, *Ey, *Ez;
    const float *dhx, *dhy, *dhz;
    float DB;
    Cpml cp;
    float dx, dy, dz;
};

struct EArgs {
    Grid g;
    float *Ex, *Ey, *Ez;
    float *Exo, *Eyo, *Ezo;   // new E of a ping-pong launch (unused in place)
    const float *Hx, *Hy, *Hz;
    const void* idx;   // uint8_t or uint16_t per the kernel's IdxT
    const float* tup;
    const float *dex, *dey, *dez;
    Cpml cp;
    float dx, dy, dz;
};

struct BatchEntry {
    int first;    // first work-group of this box in the launch
    int member;   // index into the argument array
    Box box;
};

// Entry of this work-group (binary search on `first`, uniform per group) and
// its (x, y, z-chunk) block coordinates inside the entry's box.
__device__ __forceinline__ BatchEntry entry_of(const BatchEntry* __restrict__ e, int n, int& bx, int& by, int& bz) {
    const int g = (int)blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e[mid].first <= g) lo = mid;
        else hi = mid - 1;
    }
    const BatchEntry r = e[lo];
    const int nbx = (r.box.ni + kTX - 1) / kTX, nby = (r.box.nj + kTY - 1) / kTY;
    const int l = g - r.first;
    bx = l % nbx;
    by = (l / nbx) % nby;
    bz = l / (nbx * nby);
    return r;
}

template <int C, bool PP = false, typename OffT = uint64_t>
__global__ void __launch_bounds__(kTX * kTY) k_h_bulk(HArgs a, BoxSet bs) {
    int bx, by, bz;
    const Box b = box_of(bs, bx, by, bz);
    h_bulk_body<C, PP, OffT>(a.g, b, bx, by, bz, a.Hx, a.Hy, a.Hz, a.Hxo, a.Hyo, a.Hzo, a.Ex, a.Ey, a.Ez, a.dhx, a.dhy,
                       a.dhz, a.DB, a.cp, a.dx, a.dy, a.dz);
}

template <int C, typename OffT = uint64_t>
__global__ void __launch_bounds__(kTX * kTY) k_h_bulk_batch(const HArgs* __restrict__ args,
                                                            const BatchEntry* __restrict__ ents, int n_ents) {
    int bx, by, bz;
    const BatchEntry e = entry_of(ents, n_ents, bx, by, bz);
    const HArgs& a = args[e.member];
    h_bulk_body<C, false, OffT>(a.g, e.box, bx, by, bz, a.Hx, a.Hy, a.Hz, a.Hx, a.Hy, a.Hz, a.Ex, a.Ey, a.Ez, a.dhx, a.dhy,
                          a.dhz, a.DB, a.cp, a.dx, a.dy, a.dz);
}

template <typename IdxT, int C, bool PP = false, typename OffT = uint64_t>
__global__ void __launch_bounds__(kTX * kTY) k_e_bulk(EArgs a, BoxSet bs) {
    int bx, by, bz;
    const Box b = box_of(bs, bx, by, bz);
    e_bulk_body<IdxT, C, PP, OffT>(a.g, b, bx, by, bz, a.Ex, a.Ey, a.Ez, a.Exo, a.Eyo, a.Ezo, a.Hx, a.Hy, a.Hz,
                             static_cast<const IdxT*>(a.idx), a.tup, a.dex, a.dey, a.dez, a.cp, a.dx, a.dy, a.dz);
}

template <typename IdxT, int C, typename OffT = uint64_t>
__global__ void __launch_bounds__(kTX * kTY) k_e_bulk_batch(const EArgs* __restrict__ args,
                                                            const BatchEntry* __restrict__ ents, int n_ents) {
    int bx, by, bz;
    const BatchEntry e = entry_of(ents, n_ents, bx, by, bz);
    const EArgs& a = args[e.member];
    e_bulk_body<IdxT, C, false, OffT>(a.g, e.box, bx, by, bz, a.Ex, a.Ey, a.Ez, a.Ex, a.Ey, a.Ez, a.Hx, a.Hy, a.Hz,
                                static_cast<const IdxT*>(a.idx), a.tup, a.dex, a.dey, a.dez, a.cp, a.dx, a.dy, a.dz);
}

// ---------------------------------------------------------------------------
// Single-pass step of the core box: the H and the E update of one leapfrog
// step in one launch, from the fields at the start of the step (E, H and the
// x-face psi) into a second set of buffers.  In the core box nothing touches
// a cell between its bulk H and bulk E updates when the deck has no sheets,
// TF/SF, PBC or Drude media (ports only read H there), so E(k) can follow
// H(k) at once: each field is read and written once per step instead of
// twice read and once written per phase (49 instead of 73 B per cell).
//
// Every wave is independent (no LDS, no barrier): it owns a 64-cell x row at
// one y and marches b.kc planes.  The E update needs new H from outside the
// wave: the x neighbour comes from the lane below (a lane shuffle), lane 0
// forms Hy, Hz of the cell i0-1 itself; the row j-1 (Hx, Hz) is formed by
// every lane for its column; Hx, Hy of the plane k0-1 before the first
// plane.  A value formed again this way comes from the step-start state by
// the same expression as its owner's, so it is bit-identical; only the owner
// stores H and the new psi.  The y and z CPML slabs (outside the core box)
// run after this launch as ping-pong bulk launches (k_h_bulk / k_e_bulk
// with PP).
// ---------------------------------------------------------------------------
constexpr int kSY = 4;   // rows (waves) per work-group

struct SpArgs {
    Grid g;
    const float *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;   // fields at the start of the step
    float *Exo, *Eyo, *Ezo, *Hxo, *Hyo, *Hzo;   // fields at its end
    const void* idx;                            // uint8_t or uint16_t per IdxT
    const uint32_t* idx4;                       // uint8_t indices packed 4 z planes per word (IdxT uint8_t)
    const float* tup;
    int n_tup;                                  // coefficient tuples (at most 256 with uint8_t indices)
    const float *dex, *dey, *dez, *dhx, *dhy, *dhz;
    float DB;
    Cpml cpH, cpE;                              // x faces: psi read, pso written
    float dx, dy, dz;
};

// With uint8_t indices the single pass reads the index of four z planes
// with one dword load (word (i, j, q) holds the bytes of planes 4q+1..4q+4)
// and the coefficient tuples from LDS: one byte load and six table loads
// per cell and plane took 6% of the launch at 400^3.
constexpr int kSpTupMax = 256;

// The packed copy from the byte indices (after they change on the device).
__global__ void k_pack_idx4(const uint8_t* idx, uint32_t* idx4, long long plane, int nz) {
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= plane * ((nz + 3) / 4)) return;
    const long long q = w / plane, c = w % plane;
    uint32_t v = 0u;
    for (int s = 0; s < 4; ++s) {
        const long long k = 4 * q + s;
        if (k < nz) v |= (uint32_t)idx[k * plane + c] << (8 * s);
    }
    idx4[w] = v;
}

// Field access by 32-bit byte offset from a uniform base: one VGPR per
// distinct offset, shared by all arrays (global loads with a scalar base).
__device__ __forceinline__ float ldb(const float* p, uint32_t ob) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + ob);
}
__device__ __forceinline__ void stb(float v, float* p, uint32_t ob) {
    __builtin_nontemporal_store(v, reinterpret_cast<float*>(reinterpret_cast<char*>(p) + ob));
}

// Values across the lanes of a wave (DPP wave shifts): lane L gets lane
// L-1's v (lane 0: first), or lane L+1's v (lane 63: last).
__device__ __forceinline__ float from_lane_below(float v, float first) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(first), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_lane_above(float v, float last) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(last), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

constexpr int kSpRows = 2;   // y rows marched together by one wave

template <typename IdxT, int C>
__global__ void __launch_bounds__(64 * kSY) k_step_sp(SpArgs a, Box b) {
    constexpr int R = kSpRows;
    constexpr bool kPacked = sizeof(IdxT) == 1;
    __shared__ float stup[kPacked ? kSpTupMax * 6 : 1];
    if (kPacked) {   // before any wave leaves: every wave reaches the barrier
        const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
        for (int q = tid; q < a.n_tup * 6; q += (int)(blockDim.x * blockDim.y)) stup[q] = a.tup[q];
        __syncthreads();
    }
    const Grid& g = a.g;
    const float* __restrict__ Ex = a.Ex;
    const float* __restrict__ Ey = a.Ey;
    const float* __restrict__ Ez = a.Ez;
    const float* __restrict__ Hx = a.Hx;
    const float* __restrict__ Hy = a.Hy;
    const float* __restrict__ Hz = a.Hz;
    const char* __restrict__ idx = static_cast<const char*>(a.idx);
    const float* __restrict__ tup = a.tup;
    const float DB = a.DB;
    const int nbx = (b.ni + 63) / 64, nby = (b.nj + kSY * R - 1) / (kSY * R);
    // Tiles in launch order, x fastest.  (Dealing contiguous tile ranges to
    // each XCD, so that neighbours share an L2, cut this launch's reads from
    // 2.0x to 1.2x the algorithmic bytes, but took 1% longer here.)
    const int l = (int)blockIdx.x;
    const int bx = l % nbx, by = (l / nbx) % nby, bz = l / (nbx * nby);
    const int lane = (int)threadIdx.x;
    const int i0 = b.i0 + bx * 64;
    const int imax = b.i0 + b.ni - 1, jmax = b.j0 + b.nj - 1;
    // rows jw .. jw+nr-1 of this wave (threadIdx.y is one value per wave)
    const int jw = __builtin_amdgcn_readfirstlane(b.j0 + (by * kSY + (int)threadIdx.y) * R);
    if (jw > jmax) return;
    const int nr = min(R, jmax - jw + 1);
    const int i = i0 + lane;
    const bool act = i <= imax;
    const int it = min(i, imax);              // tables and slab tests
    const int ic = min(i, imax + 1);          // loads: the lane past the last cell holds its +x neighbour
    const int k0 = b.k0 + bz * b.kc;
    const int k1 = min(k0 + b.kc, b.k0 + b.nk) - 1;
    const bool ix = i >= 2;
    const bool hrow = jw >= 2;                // the row jw-1 exists
    const int ih = max(i0 - 1, 1);            // the cell i0-1 (its H is lane 0's x neighbour)
    const int inx = min(i0 + 64, imax + 1);   // the first cell past this tile
    const float dxh = a.dhx[it - 1], dxe = a.dex[it - 1], dxc = a.dhx[ih - 1];
    const float dyhm = a.dhy[max(jw - 1, 1) - 1];
    float dyh[R], dye[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int jr = min(jw + r, jmax);
        dyh[r] = a.dhy[jr - 1];
        dye[r] = a.dey[jr - 1];
    }
    const uint32_t NY = (uint32_t)g.NY;
    // x-slab CPML of the lane's cell (H, E) and of the cell i0-1: coefficients
    // and psi offsets fixed over the march.  The two x faces of a slot share
    // one allocation (face 1 after face 0, the same layout in both buffer
    // sets), so the lane keeps an element offset from the face-0 base; psi
    // index n*((j-1) + NY*(k-1)) on top
    int qxh = 0, qxe = 0, txe = 0, qxc = 0;
    const int fxh = (C & kCx) ? slab_h(it, g.NX, a.cpH.n[0], a.cpH.n[1], qxh) : -1;
    const int fxe = (C & kCx) ? slab_e(it, g.NX, a.cpE.n[0], a.cpE.n[1], qxe, txe) : -1;
    const int fxc = (C & kCx) ? slab_h(ih, g.NX, a.cpH.n[0], a.cpH.n[1], qxc) : -1;
    const float* __restrict__ ph0 = a.cpH.psi[0][0];
    const float* __restrict__ ph1 = a.cpH.psi[0][1];
    const float* __restrict__ pe0 = a.cpE.psi[0][0];
    const float* __restrict__ pe1 = a.cpE.psi[0][1];
    float* __restrict__ qh0 = a.cpH.pso[0][0];
    float* __restrict__ qh1 = a.cpH.pso[0][1];
    float* __restrict__ qe0 = a.cpE.pso[0][0];
    float* __restrict__ qe1 = a.cpE.pso[0][1];
    const uint32_t fH1 = (uint32_t)(a.cpH.psi[1][0] - a.cpH.psi[0][0]);   // face 1 after face 0
    const uint32_t fE1 = (uint32_t)(a.cpE.psi[1][0] - a.cpE.psi[0][0]);
    float cbh = 0.f, cch = 0.f, cbe = 0.f, cce = 0.f, cbc = 0.f, ccc = 0.f;
    uint32_t nth = 0, nte = 0, ntc = 0, lh = 0, le = 0, lc = 0;
    if ((C & kCx) && fxh >= 0) {
        const int f = fxh;
        nth = (uint32_t)(a.cpH.n[f] - 1);
        cbh = a.cpH.b[f][qxh - 1];
        cch = a.cpH.c[f][qxh - 1];
        lh = (f ? fH1 : 0u) + (uint32_t)(qxh - 1);
    }
    if ((C & kCx) && fxe >= 0) {
        const int f = fxe;
        nte = (uint32_t)(a.cpE.n[f] - 1);
        cbe = a.cpE.b[f][qxe - 1];
        cce = a.cpE.c[f][qxe - 1];
        le = (f ? fE1 : 0u) + (uint32_t)txe;
    }
    if ((C & kCx) && fxc >= 0) {
        const int f = fxc;
        ntc = (uint32_t)(a.cpH.n[f] - 1);
        cbc = a.cpH.b[f][qxc - 1];
        ccc = a.cpH.c[f][qxc - 1];
        lc = (f ? fH1 : 0u) + (uint32_t)(qxc - 1);
    }
    uint32_t prow = (uint32_t)(jw - 1) + NY * (uint32_t)(k0 - 1);   // (j-1) + NY*(k-1) of the row jw
    // byte offsets (the host admits grids whose arrays stay under 4 GiB); row r at + r*PB
    const uint32_t PB = (uint32_t)g.PX * 4u, ZB = (uint32_t)g.plane * 4u;
    const uint32_t dH = hrow ? PB : 0u;       // ob - dH: the row jw-1 (the row jw itself when absent)
    uint32_t ob = (uint32_t)off(g, ic, jw, k0) * 4u;
    uint32_t oc = (uint32_t)off(g, ih, jw, k0) * 4u;    // the cell i0-1, row jw
    uint32_t on = (uint32_t)off(g, inx, jw, k0) * 4u;   // the cell past the tile, row jw
    float exk[R], eyk[R], exc[R], hxm[R], hym[R];
    int t[R], tn[R];
    const uint32_t IB = (uint32_t)sizeof(IdxT), ZI = (uint32_t)g.plane * IB;
    const uint32_t PI = (uint32_t)g.PX * IB;
    uint32_t oi = (uint32_t)off(g, ic, jw, k0) * IB;
    const uint32_t ow = (uint32_t)off(g, ic, jw, 1), PW = (uint32_t)g.PX, QW = (uint32_t)g.plane;
    uint32_t wd[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        exk[r] = eyk[r] = exc[r] = hxm[r] = hym[r] = 0.f;
        t[r] = tn[r] = 0;
        wd[r] = 0u;
        if (r < nr) {
            const uint32_t o = ob + r * PB;
            exk[r] = ldb(Ex, o);
            eyk[r] = ldb(Ey, o);
            exc[r] = ldb(Ex, oc + r * PB);
            if (kPacked) {
                wd[r] = a.idx4[ow + r * PW + (uint32_t)((k0 - 1) >> 2) * QW];
                t[r] = (int)((wd[r] >> (8 * ((k0 - 1) & 3))) & 0xffu) * 6;
            } else {
                t[r] = (int)*reinterpret_cast<const IdxT*>(idx + oi + r * PI) * 6;
            }
            if (k0 >= 2) {                    // Hx, Hy of the plane below the chunk
                const int j = jw + r;
                const uint32_t om = o - ZB;
                const float eyb = ldb(Ey, om), exb = ldb(Ex, om);
                const float ez = ldb(Ez, om), ez_xp = ldb(Ez, om + 4u), ez_yp = ldb(Ez, om + PB);
                const float dzm = a.dhz[k0 - 2];
                hxm[r] = ldb(Hx, om) + DB * ((ez - ez_yp) * dyh[r] + (eyk[r] - eyb) * dzm);
                hym[r] = ldb(Hy, om) + DB * ((ez_xp - ez) * dxh + (exb - exk[r]) * dzm);
                if ((C & kCx) && fxh >= 0)
                    hym[r] = hym[r] + DB * (cbh * ph0[lh + nth * ((uint32_t)(j - 1) + NY * (uint32_t)(k0 - 2))] +
                                            cch * (ez_xp - ez) / a.dx);
            }
        }
    }
    float eyh = ldb(Ey, ob - dH);             // Ey(k) of the row jw-1, carried
    for (int k = k0; k <= k1; ++k) {
        const float dzh = a.dhz[k - 1], dze = a.dez[k - 1];
        // every load of the plane first
        float eyk1[R], exk1[R], ez[R], hx0[R], hy0[R], hz0[R], ezn[R], eyn[R];
        float eyc[R], ezc[R], exc1[R], excy[R], hyc0[R], hzc0[R], ezi[R], eyi[R];
        float ga[R][6];
        float sh0[R], sh1[R], se0[R], se1[R], sc0[R], sc1[R];
        if (k < k1) oi += ZI;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) continue;
            const uint32_t o = ob + r * PB, c = oc + r * PB, u = on + r * PB;
            eyk1[r] = ldb(Ey, o + ZB);
            exk1[r] = ldb(Ex, o + ZB);
            ez[r] = ldb(Ez, o);
            hx0[r] = ldb(Hx, o);
            hy0[r] = ldb(Hy, o);
            hz0[r] = ldb(Hz, o);
            ezn[r] = ldb(Ez, u);              // one address per wave: the next tile's first cell
            eyn[r] = ldb(Ey, u);
            eyc[r] = ldb(Ey, c);              // the cell i0-1 and its +x, +y, +z neighbours
            ezc[r] = ldb(Ez, c);
            exc1[r] = ldb(Ex, c + ZB);
            excy[r] = ldb(Ex, c + PB);
            hyc0[r] = ldb(Hy, c);
            hzc0[r] = ldb(Hz, c);
            if (kPacked) {
#pragma unroll
                for (int q = 0; q < 6; ++q) ga[r][q] = stup[t[r] + q];
                // index of plane k+1 (0-based k): a new word starts every four planes
                if (k < k1 && (k & 3) == 0) wd[r] = a.idx4[ow + r * PW + (uint32_t)(k >> 2) * QW];
                tn[r] = (int)((wd[r] >> (8 * (k & 3))) & 0xffu) * 6;
            } else {
#pragma unroll
                for (int q = 0; q < 6; ++q) ga[r][q] = tup[t[r] + q];
                tn[r] = (int)*reinterpret_cast<const IdxT*>(idx + oi + r * PI) * 6;
            }
            const uint32_t pr = prow + (uint32_t)r;
            sh0[r] = sh1[r] = se0[r] = se1[r] = sc0[r] = sc1[r] = 0.f;
            if ((C & kCx) && fxh >= 0) {
                sh0[r] = ph0[lh + nth * pr];
                sh1[r] = ph1[lh + nth * pr];
            }
            if ((C & kCx) && fxe >= 0) {
                se0[r] = pe0[le + nte * pr];
                se1[r] = pe1[le + nte * pr];
            }
            if ((C & kCx) && fxc >= 0) {
                sc0[r] = ph0[lc + ntc * pr];
                sc1[r] = ph1[lc + ntc * pr];
            }
        }
        const uint32_t oT = ob + (uint32_t)nr * PB;   // the row past the last: +y neighbours
        const float ezT = ldb(Ez, oT), exT = ldb(Ex, oT);
        const uint32_t oH = ob - dH;                  // the row jw-1
        const float ezh = ldb(Ez, oH), eyh1 = ldb(Ey, oH + ZB), exh = ldb(Ex, oH);
        const float hxh0 = ldb(Hx, oH), hzh0 = ldb(Hz, oH), eyhn = ldb(Ey, on - dH);
        const float srh = ((C & kCx) && fxh >= 0) ? ph1[lh + nth * (prow - (hrow ? 1u : 0u))] : 0.f;
        // new Hx, Hz of the row jw-1 (from the step-start state, as its owner forms them)
        float hxp, hzp;
        {
            const float eyh_xp = from_lane_above(eyh, eyhn);
            hxp = hxh0 + DB * ((ezh - ez[0]) * dyhm + (eyh1 - eyh) * dzh);
            hzp = hzh0 + DB * ((eyh - eyh_xp) * dxh + (exk[0] - exh) * dyhm);
            if ((C & kCx) && fxh >= 0) hzp = hzp + DB * (cbh * srh + cch * (eyh - eyh_xp) / a.dx);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) continue;
            const int j = jw + r;
            const bool jy = j >= 2;
            const uint32_t o = ob + r * PB;
            const uint32_t pr = prow + (uint32_t)r;
            const float ez_xp = from_lane_above(ez[r], ezn[r]), ey_xp = from_lane_above(eyk[r], eyn[r]);
            const float ez_yp = (r + 1 < nr) ? ez[(r + 1) % R] : ezT;
            const float ex_yp = (r + 1 < nr) ? exk[(r + 1) % R] : exT;
            // new H of the cell
            float hx = hx0[r] + DB * ((ez[r] - ez_yp) * dyh[r] + (eyk1[r] - eyk[r]) * dzh);
            float hy = hy0[r] + DB * ((ez_xp - ez[r]) * dxh + (exk[r] - exk1[r]) * dzh);
            float hz = hz0[r] + DB * ((eyk[r] - ey_xp) * dxh + (ex_yp - exk[r]) * dyh[r]);
            if ((C & kCx) && fxh >= 0) {                      // psi_Hyx, psi_Hzx
                const float p0 = cbh * sh0[r] + cch * (ez_xp - ez[r]) / a.dx;
                const float p1 = cbh * sh1[r] + cch * (eyk[r] - ey_xp) / a.dx;
                if (act) {
                    qh0[lh + nth * pr] = p0;
                    qh1[lh + nth * pr] = p1;
                }
                hy = hy + DB * p0;
                hz = hz + DB * p1;
            }
            if (act) {
                stb(hx, a.Hxo, o);
                stb(hy, a.Hyo, o);
                stb(hz, a.Hzo, o);
            }
            // new Hy, Hz of the cell i0-1 (lane 0's x neighbour); E of the cell i0 is lane 0's
            ezi[r] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(ez[r])));
            eyi[r] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(eyk[r])));
            float hyc = hyc0[r] + DB * ((ezi[r] - ezc[r]) * dxc + (exc[r] - exc1[r]) * dzh);
            float hzc = hzc0[r] + DB * ((eyc[r] - eyi[r]) * dxc + (excy[r] - exc[r]) * dyh[r]);
            if ((C & kCx) && fxc >= 0) {
                hyc = hyc + DB * (cbc * sc0[r] + ccc * (ezi[r] - ezc[r]) / a.dx);
                hzc = hzc + DB * (cbc * sc1[r] + ccc * (eyc[r] - eyi[r]) / a.dx);
            }
            const float hy_xm = from_lane_below(hy, hyc), hz_xm = from_lane_below(hz, hzc);
            // new E of the cell
            const float* gq = ga[r];
            float ex = gq[0] * exk[r] + gq[1] * ((hz - hzp) * dye[r] + (hym[r] - hy) * dze);
            float ey = gq[2] * eyk[r] + gq[3] * ((hz_xm - hz) * dxe + (hx - hxm[r]) * dze);
            float ezn_ = gq[4] * ez[r] + gq[5] * ((hy - hy_xm) * dxe + (hxp - hx) * dye[r]);
            if ((C & kCx) && fxe >= 0) {                      // psi_Eyx, psi_Ezx
                if (k >= 2) {
                    const float q = cbe * se0[r] + cce * (hz_xm - hz) / a.dx;
                    if (act) qe0[le + nte * pr] = q;
                    ey = ey + gq[3] * q;
                }
                if (jy) {
                    const float q = cbe * se1[r] + cce * (hy - hy_xm) / a.dx;
                    if (act) qe1[le + nte * pr] = q;
                    ezn_ = ezn_ + gq[5] * q;
                }
            }
            if (act) {
                const bool ex_on = jy && k >= 2, ey_on = ix && k >= 2, ez_on = ix && jy;
                stb(ex_on ? ex : exk[r], a.Exo, o);
                stb(ey_on ? ey : eyk[r], a.Eyo, o);
                stb(ez_on ? ezn_ : ez[r], a.Ezo, o);
            }
            hxp = hx;
            hzp = hz;
            hxm[r] = hx;
            hym[r] = hy;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) continue;
            exk[r] = exk1[r];
            eyk[r] = eyk1[r];
            exc[r] = exc1[r];
            t[r] = tn[r];
        }
        eyh = eyh1;
        ob += ZB;
        oc += ZB;
        on += ZB;
        prow += NY;
    }
}

// ---------------------------------------------------------------------------
// Separate CPML slab kernels, used when a port, sheet, TF/SF face or Drude
// box lies inside a slab (the reference applies those before the CPML
// correction).  One launch per normal axis, in x, y, z order, which keeps
// the per-component add order (Hx: y then z, Hy: x then z, Hz: x then y).
// ---------------------------------------------------------------------------

// H, slab normal to x.  Threads: (t, j, k); t enumerates low then high layers.
__global__ void k_h_cpml_x(Grid g, Cpml p, float* Hy, float* Hz, const float* Ey, const float* Ez, float DB,
                           RDiv dx) {
    const int nl = max(p.n[0] - 1, 0), nh = max(p.n[1] - 1, 0), nt = nl + nh;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)nt * (g.NY - 1) * (g.NZ - 1);
    if (tid >= tot) return;
    const int t = (int)(tid % nt);
    const long long r = tid / nt;
    const int j = (int)(r % (g.NY - 1)) + 1, k = (int)(r / (g.NY - 1)) + 1;
    int i, f, q;
    if (t < nl) { i = t + 1; f = 0; q = i; }
    else { i = g.NX + 1 - p.n[1] + (t - nl); f = 1; q = g.NX - i; }
    const int ntf = f == 0 ? nl : nh;
    const long long po = (long long)(q - 1) + (long long)ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
    const long long o = off(g, i, j, k);
    const float b = p.b[f][q - 1], c = p.c[f][q - 1];
    float* ps = p.psi[f][0];
    ps[po] = b * ps[po] + rdiv(c * (Ez[o + 1] - Ez[o]), dx);
    Hy[o] = Hy[o] + DB * ps[po];
    ps = p.psi[f][1];
    ps[po] = b * ps[po] + rdiv(c * (Ey[o] - Ey[o + 1]), dx);
    Hz[o] = Hz[o] + DB * ps[po];
}

// H, slab normal to y.  Threads: (i, t, k).
__global__ void k_h_cpml_y(Grid g, Cpml p, float* Hx, float* Hz, const float* Ex, const float* Ez, float DB,
                           RDiv dy) {
    const int nl = max(p.n[2] - 1, 0), nh = max(p.n[3] - 1, 0), nt = nl + nh;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)(g.NX - 1) * nt * (g.NZ - 1);
    if (tid >= tot) return;
    const int i = (int)(tid % (g.NX - 1)) + 1;
    const long long r = tid / (g.NX - 1);
    const int t = (int)(r % nt), k = (int)(r / nt) + 1;
    int j, f, q;
    if (t < nl) { j = t + 1; f = 2; q = j; }
    else { j = g.NY + 1 - p.n[3] + (t - nl); f = 3; q = g.NY - j; }
    const int ntf = f == 2 ? nl : nh;
    const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)(q - 1) + (long long)ntf * (k - 1));
    const long long o = off(g, i, j, k);
    const float b = p.b[f][q - 1], c = p.c[f][q - 1];
    float* ps = p.psi[f][0];
    ps[po] = b * ps[po] + rdiv(c * (Ez[o] - Ez[o + g.PX]), dy);
    Hx[o] = Hx[o] + DB * ps[po];
    ps = p.psi[f][1];
    ps[po] = b * ps[po] + rdiv(c * (Ex[o + g.PX] - Ex[o]), dy);
    Hz[o] = Hz[o] + DB * ps[po];
}

// H, slab normal to z.  Threads: (i, j, t).
__global__ void k_h_cpml_z(Grid g, Cpml p, float* Hx, float* Hy, const float* Ex, const float* Ey, float DB,
                           RDiv dz) {
    const int nl = max(p.n[4] - 1, 0), nh = max(p.n[5] - 1, 0), nt = nl + nh;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)(g.NX - 1) * (g.NY - 1) * nt;
    if (tid >= tot) return;
    const int i = (int)(tid % (g.NX - 1)) + 1;
    const long long r = tid / (g.NX - 1);
    const int j = (int)(r % (g.NY - 1)) + 1, t = (int)(r / (g.NY - 1));
    int k, f, q;
    if (t < nl) { k = t + 1; f = 4; q = k; }
    else { k = g.NZ + 1 - p.n[5] + (t - nl); f = 5; q = g.NZ - k; }
    const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * (q - 1));
    const long long o = off(g, i, j, k);
    const float b = p.b[f][q - 1], c = p.c[f][q - 1];
    float* ps = p.psi[f][0];
    ps[po] = b * ps[po] + rdiv(c * (Ey[o + g.plane] - Ey[o]), dz);
    Hx[o] = Hx[o] + DB * ps[po];
    ps = p.psi[f][1];
    ps[po] = b * ps[po] + rdiv(c * (Ex[o] - Ex[o + g.plane]), dz);
    Hy[o] = Hy[o] + DB * ps[po];
}

template <typename IdxT>
__device__ __forceinline__ float gb_at(const IdxT* idx, const float* tup, long long o, int c) {
    return tup[(int)idx[o] * 6 + 2 * c + 1];
}

// E, slab normal to x: Ey (j 1..NY-1, k 2..NZ-1) and Ez (j 2..NY-1, k 1..NZ-1).
// Layers: low i = 2..n1 (q = i), high i = NX+1-n2..NX-1 (q = NX+1-i).
template <typename IdxT>
__global__ void k_e_cpml_x(Grid g, Cpml p, float* Ey, float* Ez, const float* Hy, const float* Hz,
                           const IdxT* idx, const float* tup, RDiv dx) {
    const int nl = max(p.n[0] - 1, 0), nh = max(p.n[1] - 1, 0), nt = nl + nh;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)nt * (g.NY - 1) * (g.NZ - 1);
    if (tid >= tot) return;
    const int t = (int)(tid % nt);
    const long long r = tid / nt;
    const int j = (int)(r % (g.NY - 1)) + 1, k = (int)(r / (g.NY - 1)) + 1;
    int i, f, q, ql;
    if (t < nl) { i = t + 2; f = 0; q = i; ql = t; }
    else { i = g.NX + 1 - p.n[1] + (t - nl); f = 1; q = g.NX + 1 - i; ql = q - 2; }
    const int ntf = f == 0 ? nl : nh;
    const long long po = (long long)ql + (long long)ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
    const long long o = off(g, i, j, k);
    const float b = p.b[f][q - 1], c = p.c[f][q - 1];
    if (k >= 2) {
        float* ps = p.psi[f][0];
        ps[po] = b * ps[po] + rdiv(c * (Hz[o - 1] - Hz[o]), dx);
        Ey[o] = Ey[o] + gb_at(idx, tup, o, 1) * ps[po];
    }
    if (j >= 2) {
        float* ps = p.psi[f][1];
        ps[po] = b * ps[po] + rdiv(c * (Hy[o] - Hy[o - 1]), dx);
        Ez[o] = Ez[o] + gb_at(idx, tup, o, 2) * ps[po];
    }
}

// E, slab normal to y: Ex (i 1..NX-1, k 2..NZ-1) and Ez (i 2..NX-1, k 1..NZ-1).
template <typename IdxT>
__global__ void k_e_cpml_y(Grid g, Cpml p, float* Ex, float* Ez, const float* Hx, const float* Hz,
                           const IdxT* idx, const float* tup, RDiv dy) {
    const int nl = max(p.n[2] - 1, 0), nh = max(p.n[3] - 1, 0), nt = nl + nh;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)(g.NX - 1) * nt * (g.NZ - 1);
    if (tid >= tot) return;
    const int i = (int)(tid % (g.NX - 1)) + 1;
    const long long r = tid / (g.NX - 1);
    const int t = (int)(r % nt), k = (int)(r / nt) + 1;
    int j, f, q, ql;
    if (t < nl) { j = t + 2; f = 2; q = j; ql = t; }
    else { j = g.NY + 1 - p.n[3] + (t - nl); f = 3; q = g.NY + 1 - j; ql = q - 2; }
    const int ntf = f == 2 ? nl : nh;
    const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)ql + (long long)ntf * (k - 1));
    const long long o = off(g, i, j, k);
    const float b = p.b[f][q - 1], c = p.c[f][q - 1];
    if (k >= 2) {
        float* ps = p.psi[f][0];
        ps[po] = b * ps[po] + rdiv(c * (Hz[o] - Hz[o - g.PX]), dy);
        Ex[o] = Ex[o] + gb_at(idx, tup, o, 0) * ps[po];
    }
    if (i >= 2) {
        float* ps = p.psi[f][1];
        ps[po] = b * ps[po] + rdiv(c * (Hx[o - g.PX] - Hx[o]), dy);
        Ez[o] = Ez[o] + gb_at(idx, tup, o, 2) * ps[po];
    }
}

// E, slab normal to z: Ex (i 1..NX-1, j 2..NY-1) and Ey (i 2..NX-1, j 1..NY-1).
template <typename IdxT>
__global__ void k_e_cpml_z(Grid g, Cpml p, float* Ex, float* Ey, const float* Hx, const float* Hy,
                           const IdxT* idx, const float* tup, RDiv dz) {
    const int nl = max(p.n[4] - 1, 0), nh = max(p.n[5] - 1, 0), nt = nl + nh;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)(g.NX - 1) * (g.NY - 1) * nt;
    if (tid >= tot) return;
    const int i = (int)(tid % (g.NX - 1)) + 1;
    const long long r = tid / (g.NX - 1);
    const int j = (int)(r % (g.NY - 1)) + 1, t = (int)(r / (g.NY - 1));
    int k, f, q, ql;
    if (t < nl) { k = t + 2; f = 4; q = k; ql = t; }
    else { k = g.NZ + 1 - p.n[5] + (t - nl); f = 5; q = g.NZ + 1 - k; ql = q - 2; }
    const long long po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * ql);
    const long long o = off(g, i, j, k);
    const float b = p.b[f][q - 1], c = p.c[f][q - 1];
    if (j >= 2) {
        float* ps = p.psi[f][0];
        ps[po] = b * ps[po] + rdiv(c * (Hy[o - g.plane] - Hy[o]), dz);
        Ex[o] = Ex[o] + gb_at(idx, tup, o, 0) * ps[po];
    }
    if (i >= 2) {
        float* ps = p.psi[f][1];
        ps[po] = b * ps[po] + rdiv(c * (Hx[o] - Hx[o - g.plane]), dz);
        Ey[o] = Ey[o] + gb_at(idx, tup, o, 1) * ps[po];
    }
}

// ---------------------------------------------------------------------------
// TF/SF plane-wave faces (H f90:4627-4710, E f90:5950-6033) with the
// analytic incident field Inc() (f90:8825-8842).
// ---------------------------------------------------------------------------
struct Tfsf {
    int lo[3], hi[3], delay[3], pbc[3];
    float WE[3], WH[3];
    float s_th, c_th, s_ph, c_ph;
    float kH[3], kE[3];
    float dx, dy, dz, dt, inv_c, spread, t_spread;
    float exp_half;            // expf(0.5) as evaluated on the host
    int pulse_type;
    float wall[6];             // face factors: 0 on an infinite ground plane's face (f90:4627-4790)
};

// FP32 exp rounded from an FP64 evaluation: the correctly rounded value
// except for arguments within ~1e-9 ulp of a rounding midpoint.
__device__ __forceinline__ float exp_rn(float x) { return (float)exp((double)x); }

__device__ __forceinline__ float inc_fn(const Tfsf& p, float w, float xf, float yf, float zf, float cf) {
    const float tx = p.s_th * p.c_ph * (xf - (float)p.delay[0]) * p.dx;
    const float ty = p.s_th * p.s_ph * (yf - (float)p.delay[1]) * p.dy;
    const float tz = p.c_th * (zf - (float)p.delay[2]) * p.dz;
    const float arg = (p.t_spread - (cf * p.dt) + p.inv_c * (tx + ty + tz)) / p.spread;
    if (p.pulse_type == 1) return w * exp_rn(-0.5f * (arg * arg));
    return -1.0f * w * arg * p.exp_half * exp_rn(-0.5f * (arg * arg));
}

// face: 0 = y faces, 1 = z faces, 2 = x faces.  (u,v) spans the face.
__device__ __forceinline__ void h_tfsf_body(const Grid& g, const Tfsf& p, int face, float* Hx, float* Hy, float* Hz,
                                            const int* step);

__global__ void k_h_tfsf(Grid g, Tfsf p, int face, float* Hx, float* Hy, float* Hz, const int* step) {
    h_tfsf_body(g, p, face, Hx, Hy, Hz, step);
}

// The three face pairs in one launch (blockIdx.z = face).  Their H updates
// touch disjoint elements (each face's ranges stop one short of the others'
// planes), so the order between faces does not matter.
__global__ void k_h_tfsf3(Grid g, Tfsf p, float* Hx, float* Hy, float* Hz, const int* step) {
    const int face = blockIdx.z;
    if (p.pbc[face == 0 ? 1 : (face == 1 ? 2 : 0)]) return;
    h_tfsf_body(g, p, face, Hx, Hy, Hz, step);
}

__device__ __forceinline__ void h_tfsf_body(const Grid& g, const Tfsf& p, int face, float* Hx, float* Hy, float* Hz,
                                            const int* step) {
    const int n = *step;
    const float cf = (float)n - 1.0f;
    const int xl = p.lo[0], xh = p.hi[0], yl = p.lo[1], yh = p.hi[1], zl = p.lo[2], zh = p.hi[2];
    const int px = p.pbc[0], py = p.pbc[1], pz = p.pbc[2];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    if (face == 0) {  // u = i, v = k
        const int i = xl + a, k = zl + b;
        if (k <= zh && i <= xh - 1 + px) {
            Hz[off(g, i, yl - 1, k)] = Hz[off(g, i, yl - 1, k)] - p.kH[1] * inc_fn(p, p.WE[0], (float)i + 0.5f, (float)yl, (float)k, cf) * p.wall[2];
            Hz[off(g, i, yh, k)] = Hz[off(g, i, yh, k)] + p.kH[1] * inc_fn(p, p.WE[0], (float)i + 0.5f, (float)yh, (float)k, cf) * p.wall[3];
        }
        if (k <= zh - 1 + pz && i <= xh) {
            Hx[off(g, i, yl - 1, k)] = Hx[off(g, i, yl - 1, k)] + p.kH[1] * inc_fn(p, p.WE[2], (float)i, (float)yl, (float)k + 0.5f, cf) * p.wall[2];
            Hx[off(g, i, yh, k)] = Hx[off(g, i, yh, k)] - p.kH[1] * inc_fn(p, p.WE[2], (float)i, (float)yh, (float)k + 0.5f, cf) * p.wall[3];
        }
    } else if (face == 1) {  // u = i, v = j
        const int i = xl + a, j = yl + b;
        if (j <= yh && i <= xh - 1 + px) {
            Hy[off(g, i, j, zl - 1)] = Hy[off(g, i, j, zl - 1)] + p.kH[2] * inc_fn(p, p.WE[0], (float)i + 0.5f, (float)j, (float)zl, cf) * p.wall[4];
            Hy[off(g, i, j, zh)] = Hy[off(g, i, j, zh)] - p.kH[2] * inc_fn(p, p.WE[0], (float)i + 0.5f, (float)j, (float)zh, cf) * p.wall[5];
        }
        if (j <= yh - 1 + py && i <= xh) {
            Hx[off(g, i, j, zl - 1)] = Hx[off(g, i, j, zl - 1)] - p.kH[2] * inc_fn(p, p.WE[1], (float)i, (float)j + 0.5f, (float)zl, cf) * p.wall[4];
            Hx[off(g, i, j, zh)] = Hx[off(g, i, j, zh)] + p.kH[2] * inc_fn(p, p.WE[1], (float)i, (float)j + 0.5f, (float)zh, cf) * p.wall[5];
        }
    } else {  // u = j, v = k
        const int j = yl + a, k = zl + b;
        if (k <= zh && j <= yh - 1 + py) {
            Hz[off(g, xl - 1, j, k)] = Hz[off(g, xl - 1, j, k)] + p.kH[0] * inc_fn(p, p.WE[1], (float)xl, (float)j + 0.5f, (float)k, cf) * p.wall[0];
            Hz[off(g, xh, j, k)] = Hz[off(g, xh, j, k)] - p.kH[0] * inc_fn(p, p.WE[1], (float)xh, (float)j + 0.5f, (float)k, cf) * p.wall[1];
        }
        if (k <= zh - 1 + pz && j <= yh) {
            Hy[off(g, xl - 1, j, k)] = Hy[off(g, xl - 1, j, k)] - p.kH[0] * inc_fn(p, p.WE[2], (float)xl, (float)j, (float)k + 0.5f, cf) * p.wall[0];
            Hy[off(g, xh, j, k)] = Hy[off(g, xh, j, k)] + p.kH[0] * inc_fn(p, p.WE[2], (float)xh, (float)j, (float)k + 0.5f, cf) * p.wall[1];
        }
    }
}

__global__ void k_e_tfsf(Grid g, Tfsf p, int face, float* Ex, float* Ey, float* Ez, const int* step) {
    const int n = *step;
    const float cf = (float)n - 0.5f;
    const int xl = p.lo[0], xh = p.hi[0], yl = p.lo[1], yh = p.hi[1], zl = p.lo[2], zh = p.hi[2];
    const int px = p.pbc[0], py = p.pbc[1], pz = p.pbc[2];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    if (face == 0) {  // y faces: u = i (from xl), v = k (from zl)
        const int i = xl + a, k = zl + b;
        if (k <= zh - 1 + pz && i >= xl + px && i <= xh) {
            Ez[off(g, i, yl, k)] = Ez[off(g, i, yl, k)] + p.kE[1] * inc_fn(p, p.WH[0], (float)i, (float)yl - 0.5f, (float)k + 0.5f, cf) * p.wall[2];
            Ez[off(g, i, yh, k)] = Ez[off(g, i, yh, k)] - p.kE[1] * inc_fn(p, p.WH[0], (float)i, (float)yh + 0.5f, (float)k + 0.5f, cf) * p.wall[3];
        }
        if (k >= zl + pz && k <= zh && i <= xh - 1 + px) {
            Ex[off(g, i, yl, k)] = Ex[off(g, i, yl, k)] - p.kE[1] * inc_fn(p, p.WH[2], (float)i + 0.5f, (float)yl - 0.5f, (float)k, cf) * p.wall[2];
            Ex[off(g, i, yh, k)] = Ex[off(g, i, yh, k)] + p.kE[1] * inc_fn(p, p.WH[2], (float)i + 0.5f, (float)yh + 0.5f, (float)k, cf) * p.wall[3];
        }
    } else if (face == 1) {  // z faces: u = i, v = j
        const int i = xl + a, j = yl + b;
        if (j <= yh - 1 + py && i >= xl + px && i <= xh) {
            Ey[off(g, i, j, zl)] = Ey[off(g, i, j, zl)] - p.kE[2] * inc_fn(p, p.WH[0], (float)i, (float)j + 0.5f, (float)zl - 0.5f, cf) * p.wall[4];
            Ey[off(g, i, j, zh)] = Ey[off(g, i, j, zh)] + p.kE[2] * inc_fn(p, p.WH[0], (float)i, (float)j + 0.5f, (float)zh + 0.5f, cf) * p.wall[5];
        }
        if (j >= yl + py && j <= yh && i <= xh - 1 + px) {
            Ex[off(g, i, j, zl)] = Ex[off(g, i, j, zl)] + p.kE[2] * inc_fn(p, p.WH[1], (float)i + 0.5f, (float)j, (float)zl - 0.5f, cf) * p.wall[4];
            Ex[off(g, i, j, zh)] = Ex[off(g, i, j, zh)] - p.kE[2] * inc_fn(p, p.WH[1], (float)i + 0.5f, (float)j, (float)zh + 0.5f, cf) * p.wall[5];
        }
    } else {  // x faces: u = j, v = k
        const int j = yl + a, k = zl + b;
        if (k <= zh - 1 + pz && j >= yl + py && j <= yh) {
            Ez[off(g, xl, j, k)] = Ez[off(g, xl, j, k)] - p.kE[0] * inc_fn(p, p.WH[1], (float)xl - 0.5f, (float)j, (float)k + 0.5f, cf) * p.wall[0];
            Ez[off(g, xh, j, k)] = Ez[off(g, xh, j, k)] + p.kE[0] * inc_fn(p, p.WH[1], (float)xh + 0.5f, (float)j, (float)k + 0.5f, cf) * p.wall[1];
        }
        if (k >= zl + pz && k <= zh && j <= yh - 1 + py) {
            Ey[off(g, xl, j, k)] = Ey[off(g, xl, j, k)] + p.kE[0] * inc_fn(p, p.WH[2], (float)xl - 0.5f, (float)j + 0.5f, (float)k, cf) * p.wall[0];
            Ey[off(g, xh, j, k)] = Ey[off(g, xh, j, k)] - p.kE[0] * inc_fn(p, p.WH[2], (float)xh + 0.5f, (float)j + 0.5f, (float)k, cf) * p.wall[1];
        }
    }
}

// The three E face pairs in one launch (blockIdx.z = face).  Unlike H, the
// E corrections of two faces meet on the TF/SF box edges (Ex on the y/z
// edges, Ez on x/y, Ey on x/z), where the reference adds the earlier face's
// term first.  Here the thread of the earlier face adds both terms, in that
// order, and the later face skips those elements, so every element gets the
// reference's sum bit for bit.
__global__ void k_e_tfsf3(Grid g, Tfsf p, float* Ex, float* Ey, float* Ez, const int* step) {
    const int face = blockIdx.z;
    const int n = *step;
    const float cf = (float)n - 0.5f;
    const int xl = p.lo[0], xh = p.hi[0], yl = p.lo[1], yh = p.hi[1], zl = p.lo[2], zh = p.hi[2];
    const int px = p.pbc[0], py = p.pbc[1], pz = p.pbc[2];
    const bool f0 = !py, f1 = !pz, f2 = !px;   // face pairs present: y, z, x
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    // x-face terms (face 2) of Ez and Ey at the low (xl) or high (xh) face
    auto ez_x = [&](int x, int j, int k) {
        return x == xl ? -(p.kE[0] * inc_fn(p, p.WH[1], (float)xl - 0.5f, (float)j, (float)k + 0.5f, cf) * p.wall[0])
                       : p.kE[0] * inc_fn(p, p.WH[1], (float)xh + 0.5f, (float)j, (float)k + 0.5f, cf) * p.wall[1];
    };
    auto ey_x = [&](int x, int j, int k) {
        return x == xl ? p.kE[0] * inc_fn(p, p.WH[2], (float)xl - 0.5f, (float)j + 0.5f, (float)k, cf) * p.wall[0]
                       : -(p.kE[0] * inc_fn(p, p.WH[2], (float)xh + 0.5f, (float)j + 0.5f, (float)k, cf) * p.wall[1]);
    };
    // z-face term (face 1) of Ex at the low (zl) or high (zh) face
    auto ex_z = [&](int i, int j, int z) {
        return z == zl ? p.kE[2] * inc_fn(p, p.WH[1], (float)i + 0.5f, (float)j, (float)zl - 0.5f, cf) * p.wall[4]
                       : -(p.kE[2] * inc_fn(p, p.WH[1], (float)i + 0.5f, (float)j, (float)zh + 0.5f, cf) * p.wall[5]);
    };
    if (face == 0) {
        if (!f0) return;
        const int i = xl + a, k = zl + b;
        if (k <= zh - 1 + pz && i >= xl + px && i <= xh) {
            const bool e2 = f2 && (i == xl || i == xh);
            long long o = off(g, i, yl, k);
            float e = Ez[o] + p.kE[1] * inc_fn(p, p.WH[0], (float)i, (float)yl - 0.5f, (float)k + 0.5f, cf) * p.wall[2];
            if (e2) e = e + ez_x(i, yl, k);
            Ez[o] = e;
            o = off(g, i, yh, k);
            e = Ez[o] - p.kE[1] * inc_fn(p, p.WH[0], (float)i, (float)yh + 0.5f, (float)k + 0.5f, cf) * p.wall[3];
            if (e2) e = e + ez_x(i, yh, k);
            Ez[o] = e;
        }
        if (k >= zl + pz && k <= zh && i <= xh - 1 + px) {
            const bool e1 = f1 && (k == zl || k == zh);
            long long o = off(g, i, yl, k);
            float e = Ex[o] - p.kE[1] * inc_fn(p, p.WH[2], (float)i + 0.5f, (float)yl - 0.5f, (float)k, cf) * p.wall[2];
            if (e1) e = e + ex_z(i, yl, k);
            Ex[o] = e;
            o = off(g, i, yh, k);
            e = Ex[o] + p.kE[1] * inc_fn(p, p.WH[2], (float)i + 0.5f, (float)yh + 0.5f, (float)k, cf) * p.wall[3];
            if (e1) e = e + ex_z(i, yh, k);
            Ex[o] = e;
        }
    } else if (face == 1) {
        if (!f1) return;
        const int i = xl + a, j = yl + b;
        if (j <= yh - 1 + py && i >= xl + px && i <= xh) {
            const bool e2 = f2 && (i == xl || i == xh);
            long long o = off(g, i, j, zl);
            float e = Ey[o] - p.kE[2] * inc_fn(p, p.WH[0], (float)i, (float)j + 0.5f, (float)zl - 0.5f, cf) * p.wall[4];
            if (e2) e = e + ey_x(i, j, zl);
            Ey[o] = e;
            o = off(g, i, j, zh);
            e = Ey[o] + p.kE[2] * inc_fn(p, p.WH[0], (float)i, (float)j + 0.5f, (float)zh + 0.5f, cf) * p.wall[5];
            if (e2) e = e + ey_x(i, j, zh);
            Ey[o] = e;
        }
        if (j >= yl + py && j <= yh && i <= xh - 1 + px && !(f0 && (j == yl || j == yh))) {
            long long o = off(g, i, j, zl);
            Ex[o] = Ex[o] + ex_z(i, j, zl);
            o = off(g, i, j, zh);
            Ex[o] = Ex[o] + ex_z(i, j, zh);
        }
    } else {
        if (!f2) return;
        const int j = yl + a, k = zl + b;
        if (k <= zh - 1 + pz && j >= yl + py && j <= yh && !(f0 && (j == yl || j == yh))) {
            long long o = off(g, xl, j, k);
            Ez[o] = Ez[o] + ez_x(xl, j, k);
            o = off(g, xh, j, k);
            Ez[o] = Ez[o] + ez_x(xh, j, k);
        }
        if (k >= zl + pz && k <= zh && j <= yh - 1 + py && !(f1 && (k == zl || k == zh))) {
            long long o = off(g, xl, j, k);
            Ey[o] = Ey[o] + ey_x(xl, j, k);
            o = off(g, xh, j, k);
            Ey[o] = Ey[o] + ey_x(xh, j, k);
        }
    }
}

// ---------------------------------------------------------------------------
// Lumped ports (f90:4800-5013 H phase, 6120-6283 E phase).  One work-group
// handles every port in order; the 1-D line updates are spread over threads.
// ---------------------------------------------------------------------------
struct PortDev {
    int dir, loc[3], size[3];
    float kI, kV, kS, dl, cur_num, cur_den, gap_div;
    int excited;
    RDiv cdv, gdv;     // the divisions by cur_den and gap_div
};

// Time-only source waveforms E_or_V_1D() (f90:8844-8857), evaluated once per
// step on the host: src[n-1] current source, src[T+n-1] voltage source,
// src[2T+n-1] plane-wave incident field.
struct Pulse {
    const float* src;
    int T;
};

// Arguments of both port kernels; the batched forms take one per member and
// run one work-group per member.
struct PortArgs {
    Grid g;
    const PortDev* ports;
    int np, M, T;
    float *V, *I, *Vout, *Iout;
    float *Ex, *Ey, *Ez;
    const float *Hx, *Hy, *Hz;
    float dx, dy, dz;
    Pulse pu;
    const int* step;
};

// Loop integral of H around a lumped gap of cells [loc, loc+size-1] along
// axis dir, in the reference's loop order (f90:4820-4857, 5456-5496), before
// the (I*num)/den scaling.
__device__ __forceinline__ float gap_loop_current(const Grid& g, int dir, const int* loc, const int* size,
                                                  const float* Hx, const float* Hy, const float* Hz, float dx,
                                                  float dy, float dz) {
    const int l1 = loc[0], l2 = loc[1], l3 = loc[2];
    const int d1 = size[0] + l1 - 1, d2 = size[1] + l2 - 1, d3 = size[2] + l3 - 1;
    float c = 0.0f;
    if (dir == 0) {
        for (int ii = l1; ii <= d1; ++ii) {
            for (int jj = l2; jj <= d2; ++jj) c = c + dy * (Hy[off(g, ii, jj, l3 - 1)] - Hy[off(g, ii, jj, d3)]);
            for (int kk = l3; kk <= d3; ++kk) c = c + dz * (Hz[off(g, ii, l2, kk)] - Hz[off(g, ii, d2 - 1, kk)]);
        }
    } else if (dir == 1) {
        for (int jj = l2; jj <= d2; ++jj) {
            for (int ii = l1; ii <= d1; ++ii) c = c + dx * (Hx[off(g, ii, jj, l3)] - Hx[off(g, ii, jj, d3 - 1)]);
            for (int kk = l3; kk <= d3; ++kk) c = c + dz * (Hz[off(g, l1 - 1, jj, kk)] - Hz[off(g, d1, jj, kk)]);
        }
    } else {
        for (int kk = l3; kk <= d3; ++kk) {
            for (int ii = l1; ii <= d1; ++ii) c = c + dx * (Hx[off(g, ii, l2 - 1, kk)] - Hx[off(g, ii, d2, kk)]);
            for (int jj = l2; jj <= d2; ++jj) c = c + dy * (Hy[off(g, l1, jj, kk)] - Hy[off(g, d1 - 1, jj, kk)]);
        }
    }
    return c;
}

__device__ __forceinline__ void ports_h_body(const Grid& g, const PortDev* ports, int np, int M, float* V, float* I,
                                             float* Iout, int T, const float* Hx, const float* Hy, const float* Hz,
                                             float dx, float dy, float dz, const Pulse& pu, const int* step) {
    const int n = *step;
    if (threadIdx.x == 0) {
        for (int q = 0; q < np; ++q) {
            const PortDev pt = ports[q];
            float c = gap_loop_current(g, pt.dir, pt.loc, pt.size, Hx, Hy, Hz, dx, dy, dz);
            c = rdiv(c * pt.cur_num, pt.cdv);
            I[(long long)q * M] = c;
            Iout[(long long)q * T + (n - 1)] = c;
        }
    }
    // 1-D line current, j = 2..M-1 (0-based j-1); reads V only
    for (int q = 0; q < np; ++q) {
        const float kI = ports[q].kI;
        float* Iq = I + (long long)q * M;
        const float* Vq = V + (long long)q * M;
        for (int j = 2 + threadIdx.x; j <= M - 1; j += blockDim.x) Iq[j - 1] = Iq[j - 1] - kI * (Vq[j] - Vq[j - 1]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < np; ++q)
            if (ports[q].excited) {
                float* Iq = I + (long long)q * M;
                Iq[1] = Iq[1] - ports[q].kI * pu.src[n - 1];
            }
    }
}

__device__ __forceinline__ void ports_e_body(const Grid& g, const PortDev* ports, int np, int M, float* V,
                                             const float* I, float* Vout, int T, float* Ex, float* Ey, float* Ez,
                                             const Pulse& pu, const int* step) {
    const int n = *step;
    for (int q = 0; q < np; ++q) {
        const float kV = ports[q].kV;
        float* Vq = V + (long long)q * M;
        const float* Iq = I + (long long)q * M;
        for (int j = 2 + threadIdx.x; j <= M - 1; j += blockDim.x) Vq[j - 1] = Vq[j - 1] - kV * (Iq[j - 1] - Iq[j - 2]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < np; ++q)
            if (ports[q].excited) {
                float* Vq = V + (long long)q * M;
                Vq[1] = Vq[1] + ports[q].kS * pu.src[pu.T + n - 1];
            }
        for (int q = 0; q < np; ++q) {
            const PortDev pt = ports[q];
            const float v2 = V[(long long)q * M + 1];
            float* F = pt.dir == 0 ? Ex : (pt.dir == 1 ? Ey : Ez);
            const float val = rdiv(1.0f * v2, pt.gdv);
            for (int ii = pt.loc[0]; ii <= pt.loc[0] + pt.size[0] - 1; ++ii)
                for (int jj = pt.loc[1]; jj <= pt.loc[1] + pt.size[1] - 1; ++jj)
                    for (int kk = pt.loc[2]; kk <= pt.loc[2] + pt.size[2] - 1; ++kk) F[off(g, ii, jj, kk)] = val;
            Vout[(long long)q * T + (n - 1)] = v2;
        }
    }
}

__global__ void __launch_bounds__(256) k_ports_h(PortArgs a) {
    ports_h_body(a.g, a.ports, a.np, a.M, a.V, a.I, a.Iout, a.T, a.Hx, a.Hy, a.Hz, a.dx, a.dy, a.dz, a.pu, a.step);
}
__global__ void __launch_bounds__(256) k_ports_e(PortArgs a) {
    ports_e_body(a.g, a.ports, a.np, a.M, a.V, a.I, a.Vout, a.T, a.Ex, a.Ey, a.Ez, a.pu, a.step);
}
__global__ void __launch_bounds__(256) k_ports_h_batch(const PortArgs* __restrict__ args) {
    const PortArgs& a = args[blockIdx.x];
    ports_h_body(a.g, a.ports, a.np, a.M, a.V, a.I, a.Iout, a.T, a.Hx, a.Hy, a.Hz, a.dx, a.dy, a.dz, a.pu, a.step);
}
__global__ void __launch_bounds__(256) k_ports_e_batch(const PortArgs* __restrict__ args) {
    const PortArgs& a = args[blockIdx.x];
    ports_e_body(a.g, a.ports, a.np, a.M, a.V, a.I, a.Vout, a.T, a.Ex, a.Ey, a.Ez, a.pu, a.step);
}

// ---------------------------------------------------------------------------
// SPICE ports (f90:5435-5665 current, 6285-6520 gap field).  After the H
// phase, k_spice_h forms each port's gap current as a lumped port does; a
// host circuit answers with node voltages, and in the E phase (after the
// lumped ports, before the separate CPML) k_spice_e hard-sets each gap to
// (V - bias) / gap_div.  One thread: a few ports of a few cells.
// ---------------------------------------------------------------------------
struct SpiceDev {
    int dir, loc[3], size[3];
    float cur_num, cur_den, gap_div, bias;
};

struct SpiceArgs {
    Grid g;
    const SpiceDev* ports;
    int np, nextra, T;
    float* I;                  // this step's gap currents [np], read by the host
    const float* Vnode;        // the host's node voltages [np + nextra]
    float *Iout, *Vout;        // series [np][T] and [np + nextra][T]
    float *Ex, *Ey, *Ez;
    const float *Hx, *Hy, *Hz;
    float dx, dy, dz;
    const int* step;
};

__global__ void k_spice_h(SpiceArgs a) {
    if (threadIdx.x != 0) return;
    const int n = *a.step;
    for (int q = 0; q < a.np; ++q) {
        const SpiceDev pt = a.ports[q];
        float c = gap_loop_current(a.g, pt.dir, pt.loc, pt.size, a.Hx, a.Hy, a.Hz, a.dx, a.dy, a.dz);
        c = c * pt.cur_num / pt.cur_den;
        a.I[q] = c;
        a.Iout[(long long)q * a.T + (n - 1)] = c;
    }
}

__global__ void k_spice_e(SpiceArgs a) {
    if (threadIdx.x != 0) return;
    const int n = *a.step;
    for (int q = 0; q < a.np; ++q) {
        const SpiceDev pt = a.ports[q];
        const float vs = a.Vnode[q] - pt.bias;
        float* F = pt.dir == 0 ? a.Ex : (pt.dir == 1 ? a.Ey : a.Ez);
        const float val = 1.0f * vs / pt.gap_div;
        for (int ii = pt.loc[0]; ii <= pt.loc[0] + pt.size[0] - 1; ++ii)
            for (int jj = pt.loc[1]; jj <= pt.loc[1] + pt.size[1] - 1; ++jj)
                for (int kk = pt.loc[2]; kk <= pt.loc[2] + pt.size[2] - 1; ++kk) F[off(a.g, ii, jj, kk)] = val;
        a.Vout[(long long)q * a.T + (n - 1)] = vs;
    }
    for (int e = 0; e < a.nextra; ++e) a.Vout[(long long)(a.np + e) * a.T + (n - 1)] = a.Vnode[a.np + e];
}

// ---------------------------------------------------------------------------
// Thin sheets (H f90:4486-4554, E f90:5712-5780).  One launch per normal
// axis handles every plane of that axis; (p,q) span the plane.
// ---------------------------------------------------------------------------
struct Sheets {
    int np;                // planes on this axis
    const int* pos;
    const float* sig;
    const float* ep;
    float* es;             // special E of each plane
    int P, Q;
};

__device__ __forceinline__ long long sh_off(const Sheets& s, int pl, int p, int q) {
    return (long long)pl * s.P * s.Q + (p - 1) + (long long)s.P * (q - 1);
}
__device__ __forceinline__ bool sh_on(const Sheets& s, int pl, int p, int q) {
    const long long o = sh_off(s, pl, p, q);
    return s.sig[o] > 0 || s.ep[o] > 1;
}

__global__ void k_h_sheets(Grid g, Sheets s, int axis, float* Hx, float* Hy, float* Hz, const float* Ex,
                           const float* Ey, const float* Ez, const float* dhx, const float* dhy, const float* dhz,
                           float f) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int q = blockIdx.y * blockDim.y + threadIdx.y + 1;
    for (int pl = 0; pl < s.np; ++pl) {
        const int c = s.pos[pl];
        if (axis == 0) {           // (p,q) = (j,k), j,k in 1..N-1
            const int i = c, j = p, k = q;
            if (j > g.NY - 1 || k > g.NZ - 1 || !sh_on(s, pl, j, k)) continue;
            const long long o = off(g, i, j, k);
            Hy[o] = Hy[o] + f * ((Ex[o + g.plane] - Ex[o] + s.es[sh_off(s, pl, j, k)] - s.es[sh_off(s, pl, j, k + 1)]) * dhz[k - 1]);
            Hz[o] = Hz[o] + f * ((Ex[o] - Ex[o + g.PX] + s.es[sh_off(s, pl, j + 1, k)] - s.es[sh_off(s, pl, j, k)]) * dhy[j - 1]);
        } else if (axis == 1) {    // (p,q) = (i,k)
            const int i = p, j = c, k = q;
            if (i > g.NX - 1 || k > g.NZ - 1 || !sh_on(s, pl, i, k)) continue;
            const long long o = off(g, i, j, k);
            Hx[o] = Hx[o] + f * (Ey[o] - Ey[o + g.plane] + s.es[sh_off(s, pl, i, k + 1)] - s.es[sh_off(s, pl, i, k)]) * dhz[k - 1];
            Hz[o] = Hz[o] + f * ((Ey[o + 1] - Ey[o] - s.es[sh_off(s, pl, i + 1, k)] + s.es[sh_off(s, pl, i, k)]) * dhx[i - 1]);
        } else {                   // (p,q) = (i,j)
            const int i = p, j = q, k = c;
            if (i > g.NX - 1 || j > g.NY - 1 || !sh_on(s, pl, i, j)) continue;
            const long long o = off(g, i, j, k);
            Hx[o] = Hx[o] + f * (Ez[o + g.PX] - Ez[o] - s.es[sh_off(s, pl, i, j + 1)] + s.es[sh_off(s, pl, i, j)]) * dhy[j - 1];
            Hy[o] = Hy[o] + f * ((Ez[o] - Ez[o + 1] + s.es[sh_off(s, pl, i + 1, j)] - s.es[sh_off(s, pl, i, j)]) * dhx[i - 1]);
        }
    }
}

__global__ void k_e_sheets(Grid g, Sheets s, int axis, const float* Hx, const float* Hy, const float* Hz,
                           const float* dex, const float* dey, const float* dez, float dt, float ep0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int q = blockIdx.y * blockDim.y + threadIdx.y + 2;
    for (int pl = 0; pl < s.np; ++pl) {
        const int c = s.pos[pl];
        int i, j, k;
        float d1, d2;
        if (axis == 0) { i = c; j = p; k = q; if (j > g.NY - 1 || k > g.NZ - 1) continue; }
        else if (axis == 1) { i = p; j = c; k = q; if (i > g.NX - 1 || k > g.NZ - 1) continue; }
        else { i = p; j = q; k = c; if (i > g.NX - 1 || j > g.NY - 1) continue; }
        if (!sh_on(s, pl, p, q)) continue;
        const long long so = sh_off(s, pl, p, q);
        const float sg = s.sig[so], ep = s.ep[so];
        const long long o = off(g, i, j, k);
        float curl;
        if (axis == 0) curl = (Hz[o] - Hz[o - g.PX]) * dey[j - 1] + (Hy[o - g.plane] - Hy[o]) * dez[k - 1];
        else if (axis == 1) curl = (Hz[o - 1] - Hz[o]) * dex[i - 1] + (Hx[o] - Hx[o - g.plane]) * dez[k - 1];
        else curl = (Hy[o] - Hy[o - 1]) * dex[i - 1] + (Hx[o - g.PX] - Hx[o]) * dey[j - 1];
        (void)d1; (void)d2;
        s.es[so] = ((1.0f - sg * dt / (2.0f * ep0 * ep)) / (1.0f + sg * dt / (2.0f * ep0 * ep))) * s.es[so] +
                   ((dt / (ep0 * ep)) / (1.0f + sg * dt / (2.0f * ep0 * ep))) * curl;
    }
}

// ---------------------------------------------------------------------------
// Drude ADE on the plasma bounding box (f90:5782-5854) and the E_old save
// (f90:6812-6826, restricted to the box).
// ---------------------------------------------------------------------------
struct DrudeBox {
    int lo[3], hi[3];
    long long nb;
    const float* Ja;
    const float* Jb;
    float* J;
    float* Eold;
};

template <typename IdxT>
__global__ void k_drude(Grid g, DrudeBox b, int comp, int P, float* E, const IdxT* idx, const float* tup) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= b.nb) return;
    const int bx = b.hi[0] - b.lo[0] + 1, by = b.hi[1] - b.lo[1] + 1;
    const int i = b.lo[0] + (int)(tid % bx);
    const int j = b.lo[1] + (int)((tid / bx) % by);
    const int k = b.lo[2] + (int)(tid / ((long long)bx * by));
    const long long o = off(g, i, j, k);
    float t = 0.0f;
    for (int q = 0; q < P; ++q) t = t + (1.0f + b.Ja[q * b.nb + tid]) * b.J[q * b.nb + tid];
    const float e = E[o] - gb_at(idx, tup, o, comp) * t / 2.0f;
    E[o] = e;
    const float eo = b.Eold[tid];
    for (int q = 0; q < P; ++q) b.J[q * b.nb + tid] = b.Ja[q * b.nb + tid] * b.J[q * b.nb + tid] + b.Jb[q * b.nb + tid] * (e + eo);
}

__global__ void k_save_eold(Grid g, DrudeBox b, const float* E) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= b.nb) return;
    const int bx = b.hi[0] - b.lo[0] + 1, by = b.hi[1] - b.lo[1] + 1;
    const int i = b.lo[0] + (int)(tid % bx);
    const int j = b.lo[1] + (int)((tid / bx) % by);
    const int k = b.lo[2] + (int)(tid / ((long long)bx * by));
    b.Eold[tid] = E[off(g, i, j, k)];
}

// Transposed Drude update of one component: the adjoint of k_drude followed
// by k_save_eold, on the adjoint state after step n, before the transposed
// bulk E update.  Forward, per cell: E2 = E1 - gb/2 sum_q (1 + Ja_q) J_q,
// J_q' = Ja_q J_q + Jb_q (E2 + Eold), and Eold' = E2 at the end of the step.
// Here b.J holds a_q, the adjoint of J_q, and b.Eold holds o, the adjoint of
// Eold:
//   s = sum_q Jb_q a_q,   lambda_E2 = lambda_E + o + s,
//   a_q <- Ja_q a_q - gb/2 (1 + Ja_q) lambda_E2,   o <- s,
// and lambda_E1 = lambda_E2 with mu = gb lambda_E1 beside it.
template <typename IdxT>
__global__ void k_adj_drude(Grid g, DrudeBox b, int comp, int P, float* lE, float* mE, const IdxT* idx,
                            const float* tup) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= b.nb) return;
    const int bx = b.hi[0] - b.lo[0] + 1, by = b.hi[1] - b.lo[1] + 1;
    const int i = b.lo[0] + (int)(tid % bx);
    const int j = b.lo[1] + (int)((tid / bx) % by);
    const int k = b.lo[2] + (int)(tid / ((long long)bx * by));
    const long long o = off(g, i, j, k);
    float s = 0.0f;
    for (int q = 0; q < P; ++q) s = s + b.Jb[q * b.nb + tid] * b.J[q * b.nb + tid];
    const float gb = gb_at(idx, tup, o, comp);
    const float le = lE[o] + b.Eold[tid] + s;
    lE[o] = le;
    mE[o] = gb * le;
    const float h = gb / 2.0f * le;
    for (int q = 0; q < P; ++q) {
        const float ja = b.Ja[q * b.nb + tid];
        b.J[q * b.nb + tid] = ja * b.J[q * b.nb + tid] - (1.0f + ja) * h;
    }
    b.Eold[tid] = s;
}

// ---------------------------------------------------------------------------
// Periodic boundary copies (H f90:5377-5433, E f90:6672-6741).  One launch per
// axis, in x, y, z order; the special E of sheet planes is copied alongside.
// ---------------------------------------------------------------------------
__global__ void k_h_pbc(Grid g, int axis, float* Ha, float* Hb) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int v = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (axis == 0) {  // Hz, Hy: (NX-1,j,k) <- (1,j,k)
        if (u > g.NY - 1 || v > g.NZ - 1) return;
        Ha[off(g, g.NX - 1, u, v)] = Ha[off(g, 1, u, v)];
        Hb[off(g, g.NX - 1, u, v)] = Hb[off(g, 1, u, v)];
    } else if (axis == 1) {  // Hz, Hx
        if (u > g.NX - 1 || v > g.NZ - 1) return;
        Ha[off(g, u, g.NY - 1, v)] = Ha[off(g, u, 1, v)];
        Hb[off(g, u, g.NY - 1, v)] = Hb[off(g, u, 1, v)];
    } else {  // Hy, Hx
        if (u > g.NX - 1 || v > g.NY - 1) return;
        Ha[off(g, u, v, g.NZ - 1)] = Ha[off(g, u, v, 1)];
        Hb[off(g, u, v, g.NZ - 1)] = Hb[off(g, u, v, 1)];
    }
}

__global__ void k_e_pbc(Grid g, int axis, float* Ea, float* Eb) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int v = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (axis == 0) {  // Ez, Ey: (1,j,k) <- (NX-1,j,k)
        if (u > g.NY - 1 || v > g.NZ - 1) return;
        Ea[off(g, 1, u, v)] = Ea[off(g, g.NX - 1, u, v)];
        Eb[off(g, 1, u, v)] = Eb[off(g, g.NX - 1, u, v)];
    } else if (axis == 1) {  // Ez, Ex
        if (u > g.NX - 1 || v > g.NZ - 1) return;
        Ea[off(g, u, 1, v)] = Ea[off(g, u, g.NY - 1, v)];
        Eb[off(g, u, 1, v)] = Eb[off(g, u, g.NY - 1, v)];
    } else {  // Ey, Ex
        if (u > g.NX - 1 || v > g.NY - 1) return;
        Ea[off(g, u, v, 1)] = Ea[off(g, u, v, g.NZ - 1)];
        Eb[off(g, u, v, 1)] = Eb[off(g, u, v, g.NZ - 1)];
    }
}

// In-plane PBC copy of a special-E plane array: dst row/column <- src.
// mode 0: p index 1 <- P_src (copy along p); mode 1: q index 1 <- Q_src.
__global__ void k_sheet_pbc(Sheets s, int mode, int src, int n_other) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int pl = blockIdx.y;
    if (u > n_other || pl >= s.np) return;
    if (mode == 0) s.es[sh_off(s, pl, 1, u)] = s.es[sh_off(s, pl, src, u)];
    else s.es[sh_off(s, pl, u, 1)] = s.es[sh_off(s, pl, u, src)];
}

// ---------------------------------------------------------------------------
// Outputs
// ---------------------------------------------------------------------------

// Reflection/transmission plane sums (f90:7005-7103): one work-group, FP64
// accumulation in a fixed order (deterministic), stored per step.
__global__ void __launch_bounds__(256) k_plane_sums(Grid g, int axis, int lo_pos, int hi_pos, int u0, int u1, int v0,
                                                    int v1, const float* Ea, const float* Eb, float c_a, float c_b1,
                                                    float c_b2, int minus_b, double* refl, double* trans,
                                                    const int* step) {
    // axis 0: planes x=lo/hi, (u,v)=(j,k), Ea=Ez, Eb=Ey;  axis 1: (u,v)=(i,k), Ea=Ez, Eb=Ex;
    // axis 2: (u,v)=(i,j), Ea=Ey, Eb=Ex
    __shared__ double sr[256], st[256];
    const int n = *step;
    const int nu = u1 - u0 + 1, nv = v1 - v0 + 1;
    double ar = 0.0, at = 0.0;
    for (int t = threadIdx.x; t < nu * nv; t += blockDim.x) {
        const int u = u0 + t % nu, v = v0 + t / nu;
        long long ol, oh;
        if (axis == 0) { ol = off(g, lo_pos, u, v); oh = off(g, hi_pos, u, v); }
        else if (axis == 1) { ol = off(g, u, lo_pos, v); oh = off(g, u, hi_pos, v); }
        else { ol = off(g, u, v, lo_pos); oh = off(g, u, v, hi_pos); }
        // per-cell terms exactly as the reference forms them: (a*c_a) -/+ ((b*c_b1)*c_b2)
        const float ra = Ea[ol] * c_a, rb = Eb[ol] * c_b1 * c_b2;
        const float ta = Ea[oh] * c_a, tb = Eb[oh] * c_b1 * c_b2;
        ar = ar + (double)ra;
        at = at + (double)ta;
        if (minus_b) { ar = ar - (double)rb; at = at - (double)tb; }
        else { ar = ar + (double)rb; at = at + (double)tb; }
    }
    sr[threadIdx.x] = ar;
    st[threadIdx.x] = at;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            st[threadIdx.x] += st[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        refl[n - 1] = sr[0];
        trans[n - 1] = st[0];
    }
}

// NTFF: face currents.  Faces 0..5 = x-low, x-high, y-low, y-high, z-low,
// z-high; per face two magnetic (M) and two electric (J) components with
// their previous-step values (f90:7512-7641).
struct NtffFaces {
    float* M[6][2];
    float* Mo[6][2];
    float* J[6][2];
    float* Jo[6][2];
    int fd0[6];            // row length of each face array
    int lo[3], hi[3];
    int pbc[3];
    float wall[6];         // face factors (0 on an infinite ground plane's face, f90:7223-7325)
};

__device__ __forceinline__ void ntff_currents_body(const Grid& g, const NtffFaces& f, int face, const float* Ex,
                                                   const float* Ey, const float* Ez, const float* Hx,
                                                   const float* Hy, const float* Hz) {
    const int a = face / 2;
    if (f.pbc[a]) return;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    const int u = f.lo[b] + blockIdx.x * blockDim.x + threadIdx.x;
    const int v = f.lo[c] + blockIdx.y * blockDim.y + threadIdx.y;
    if (u > f.hi[b] || v > f.hi[c]) return;
    const bool low = (face & 1) == 0;
    const int pos = low ? f.lo[a] : f.hi[a];
    const float s = low ? 1.0f : -1.0f;      // sign flip between low and high faces
    const float wl = f.wall[face];           // 0 on an infinite ground plane's face
    const long long fo = (long long)u + (long long)f.fd0[face] * v;
    float m0, m1, j0, j1;
    if (a == 0) {
        const int i = pos, j = u, k = v;
        const long long o = off(g, i, j, k);
        m0 = (-1.0f * s) * (Ez[o] + Ez[o + g.PX] + Ez[o + 1] + Ez[o + 1 + g.PX]) / 4.0f * wl;
        m1 = s * (Ey[o] + Ey[o + 1] + Ey[o + g.plane] + Ey[o + 1 + g.plane]) / 4.0f * wl;
        j0 = s * (Hz[o] + Hz[o + g.plane]) / 2.0f * wl;
        j1 = (-1.0f * s) * (Hy[o] + Hy[o + g.PX]) / 2.0f * wl;
    } else if (a == 1) {
        const int i = u, j = pos, k = v;
        const long long o = off(g, i, j, k);
        m0 = s * (Ez[o] + Ez[o + 1] + Ez[o + g.PX] + Ez[o + 1 + g.PX]) / 4.0f * wl;
        m1 = (-1.0f * s) * (Ex[o] + Ex[o + g.PX] + Ex[o + g.plane] + Ex[o + g.PX + g.plane]) / 4.0f * wl;
        j0 = (-1.0f * s) * (Hz[o] + Hz[o + g.plane]) / 2.0f * wl;
        j1 = s * (Hx[o] + Hx[o + 1]) / 2.0f * wl;
    } else {
        const int i = u, j = v, k = pos;
        const long long o = off(g, i, j, k);
        m0 = (-1.0f * s) * (Ey[o] + Ey[o + 1] + Ey[o + g.plane] + Ey[o + 1 + g.plane]) / 4.0f * wl;
        m1 = s * (Ex[o] + Ex[o + g.PX] + Ex[o + g.plane] + Ex[o + g.PX + g.plane]) / 4.0f * wl;
        j0 = s * (Hy[o] + Hy[o + g.PX]) / 2.0f * wl;
        j1 = (-1.0f * s) * (Hx[o] + Hx[o + 1]) / 2.0f * wl;
    }
    f.Mo[face][0][fo] = f.M[face][0][fo];
    f.M[face][0][fo] = m0;
    f.Mo[face][1][fo] = f.M[face][1][fo];
    f.M[face][1][fo] = m1;
    f.Jo[face][0][fo] = f.J[face][0][fo];
    f.J[face][0][fo] = j0;
    f.Jo[face][1][fo] = f.J[face][1][fo];
    f.J[face][1][fo] = j1;
}

__global__ void k_ntff_currents(Grid g, NtffFaces f, const float* Ex, const float* Ey, const float* Ez,
                                const float* Hx, const float* Hy, const float* Hz) {
    ntff_currents_body(g, f, blockIdx.z, Ex, Ey, Ez, Hx, Hy, Hz);
}

// NTFF retarded-time binning (f90:7643-7927).  Index and weight arithmetic is
// the reference's FP32; accumulation is FP64 (atomic adds).
struct NtffBins {
    const float* ang;      // [n_angles][6]
    int n_angles;
    int ic, jc, kc;
    float r_relay, dx, dy, dz, c_dt;
    float K[3];
    long long LL;          // bins per angle (len_ff + 2)
    double* U[3];          // FP64 accumulators (atomic binning)
    double* W[3];
    float* Uf[3];          // FP32 accumulators (ordered binning, the reference's arithmetic)
    float* Wf[3];
    int exact;             // 1: Uf/Wf hold the result, 0: U/W
};

__global__ void k_ntff_bins(NtffFaces f, NtffBins bn, const int* step) {
    const int n = *step;
    const int face = blockIdx.z % 6;
    const int kk = blockIdx.z / 6;
    const int a = face / 2;
    if (f.pbc[a]) return;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    const int u = f.lo[b] + blockIdx.x * blockDim.x + threadIdx.x;
    const int v = f.lo[c] + blockIdx.y * blockDim.y + threadIdx.y;
    if (u > f.hi[b] || v > f.hi[c]) return;
    const bool low = (face & 1) == 0;
    const int pos = low ? f.lo[a] : f.hi[a];
    int i, j, k;
    if (a == 0) { i = pos; j = u; k = v; }
    else if (a == 1) { i = u; j = pos; k = v; }
    else { i = u; j = v; k = pos; }
    const float* an = bn.ang + 6 * kk;
    const float s1 = an[2], c1 = an[3], s2 = an[4], c2 = an[5];
    const float tfv = (bn.r_relay - ((float)(i - bn.ic) * bn.dx * s1 * c2 + (float)(j - bn.jc) * bn.dy * s1 * s2 +
                                     (float)(k - bn.kc) * bn.dz * c1)) / bn.c_dt;
    const long long fo = (long long)u + (long long)f.fd0[face] * v;
    const double K = (double)bn.K[a];
    // slots: (Ua,Ub) receive (M slot 0, slot 1) except on z faces where the
    // reference orders y before x -- irrelevant for FP64 accumulation.
    const int ua = a == 0 ? 1 : 0, ub = a == 2 ? 1 : 2;
    const float dM0 = f.M[face][0][fo] - f.Mo[face][0][fo];
    const float dM1 = f.M[face][1][fo] - f.Mo[face][1][fo];
    const float dJ0 = f.J[face][0][fo] - f.Jo[face][0][fo];
    const float dJ1 = f.J[face][1][fo] - f.Jo[face][1][fo];
    float t = (float)n + 0.5f + tfv - 1.0f;
    int ii = (int)t;
    float w = t - (float)ii;
    const float base = (1.0f - w) * bn.K[a], wk = w * bn.K[a];
    (void)K;
    double* Ua = bn.U[ua] + kk * bn.LL;
    double* Ub = bn.U[ub] + kk * bn.LL;
    atomicAdd(Ua + ii, (double)(base * dM0));
    atomicAdd(Ua + ii + 1, (double)(wk * dM0));
    atomicAdd(Ub + ii, (double)(base * dM1));
    atomicAdd(Ub + ii + 1, (double)(wk * dM1));
    t = (float)n + tfv - 1.0f;
    ii = (int)t;
    w = t - (float)ii;
    const float base2 = (1.0f - w) * bn.K[a], wk2 = w * bn.K[a];
    double* Wa = bn.W[ua] + kk * bn.LL;
    double* Wb = bn.W[ub] + kk * bn.LL;
    atomicAdd(Wa + ii, (double)(base2 * dJ0));
    atomicAdd(Wa + ii + 1, (double)(wk2 * dJ0));
    atomicAdd(Wb + ii, (double)(base2 * dJ1));
    atomicAdd(Wb + ii + 1, (double)(wk2 * dJ1));
}

// NTFF retarded-time binning with the serial reference's FP32 accumulation
// order (f90:7643-7927).  One work-group per (angle, accumulator, bin
// window); thread t owns bin lo+t and walks the step's contributions in the
// reference's loop order -- face pairs x, y, z; (v outer, u inner); low then
// high face -- adding those that target its bin.  A cell adds to a bin at
// most once per accumulator, so each bin sees exactly the reference's
// sequence of FP32 additions.
struct NtffExact {
    float* acc[6];          // Ux,Uy,Uz,Wx,Wy,Wz as FP32 [angle][LL]
    const int* win;         // per angle: {lowest bin offset relative to n, window length} (host bound)
};

// Ordered binning, one thread per (bin of the step's window, accumulator,
// angle), 256 bins per work-group.  Each thread walks the face cells in the
// reference's loop order -- face pairs x, y, z; rows v (outer) and u (inner);
// low then high face of each (u, v) -- and adds the contributions that land
// in its bin, so each FP32 sum is the serial reference's bit for bit.
// The work-group stages chunks of rows in LDS: every cell's bin and weight
// (the reference's exact FP32 expressions) and its current difference.  A
// thread finds a row's candidate cells from an FP64 linear model of the
// retarded time, widened by a margin that covers the FP32 evaluation error
// (at most a tenth of it is used on the tested geometries), and checks each
// candidate's staged bin.
constexpr int kExactCells = 4096;   // staged cells per chunk (48 KB of LDS); rows of <= 2048 cells

__global__ void __launch_bounds__(256) k_ntff_exact(NtffFaces f, NtffBins bn, NtffExact ex, const int* step) {
    __shared__ int sii[kExactCells];
    __shared__ float sw[kExactCells], sd[kExactCells];
    __shared__ int s_vr[2];
    const int n = *step;
    const int kk = blockIdx.z;
    const int arr = blockIdx.y;               // 0..2 U, 3..5 W
    const int comp = arr % 3;                 // x, y, z component of U or W
    const bool isW = arr >= 3;
    const int slot_in_win = blockIdx.x * blockDim.x + threadIdx.x;
    const int bin = n + ex.win[2 * kk] + slot_in_win;
    const bool own = slot_in_win < ex.win[2 * kk + 1] && bin >= 0 && bin < bn.LL;
    float* A = ex.acc[arr] + kk * bn.LL;
    float acc = own ? A[bin] : 0.0f;
    const float* an = bn.ang + 6 * kk;
    const float s1 = an[2], c1 = an[3], s2 = an[4], c2 = an[5];
    const double cdt = (double)bn.c_dt;
    const double tn = (double)n + (isW ? 0.0 : 0.5) - 1.0;
    const double gg[3] = {(double)bn.dx * s1 * c2, (double)bn.dy * s1 * s2, (double)bn.dz * c1};
    for (int a = 0; a < 3; ++a) {             // face pairs in the reference's order
        if (a == comp || f.pbc[a]) continue;  // faces normal to a feed the other two components
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        const int slot = comp == b ? 0 : 1;   // slot of this component on faces normal to a
        const float K = bn.K[a];
        const int nu = f.hi[b] - f.lo[b] + 1, nv = f.hi[c] - f.lo[c] + 1;
        const int rows = max(1, kExactCells / (2 * nu));
        // FP64 model of the retarded time on side s: t = base[s] + mu * u + mv * v
        const double mu = -(b == 0 ? gg[0] : gg[1]) / cdt, mv = -(c == 1 ? gg[1] : gg[2]) / cdt;
        const double ga = a == 0 ? gg[0] : (a == 1 ? gg[1] : gg[2]);
        const int ica = a == 0 ? bn.ic : (a == 1 ? bn.jc : bn.kc);
        const int icb = b == 0 ? bn.ic : bn.jc, icc = c == 1 ? bn.jc : bn.kc;
        double base[2];
        for (int side = 0; side < 2; ++side) {
            const int pos = side ? f.hi[a] : f.lo[a];
            const double S0 = (double)(pos - ica) * ga - (double)icb * (b == 0 ? gg[0] : gg[1]) -
                              (double)icc * (c == 1 ? gg[1] : gg[2]);
            base[side] = tn + ((double)bn.r_relay - S0) / cdt;
        }
        const double eps = 4e-3 + 4e-7 * fabs(base[0] + mu * f.lo[b] + mv * f.lo[c]);
        const double lo_t = (double)bin - 1.0 - eps, hi_t = (double)bin + 1.0 + eps;
        const double umin = fmin(mu * f.lo[b], mu * f.hi[b]), umax = fmax(mu * f.lo[b], mu * f.hi[b]);
        // rows (per side) whose t range over u meets [lo_t, hi_t]
        int sa[2] = {f.hi[c] + 1, f.hi[c] + 1}, sb[2] = {f.lo[c] - 1, f.lo[c] - 1};
        if (own)
            for (int side = 0; side < 2; ++side) {
                if (fabs(mv) < 1e-12) {
                    const bool in = base[side] + umax >= lo_t && base[side] + umin <= hi_t;
                    sa[side] = in ? f.lo[c] : f.hi[c] + 1;
                    sb[side] = in ? f.hi[c] : f.lo[c] - 1;
                } else {
                    const double v1 = (lo_t - base[side] - umax) / mv, v2 = (hi_t - base[side] - umin) / mv;
                    sa[side] = max(f.lo[c], (int)ceil(fmax(fmin(v1, v2), -1e9)));
                    sb[side] = min(f.hi[c], (int)floor(fmin(fmax(v1, v2), 1e9)));
                }
            }
        const int va = min(sa[0], sa[1]), vb = max(sb[0], sb[1]);
        // rows the work-group needs
        if (threadIdx.x == 0) { s_vr[0] = f.hi[c] + 1; s_vr[1] = f.lo[c] - 1; }
        __syncthreads();
        if (va <= vb) {
            atomicMin(&s_vr[0], va);
            atomicMax(&s_vr[1], vb);
        }
        __syncthreads();
        const int gva = s_vr[0], gvb = s_vr[1];
        const double inv_mu = fabs(mu) < 1e-12 ? 0.0 : 1.0 / mu;
        for (int v0 = gva; v0 <= gvb; v0 += rows) {
            const int nr = min(rows, gvb - v0 + 1);
            __syncthreads();
            // stage rows v0 .. v0+nr-1 of both faces: index (r * 2 + side) * nu + (u - lo)
            for (int q = threadIdx.x; q < nr * 2 * nu; q += blockDim.x) {
                const int uu = q % nu, rs = q / nu, side = rs & 1, r = rs >> 1;
                const int u = f.lo[b] + uu, v = v0 + r;
                const int face = 2 * a + side;
                const int pos = side ? f.hi[a] : f.lo[a];
                int i, j, k;
                if (a == 0) { i = pos; j = u; k = v; }
                else if (a == 1) { i = u; j = pos; k = v; }
                else { i = u; j = v; k = pos; }
                const float tfv = (bn.r_relay - ((float)(i - bn.ic) * bn.dx * s1 * c2 +
                                                 (float)(j - bn.jc) * bn.dy * s1 * s2 +
                                                 (float)(k - bn.kc) * bn.dz * c1)) / bn.c_dt;
                const float t = isW ? ((float)n + tfv - 1.0f) : ((float)n + 0.5f + tfv - 1.0f);
                const int ii = (int)t;
                const long long fo = (long long)u + (long long)f.fd0[face] * v;
                sii[q] = ii;
                sw[q] = t - (float)ii;
                sd[q] = isW ? (f.J[face][slot][fo] - f.Jo[face][slot][fo]) : (f.M[face][slot][fo] - f.Mo[face][slot][fo]);
            }
            __syncthreads();
            // this thread's rows in the chunk: the union of the two sides'
            // ranges, walked in order (the wave then runs as many row
            // iterations as its busiest thread, not the union of all rows)
            const int ra0 = max(sa[0], v0), rb0 = min(sb[0], v0 + nr - 1);
            const int ra1 = max(sa[1], v0), rb1 = min(sb[1], v0 + nr - 1);
            const int vend = max(ra0 <= rb0 ? rb0 : v0 - 1, ra1 <= rb1 ? rb1 : v0 - 1);
            int v = min(ra0 <= rb0 ? ra0 : v0 + nr, ra1 <= rb1 ? ra1 : v0 + nr);
            for (; v <= vend; ++v) {
                if (ra0 <= rb0 && ra1 <= rb1) {          // jump the gap between the ranges
                    if (v > rb0 && v < ra1) v = ra1;
                    if (v > rb1 && v < ra0) v = ra0;
                }
                const int r = v - v0;
                int ulo[2], uhi[2];
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const double bv = base[side] + mv * v;
                    const bool row_in = side ? (v >= ra1 && v <= rb1) : (v >= ra0 && v <= rb0);
                    if (!row_in) {
                        ulo[side] = f.hi[b] + 1;
                        uhi[side] = f.lo[b] - 1;
                    } else if (inv_mu == 0.0) {
                        const bool in = bv >= lo_t && bv <= hi_t;
                        ulo[side] = in ? f.lo[b] : f.hi[b] + 1;
                        uhi[side] = in ? f.hi[b] : f.lo[b] - 1;
                    } else {
                        const double u1 = (lo_t - bv) * inv_mu, u2 = (hi_t - bv) * inv_mu;
                        ulo[side] = max(f.lo[b], (int)ceil(fmax(fmin(u1, u2), -1e9)));
                        uhi[side] = min(f.hi[b], (int)floor(fmin(fmax(u1, u2), 1e9)));
                    }
                }
                // Branch-free over the merged candidate range (the LDS loads of
                // successive cells can then issue together); a cell outside
                // its side's range or bin leaves acc unchanged through a select.
                const int row0 = r * 2 * nu - f.lo[b];
                const bool e0 = ulo[0] <= uhi[0], e1 = ulo[1] <= uhi[1];
                if (e0 && e1) {
                    const int ua = min(ulo[0], ulo[1]), ub = max(uhi[0], uhi[1]);
#pragma unroll 4
                    for (int u = ua; u <= ub; ++u) {
#pragma unroll
                        for (int side = 0; side < 2; ++side) {
                            const int q = row0 + side * nu + u;
                            const int ii = sii[q];
                            const float w = sw[q], d = sd[q];
                            const bool in = u >= ulo[side] && u <= uhi[side];
                            const float a_lo = acc + (1.0f - w) * K * d, a_hi = acc + w * K * d;
                            acc = (in && ii == bin) ? a_lo : ((in && ii + 1 == bin) ? a_hi : acc);
                        }
                    }
                } else if (e0 || e1) {                   // one face only: its own run
                    const int side = e0 ? 0 : 1;
                    const int q0 = row0 + side * nu;
#pragma unroll 4
                    for (int u = ulo[side]; u <= uhi[side]; ++u) {
                        const int ii = sii[q0 + u];
                        const float w = sw[q0 + u], d = sd[q0 + u];
                        const float a_lo = acc + (1.0f - w) * K * d, a_hi = acc + w * K * d;
                        acc = ii == bin ? a_lo : (ii + 1 == bin ? a_hi : acc);
                    }
                }
            }
        }
    }
    if (own) A[bin] = acc;
}

// Ordered binning from contributor lists built at setup (aem_sim_options
// ntff_lists).  A face cell's retarded time tfv does not change from step
// to step, so the bins it can reach are fixed relative to n: offset o =
// int(n + 0.5 + tfv - 1) - n (W: without the 0.5), up to the FP32 rounding
// of that sum near a bin edge.  For every (angle, accumulator) the host
// lists, per bin offset k, the cells whose o can be k or k - 1, in the
// reference's loop order (face pairs x, y, z; rows v outer, u inner; low
// then high face).  The bin's sum then walks its list, forming t, ii and
// the weight with the reference's FP32 expressions and adding the
// contribution when ii or ii + 1 is the bin, so every FP32 sum is the serial
// reference's bit for bit.
//  - Short lists: a wave owns 64 consecutive bins, one per lane, and reads
//    the lists interleaved ([entry][lane]), padded to the wave's longest.
//  - Long lists (a face whose cells share one retarded time, as when the
//    direction is along an axis): a wave owns one bin; its lanes compute 64
//    consecutive contributions at once and the sum takes them in lane
//    order, so the serial chain waits on arithmetic, not on memory.
// ---------------------------------------------------------------------------
struct NtffList {
    const uint2* ent;       // x: element index of the face cell (0xffffffff: padding), y: tfv bits
    const int4* waves;      // x: accumulator row (arr * n_angles + angle), y: first bin offset,
                            // z: first entry, w: entries per lane (>= 0) or -(entries of one long list)
    const unsigned long long* skip;   // per wave of short lists: lanes whose bin a long-list wave owns
    int n_waves;
    const float* Mbase;     // M of every (face, slot) at (2 face + slot) * stride; Mo at + off_old
    const float* Jbase;     // J, likewise; Jo at + off_old
    unsigned off_old;
    int face_shift;         // face of an element index: idx >> face_shift
    float* acc;             // FP32 accumulators [6][n_angles][LL]: Ux,Uy,Uz,Wx,Wy,Wz
};

// Contribution of entry q to bin `bin` at step n (the reference's FP32
// expressions), without branches: the data
// loads are unconditional (a padding or non-contributing entry reads element
// 0), so the loads of a batch of entries are all in flight together; `hit`
// says whether the value counts.
__device__ __forceinline__ float list_term_nb(const NtffList& L, const NtffBins& bn, uint2 q, bool isW, float nf,
                                              int bin, bool& hit) {
    const float tfv = __uint_as_float(q.y);
    const float t = isW ? (nf + tfv - 1.0f) : (nf + 0.5f + tfv - 1.0f);
    const int ii = (int)t;
    hit = q.x != 0xffffffffu && (ii == bin || ii + 1 == bin);
    // the top bit marks an image current of sign -1 (infinite ground plane)
    const unsigned e = hit ? (q.x & 0x7fffffffu) : 0u;
    const float sg = (hit && (q.x >> 31)) ? -1.0f : 1.0f;
    const float wt = t - (float)ii;
    const float* B = isW ? L.Jbase : L.Mbase;
    const float dd = B[e] - B[e + L.off_old];
    const int ax = (int)(e >> L.face_shift) >> 1;
    const float K = ax == 0 ? bn.K[0] : (ax == 1 ? bn.K[1] : bn.K[2]);
    return (ii == bin ? (1.0f - wt) * K * dd : wt * K * dd) * sg;
}

// Entries per batch: their loads are issued together, then their terms are
// added in list order (the list walk is bound by load latency otherwise, with
// about one wave per SIMD in flight).
constexpr int kListBatch = 8;
constexpr int kLongBatch = 4;   // 64-entry chunks of a long list formed together

__device__ __forceinline__ void ntff_list_body(const NtffList& L, const NtffBins& bn, const int* step, int blk) {
    const int w = blk * 4 + (int)(threadIdx.x >> 6);
    if (w >= L.n_waves) return;
    const int lane = (int)(threadIdx.x & 63);
    const int4 d = L.waves[w];
    const int arr = d.x / bn.n_angles;
    const bool isW = arr >= 3;
    const int n = *step;
    const float nf = (float)n;
    float* A = L.acc + (long long)d.x * bn.LL;
    if (d.w >= 0) {                                   // 64 bins, one per lane
        const int bin = n + d.y + lane;
        const bool own = bin >= 0 && bin < bn.LL && !((L.skip[w] >> lane) & 1ull);
        float acc = own ? A[bin] : 0.0f;
        const uint2* e = L.ent + d.z + lane;
        int p = 0;
        for (; p + kListBatch <= d.w; p += kListBatch) {
            uint2 q[kListBatch];
#pragma unroll
            for (int u = 0; u < kListBatch; ++u) q[u] = e[(long long)(p + u) * 64];
            float c[kListBatch];
            bool hit[kListBatch];
#pragma unroll
            for (int u = 0; u < kListBatch; ++u) c[u] = list_term_nb(L, bn, q[u], isW, nf, bin, hit[u]);
#pragma unroll
            for (int u = 0; u < kListBatch; ++u)
                if (hit[u]) acc = acc + c[u];
        }
        for (; p < d.w; ++p) {
            bool hit;
            const float c = list_term_nb(L, bn, e[(long long)p * 64], isW, nf, bin, hit);
            if (hit) acc = acc + c;
        }
        if (own) A[bin] = acc;
        return;
    }
    // one long list: lanes form 64 contributions at a time (kLongBatch
    // chunks together), the sum takes them in order
    const int bin = n + d.y;
    if (bin < 0 || bin >= bn.LL) return;
    float acc = A[bin];
    const int cnt = -d.w;
    for (int p0 = 0; p0 < cnt; p0 += 64 * kLongBatch) {
        float c[kLongBatch];
        bool hit[kLongBatch];
#pragma unroll
        for (int u = 0; u < kLongBatch; ++u) {
            const int p = p0 + u * 64 + lane;
            const uint2 q = p < cnt ? L.ent[d.z + p] : make_uint2(0xffffffffu, 0u);
            c[u] = list_term_nb(L, bn, q, isW, nf, bin, hit[u]);
        }
#pragma unroll
        for (int u = 0; u < kLongBatch; ++u) {
            const unsigned long long m = __ballot(hit[u]);
            if (m == ~0ull) {   // every entry counts (the usual case): no per-entry test
#pragma unroll
                for (int l = 0; l < 64; ++l) acc = acc + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c[u]), l));
            } else {
                for (int l = 0; l < 64; ++l)
                    if ((m >> l) & 1ull) acc = acc + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c[u]), l));
            }
        }
    }
    if (lane == 0) A[bin] = acc;
}

__global__ void __launch_bounds__(256) k_ntff_list(NtffList L, NtffBins bn, const int* step) {
    ntff_list_body(L, bn, step, (int)blockIdx.x);
}

// Per-member arguments of the batched list binning: grid (blocks of the
// member with the most waves, members); a member's surplus blocks exit.
// Each member's bins take the same sums in the same order as its solo run.
struct NtffListArgs {
    NtffList L;
    NtffBins bn;
    const int* step;
};

__global__ void __launch_bounds__(256) k_ntff_list_batch(const NtffListArgs* __restrict__ args) {
    const NtffListArgs& a = args[blockIdx.y];
    ntff_list_body(a.L, a.bn, a.step, (int)blockIdx.x);
}

// NTFF retarded-time binning through an LDS histogram (f90:7643-7927 with
// FP64 accumulation).  One work-group per (angle, accumulator): the step's
// contributions all fall in a window of bins [n + win[0], n + win[0] +
// win[1]) fixed by the NTFF box geometry, so the work-group adds them into an
// FP64 histogram of that window in LDS and then adds the window into its own
// accumulator row -- one global read-modify-write per bin instead of eight
// global atomics per face cell.  Index and weight arithmetic per contribution
// is the reference's FP32, as in k_ntff_bins; only the FP64 summation order
// differs from the serial reference (agreement at FP32 rounding level).
constexpr int kHistMax = 4096;   // bins of one window held in LDS (32 KB)

__device__ __forceinline__ void ntff_hist_body(const NtffFaces& f, const NtffBins& bn, const int* win, int kk,
                                               int arr, const int* step) {
    __shared__ double h[kHistMax];
    const int n = *step;
    const int lo = n + win[2 * kk], len = min(win[2 * kk + 1], kHistMax);
    for (int t = threadIdx.x; t < len; t += blockDim.x) h[t] = 0.0;
    __syncthreads();
    const int comp = arr % 3;
    const bool isW = arr >= 3;
    const float* an = bn.ang + 6 * kk;
    const float s1 = an[2], c1 = an[3], s2 = an[4], c2 = an[5];
    for (int a = 0; a < 3; ++a) {
        if (a == comp || f.pbc[a]) continue;
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        const int slot = comp == b ? 0 : 1;
        const int nu = f.hi[b] - f.lo[b] + 1, nv = f.hi[c] - f.lo[c] + 1;
        // each wave takes 64 consecutive cells of one face side, so at
        // angles where a row of cells shares one time bin the whole wave
        // does (below); the loop bound is a multiple of 128 and whole waves
        // stay converged through it
        const long long nuv = (long long)nu * nv;
        const long long ncell = ((nuv + 63) >> 6) << 7;
        const float K = bn.K[a];
        for (long long cid = threadIdx.x; cid - (long long)(threadIdx.x & 63) < ncell; cid += blockDim.x) {
            const int side = (int)((cid >> 6) & 1);
            const long long uvr = ((cid >> 7) << 6) | (cid & 63);
            const bool live = uvr < nuv;
            const long long uv = live ? uvr : 0;
            const int u = f.lo[b] + (int)(uv % nu), v = f.lo[c] + (int)(uv / nu);
            const int face = 2 * a + side;
            const int pos = side ? f.hi[a] : f.lo[a];
            int i, j, k;
            if (a == 0) { i = pos; j = u; k = v; }
            else if (a == 1) { i = u; j = pos; k = v; }
            else { i = u; j = v; k = pos; }
            const float tfv = (bn.r_relay - ((float)(i - bn.ic) * bn.dx * s1 * c2 + (float)(j - bn.jc) * bn.dy * s1 * s2 +
                                             (float)(k - bn.kc) * bn.dz * c1)) / bn.c_dt;
            const long long fo = (long long)u + (long long)f.fd0[face] * v;
            const float d = isW ? (f.J[face][slot][fo] - f.Jo[face][slot][fo])
                                : (f.M[face][slot][fo] - f.Mo[face][slot][fo]);
            const float t = isW ? ((float)n + tfv - 1.0f) : ((float)n + 0.5f + tfv - 1.0f);
            const int ii = (int)t;
            const float w = t - (float)ii;
            int q = ii - lo;
            if (!live || q < 0 || q + 1 >= len) q = -1;
            const float clo = (1.0f - w) * K * d, chi = w * K * d;
            // a wave whose cells all fall in one bin adds its FP64 wave sum
            // once instead of 64 contending LDS atomics on one address
            const int q0 = __builtin_amdgcn_readfirstlane(q);
            if (__all(q == q0)) {
                if (q0 >= 0) {
                    double slo = (double)clo, shi = (double)chi;
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) {
                        slo += __shfl_xor(slo, o);
                        shi += __shfl_xor(shi, o);
                    }
                    if ((threadIdx.x & 63) == 0) {
                        atomicAdd(&h[q0], slo);
                        atomicAdd(&h[q0 + 1], shi);
                    }
                }
            } else if (q >= 0) {
                atomicAdd(&h[q], (double)clo);
                atomicAdd(&h[q + 1], (double)chi);
            }
        }
    }
    __syncthreads();
    double* A = (isW ? bn.W[comp] : bn.U[comp]) + (long long)kk * bn.LL;
    for (int t = threadIdx.x; t < len; t += blockDim.x) {
        const long long bin = (long long)lo + t;
        if (bin >= 0 && bin < bn.LL) A[bin] += h[t];
    }
}

__global__ void __launch_bounds__(256) k_ntff_hist(NtffFaces f, NtffBins bn, const int* win, const int* step) {
    ntff_hist_body(f, bn, win, blockIdx.x, blockIdx.y, step);
}

// Per-member arguments of the batched far-field kernels.
struct NtffArgs {
    Grid g;
    NtffFaces f;
    NtffBins bn;
    const int* win;
    const float *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
    const int* step;
};

// grid (x-blocks, y-blocks, 6 * members); blocks beyond a member's faces exit
__global__ void k_ntff_currents_batch(const NtffArgs* __restrict__ args) {
    const NtffArgs& a = args[blockIdx.z / 6];
    ntff_currents_body(a.g, a.f, blockIdx.z % 6, a.Ex, a.Ey, a.Ez, a.Hx, a.Hy, a.Hz);
}

// grid (max angles, 6, members)
__global__ void __launch_bounds__(256) k_ntff_hist_batch(const NtffArgs* __restrict__ args) {
    const NtffArgs& a = args[blockIdx.z];
    if ((int)blockIdx.x >= a.bn.n_angles) return;
    ntff_hist_body(a.f, a.bn, a.win, blockIdx.x, blockIdx.y, a.step);
}

__global__ void k_step_inc_batch(int* const* __restrict__ steps, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) *steps[m] = *steps[m] + 1;
}

// Phase-centre samples (f90:8259-8290) and the step counter.
__global__ void k_phase_centre(Grid g, int ic, int jc, int kc, int T, const float* Ex, const float* Ey,
                               const float* Ez, const float* Hx, const float* Hy, const float* Hz, float* pc,
                               const int* step) {
    const int n = *step;
    const long long o = off(g, ic, jc, kc);
    float* P[6];
    for (int q = 0; q < 6; ++q) P[q] = pc + (long long)q * (T + 1);
    P[0][n - 1] = (Ex[o] + Ex[o + g.PX] + Ex[o + g.plane] + Ex[o + g.PX + g.plane]) / 4.0f;
    P[1][n - 1] = (Ey[o] + Ey[o + 1] + Ey[o + g.plane] + Ey[o + 1 + g.plane]) / 4.0f;
    P[2][n - 1] = (Ez[o] + Ez[o + 1] + Ez[o + g.PX] + Ez[o + 1 + g.PX]) / 4.0f;
    const float hx = (Hx[o] + Hx[o + 1]) / 4.0f, hy = (Hy[o] + Hy[o + g.PX]) / 4.0f,
                hz = (Hz[o] + Hz[o + g.plane]) / 4.0f;
    P[3][n - 1] = hx; P[4][n - 1] = hy; P[5][n - 1] = hz;
    if (n < T) { P[3][n] = hx; P[4][n] = hy; P[5][n] = hz; }
}

__global__ void k_step_inc(int* step) { *step = *step + 1; }

// Phase-centre samples (when do_pc) and the step counter in one launch.
__global__ void k_step_end(Grid g, int ic, int jc, int kc, int T, const float* Ex, const float* Ey, const float* Ez,
                           const float* Hx, const float* Hy, const float* Hz, float* pc, int* step, int do_pc) {
    const int n = *step;
    if (do_pc) {
        const long long o = off(g, ic, jc, kc);
        float* P0 = pc;
        float* P1 = pc + (long long)(T + 1);
        float* P2 = pc + 2 * (long long)(T + 1);
        float* P3 = pc + 3 * (long long)(T + 1);
        float* P4 = pc + 4 * (long long)(T + 1);
        float* P5 = pc + 5 * (long long)(T + 1);
        P0[n - 1] = (Ex[o] + Ex[o + g.PX] + Ex[o + g.plane] + Ex[o + g.PX + g.plane]) / 4.0f;
        P1[n - 1] = (Ey[o] + Ey[o + 1] + Ey[o + g.plane] + Ey[o + 1 + g.plane]) / 4.0f;
        P2[n - 1] = (Ez[o] + Ez[o + 1] + Ez[o + g.PX] + Ez[o + 1 + g.PX]) / 4.0f;
        const float hx = (Hx[o] + Hx[o + 1]) / 4.0f, hy = (Hy[o] + Hy[o + g.PX]) / 4.0f,
                    hz = (Hz[o] + Hz[o + g.plane]) / 4.0f;
        P3[n - 1] = hx; P4[n - 1] = hy; P5[n - 1] = hz;
        if (n < T) { P3[n] = hx; P4[n] = hy; P5[n] = hz; }
    }
    *step = n + 1;
}

// ---------------------------------------------------------------------------
// Post-loop products on the device
// ---------------------------------------------------------------------------

// Clear 1-D line of length 2M for the antenna incident wave (f90:8373-8407);
// C_inc(0) (read one element before the array by the reference) is 0.
__global__ void __launch_bounds__(1024) k_clear_line(int M, int T, float kI, float kV, float kS, Pulse pu, float* Vi,
                                                     float* Ci, float* incident) {
    const int M2 = 2 * M;
    for (int n = 1; n <= T; ++n) {
        for (int j = 1 + threadIdx.x; j <= M2 - 1; j += blockDim.x) Ci[j] = Ci[j] - kI * (Vi[j + 1] - Vi[j]);
        __syncthreads();
        if (threadIdx.x == 0) Ci[M] = Ci[M] - kI * pu.src[n - 1];
        __syncthreads();
        for (int j = 1 + threadIdx.x; j <= M2 - 1; j += blockDim.x) Vi[j] = Vi[j] - kV * (Ci[j] - Ci[j - 1]);
        __syncthreads();
        if (threadIdx.x == 0) {
            Vi[M] = Vi[M] + kS * pu.src[T + n - 1];
            incident[n - 1] = Vi[M];
        }
        __syncthreads();
    }
}

__global__ void k_plane_incident(int T, Pulse pu, float* incident) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (n > T) return;
    incident[n - 1] = pu.src[2 * T + n - 1];
}

// NTFF combine and output window (f90:8483-8534): one work-group per angle.
__global__ void __launch_bounds__(256) k_ntff_combine(NtffBins bn, int L, int T, float eta, float dt, float* Eth,
                                                      float* Eph, float* fftime, float* scratch) {
    const int kk = blockIdx.x;
    const float* an = bn.ang + 6 * kk;
    const float s1 = an[2], c1 = an[3], s2 = an[4], c2 = an[5];
    float* Et = scratch + (long long)kk * 2 * (L + 1);
    float* Ep = Et + (L + 1);
    __shared__ int first_t, first_p;
    if (threadIdx.x == 0) { first_t = L; first_p = L; }
    __syncthreads();
    for (int j = 1 + threadIdx.x; j <= L; j += blockDim.x) {
        const long long q = kk * bn.LL + j;
        float wx, wy, wz, ux, uy, uz;
        if (bn.exact) {
            wx = bn.Wf[0][q]; wy = bn.Wf[1][q]; wz = bn.Wf[2][q];
            ux = bn.Uf[0][q]; uy = bn.Uf[1][q]; uz = bn.Uf[2][q];
        } else {
            wx = (float)bn.W[0][q]; wy = (float)bn.W[1][q]; wz = (float)bn.W[2][q];
            ux = (float)bn.U[0][q]; uy = (float)bn.U[1][q]; uz = (float)bn.U[2][q];
        }
        const float wt = wx * c1 * c2 + wy * c1 * s2 - wz * s1;
        const float wph = -1.0f * wx * s2 + wy * c2;
        const float ut = ux * c1 * c2 + uy * c1 * s2 - uz * s1;
        const float uph = -1.0f * ux * s2 + uy * c2;
        const float et = -1.0f * eta * wt - uph;
        const float ep = -1.0f * eta * wph + ut;
        Et[j] = et;
        Ep[j] = ep;
        if (et != 0.0f) atomicMin(&first_t, j);
        if (ep != 0.0f) atomicMin(&first_p, j);
    }
    __syncthreads();
    int dot = min(first_p, first_t);
    if (dot > L - T + 1) dot = L - T + 1;
    for (int j = 1 + threadIdx.x; j <= T; j += blockDim.x) {
        Eth[(long long)kk * T + j - 1] = Et[j + dot - 1];
        Eph[(long long)kk * T + j - 1] = Ep[j + dot - 1];
    }
    if (threadIdx.x == 0) fftime[kk] = (bn.r_relay / 299792448.0f) - (float)(dot - 1) * dt;
}

// Periodic-slab normalisation (f90:8342-8361) to FP32.
__global__ void k_plane_norm(int T, const double* refl, const double* trans, RDiv den, float* rf, float* tf) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= T) return;
    rf[n] = rdiv((float)refl[n], den);
    tf[n] = rdiv((float)trans[n], den);
}

// ---------------------------------------------------------------------------
// Discrete Fourier transform of one output series at arbitrary frequencies,
// in the convention of the reference's post-processor (numpy's FFT,
// utility_scripts/post_processor.py:191-215):  X(f) = sum_n x[n] e^{-2 pi i f n dt},
// n = 0..T-1.  Zero padding adds nothing to the sum, so X at f = k/(P dt)
// equals bin k of the FFT of the series padded to P samples.
//
// FP64 throughout.  One work-group per frequency; thread t sums samples
// t, t+256, ... with Kahan compensation, and the 256 partial sums are added
// in a fixed tree order, so the result does not depend on scheduling.  The
// phase f*dt*n is reduced to [0,1) before sincospi, which keeps it exact to
// FP64 rounding for any series length.  freq holds the frequencies in Hz;
// out receives (re, im) per frequency.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_dft(const float* __restrict__ x, int n, const double* __restrict__ freq,
                                             double dt, double* __restrict__ out) {
    __shared__ double sr[256], si[256];
    const double a = freq[blockIdx.x] * dt;
    double re = 0.0, im = 0.0, cre = 0.0, cim = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) {
        double p = a * (double)t;
        p = p - floor(p);
        double s, c;
        sincospi(2.0 * p, &s, &c);
        const double v = (double)x[t];
        double y = v * c - cre;
        double u = re + y;
        cre = (u - re) - y;
        re = u;
        y = -(v * s) - cim;
        u = im + y;
        cim = (u - im) - y;
        im = u;
    }
    sr[threadIdx.x] = re - cre;
    si[threadIdx.x] = im - cim;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sr[threadIdx.x] += sr[threadIdx.x + w];
            si[threadIdx.x] += si[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sr[0];
        out[2 * blockIdx.x + 1] = si[0];
    }
}

// ===========================================================================
// Differentiation.  A frequency-domain far-field objective and the discrete
// adjoint of the field update (bulk H/E and CPML; the update is linear in
// the fields, so its adjoint is the transposed update run backwards).
// ===========================================================================

// Far-field objective F = r E_pol(theta, phi, f) (e^{jkr} dropped) by the
// frequency-domain near-to-far-field transform of the surface currents that
// k_ntff_currents samples (same faces, same averages):
//   L = sum M e^{jk r^.r'} dS,  N = sum J e^{jk r^.r'} dS,
//   E_theta r = -jk/(4 pi) (L_phi + eta N_theta),  E_phi r = jk/(4 pi) (L_theta - eta N_phi),
// M sampled at t = n dt (E after step n), J at t = (n - 1/2) dt, r' the
// face cell's offset from the phase centre.  Accumulated in FP64; the
// spatial factor pre dS e^{jk r^.r'} of every face cell and the time factors
// of every step are tables built on the host.
struct Objective {
    int lo[3], hi[3], pbc[3];
    int mu, mv;               // face table extents
    double gM[3], gJ[3];      // real projections of M and J components
    const double2* g;         // [6][mv][mu]: pre dS e^{jk r^.r'} of each face cell
    const double2* tE;        // [T]: e^{-jw n dt}
    const double2* tH;        // [T]: e^{-jw (n - 1/2) dt}
    double* part;             // [blocks][2]: per-work-group partial sums of a step
    unsigned* done;           // work-groups finished this step (reset by the last)
    double* F;                // (re, im), device
};

// One surface term of a face cell: value = w * sum_q field[comp][o + offs[q]].
struct FaceTerm {
    int comp, n;
    long long offs[4];
    double w;
};

// The four current terms of a face cell (M slot 0/1, J slot 0/1) exactly as
// k_ntff_currents forms them, for faces normal to axis A; s = +1 on the low
// face, -1 on the high face.  Slot q belongs to component (b, c)[q % 2].
template <int A>
__device__ __forceinline__ void face_terms(const Grid& g, double s, FaceTerm t[4]) {
    const long long X = g.PX, P = g.plane;
    if (A == 0) {
        t[0] = FaceTerm{2, 4, {0, X, 1, 1 + X}, -s / 4};
        t[1] = FaceTerm{1, 4, {0, 1, P, 1 + P}, s / 4};
        t[2] = FaceTerm{5, 2, {0, P, 0, 0}, s / 2};
        t[3] = FaceTerm{4, 2, {0, X, 0, 0}, -s / 2};
    } else if (A == 1) {
        t[0] = FaceTerm{2, 4, {0, 1, X, 1 + X}, s / 4};
        t[1] = FaceTerm{0, 4, {0, X, P, X + P}, -s / 4};
        t[2] = FaceTerm{5, 2, {0, P, 0, 0}, -s / 2};
        t[3] = FaceTerm{3, 2, {0, 1, 0, 0}, s / 2};
    } else {
        t[0] = FaceTerm{1, 4, {0, 1, P, 1 + P}, -s / 4};
        t[1] = FaceTerm{0, 4, {0, X, P, X + P}, s / 4};
        t[2] = FaceTerm{4, 2, {0, X, 0, 0}, s / 2};
        t[3] = FaceTerm{3, 2, {0, 1, 0, 0}, -s / 2};
    }
}

// Complex weight of each face term at step n:  dF/d(term value), for a
// face normal to axis A (axes as template parameters keep the kernel
// argument struct out of scratch).
template <int A>
__device__ __forceinline__ void face_weights(const Objective& ob, int face, int uu, int vv, int n, double wr[4],
                                             double wi[4]) {
    constexpr int b = A == 0 ? 1 : 0, c = A == 2 ? 1 : 2;
    const double2 g = ob.g[((long long)face * ob.mv + vv) * ob.mu + uu];
    const double2 e = ob.tE[n - 1], h = ob.tH[n - 1];
    const double eMr = g.x * e.x - g.y * e.y, eMi = g.x * e.y + g.y * e.x;
    const double eJr = g.x * h.x - g.y * h.y, eJi = g.x * h.y + g.y * h.x;
    wr[0] = eMr * ob.gM[b]; wi[0] = eMi * ob.gM[b];
    wr[1] = eMr * ob.gM[c]; wi[1] = eMi * ob.gM[c];
    wr[2] = eJr * ob.gJ[b]; wi[2] = eJi * ob.gJ[b];
    wr[3] = eJr * ob.gJ[c]; wi[3] = eJi * ob.gJ[c];
}

// Face cell (u, v) of the NTFF box face normal to A (high face if `high`)
// -> cell offset; false outside.
template <int A>
__device__ __forceinline__ bool face_cell(const Grid& g, const Objective& ob, bool high, int uu, int vv, long long& o) {
    constexpr int b = A == 0 ? 1 : 0, c = A == 2 ? 1 : 2;
    if (ob.pbc[A]) return false;
    const int u = ob.lo[b] + uu, v = ob.lo[c] + vv;
    if (u > ob.hi[b] || v > ob.hi[c]) return false;
    const int pos = high ? ob.hi[A] : ob.lo[A];
    if (A == 0) o = off(g, pos, u, v);
    else if (A == 1) o = off(g, u, pos, v);
    else o = off(g, u, v, pos);
    return true;
}

// Field component c (0..5 = Ex..Hz) by selects rather than an indexed array
// of pointers (which would live in scratch).
__device__ __forceinline__ const float* pick6(int c, const float* Ex, const float* Ey, const float* Ez,
                                              const float* Hx, const float* Hy, const float* Hz) {
    return c == 0 ? Ex : c == 1 ? Ey : c == 2 ? Ez : c == 3 ? Hx : c == 4 ? Hy : Hz;
}
__device__ __forceinline__ float* pick6(int c, float* Ex, float* Ey, float* Ez, float* Hx, float* Hy, float* Hz) {
    return c == 0 ? Ex : c == 1 ? Ey : c == 2 ? Ez : c == 3 ? Hx : c == 4 ? Hy : Hz;
}

template <int A>
__device__ __forceinline__ void objective_terms(const Grid& g, const Objective& ob, int face, int uu, int vv,
                                                const int* step, const float* Ex, const float* Ey, const float* Ez,
                                                const float* Hx, const float* Hy, const float* Hz, double& re,
                                                double& im) {
    long long o;
    if (!face_cell<A>(g, ob, face & 1, uu, vv, o)) return;
    double wr[4], wi[4];
    face_weights<A>(ob, face, uu, vv, *step, wr, wi);
    FaceTerm t[4];
    face_terms<A>(g, (face & 1) ? -1.0 : 1.0, t);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float* Fq = pick6(t[q].comp, Ex, Ey, Ez, Hx, Hy, Hz) + o;
        double v = (double)Fq[t[q].offs[0]] + (double)Fq[t[q].offs[1]];
        if (t[q].n == 4) v = (v + (double)Fq[t[q].offs[2]]) + (double)Fq[t[q].offs[3]];
        const double val = t[q].w * v;
        re += wr[q] * val;
        im += wi[q] * val;
    }
}

// Per-work-group partial sums of this step's terms; the last work-group to
// finish adds them to F in a fixed order.  grid (x, y, 6 faces), block 256.
__global__ void __launch_bounds__(256) k_objective(Grid g, Objective ob, const float* Ex, const float* Ey,
                                                   const float* Ez, const float* Hx, const float* Hy,
                                                   const float* Hz, const int* step) {
    const int face = blockIdx.z;
    const int tid = threadIdx.x;
    double re = 0.0, im = 0.0;
    const int uu = blockIdx.x * 32 + (tid & 31), vv = blockIdx.y * 8 + (tid >> 5);
    switch (face / 2) {
        case 0: objective_terms<0>(g, ob, face, uu, vv, step, Ex, Ey, Ez, Hx, Hy, Hz, re, im); break;
        case 1: objective_terms<1>(g, ob, face, uu, vv, step, Ex, Ey, Ez, Hx, Hy, Hz, re, im); break;
        default: objective_terms<2>(g, ob, face, uu, vv, step, Ex, Ey, Ez, Hx, Hy, Hz, re, im); break;
    }
    __shared__ double sr[256], si[256];
    __shared__ bool last;
    sr[tid] = re;
    si[tid] = im;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            sr[tid] += sr[tid + w];
            si[tid] += si[tid + w];
        }
        __syncthreads();
    }
    const int nblk = gridDim.x * gridDim.y * gridDim.z;
    if (tid == 0) {
        const long long blk = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
        ob.part[2 * blk] = sr[0];
        ob.part[2 * blk + 1] = si[0];
        __threadfence();
        last = atomicAdd(ob.done, 1u) == (unsigned)nblk - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    re = 0.0;
    im = 0.0;
    for (int q = tid; q < nblk; q += 256) {
        re += ob.part[2 * q];
        im += ob.part[2 * q + 1];
    }
    sr[tid] = re;
    si[tid] = im;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            sr[tid] += sr[tid + w];
            si[tid] += si[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        ob.F[0] = ob.F[0] + sr[0];
        ob.F[1] = ob.F[1] + si[0];
        *ob.done = 0u;
    }
}

// ---------------------------------------------------------------------------
// Running-DFT far field: the frequency-domain near-to-far-field transform of
// the same face currents, at n_ang directions and n_freq frequencies
// f0 + q df, accumulated every step.  Per (direction, frequency) pair:
//   S_theta = e_M U_theta + e_J V_theta,   U_theta = sum dS e^{jk r^.r'} M.phi^,
//                                          V_theta = sum dS e^{jk r^.r'} eta J.theta^
//   S_phi   = e_M U_phi   - e_J V_phi,     U_phi = ... M.theta^, V_phi = ... eta J.phi^
// with e_M = e^{-jw n dt}, e_J = e^{-jw (n-1/2) dt}; r E_theta = -jk/(4 pi)
// S_theta and r E_phi = jk/(4 pi) S_phi are formed when read.  r^.r' is
// linear along a face, so within a work-group's 32 x 8 tile the spatial
// factor advances by one complex product per cell from a host-built table
// (the tile's first cell and the per-row and per-column steps).  A
// work-group stages its tile's currents in LDS, each thread then owns whole
// (direction, frequency) pairs and writes the tile's sums; k_ffdft_sum adds
// the tiles in order, applies the step's time factors and folds the result
// into FP64 totals with Kahan compensation.
// ---------------------------------------------------------------------------
struct FfDft {
    int lo[3], hi[3], pbc[3];
    int mu, mv;              // face extents of the launch grid
    int n_ang, n_freq;
    double dt, w0, dw, eta;
    const double* dir;       // [n_ang][9]: r^, theta^, phi^
    const double* step;      // [6 faces][n_pairs][4]: e^{jk r^.d_u} (re, im), e^{jk r^.d_v} (re, im)
    const double* z0;        // [tiles][n_pairs][2]: dS e^{jk r^.r'} at each tile's first cell
    double* part;            // [n_pairs][tiles][8]: U_theta, V_theta, U_phi, V_phi (re, im)
    double* tot;             // [n_pairs][4]: S_theta (re, im), S_phi (re, im)
    double* comp;            // Kahan compensation of tot
};

template <int A>
__device__ __forceinline__ bool ffdft_cell(const Grid& g, const FfDft& p, int face, int uu, int vv, const float* Ex,
                                           const float* Ey, const float* Ez, const float* Hx, const float* Hy,
                                           const float* Hz, double cur[4]) {
    constexpr int b = A == 0 ? 1 : 0, c = A == 2 ? 1 : 2;
    if (p.pbc[A]) return false;
    const int u = p.lo[b] + uu, v = p.lo[c] + vv;
    if (u > p.hi[b] || v > p.hi[c]) return false;
    const int pos = (face & 1) ? p.hi[A] : p.lo[A];
    int ijk[3];
    ijk[A] = pos;
    ijk[b] = u;
    ijk[c] = v;
    const long long o = off(g, ijk[0], ijk[1], ijk[2]);
    FaceTerm t[4];
    face_terms<A>(g, (face & 1) ? -1.0 : 1.0, t);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float* Fq = pick6(t[q].comp, Ex, Ey, Ez, Hx, Hy, Hz) + o;
        double s = (double)Fq[t[q].offs[0]] + (double)Fq[t[q].offs[1]];
        if (t[q].n == 4) s = (s + (double)Fq[t[q].offs[2]]) + (double)Fq[t[q].offs[3]];
        cur[q] = t[q].w * s;
    }
    return true;
}

// grid (x, y, 6 faces) of 32 x 8 tiles over the face cells, block 256.
__global__ void __launch_bounds__(256) k_ffdft(Grid g, FfDft p, const float* Ex, const float* Ey, const float* Ez,
                                               const float* Hx, const float* Hy, const float* Hz) {
    __shared__ double sc[256][4];
    const int face = blockIdx.z, a = face / 2;
    const int tid = threadIdx.x;
    {
        const int uu = blockIdx.x * 32 + (tid & 31), vv = blockIdx.y * 8 + (tid >> 5);
        double cur[4] = {0, 0, 0, 0};
        switch (a) {
            case 0: ffdft_cell<0>(g, p, face, uu, vv, Ex, Ey, Ez, Hx, Hy, Hz, cur); break;
            case 1: ffdft_cell<1>(g, p, face, uu, vv, Ex, Ey, Ez, Hx, Hy, Hz, cur); break;
            default: ffdft_cell<2>(g, p, face, uu, vv, Ex, Ey, Ez, Hx, Hy, Hz, cur); break;
        }
        sc[tid][0] = cur[0];   // outside the face: 0
        sc[tid][1] = cur[1];
        sc[tid][2] = cur[2];
        sc[tid][3] = cur[3];
    }
    __syncthreads();
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    const int npair = p.n_ang * p.n_freq;
    const long long tile = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
    for (int pr = tid; pr < npair; pr += blockDim.x) {
        const double* dv = p.dir + 9 * (pr / p.n_freq);
        const double tb = dv[3 + b], tc = dv[3 + c], pb = dv[6 + b], pc = dv[6 + c];
        const double* st = p.step + ((long long)face * npair + pr) * 4;
        const double ar = st[0], ai = st[1], br = st[2], bi = st[3];
        double zr0 = p.z0[(tile * npair + pr) * 2], zi0 = p.z0[(tile * npair + pr) * 2 + 1];
        double uth_r = 0, uth_i = 0, vth_r = 0, vth_i = 0, uph_r = 0, uph_i = 0, vph_r = 0, vph_i = 0;
        for (int vv = 0; vv < 8; ++vv) {
            double zr = zr0, zi = zi0;
#pragma unroll 4
            for (int uu = 0; uu < 32; ++uu) {
                const double* cl = sc[vv * 32 + uu];
                const double mth = cl[0] * tb + cl[1] * tc, mph = cl[0] * pb + cl[1] * pc;
                const double jth = p.eta * (cl[2] * tb + cl[3] * tc), jph = p.eta * (cl[2] * pb + cl[3] * pc);
                uth_r += zr * mph; uth_i += zi * mph;
                vth_r += zr * jth; vth_i += zi * jth;
                uph_r += zr * mth; uph_i += zi * mth;
                vph_r += zr * jph; vph_i += zi * jph;
                const double t0 = zr * ar - zi * ai;
                zi = zr * ai + zi * ar;
                zr = t0;
            }
            const double t0 = zr0 * br - zi0 * bi;
            zi0 = zr0 * bi + zi0 * br;
            zr0 = t0;
        }
        double* out = p.part + ((long long)pr * gridDim.x * gridDim.y * gridDim.z + tile) * 8;
        out[0] = uth_r; out[1] = uth_i; out[2] = vth_r; out[3] = vth_i;
        out[4] = uph_r; out[5] = uph_i; out[6] = vph_r; out[7] = vph_i;
    }
}

// One work-group of 256 per (direction, frequency) pair: the tiles' sums in
// a fixed order (strided per thread, then a tree), the step's time factors,
// and the Kahan-compensated totals.
__global__ void __launch_bounds__(256) k_ffdft_sum(FfDft p, int ntile, const int* step) {
    __shared__ double red[8][256];
    const int pr = blockIdx.x, tid = threadIdx.x;
    double u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = tid; t < ntile; t += 256) {
        const double* in = p.part + ((long long)pr * ntile + t) * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) u[e] += in[e];
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) red[e][tid] = u[e];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
#pragma unroll
            for (int e = 0; e < 8; ++e) red[e][tid] += red[e][tid + w];
        __syncthreads();
    }
    if (tid != 0) return;
#pragma unroll
    for (int e = 0; e < 8; ++e) u[e] = red[e][0];
    const int n = *step;
    const double w = p.w0 + (pr % p.n_freq) * p.dw;
    double sM, cM, sJ, cJ;
    sincos(-w * (n * p.dt), &sM, &cM);
    sincos(-w * ((n - 0.5) * p.dt), &sJ, &cJ);
    // S_theta = e_M U_theta + e_J V_theta,  S_phi = e_M U_phi - e_J V_phi
    const double add[4] = {cM * u[0] - sM * u[1] + (cJ * u[2] - sJ * u[3]),
                           cM * u[1] + sM * u[0] + (cJ * u[3] + sJ * u[2]),
                           cM * u[4] - sM * u[5] - (cJ * u[6] - sJ * u[7]),
                           cM * u[5] + sM * u[4] - (cJ * u[7] + sJ * u[6])};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const long long v = (long long)pr * 4 + e;
        const double y = add[e] - p.comp[v];
        const double t = p.tot[v] + y;
        p.comp[v] = (t - p.tot[v]) - y;
        p.tot[v] = t;
    }
}

// Adjoint source of J = |F|^2 at step n: dJ/d(term) = 2 Re(conj(F) w),
// spread onto the adjoint fields through each term's average, in a fixed
// summation order (run to run bit-identical):
// k_adj_source_terms stores each face cell's four increments dj, and
// k_adj_source_gather gives every adjoint element near the far-field box to
// one thread, which adds the increments that reach it in face, then term
// offset order.  D is [6 faces][4 terms][mv][mu].
template <int A>
__device__ __forceinline__ void source_dj(const Grid& g, const Objective& ob, int face, int uu, int vv,
                                          const int* step, double Fr, double Fi, float* D) {
    long long o;
    if (!face_cell<A>(g, ob, face & 1, uu, vv, o)) return;
    double wr[4], wi[4];
    face_weights<A>(ob, face, uu, vv, *step, wr, wi);
    FaceTerm t[4];
    face_terms<A>(g, (face & 1) ? -1.0 : 1.0, t);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        D[(((long long)face * 4 + q) * ob.mv + vv) * ob.mu + uu] = (float)(2.0 * (Fr * wr[q] + Fi * wi[q]) * t[q].w);
}

__global__ void __launch_bounds__(256) k_adj_source_terms(Grid g, Objective ob, double Fr, double Fi, float* D,
                                                          const int* step) {
    const int face = blockIdx.z;
    const int uu = blockIdx.x * 32 + (threadIdx.x & 31), vv = blockIdx.y * 8 + (threadIdx.x >> 5);
    switch (face / 2) {
        case 0: source_dj<0>(g, ob, face, uu, vv, step, Fr, Fi, D); break;
        case 1: source_dj<1>(g, ob, face, uu, vv, step, Fr, Fi, D); break;
        default: source_dj<2>(g, ob, face, uu, vv, step, Fr, Fi, D); break;
    }
}

// Term q of a face normal to A: its field component and its n cell offsets
// (di, dj, dk) -- the same stencils as face_terms.
__device__ __forceinline__ int term_stencil(int A, int q, int d[4][3]) {
    for (int r = 0; r < 4; ++r) d[r][0] = d[r][1] = d[r][2] = 0;
    // the two tangential axes (b, c) and the normal A
    const int b = A == 0 ? 1 : 0, c = A == 2 ? 1 : 2;
    if (q == 0 || q == 1) {          // averages of an E component over 4 edges
        // q0: component c-axis E with offsets {0, +b, +A, +A+b}; q1: component b-axis E, {0, +A, +c, +A+c}
        // (the order of face_terms for each A)
        if (A == 0) {
            if (q == 0) { d[1][1] = 1; d[2][0] = 1; d[3][0] = 1; d[3][1] = 1; return 2 * 10 + 4; }   // Ez {0,X,1,1+X}
            d[1][0] = 1; d[2][2] = 1; d[3][0] = 1; d[3][2] = 1; return 1 * 10 + 4;                    // Ey {0,1,P,1+P}
        } else if (A == 1) {
            if (q == 0) { d[1][0] = 1; d[2][1] = 1; d[3][0] = 1; d[3][1] = 1; return 2 * 10 + 4; }   // Ez {0,1,X,1+X}
            d[1][1] = 1; d[2][2] = 1; d[3][1] = 1; d[3][2] = 1; return 0 * 10 + 4;                    // Ex {0,X,P,X+P}
        } else {
            if (q == 0) { d[1][0] = 1; d[2][2] = 1; d[3][0] = 1; d[3][2] = 1; return 1 * 10 + 4; }   // Ey {0,1,P,1+P}
            d[1][1] = 1; d[2][2] = 1; d[3][1] = 1; d[3][2] = 1; return 0 * 10 + 4;                    // Ex {0,X,P,X+P}
        }
    }
    (void)b; (void)c;
    if (A == 0) {
        if (q == 2) { d[1][2] = 1; return 5 * 10 + 2; }   // Hz {0,P}
        d[1][1] = 1; return 4 * 10 + 2;                   // Hy {0,X}
    } else if (A == 1) {
        if (q == 2) { d[1][2] = 1; return 5 * 10 + 2; }   // Hz {0,P}
        d[1][0] = 1; return 3 * 10 + 2;                   // Hx {0,1}
    } else {
        if (q == 2) { d[1][1] = 1; return 4 * 10 + 2; }   // Hy {0,X}
        d[1][0] = 1; return 3 * 10 + 2;                   // Hx {0,1}
    }
}

// One thread per cell of the shells {lo, lo+1} and {hi, hi+1} of each axis
// of the far-field box (tangentially lo..hi+1), each cell taken by the first
// shell holding it; grid z = shell * 2 + layer.
template <typename IdxT>
__global__ void __launch_bounds__(256) k_adj_source_gather(Grid g, Objective ob, const float* __restrict__ D,
                                                           float* lEx, float* lEy, float* lEz, float* lHx,
                                                           float* lHy, float* lHz, float* mEx, float* mEy,
                                                           float* mEz, const IdxT* idx, const float* tup) {
    const int sh = (int)blockIdx.z >> 1, layer = (int)blockIdx.z & 1;
    const int A = sh / 2;
    if (ob.pbc[A]) return;
    const int b = A == 0 ? 1 : 0, c = A == 2 ? 1 : 2;
    const int uu = blockIdx.x * 32 + (threadIdx.x & 31), vv = blockIdx.y * 8 + (threadIdx.x >> 5);
    if (uu > ob.hi[b] - ob.lo[b] + 1 || vv > ob.hi[c] - ob.lo[c] + 1) return;
    int x[3];
    x[A] = ((sh & 1) ? ob.hi[A] : ob.lo[A]) + layer;
    x[b] = ob.lo[b] + uu;
    x[c] = ob.lo[c] + vv;
    // taken by an earlier shell?
    for (int s2 = 0; s2 < sh; ++s2) {
        const int A2 = s2 / 2;
        if (ob.pbc[A2]) continue;
        const int p2 = (s2 & 1) ? ob.hi[A2] : ob.lo[A2];
        bool in = x[A2] == p2 || x[A2] == p2 + 1;
        for (int a = 0; a < 3; ++a)
            if (a != A2 && (x[a] < ob.lo[a] || x[a] > ob.hi[a] + 1)) in = false;
        if (in) return;
    }
    const long long o = off(g, x[0], x[1], x[2]);
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool hit[6] = {false, false, false, false, false, false};
    for (int f = 0; f < 6; ++f) {
        const int Af = f / 2;
        if (ob.pbc[Af]) continue;
        const int bf = Af == 0 ? 1 : 0, cf = Af == 2 ? 1 : 2;
        const int pf = (f & 1) ? ob.hi[Af] : ob.lo[Af];
        // a face reaches only the two planes pf, pf+1 and lo..hi+1 across
        if ((x[Af] != pf && x[Af] != pf + 1) || x[bf] < ob.lo[bf] || x[bf] > ob.hi[bf] + 1 || x[cf] < ob.lo[cf] ||
            x[cf] > ob.hi[cf] + 1)
            continue;
        for (int q = 0; q < 4; ++q) {
            int d[4][3];
            const int code = term_stencil(Af, q, d);
            const int comp = code / 10, n = code % 10;
            for (int r = 0; r < n; ++r) {
                const int y0 = x[0] - d[r][0], y1 = x[1] - d[r][1], y2 = x[2] - d[r][2];
                const int y[3] = {y0, y1, y2};
                if (y[Af] != pf || y[bf] < ob.lo[bf] || y[bf] > ob.hi[bf] || y[cf] < ob.lo[cf] || y[cf] > ob.hi[cf])
                    continue;
                const float dj = D[(((long long)f * 4 + q) * ob.mv + (y[cf] - ob.lo[cf])) * ob.mu + (y[bf] - ob.lo[bf])];
                acc[comp] = acc[comp] + dj;
                hit[comp] = true;
            }
        }
    }
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        if (!hit[comp]) continue;
        float* L = pick6(comp, lEx, lEy, lEz, lHx, lHy, lHz);
        L[o] = L[o] + acc[comp];
        if (comp < 3) {   // keep mu = gb lambda_E of the E components
            float* M = pick6(comp, mEx, mEy, mEz, mEx, mEy, mEz);
            M[o] = M[o] + tup[(int)idx[o] * 6 + 2 * comp + 1] * acc[comp];
        }
    }
}

// Design-box record of one step (between the H and E phases): the E value
// each design edge starts the step with and the curl term S of its update
// E' = ga E + gb S, both in the bulk kernel's FP32 arithmetic.
struct AdjRec {
    int lo[3];        // first edge (1-based) of the record box
    int n[3];         // extents
    long long nedge;  // n[0] * n[1] * n[2]
    float* hist;      // [slot][2 (E, S)][3 comps][nedge]; step n in slot (n-1) % slots
    int slots;        // time_steps, or the checkpoint interval
};

__global__ void k_adj_record(Grid g, AdjRec r, const float* Ex, const float* Ey, const float* Ez, const float* Hx,
                             const float* Hy, const float* Hz, const float* dex, const float* dey, const float* dez,
                             const int* step) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= r.nedge) return;
    const int i = r.lo[0] + (int)(e % r.n[0]);
    const long long rr = e / r.n[0];
    const int j = r.lo[1] + (int)(rr % r.n[1]), k = r.lo[2] + (int)(rr / r.n[1]);
    const long long o = off(g, i, j, k);
    const float dx_i = dex[i - 1], dy_j = dey[j - 1], dz_k = dez[k - 1];
    const float hx = Hx[o], hy = Hy[o], hz = Hz[o];
    const float s0 = (hz - Hz[o - g.PX]) * dy_j + (Hy[o - g.plane] - hy) * dz_k;
    const float s1 = (Hz[o - 1] - hz) * dx_i + (hx - Hx[o - g.plane]) * dz_k;
    const float s2 = (hy - Hy[o - 1]) * dx_i + (Hx[o - g.PX] - hx) * dy_j;
    float* h = r.hist + (long long)((*step - 1) % r.slots) * 6 * r.nedge;
    h[e] = Ex[o];
    h[r.nedge + e] = Ey[o];
    h[2 * r.nedge + e] = Ez[o];
    h[3 * r.nedge + e] = s0;
    h[4 * r.nedge + e] = s1;
    h[5 * r.nedge + e] = s2;
}

// dJ/dga and dJ/dgb of every design edge from step n: lambda_E' * E and lambda_E' * S.
__global__ void k_adj_grad(Grid g, AdjRec r, const float* lEx, const float* lEy, const float* lEz, double* dga,
                           double* dgb, const int* step) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= r.nedge) return;
    const int i = r.lo[0] + (int)(e % r.n[0]);
    const long long rr = e / r.n[0];
    const int j = r.lo[1] + (int)(rr % r.n[1]), k = r.lo[2] + (int)(rr / r.n[1]);
    const long long o = off(g, i, j, k);
    const float* h = r.hist + (long long)((*step - 1) % r.slots) * 6 * r.nedge;
    const float lam[3] = {lEx[o], lEy[o], lEz[o]};
    for (int c = 0; c < 3; ++c) {
        dga[c * r.nedge + e] += (double)lam[c] * (double)h[c * r.nedge + e];
        dgb[c * r.nedge + e] += (double)lam[c] * (double)h[(3 + c) * r.nedge + e];
    }
}

// Is E component c at (i,j,k) updated by the bulk loop (outer PEC tangential
// E is not)?  Ex: j,k >= 2;  Ey: i,k >= 2;  Ez: i,j >= 2; all <= N-1.
__device__ __forceinline__ bool e_upd(const Grid& g, int c, int i, int j, int k) {
    if (i < 1 || j < 1 || k < 1 || i > g.NX - 1 || j > g.NY - 1 || k > g.NZ - 1) return false;
    if (c == 0) return j >= 2 && k >= 2;
    if (c == 1) return i >= 2 && k >= 2;
    return i >= 2 && j >= 2;
}
__device__ __forceinline__ bool h_in(const Grid& g, int i, int j, int k) {
    return i >= 1 && j >= 1 && k >= 1 && i <= g.NX - 1 && j <= g.NY - 1 && k <= g.NZ - 1;
}

// CPML psi cell of the E (slab_e) or H (slab_h) update normal to axis A at
// (i,j,k): face f, local layer q and psi offset po (as in k_e_cpml_* /
// k_h_cpml_*); false outside the slab or its transverse range.
__device__ __forceinline__ bool e_psi(const Grid& g, const Cpml& p, int A, int i, int j, int k, int& f, int& q,
                                      long long& po) {
    const int N[3] = {g.NX, g.NY, g.NZ};
    const int x[3] = {i, j, k};
    const int o1 = A == 0 ? 1 : 0, o2 = A == 2 ? 1 : 2;
    if (x[o1] < 1 || x[o1] > N[o1] - 1 || x[o2] < 1 || x[o2] > N[o2] - 1) return false;
    int t;
    const int side = slab_e(x[A], N[A], p.n[2 * A], p.n[2 * A + 1], q, t);
    if (side < 0) return false;
    f = 2 * A + side;
    const long long ntf = p.n[f] - 1;
    if (A == 0) po = t + ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
    else if (A == 1) po = (long long)(i - 1) + (long long)g.PX * (t + ntf * (k - 1));
    else po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * t);
    return true;
}
__device__ __forceinline__ bool h_psi(const Grid& g, const Cpml& p, int A, int i, int j, int k, int& f, int& q,
                                      long long& po) {
    const int N[3] = {g.NX, g.NY, g.NZ};
    const int x[3] = {i, j, k};
    const int o1 = A == 0 ? 1 : 0, o2 = A == 2 ? 1 : 2;
    if (x[o1] < 1 || x[o1] > N[o1] - 1 || x[o2] < 1 || x[o2] > N[o2] - 1 || x[A] < 1) return false;
    const int side = slab_h(x[A], N[A], p.n[2 * A], p.n[2 * A + 1], q);
    if (side < 0) return false;
    f = 2 * A + side;
    const long long ntf = p.n[f] - 1;
    if (A == 0) po = (q - 1) + ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
    else if (A == 1) po = (long long)(i - 1) + (long long)g.PX * ((q - 1) + ntf * (k - 1));
    else po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * (q - 1));
    return true;
}

// Slot conditions of the E psi (slot 0/1 of axis A): x: (Ey, k>=2), (Ez, j>=2);
// y: (Ex, k>=2), (Ez, i>=2); z: (Ex, j>=2), (Ey, i>=2).
__device__ __forceinline__ bool e_slot_ok(int A, int slot, int i, int j, int k) {
    if (A == 0) return slot == 0 ? k >= 2 : j >= 2;
    if (A == 1) return slot == 0 ? k >= 2 : i >= 2;
    return slot == 0 ? j >= 2 : i >= 2;
}
__device__ __forceinline__ int e_slot_comp(int A, int slot) { return A == 0 ? 1 + slot : (A == 1 ? 2 * slot : slot); }
__device__ __forceinline__ int h_slot_comp(int A, int slot) { return A == 0 ? 1 + slot : (A == 1 ? 2 * slot : slot); }

// c * nu of an E / H psi cell, 0 where there is none.  The adjoint CPML
// state holds nu = lambda_psi' + (gb or DB) lambda_field' of the step being
// transposed; lambda_psi of the step before is b nu, folded into the next
// update of nu.
__device__ __forceinline__ float cnu_e(const Grid& g, const Cpml& p, const Cpml& lp, int A, int slot, int i, int j,
                                       int k) {
    int f, q;
    long long po;
    if (!e_psi(g, p, A, i, j, k, f, q, po) || !e_slot_ok(A, slot, i, j, k)) return 0.0f;
    return p.c[f][q - 1] * lp.psi[f][slot][po];
}
__device__ __forceinline__ float cnu_h(const Grid& g, const Cpml& p, const Cpml& lp, int A, int slot, int i, int j,
                                       int k) {
    int f, q;
    long long po;
    if (!h_psi(g, p, A, i, j, k, f, q, po)) return 0.0f;
    return p.c[f][q - 1] * lp.psi[f][slot][po];
}

// c * nu of both slots of the E psi normal to axis A at the cell and at its
// +1 neighbour along A (cn[slot][0: cell, 1: neighbour]); 0 where absent.
// The transverse coordinates are those of an H cell, inside the psi range.
__device__ __forceinline__ void cnu_e_pair(const Grid& g, const Cpml& p, const Cpml& lp, int A, int i, int j, int k,
                                           float cn[2][2]) {
    const int N[3] = {g.NX, g.NY, g.NZ};
    const int x[3] = {i, j, k};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int q, t;
        const int side = slab_e(x[A] + h, N[A], p.n[2 * A], p.n[2 * A + 1], q, t);
        if (side < 0) {
            cn[0][h] = cn[1][h] = 0.0f;
            continue;
        }
        const int f = 2 * A + side;
        const long long ntf = p.n[f] - 1;
        long long po;
        if (A == 0) po = t + ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
        else if (A == 1) po = (long long)(i - 1) + (long long)g.PX * (t + ntf * (k - 1));
        else po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * t);
        const float cc = p.c[f][q - 1];
        int xx[3] = {i, j, k};
        xx[A] += h;
        cn[0][h] = e_slot_ok(A, 0, xx[0], xx[1], xx[2]) ? cc * lp.psi[f][0][po] : 0.0f;
        cn[1][h] = e_slot_ok(A, 1, xx[0], xx[1], xx[2]) ? cc * lp.psi[f][1][po] : 0.0f;
    }
}

// c * nu of both slots of the H psi normal to axis A at the cell and at its
// -1 neighbour along A (cn[slot][0: cell, 1: neighbour]); 0 where absent.
// The transverse coordinates are those of an updated E cell.
__device__ __forceinline__ void cnu_h_pair(const Grid& g, const Cpml& p, const Cpml& lp, int A, int i, int j, int k,
                                           float cn[2][2]) {
    const int N[3] = {g.NX, g.NY, g.NZ};
    const int x[3] = {i, j, k};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int xa = x[A] - h;
        int q;
        const int side = xa >= 1 ? slab_h(xa, N[A], p.n[2 * A], p.n[2 * A + 1], q) : -1;
        if (side < 0) {
            cn[0][h] = cn[1][h] = 0.0f;
            continue;
        }
        const int f = 2 * A + side;
        const long long ntf = p.n[f] - 1;
        long long po;
        if (A == 0) po = (q - 1) + ntf * ((long long)(j - 1) + (long long)g.NY * (k - 1));
        else if (A == 1) po = (long long)(i - 1) + (long long)g.PX * ((q - 1) + ntf * (k - 1));
        else po = (long long)(i - 1) + (long long)g.PX * ((long long)(j - 1) + (long long)g.NY * (q - 1));
        const float cc = p.c[f][q - 1];
        cn[0][h] = cc * lp.psi[f][0][po];
        cn[1][h] = cc * lp.psi[f][1][po];
    }
}

// Within one cell of a CPML slab (either field's layers)?
__device__ __forceinline__ bool near_slab(const Grid& g, const Cpml& p, int i, int j, int k) {
    return i <= p.n[0] + 1 || i >= g.NX - p.n[1] - 1 || j <= p.n[2] + 1 || j >= g.NY - p.n[3] - 1 ||
           k <= p.n[4] + 1 || k >= g.NZ - p.n[5] - 1;
}

// The adjoint state keeps mu_c = gb_c lambda_E_c beside lambda_E (0 on the
// fixed outer tangential components): k_adj_e writes it with the new
// lambda_E, k_adj_source_gather adds gb times its own increments, and k_adj_mu
// rebuilds it after a host write.  The H gather then needs no tuple lookups.

// Transposed E phase, gathered per H cell (one thread per cell of
// [1, N-1]^3, block 64 x 4):
//   lambda_H += S^T mu                                 (bulk E update)
//             + D^T (c nu_E) / d                        (E-phase CPML)
// then nu_H <- b nu_H + DB lambda_H on the H-phase psi cells.
__global__ void __launch_bounds__(256) k_adj_h(Grid g, float* lHx, float* lHy, float* lHz, const float* mEx,
                                               const float* mEy, const float* mEz, const float* dex,
                                               const float* dey, const float* dez, Cpml pe, Cpml lpe, Cpml ph,
                                               Cpml lph, float DB, float dx, float dy, float dz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.NX - 1 || j > g.NY - 1) return;
    const long long o = off(g, i, j, k), X = g.PX, P = g.plane;
    const float mx = mEx[o], mx_j = mEx[o + X], mx_k = mEx[o + P];
    const float my = mEy[o], my_i = mEy[o + 1], my_k = mEy[o + P];
    const float mz = mEz[o], mz_i = mEz[o + 1], mz_j = mEz[o + X];
    const float dx_i = dex[i - 1], dy_j = dey[j - 1], dz_k = dez[k - 1];
    const float dx_i1 = i + 1 <= g.NX - 1 ? dex[i] : 0.0f, dy_j1 = j + 1 <= g.NY - 1 ? dey[j] : 0.0f,
                dz_k1 = k + 1 <= g.NZ - 1 ? dez[k] : 0.0f;
    float hx = lHx[o] + (my * dz_k - my_k * dz_k1 + mz_j * dy_j1 - mz * dy_j);
    float hy = lHy[o] + (mx_k * dz_k1 - mx * dz_k + mz * dx_i - mz_i * dx_i1);
    float hz = lHz[o] + (mx * dy_j - mx_j * dy_j1 + my_i * dx_i1 - my * dx_i);
    if (near_slab(g, pe, i, j, k)) {
        float cx[2][2], cy[2][2], cz[2][2];   // [slot][cell, +1]: x (Eyx, Ezx), y (Exy, Ezy), z (Exz, Eyz)
        cnu_e_pair(g, pe, lpe, 0, i, j, k, cx);
        cnu_e_pair(g, pe, lpe, 1, i, j, k, cy);
        cnu_e_pair(g, pe, lpe, 2, i, j, k, cz);
        hx += (cy[1][1] - cy[1][0]) / dy + (cz[1][0] - cz[1][1]) / dz;
        hy += (cx[1][0] - cx[1][1]) / dx + (cz[0][1] - cz[0][0]) / dz;
        hz += (cx[0][1] - cx[0][0]) / dx + (cy[0][0] - cy[0][1]) / dy;
#pragma unroll
        for (int A = 0; A < 3; ++A) {
            int f, q;
            long long po;
            if (!h_psi(g, ph, A, i, j, k, f, q, po)) continue;
            const float b = ph.b[f][q - 1];
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int c = h_slot_comp(A, sl);
                float* nu = lph.psi[f][sl];
                nu[po] = b * nu[po] + DB * (c == 0 ? hx : (c == 1 ? hy : hz));
            }
        }
    }
    lHx[o] = hx;
    lHy[o] = hy;
    lHz[o] = hz;
}

// Transposed H phase, gathered per E cell (one thread per cell of
// [1, NX] x [1, NY] x [1, NZ], block 64 x 4) on the updated components:
//   lambda_E <- ga lambda_E + DB C^T lambda_H + D^T (c nu_H) / d,
// 0 on the fixed outer tangential components, and mu = gb lambda_E; then
// nu_E <- b nu_E + mu on the E-phase psi cells.  Every lambda_H an updated
// component reads lies in [1, N-1]^3; the i-1 / j-1 / k-1 offsets are
// clamped so the other cells' loads stay in bounds.
template <typename IdxT>
__global__ void __launch_bounds__(256) k_adj_e(Grid g, float* lEx, float* lEy, float* lEz, float* mEx, float* mEy,
                                               float* mEz, const float* lHx, const float* lHy, const float* lHz,
                                               const IdxT* idx, const float* tup, const float* dhx, const float* dhy,
                                               const float* dhz, Cpml pe, Cpml lpe, Cpml ph, Cpml lph, float DB,
                                               float dx, float dy, float dz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.NX || j > g.NY) return;
    const long long o = off(g, i, j, k);
    const bool in = i <= g.NX - 1 && j <= g.NY - 1 && k <= g.NZ - 1;
    const bool ux = in && j >= 2 && k >= 2, uy = in && i >= 2 && k >= 2, uz = in && i >= 2 && j >= 2;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    int t = 0;
    if (ux || uy || uz) {
        const long long xm = i >= 2 ? 1 : 0, ym = j >= 2 ? g.PX : 0, zm = k >= 2 ? g.plane : 0;
        const float hx = lHx[o], hy = lHy[o], hz = lHz[o];
        const float hx_ym = lHx[o - ym], hx_km = lHx[o - zm], hy_xm = lHy[o - xm], hy_km = lHy[o - zm];
        const float hz_xm = lHz[o - xm], hz_ym = lHz[o - ym];
        const float dh_x = in ? dhx[i - 1] : 0.0f, dh_y = in ? dhy[j - 1] : 0.0f, dh_z = in ? dhz[k - 1] : 0.0f;
        const float dh_xm = i >= 2 ? dhx[i - 2] : 0.0f, dh_ym = j >= 2 ? dhy[j - 2] : 0.0f,
                    dh_zm = k >= 2 ? dhz[k - 2] : 0.0f;
        t = (int)idx[o] * 6;
        if (ux) ex = tup[t] * lEx[o] + DB * (hy * dh_z - hy_km * dh_zm + hz_ym * dh_ym - hz * dh_y);
        if (uy) ey = tup[t + 2] * lEy[o] + DB * (hx_km * dh_zm - hx * dh_z + hz * dh_x - hz_xm * dh_xm);
        if (uz) ez = tup[t + 4] * lEz[o] + DB * (hx * dh_y - hx_ym * dh_ym + hy_xm * dh_xm - hy * dh_x);
    }
    const bool ns = near_slab(g, pe, i, j, k);
    if (ns && (ux || uy || uz)) {
        float cx[2][2], cy[2][2], cz[2][2];   // [slot][cell, -1]: x (Hyx, Hzx), y (Hxy, Hzy), z (Hxz, Hyz)
        cnu_h_pair(g, ph, lph, 0, i, j, k, cx);
        cnu_h_pair(g, ph, lph, 1, i, j, k, cy);
        cnu_h_pair(g, ph, lph, 2, i, j, k, cz);
        if (ux) ex += (cz[1][0] - cz[1][1]) / dz + (cy[1][1] - cy[1][0]) / dy;
        if (uy) ey += (cz[0][1] - cz[0][0]) / dz + (cx[1][0] - cx[1][1]) / dx;
        if (uz) ez += (cx[0][1] - cx[0][0]) / dx + (cy[0][0] - cy[0][1]) / dy;
    }
    if (ux) mx = tup[t + 1] * ex;
    if (uy) my = tup[t + 3] * ey;
    if (uz) mz = tup[t + 5] * ez;
    lEx[o] = ex;
    lEy[o] = ey;
    lEz[o] = ez;
    mEx[o] = mx;
    mEy[o] = my;
    mEz[o] = mz;
    if (ns) {
#pragma unroll
        for (int A = 0; A < 3; ++A) {
            int f, q;
            long long po;
            if (!e_psi(g, pe, A, i, j, k, f, q, po)) continue;
            const float b = pe.b[f][q - 1];
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                if (!e_slot_ok(A, sl, i, j, k)) continue;
                const int c = e_slot_comp(A, sl);
                float* nu = lpe.psi[f][sl];
                nu[po] = b * nu[po] + (c == 0 ? mx : (c == 1 ? my : mz));
            }
        }
    }
}

// mu_c = gb_c lambda_E_c of one component over the grid (0 where the
// component is fixed); after a host write of lambda_E.
template <typename IdxT>
__global__ void k_adj_mu(Grid g, const float* lE, float* m, int c, const IdxT* idx, const float* tup) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.NX || j > g.NY) return;
    const long long o = off(g, i, j, k);
    m[o] = e_upd(g, c, i, j, k) ? tup[(int)idx[o] * 6 + 2 * c + 1] * lE[o] : 0.0f;
}

// nu_E += sign * gb lambda_E of component `comp` on the E psi cells: keeps
// nu consistent when the adjoint field is overwritten from the host.
template <typename IdxT>
__global__ void k_adj_nu_adjust(Grid g, const float* lE, int comp, float sign, const IdxT* idx, const float* tup,
                                Cpml pe, Cpml lpe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.NX || j > g.NY) return;
    if (!near_slab(g, pe, i, j, k)) return;
    const long long o = off(g, i, j, k);
    for (int A = 0; A < 3; ++A) {
        int f, q;
        long long po;
        if (!e_psi(g, pe, A, i, j, k, f, q, po)) continue;
        for (int sl = 0; sl < 2; ++sl) {
            if (e_slot_comp(A, sl) != comp || !e_slot_ok(A, sl, i, j, k)) continue;
            float* nu = lpe.psi[f][sl];
            nu[po] = nu[po] + sign * (tup[(int)idx[o] * 6 + 2 * comp + 1] * lE[o]);
        }
    }
}

// Adjoint of the lumped ports (k_ports_h, k_ports_e).  Forward, per port and
// step, with the sources left out (they are affine):
//   H phase: I_0 = L(H) num/den, L the gap's loop integral;
//            I_i -= kI (V_{i+1} - V_i), i = 1..M-2
//   E phase: V_i -= kV (I_i - I_{i-1}), i = 1..M-2;  E = V_1 / gap_div on the gap
// and the series record V_1 after the E phase and I_0 after the H phase.
// lV, lI ([np][M]) are the adjoints of the line; wV, wI ([np][T], or null)
// are dJ/dV_1 and dJ/dI_0 of every step.  One thread runs the ports one
// after another, in a fixed order (no two ports race on a shared H cell);
// the lines are a few hundred nodes.
struct AdjPorts {
    const PortDev* ports;
    int np, M, T;
    float *lV, *lI;
    const double *wV, *wI;
};

// Transposed E phase (before the transposed bulk update): dJ/dV_1 of step n
// and the gap cells' lambda_E / gap_div into lambda_V_1; the gap cells' own
// update is overwritten, so their lambda_E and mu become 0; then the V
// recursion transposed into lambda_I.
__global__ void k_adj_ports_e(Grid g, AdjPorts a, float* lEx, float* lEy, float* lEz, float* mEx, float* mEy,
                              float* mEz, const int* step) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int n = *step, M = a.M;
    for (int q = 0; q < a.np; ++q) {
        const PortDev pt = a.ports[q];
        float* L = pt.dir == 0 ? lEx : (pt.dir == 1 ? lEy : lEz);
        float* Mu = pt.dir == 0 ? mEx : (pt.dir == 1 ? mEy : mEz);
        float acc = 0.0f;
        for (int ii = pt.loc[0]; ii <= pt.loc[0] + pt.size[0] - 1; ++ii)
            for (int jj = pt.loc[1]; jj <= pt.loc[1] + pt.size[1] - 1; ++jj)
                for (int kk = pt.loc[2]; kk <= pt.loc[2] + pt.size[2] - 1; ++kk) {
                    const long long o = off(g, ii, jj, kk);
                    acc = acc + L[o];
                    L[o] = 0.0f;
                    Mu[o] = 0.0f;
                }
        float* lV = a.lV + (long long)q * M;
        float* lI = a.lI + (long long)q * M;
        float v1 = lV[1] + acc / pt.gap_div;
        if (a.wV) v1 = v1 + (float)a.wV[(long long)q * a.T + (n - 1)];
        lV[1] = v1;
        const float kV = pt.kV;
        for (int i = 1; i <= M - 2; ++i) {
            lI[i] = lI[i] - kV * lV[i];
            lI[i - 1] = lI[i - 1] + kV * lV[i];
        }
    }
}

// Transposed H phase (between the transposed E and H bulk updates): the I
// recursion transposed into lambda_V, dJ/dI_0 of step n, and lambda_I_0
// num/den spread over the H of the gap's loop in gap_loop_current's terms;
// I_0 is overwritten every step, so lambda_I_0 becomes 0.
__global__ void k_adj_ports_h(Grid g, AdjPorts a, float* lHx, float* lHy, float* lHz, float dx, float dy, float dz,
                              const int* step) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int n = *step, M = a.M;
    for (int q = 0; q < a.np; ++q) {
        const PortDev pt = a.ports[q];
        float* lV = a.lV + (long long)q * M;
        float* lI = a.lI + (long long)q * M;
        const float kI = pt.kI;
        for (int i = 1; i <= M - 2; ++i) {
            lV[i + 1] = lV[i + 1] - kI * lI[i];
            lV[i] = lV[i] + kI * lI[i];
        }
        float i0 = lI[0];
        if (a.wI) i0 = i0 + (float)a.wI[(long long)q * a.T + (n - 1)];
        lI[0] = 0.0f;
        const float w = i0 * pt.cur_num / pt.cur_den;
        const int l1 = pt.loc[0], l2 = pt.loc[1], l3 = pt.loc[2];
        const int d1 = pt.size[0] + l1 - 1, d2 = pt.size[1] + l2 - 1, d3 = pt.size[2] + l3 - 1;
        if (pt.dir == 0) {
            for (int ii = l1; ii <= d1; ++ii) {
                for (int jj = l2; jj <= d2; ++jj) {
                    lHy[off(g, ii, jj, l3 - 1)] += w * dy;
                    lHy[off(g, ii, jj, d3)] -= w * dy;
                }
                for (int kk = l3; kk <= d3; ++kk) {
                    lHz[off(g, ii, l2, kk)] += w * dz;
                    lHz[off(g, ii, d2 - 1, kk)] -= w * dz;
                }
            }
        } else if (pt.dir == 1) {
            for (int jj = l2; jj <= d2; ++jj) {
                for (int ii = l1; ii <= d1; ++ii) {
                    lHx[off(g, ii, jj, l3)] += w * dx;
                    lHx[off(g, ii, jj, d3 - 1)] -= w * dx;
                }
                for (int kk = l3; kk <= d3; ++kk) {
                    lHz[off(g, l1 - 1, jj, kk)] += w * dz;
                    lHz[off(g, d1, jj, kk)] -= w * dz;
                }
            }
        } else {
            for (int kk = l3; kk <= d3; ++kk) {
                for (int ii = l1; ii <= d1; ++ii) {
                    lHx[off(g, ii, l2 - 1, kk)] += w * dx;
                    lHx[off(g, ii, d2, kk)] -= w * dx;
                }
                for (int jj = l2; jj <= d2; ++jj) {
                    lHy[off(g, l1, jj, kk)] += w * dy;
                    lHy[off(g, d1 - 1, jj, kk)] -= w * dy;
                }
            }
        }
    }
}

// Design region: the edges of [lo, lo+n-1] (cells [lo, lo+n-2] and their
// far edges) carry per-edge Ga/Gb instead of a tuple.  The bulk kernel sees
// the neutral tuple (Ga=1, Gb=0) there, which leaves E unchanged bit for
// bit, and k_e_design applies the update in the bulk kernel's arithmetic.
struct DesignDev {
    int lo[3], n[3];
    long long nedge;
    const float* coef;   // [6][nedge]: ga_x, gb_x, ga_y, gb_y, ga_z, gb_z
};

template <typename IdxT>
__global__ void k_set_idx(Grid g, DesignDev d, IdxT* idx, IdxT value) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.nedge) return;
    const int i = d.lo[0] + (int)(e % d.n[0]);
    const long long rr = e / d.n[0];
    const int j = d.lo[1] + (int)(rr % d.n[1]), k = d.lo[2] + (int)(rr / d.n[1]);
    idx[off(g, i, j, k)] = value;
}

// With `record`, also stores the adjoint record of the step (as k_adj_record:
// E before the update and the curl terms), the box being the record box.
__global__ void k_e_design(Grid g, DesignDev d, float* Ex, float* Ey, float* Ez, const float* Hx, const float* Hy,
                           const float* Hz, const float* dex, const float* dey, const float* dez, AdjRec r,
                           int record, const int* step) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.nedge) return;
    const int i = d.lo[0] + (int)(e % d.n[0]);
    const long long rr = e / d.n[0];
    const int j = d.lo[1] + (int)(rr % d.n[1]), k = d.lo[2] + (int)(rr / d.n[1]);
    const long long o = off(g, i, j, k), n = d.nedge;
    const float dx_i = dex[i - 1], dy_j = dey[j - 1], dz_k = dez[k - 1];
    const float hx = Hx[o], hy = Hy[o], hz = Hz[o];
    const float hz_ym = Hz[o - g.PX], hz_xm = Hz[o - 1], hy_xm = Hy[o - 1], hx_ym = Hx[o - g.PX];
    const float hx_km = Hx[o - g.plane], hy_km = Hy[o - g.plane];
    const float* c = d.coef;
    const float ex0 = Ex[o], ey0 = Ey[o], ez0 = Ez[o];
    const float s0 = (hz - hz_ym) * dy_j + (hy_km - hy) * dz_k;
    const float s1 = (hz_xm - hz) * dx_i + (hx - hx_km) * dz_k;
    const float s2 = (hy - hy_xm) * dx_i + (hx_ym - hx) * dy_j;
    Ex[o] = c[e] * ex0 + c[n + e] * s0;
    Ey[o] = c[2 * n + e] * ey0 + c[3 * n + e] * s1;
    Ez[o] = c[4 * n + e] * ez0 + c[5 * n + e] * s2;
    if (record) {
        float* h = r.hist + (long long)((*step - 1) % r.slots) * 6 * n;
        h[e] = ex0;
        h[n + e] = ey0;
        h[2 * n + e] = ez0;
        h[3 * n + e] = s0;
        h[4 * n + e] = s1;
        h[5 * n + e] = s2;
    }
}

// Transposed k_e_design, lambda_H += S^T (gb lambda_E'), lambda_E <- ga lambda_E',
// in a fixed order: k_adj_design_m stores m = gb lambda_E
// of every design edge ([3][nedge]) and scales lambda_E by ga;
// k_adj_design_gather gives every H cell of the box extended by one below
// to one thread, which adds the terms of the edges that reach it.
__global__ void k_adj_design_m(Grid g, DesignDev d, float* lEx, float* lEy, float* lEz, float* m) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= d.nedge) return;
    const int i = d.lo[0] + (int)(e % d.n[0]);
    const long long rr = e / d.n[0];
    const int j = d.lo[1] + (int)(rr % d.n[1]), k = d.lo[2] + (int)(rr / d.n[1]);
    const long long o = off(g, i, j, k), n = d.nedge;
    const float* c = d.coef;
    const float lx = lEx[o], ly = lEy[o], lz = lEz[o];
    m[e] = c[n + e] * lx;
    m[n + e] = c[3 * n + e] * ly;
    m[2 * n + e] = c[5 * n + e] * lz;
    lEx[o] = c[e] * lx;
    lEy[o] = c[2 * n + e] * ly;
    lEz[o] = c[4 * n + e] * lz;
}

__global__ void k_adj_design_gather(Grid g, DesignDev d, const float* __restrict__ m, float* lHx, float* lHy,
                                    float* lHz, const float* dex, const float* dey, const float* dez) {
    const int ex = d.n[0] + 1, ey = d.n[1] + 1, ez = d.n[2] + 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ex * ey * ez) return;
    const int i = d.lo[0] - 1 + (int)(t % ex);
    const long long rr = t / ex;
    const int j = d.lo[1] - 1 + (int)(rr % ey), k = d.lo[2] - 1 + (int)(rr / ey);
    const long long n = d.nedge;
    // edge index of cell (a, b, c), -1 outside the design box
    auto edge = [&](int a, int b2, int c2) -> long long {
        const int u = a - d.lo[0], v = b2 - d.lo[1], w = c2 - d.lo[2];
        if (u < 0 || v < 0 || w < 0 || u >= d.n[0] || v >= d.n[1] || w >= d.n[2]) return -1;
        return (long long)u + (long long)d.n[0] * ((long long)v + (long long)d.n[1] * w);
    };
    const long long e0 = edge(i, j, k), ei = edge(i + 1, j, k), ej = edge(i, j + 1, k), ek = edge(i, j, k + 1);
    float sx = 0.f, sy = 0.f, sz = 0.f;
    bool hx = false, hy = false, hz = false;
    // Hx: + my dz (edge t), - my dz (edge k+1), + mz dy (edge j+1), - mz dy (edge t)
    if (e0 >= 0) { sx = sx + m[n + e0] * dez[k - 1]; hx = true; }
    if (ek >= 0) { sx = sx + -(m[n + ek] * dez[k]); hx = true; }
    if (ej >= 0) { sx = sx + m[2 * n + ej] * dey[j]; hx = true; }
    if (e0 >= 0) { sx = sx + -(m[2 * n + e0] * dey[j - 1]); }
    // Hy: + mx dz (edge k+1), - mx dz (edge t), + mz dx (edge t), - mz dx (edge i+1)
    if (ek >= 0) { sy = sy + m[ek] * dez[k]; hy = true; }
    if (e0 >= 0) { sy = sy + -(m[e0] * dez[k - 1]); hy = true; }
    if (e0 >= 0) { sy = sy + m[2 * n + e0] * dex[i - 1]; }
    if (ei >= 0) { sy = sy + -(m[2 * n + ei] * dex[i]); hy = true; }
    // Hz: + mx dy (edge t), - mx dy (edge j+1), + my dx (edge i+1), - my dx (edge t)
    if (e0 >= 0) { sz = sz + m[e0] * dey[j - 1]; hz = true; }
    if (ej >= 0) { sz = sz + -(m[ej] * dey[j]); hz = true; }
    if (ei >= 0) { sz = sz + m[n + ei] * dex[i]; hz = true; }
    if (e0 >= 0) { sz = sz + -(m[n + e0] * dex[i - 1]); }
    const long long o = off(g, i, j, k);
    if (hx) lHx[o] = lHx[o] + sx;
    if (hy) lHy[o] = lHy[o] + sy;
    if (hz) lHz[o] = lHz[o] + sz;
}

__global__ void k_step_dec(int* step) { *step = *step - 1; }

}  // namespace aemk
