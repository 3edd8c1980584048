/*
 * oracle.c -- CPU restatement of the reference FDTD leapfrog loop.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for libaem, never part
 * of the product.
 *
 * Every region below restates one loop region of
 * main_fdtd/fdtd_solver.f90 (cited per function) for the non-kmax build,
 * with the SPICE build's gap exchange (the circuit itself is the caller's)
 * and the kmax build (complex fields, Bloch-periodic walls, modal plane
 * source: a kmax plan makes two states, the real and the imaginary part,
 * stepped together by oracle_step; its sources are evaluated here from the
 * plan's scalars, not from the setup's tables),
 * in the reference's sequential loop order and FP32 left-to-right operation
 * order.  Build with -ffp-contract=off.
 *
 * Deliberate restatement choices (documented, value-preserving):
 *  - Fields use the common layout of include/aem_plan.h instead of six
 *    differently-shaped arrays; every index the reference touches maps 1:1.
 *  - E_oldt is copied on the plasma bounding boxes only: the Drude update is
 *    the only reader and it reads inside the boxes (f90:5791-5847).
 *  - The PBC copies of J_source (f90:6742-6806) write entries outside the
 *    bounding boxes that are never read, so they are omitted.
 *  - The clear 1-D line of the antenna incident wave reads C_inc(0), one
 *    element before the array (f90:8397-8400); it is taken as 0.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_REAL
#define ORACLE_REAL float
#endif
typedef ORACLE_REAL real;
#define R(x) ((real)(x))

#if defined(ORACLE_FP64)
#define EXP(x) exp(x)
#define SQRT(x) sqrt(x)
#define COS(x) cos(x)
#define SIN(x) sin(x)
#define ILOGB(x) ilogb(x)
#define SCALBN(x, e) scalbn(x, e)
#else
#define EXP(x) expf(x)
#define SQRT(x) sqrtf(x)
#define COS(x) cosf(x)
#define SIN(x) sinf(x)
#define ILOGB(x) ilogbf(x)
#define SCALBN(x, e) scalbnf(x, e)
#endif
/* pi, mu_0, eps_0 as the reference declares them (FP32, f90:24-26, 44) */
#define PI_F 3.14159265358979323846f
#define MU0_F 1.25663706E-6f
#define EP0_F 8.85418782E-12f

#define NFACE 6

struct oracle_state {
    aem_plan_view p;
    int n;                     /* steps done */
    int64_t nc;
    real *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
    real *F[6];
    /* CPML psi, per face and per field: [face][0..1] for the two components
     * the face affects (H and E separately). */
    real *psiH[NFACE][2], *psiE[NFACE][2];
    int64_t psi_dim[NFACE][3];
    /* sheets: special E per plane */
    real *Es[3];
    /* drude */
    real *J[3], *Eold[3];
    int64_t nbox[3];
    /* ports */
    real *V, *I;               /* [port][M] and [port][M-1] (0-based j-1) */
    real *Vout, *Iout;         /* [port][T] */
    /* SPICE ports: this step's gap currents, the node voltages less bias
     * (ports, then extra nodes as given), and their series */
    real *spI, *spV, *spIout, *spVout;
    int x_pending;             /* a step waits at the exchange */
    /* NTFF */
    real *Mcur[NFACE][2], *Mold[NFACE][2], *Jcur[NFACE][2], *Jold[NFACE][2];
    int64_t fdim[NFACE][2];
    real *U[3], *W[3];         /* [angle][len] */
    /* plane sums and phase centre */
    real *refl, *trans;
    real *pc[6];
    /* video */
    real *vid[6];
    /* post products */
    real *incident, *Eth, *Eph;
    real *fftime;
    int finalized;
    /* kmax: the TM plane sums (refl/trans hold TE), which part this state
     * is (0 real, 1 imaginary) and, on the real part, the imaginary part */
    real *refl2, *trans2;
    int part;
    struct oracle_state* kim;
};

/* ---- helpers ----------------------------------------------------------- */

#define IDX(s, i, j, k) ((int64_t)((i)-1) + (int64_t)(s)->p.PX * ((int64_t)((j)-1) + (int64_t)(s)->p.NY * (int64_t)((k)-1)))
#define EX(i, j, k) s->Ex[IDX(s, i, j, k)]
#define EY(i, j, k) s->Ey[IDX(s, i, j, k)]
#define EZ(i, j, k) s->Ez[IDX(s, i, j, k)]
#define HX(i, j, k) s->Hx[IDX(s, i, j, k)]
#define HY(i, j, k) s->Hy[IDX(s, i, j, k)]
#define HZ(i, j, k) s->Hz[IDX(s, i, j, k)]

static real* zalloc(int64_t n) {
    if (n <= 0) n = 1;
    return (real*)calloc((size_t)n, sizeof(real));
}

/* A complex value over a real one in the kmax build.  Fortran divides a
 * complex by a real as complex by complex (the real promoted, imaginary part
 * 0), and the flang build does that with the C99 complex division routine
 * __divsc3: with c' = scalbn(c, -logb c) it forms (x c') / (c' c') and scales
 * back, which is not x / c to the last bit.  Each part of the complex value
 * goes through this formula on its own (the imaginary divisor part is 0). */
static inline real cdivr(real x, real c) {
    const int e = ILOGB(c);
    const real cs = SCALBN(c, -e);
    return SCALBN((x * cs) / (cs * cs), -e);
}
#define DIVR(st, x, d) ((st)->p.kmax ? cdivr((x), (d)) : (x) / (d))

static inline int tidx(const oracle_state* s, int64_t o) {
    return s->p.idx_bytes == 1 ? (int)s->p.tuple_idx8[o] : (int)s->p.tuple_idx16[o];
}
/* gax..gbz of the cell at offset o */
static inline real GA(const oracle_state* s, int c, int64_t o) { return R(s->p.tuples[tidx(s, o) * 6 + 2 * c]); }
static inline real GB(const oracle_state* s, int c, int64_t o) { return R(s->p.tuples[tidx(s, o) * 6 + 2 * c + 1]); }

/* Inc() of f90:8825-8842: analytic incident plane wave at a retarded point */
static real inc_fn(const oracle_state* s, real w, real xf, real yf, real zf, real cf) {
    const aem_plan_view* p = &s->p;
    real tx = R(p->s_th) * R(p->c_ph) * (xf - R(p->delay[0])) * R(p->dx);
    real ty = R(p->s_th) * R(p->s_ph) * (yf - R(p->delay[1])) * R(p->dy);
    real tz = R(p->c_th) * (zf - R(p->delay[2])) * R(p->dz);
    real arg = (R(p->t_spread) - (cf * R(p->dt)) + R(p->inv_c) * (tx + ty + tz)) / R(p->spread);
    if (p->pulse_type == 1) return w * EXP(R(-0.5) * (arg * arg));
    return R(-1.0) * w * arg * EXP(R(0.5)) * EXP(R(-0.5) * (arg * arg));
}

/* Inc() of the incident (m = 0) or the ground plane's mirrored (m = 1)
 * plane wave: the same function with theta_mirror, phi_mirror and the
 * mirror delays (f90:4712-4790) */
static real inc_w(const oracle_state* s, int m, real w, real xf, real yf, real zf, real cf) {
    const aem_plan_view* p = &s->p;
    if (!m) return inc_fn(s, w, xf, yf, zf, cf);
    real tx = R(p->s_thm) * R(p->c_phm) * (xf - R(p->delay_m[0])) * R(p->dx);
    real ty = R(p->s_thm) * R(p->s_phm) * (yf - R(p->delay_m[1])) * R(p->dy);
    real tz = R(p->c_thm) * (zf - R(p->delay_m[2])) * R(p->dz);
    real arg = (R(p->t_spread) - (cf * R(p->dt)) + R(p->inv_c) * (tx + ty + tz)) / R(p->spread);
    if (p->pulse_type == 1) return w * EXP(R(-0.5) * (arg * arg));
    return R(-1.0) * w * arg * EXP(R(0.5)) * EXP(R(-0.5) * (arg * arg));
}

/* E_or_V_1D() of f90:8844-8857; the kmax build's complex one (f90:8888-8900),
 * amp exp(-x^2/2) exp(i 2 pi f_adj ((tc-1) dt - pos/c)), gives this state's part */
static real e_or_v(const oracle_state* s, real amp, real tc, real pos) {
    const aem_plan_view* p = &s->p;
    real x = ((R(p->t_spread) - ((tc - R(1.0)) * R(p->dt))) - R(p->inv_c) * pos) / R(p->spread);
    if (p->kmax) {
        const real m = amp * EXP(R(-0.5) * (x * x));
        const real th = R(2.0) * R(PI_F) * R(p->f_adj) * (((tc - R(1.0)) * R(p->dt)) - R(p->inv_c) * pos);
        return s->part ? m * SIN(th) : m * COS(th);
    }
    if (p->pulse_type == 1) return amp * EXP(R(-0.5) * (x * x));
    return amp * R(-1.0) * x * EXP(R(0.5)) * EXP(R(-0.5) * (x * x));
}

/* ---- lifetime ------------------------------------------------------------ */

static oracle_state* create_state(const aem_plan_view* plan) {
    oracle_state* s = (oracle_state*)calloc(1, sizeof(oracle_state));
    if (!s) return NULL;
    s->p = *plan;
    const aem_plan_view* p = &s->p;
    s->nc = p->ncell;
    s->Ex = zalloc(s->nc); s->Ey = zalloc(s->nc); s->Ez = zalloc(s->nc);
    s->Hx = zalloc(s->nc); s->Hy = zalloc(s->nc); s->Hz = zalloc(s->nc);
    s->F[0] = s->Ex; s->F[1] = s->Ey; s->F[2] = s->Ez; s->F[3] = s->Hx; s->F[4] = s->Hy; s->F[5] = s->Hz;
    /* psi arrays: face f normal to axis a=f/2; stored (thickness, full extents)
     * as [t][dim1][dim2] flattened with the two transverse extents PX/NY/NZ */
    for (int f = 0; f < NFACE; ++f) {
        int a = f / 2;
        int64_t n = p->npml[f] + 1;
        int64_t d1 = (a == 0) ? p->NY : p->PX;
        int64_t d2 = (a == 2) ? p->NY : p->NZ;
        s->psi_dim[f][0] = n; s->psi_dim[f][1] = d1; s->psi_dim[f][2] = d2;
        for (int c = 0; c < 2; ++c) {
            s->psiH[f][c] = zalloc(n * d1 * d2);
            s->psiE[f][c] = zalloc(n * d1 * d2);
        }
    }
    for (int a = 0; a < 3; ++a)
        s->Es[a] = zalloc((int64_t)p->n_sheet_planes[a] * p->sheet_dim[a][0] * p->sheet_dim[a][1]);
    if (p->has_drude) {
        for (int c = 0; c < 3; ++c) {
            const int* b = p->dbox[c];
            int64_t bx = b[1] - b[0] + 1, by = b[3] - b[2] + 1, bz = b[5] - b[4] + 1;
            if (bx < 0) bx = 0;
            if (by < 0) by = 0;
            if (bz < 0) bz = 0;
            s->nbox[c] = bx * by * bz;
            s->J[c] = zalloc(s->nbox[c] * p->n_poles);
            s->Eold[c] = zalloc(s->nbox[c]);
        }
    }
    int np = p->n_ports > 0 ? p->n_ports : 1;
    s->V = zalloc((int64_t)np * p->M);
    s->I = zalloc((int64_t)np * p->M);
    s->Vout = zalloc((int64_t)np * p->time_steps);
    s->Iout = zalloc((int64_t)np * p->time_steps);
    {
        int64_t ns = p->n_spice > 0 ? p->n_spice : 1, nv = (int64_t)p->n_spice + p->n_spice_extra;
        if (nv < 1) nv = 1;
        s->spI = zalloc(ns);
        s->spV = zalloc(nv);
        s->spIout = zalloc(ns * p->time_steps);
        s->spVout = zalloc(nv * p->time_steps);
    }
    if (p->n_angles > 0) {
        for (int f = 0; f < NFACE; ++f) {
            int a = f / 2;
            int b = (a == 0) ? 1 : 0, c = (a == 2) ? 1 : 2;
            s->fdim[f][0] = p->ff_hi[b] + 2;
            s->fdim[f][1] = p->ff_hi[c] + 2;
            int64_t n = s->fdim[f][0] * s->fdim[f][1];
            for (int q = 0; q < 2; ++q) {
                s->Mcur[f][q] = zalloc(n); s->Mold[f][q] = zalloc(n);
                s->Jcur[f][q] = zalloc(n); s->Jold[f][q] = zalloc(n);
            }
        }
        for (int q = 0; q < 3; ++q) {
            s->U[q] = zalloc((int64_t)p->n_angles * (p->len_ff + 2));
            s->W[q] = zalloc((int64_t)p->n_angles * (p->len_ff + 2));
        }
    }
    s->refl = zalloc(p->time_steps);
    s->trans = zalloc(p->time_steps);
    for (int q = 0; q < 6; ++q) s->pc[q] = zalloc(p->time_steps + 1);
    if (p->video_on)
        for (int q = 0; q < 6; ++q) s->vid[q] = zalloc((int64_t)p->vid_size1 * p->vid_size2 * (p->time_steps + 1));
    s->incident = zalloc(p->time_steps);
    s->Eth = zalloc((int64_t)(p->n_angles > 0 ? p->n_angles : 1) * p->time_steps);
    s->Eph = zalloc((int64_t)(p->n_angles > 0 ? p->n_angles : 1) * p->time_steps);
    s->fftime = zalloc(p->n_angles > 0 ? p->n_angles : 1);
    if (p->kmax) {
        s->refl2 = zalloc(p->time_steps);
        s->trans2 = zalloc(p->time_steps);
    }
    return s;
}

oracle_state* oracle_create(const aem_plan_view* plan) {
    oracle_state* s = create_state(plan);
    if (s && plan->kmax) {
        s->kim = create_state(plan);
        if (!s->kim) {
            oracle_destroy(s);
            return NULL;
        }
        s->kim->part = 1;
    }
    return s;
}

void oracle_destroy(oracle_state* s) {
    if (!s) return;
    oracle_destroy(s->kim);
    free(s->refl2); free(s->trans2);
    for (int q = 0; q < 6; ++q) free(s->F[q]);
    for (int f = 0; f < NFACE; ++f)
        for (int c = 0; c < 2; ++c) {
            free(s->psiH[f][c]); free(s->psiE[f][c]);
            free(s->Mcur[f][c]); free(s->Mold[f][c]); free(s->Jcur[f][c]); free(s->Jold[f][c]);
        }
    for (int a = 0; a < 3; ++a) { free(s->Es[a]); free(s->J[a]); free(s->Eold[a]); free(s->U[a]); free(s->W[a]); }
    free(s->V); free(s->I); free(s->Vout); free(s->Iout);
    free(s->spI); free(s->spV); free(s->spIout); free(s->spVout);
    free(s->refl); free(s->trans);
    for (int q = 0; q < 6; ++q) { free(s->pc[q]); free(s->vid[q]); }
    free(s->incident); free(s->Eth); free(s->Eph); free(s->fftime);
    free(s);
}

/* psi accessors: face f, local index t (1-based along the normal), transverse (u,v) */
static inline real* psiH_at(oracle_state* s, int f, int c, int t, int u, int v) {
    int64_t d1 = s->psi_dim[f][1], d2 = s->psi_dim[f][2];
    (void)d2;
    return &s->psiH[f][c][(int64_t)t + s->psi_dim[f][0] * ((int64_t)(u - 1) + d1 * (int64_t)(v - 1))];
}
static inline real* psiE_at(oracle_state* s, int f, int c, int t, int u, int v) {
    int64_t d1 = s->psi_dim[f][1];
    return &s->psiE[f][c][(int64_t)t + s->psi_dim[f][0] * ((int64_t)(u - 1) + d1 * (int64_t)(v - 1))];
}

/* special E on sheet plane q of axis a at in-plane (pp,qq) */
static inline real* ES(oracle_state* s, int a, int q, int pp, int qq) {
    int P = s->p.sheet_dim[a][0], Q = s->p.sheet_dim[a][1];
    return &s->Es[a][(int64_t)q * P * Q + (pp - 1) + (int64_t)P * (qq - 1)];
}
static inline int sheet_on(const oracle_state* s, int a, int q, int pp, int qq) {
    int P = s->p.sheet_dim[a][0], Q = s->p.sheet_dim[a][1];
    int64_t o = (int64_t)q * P * Q + (pp - 1) + (int64_t)P * (qq - 1);
    return (s->p.sheet_sig[a][o] > 0) || (s->p.sheet_ep[a][o] > 1);
}
static inline real sheet_sig(const oracle_state* s, int a, int q, int pp, int qq) {
    int P = s->p.sheet_dim[a][0], Q = s->p.sheet_dim[a][1];
    return R(s->p.sheet_sig[a][(int64_t)q * P * Q + (pp - 1) + (int64_t)P * (qq - 1)]);
}
static inline real sheet_ep(const oracle_state* s, int a, int q, int pp, int qq) {
    int P = s->p.sheet_dim[a][0], Q = s->p.sheet_dim[a][1];
    return R(s->p.sheet_ep[a][(int64_t)q * P * Q + (pp - 1) + (int64_t)P * (qq - 1)]);
}

/* ---- H phase ------------------------------------------------------------ */

/* bulk H, f90:4442-4484 (DA == 1) */
static void h_bulk(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const real DB = R(p->DB);
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    for (int k = 1; k <= NZ - 1; ++k)
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i)
                HX(i, j, k) = R(1.0) * HX(i, j, k) + DB * ((EZ(i, j, k) - EZ(i, j + 1, k)) * R(p->den_hy[j - 1]) +
                                                         (EY(i, j, k + 1) - EY(i, j, k)) * R(p->den_hz[k - 1]));
    for (int k = 1; k <= NZ - 1; ++k)
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i)
                HY(i, j, k) = R(1.0) * HY(i, j, k) + DB * ((EZ(i + 1, j, k) - EZ(i, j, k)) * R(p->den_hx[i - 1]) +
                                                         (EX(i, j, k) - EX(i, j, k + 1)) * R(p->den_hz[k - 1]));
    for (int k = 1; k <= NZ - 1; ++k)
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i)
                HZ(i, j, k) = R(1.0) * HZ(i, j, k) + DB * ((EY(i, j, k) - EY(i + 1, j, k)) * R(p->den_hx[i - 1]) +
                                                         (EX(i, j + 1, k) - EX(i, j, k)) * R(p->den_hy[j - 1]));
}

/* thin-sheet H corrections, f90:4486-4554 */
static void h_sheets(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const real DB = R(p->DB), t = R(p->sheet_thickness);
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    for (int q = 0; q < p->n_sheet_planes[0]; ++q) {
        int i = p->sheet_pos[0][q];
        real f = DB * (t / R(p->dx));
        for (int k = 1; k <= NZ - 1; ++k)
            for (int j = 1; j <= NY - 1; ++j)
                if (sheet_on(s, 0, q, j, k)) {
                    HY(i, j, k) = HY(i, j, k) + f * ((EX(i, j, k + 1) - EX(i, j, k) + *ES(s, 0, q, j, k) - *ES(s, 0, q, j, k + 1)) *
                                                     R(p->den_hz[k - 1]));
                    HZ(i, j, k) = HZ(i, j, k) + f * ((EX(i, j, k) - EX(i, j + 1, k) + *ES(s, 0, q, j + 1, k) - *ES(s, 0, q, j, k)) *
                                                     R(p->den_hy[j - 1]));
                }
    }
    for (int q = 0; q < p->n_sheet_planes[1]; ++q) {
        int j = p->sheet_pos[1][q];
        real f = DB * (t / R(p->dy));
        for (int k = 1; k <= NZ - 1; ++k)
            for (int i = 1; i <= NX - 1; ++i)
                if (sheet_on(s, 1, q, i, k)) {
                    HX(i, j, k) = HX(i, j, k) + f * (EY(i, j, k) - EY(i, j, k + 1) + *ES(s, 1, q, i, k + 1) - *ES(s, 1, q, i, k)) *
                                                    R(p->den_hz[k - 1]);
                    HZ(i, j, k) = HZ(i, j, k) + f * ((EY(i + 1, j, k) - EY(i, j, k) - *ES(s, 1, q, i + 1, k) + *ES(s, 1, q, i, k)) *
                                                     R(p->den_hx[i - 1]));
                }
    }
    for (int q = 0; q < p->n_sheet_planes[2]; ++q) {
        int k = p->sheet_pos[2][q];
        real f = DB * (t / R(p->dz));
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i)
                if (sheet_on(s, 2, q, i, j)) {
                    HX(i, j, k) = HX(i, j, k) + f * (EZ(i, j + 1, k) - EZ(i, j, k) - *ES(s, 2, q, i, j + 1) + *ES(s, 2, q, i, j)) *
                                                    R(p->den_hy[j - 1]);
                    HY(i, j, k) = HY(i, j, k) + f * ((EZ(i, j, k) - EZ(i + 1, j, k) + *ES(s, 2, q, i + 1, j) - *ES(s, 2, q, i, j)) *
                                                     R(p->den_hx[i - 1]));
                }
    }
}

/* TF/SF H faces, f90:4627-4710 (t = n-1) */
static void h_tfsf_wave(oracle_state* s, int n, int m) {
    const aem_plan_view* p = &s->p;
    const int xl = p->tf_lo[0], xh = p->tf_hi[0], yl = p->tf_lo[1], yh = p->tf_hi[1], zl = p->tf_lo[2], zh = p->tf_hi[2];
    const int px = p->pbc[0], py = p->pbc[1], pz = p->pbc[2];
    const real cf = R(n) - R(1.0);
    const real kx = R(p->kHinc[0]), ky = R(p->kHinc[1]), kz = R(p->kHinc[2]);
    const float* W = m ? p->WE_m : p->WE;
    const real WEx = R(W[0]), WEy = R(W[1]), WEz = R(W[2]);
    if (!py) {
        for (int k = zl; k <= zh; ++k)
            for (int i = xl; i <= xh - 1 + px; ++i) {
                HZ(i, yl - 1, k) = HZ(i, yl - 1, k) - ky * inc_w(s, m, WEx, R(i) + R(0.5), R(yl), R(k), cf) * R(p->wall[2]);
                HZ(i, yh, k) = HZ(i, yh, k) + ky * inc_w(s, m, WEx, R(i) + R(0.5), R(yh), R(k), cf) * R(p->wall[3]);
            }
        for (int k = zl; k <= zh - 1 + pz; ++k)
            for (int i = xl; i <= xh; ++i) {
                HX(i, yl - 1, k) = HX(i, yl - 1, k) + ky * inc_w(s, m, WEz, R(i), R(yl), R(k) + R(0.5), cf) * R(p->wall[2]);
                HX(i, yh, k) = HX(i, yh, k) - ky * inc_w(s, m, WEz, R(i), R(yh), R(k) + R(0.5), cf) * R(p->wall[3]);
            }
    }
    if (!pz) {
        for (int j = yl; j <= yh; ++j)
            for (int i = xl; i <= xh - 1 + px; ++i) {
                HY(i, j, zl - 1) = HY(i, j, zl - 1) + kz * inc_w(s, m, WEx, R(i) + R(0.5), R(j), R(zl), cf) * R(p->wall[4]);
                HY(i, j, zh) = HY(i, j, zh) - kz * inc_w(s, m, WEx, R(i) + R(0.5), R(j), R(zh), cf) * R(p->wall[5]);
            }
        for (int j = yl; j <= yh - 1 + py; ++j)
            for (int i = xl; i <= xh; ++i) {
                HX(i, j, zl - 1) = HX(i, j, zl - 1) - kz * inc_w(s, m, WEy, R(i), R(j) + R(0.5), R(zl), cf) * R(p->wall[4]);
                HX(i, j, zh) = HX(i, j, zh) + kz * inc_w(s, m, WEy, R(i), R(j) + R(0.5), R(zh), cf) * R(p->wall[5]);
            }
    }
    if (!px) {
        for (int k = zl; k <= zh; ++k)
            for (int j = yl; j <= yh - 1 + py; ++j) {
                HZ(xl - 1, j, k) = HZ(xl - 1, j, k) + kx * inc_w(s, m, WEy, R(xl), R(j) + R(0.5), R(k), cf) * R(p->wall[0]);
                HZ(xh, j, k) = HZ(xh, j, k) - kx * inc_w(s, m, WEy, R(xh), R(j) + R(0.5), R(k), cf) * R(p->wall[1]);
            }
        for (int k = zl; k <= zh - 1 + pz; ++k)
            for (int j = yl; j <= yh; ++j) {
                HY(xl - 1, j, k) = HY(xl - 1, j, k) - kx * inc_w(s, m, WEz, R(xl), R(j), R(k) + R(0.5), cf) * R(p->wall[0]);
                HY(xh, j, k) = HY(xh, j, k) + kx * inc_w(s, m, WEz, R(xh), R(j), R(k) + R(0.5), cf) * R(p->wall[1]);
            }
    }
}

static void h_tfsf(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    if (!p->has_tfsf) return;
    h_tfsf_wave(s, n, 0);
    if (p->mirror_type) h_tfsf_wave(s, n, 1);   /* after every face of the incident wave */
}

/* ports, H phase: loop integral, 1-D line current, source (f90:4800-5013) */
static void h_ports(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    const int M = p->M;
    for (int q = 0; q < p->n_ports; ++q) {
        const aem_port_view* pt = &p->ports[q];
        int l1 = pt->loc[0], l2 = pt->loc[1], l3 = pt->loc[2];
        int d1 = pt->size[0] + l1 - 1, d2 = pt->size[1] + l2 - 1, d3 = pt->size[2] + l3 - 1;
        real c = R(0.0);
        if (pt->dir == 0) {
            for (int ii = l1; ii <= d1; ++ii) {
                for (int jj = l2; jj <= d2; ++jj) c = c + R(p->dy) * (HY(ii, jj, l3 - 1) - HY(ii, jj, d3));
                for (int kk = l3; kk <= d3; ++kk) c = c + R(p->dz) * (HZ(ii, l2, kk) - HZ(ii, d2 - 1, kk));
            }
        } else if (pt->dir == 1) {
            for (int jj = l2; jj <= d2; ++jj) {
                for (int ii = l1; ii <= d1; ++ii) c = c + R(p->dx) * (HX(ii, jj, l3) - HX(ii, jj, d3 - 1));
                for (int kk = l3; kk <= d3; ++kk) c = c + R(p->dz) * (HZ(l1 - 1, jj, kk) - HZ(d1, jj, kk));
            }
        } else {
            for (int kk = l3; kk <= d3; ++kk) {
                for (int ii = l1; ii <= d1; ++ii) c = c + R(p->dx) * (HX(ii, l2 - 1, kk) - HX(ii, d2, kk));
                for (int jj = l2; jj <= d2; ++jj) c = c + R(p->dy) * (HY(l1, jj, kk) - HY(d1 - 1, jj, kk));
            }
        }
        s->I[(int64_t)q * M + 0] = DIVR(s, c * R(pt->cur_num), R(pt->cur_den));
        s->Iout[(int64_t)q * p->time_steps + (n - 1)] = s->I[(int64_t)q * M + 0];
    }
    for (int q = 0; q < p->n_ports; ++q) {
        real* I = s->I + (int64_t)q * M;
        real* V = s->V + (int64_t)q * M;
        const real kI = R(p->ports[q].kI);
        for (int j = 2; j <= M - 1; ++j) I[j - 1] = I[j - 1] - kI * (V[j] - V[j - 1]);
    }
    for (int q = 0; q < p->n_ports; ++q)
        if (q + 1 == p->excitation_port) {
            real* I = s->I + (int64_t)q * M;
            const aem_port_view* pt = &p->ports[q];
            I[1] = I[1] - R(pt->kI) * e_or_v(s, R(p->antenna_amp), R(n) + R(0.0), R(0.0) * R(pt->dl));
        }
}

/* SPICE gap currents after the H phase, f90:5442-5498 (basic ports): the
 * loop integral of H around each gap, scaled as a lumped port's. */
static void h_spice(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    for (int q = 0; q < p->n_spice; ++q) {
        const aem_spice_view* pt = &p->spice[q];
        int l1 = pt->loc[0], l2 = pt->loc[1], l3 = pt->loc[2];
        int d1 = pt->size[0] + l1 - 1, d2 = pt->size[1] + l2 - 1, d3 = pt->size[2] + l3 - 1;
        real c = R(0.0);
        if (pt->dir == 0) {
            for (int ii = l1; ii <= d1; ++ii) {
                for (int jj = l2; jj <= d2; ++jj) c = c + R(p->dy) * (HY(ii, jj, l3 - 1) - HY(ii, jj, d3));
                for (int kk = l3; kk <= d3; ++kk) c = c + R(p->dz) * (HZ(ii, l2, kk) - HZ(ii, d2 - 1, kk));
            }
        } else if (pt->dir == 1) {
            for (int jj = l2; jj <= d2; ++jj) {
                for (int ii = l1; ii <= d1; ++ii) c = c + R(p->dx) * (HX(ii, jj, l3) - HX(ii, jj, d3 - 1));
                for (int kk = l3; kk <= d3; ++kk) c = c + R(p->dz) * (HZ(l1 - 1, jj, kk) - HZ(d1, jj, kk));
            }
        } else {
            for (int kk = l3; kk <= d3; ++kk) {
                for (int ii = l1; ii <= d1; ++ii) c = c + R(p->dx) * (HX(ii, l2 - 1, kk) - HX(ii, d2, kk));
                for (int jj = l2; jj <= d2; ++jj) c = c + R(p->dy) * (HY(l1, jj, kk) - HY(d1 - 1, jj, kk));
            }
        }
        s->spI[q] = c * R(pt->cur_num) / R(pt->cur_den);
        s->spIout[(int64_t)q * p->time_steps + (n - 1)] = s->spI[q];
    }
}

/* CPML H, f90:5227-5375 */
static void h_cpml(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const real DB = R(p->DB);
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    const int nx1 = p->npml[0], nx2 = p->npml[1], ny1 = p->npml[2], ny2 = p->npml[3], nz1 = p->npml[4], nz2 = p->npml[5];
    const real dx = R(p->dx), dy = R(p->dy), dz = R(p->dz);
    /* Hx, y faces (psi_Hxy): face 2/3, component slot 0; transverse (u=i, v=k) */
    for (int k = 1; k <= NZ - 1; ++k)
        for (int i = 1; i <= NX - 1; ++i) {
            for (int j = 1; j <= ny1 - 1; ++j) {
                real* ps = psiH_at(s, 2, 0, j, i, k);
                *ps = R(p->bh[2][j - 1]) * *ps + DIVR(s, R(p->ch[2][j - 1]) * (EZ(i, j, k) - EZ(i, j + 1, k)), dy);
                HX(i, j, k) = HX(i, j, k) + DB * *ps;
            }
            for (int j = NY + 1 - ny2; j <= NY - 1; ++j) {
                int jj = NY - j;
                real* ps = psiH_at(s, 3, 0, jj, i, k);
                *ps = R(p->bh[3][jj - 1]) * *ps + DIVR(s, R(p->ch[3][jj - 1]) * (EZ(i, j, k) - EZ(i, j + 1, k)), dy);
                HX(i, j, k) = HX(i, j, k) + DB * *ps;
            }
        }
    /* Hx, z faces (psi_Hxz): transverse (u=i, v=j) */
    for (int j = 1; j <= NY - 1; ++j)
        for (int i = 1; i <= NX - 1; ++i) {
            for (int k = 1; k <= nz1 - 1; ++k) {
                real* ps = psiH_at(s, 4, 0, k, i, j);
                *ps = R(p->bh[4][k - 1]) * *ps + DIVR(s, R(p->ch[4][k - 1]) * (EY(i, j, k + 1) - EY(i, j, k)), dz);
                HX(i, j, k) = HX(i, j, k) + DB * *ps;
            }
            for (int k = NZ + 1 - nz2; k <= NZ - 1; ++k) {
                int kk = NZ - k;
                real* ps = psiH_at(s, 5, 0, kk, i, j);
                *ps = R(p->bh[5][kk - 1]) * *ps + DIVR(s, R(p->ch[5][kk - 1]) * (EY(i, j, k + 1) - EY(i, j, k)), dz);
                HX(i, j, k) = HX(i, j, k) + DB * *ps;
            }
        }
    /* Hy, x faces (psi_Hyx): transverse (u=j, v=k) */
    for (int k = 1; k <= NZ - 1; ++k)
        for (int j = 1; j <= NY - 1; ++j) {
            for (int i = 1; i <= nx1 - 1; ++i) {
                real* ps = psiH_at(s, 0, 0, i, j, k);
                *ps = R(p->bh[0][i - 1]) * *ps + DIVR(s, R(p->ch[0][i - 1]) * (EZ(i + 1, j, k) - EZ(i, j, k)), dx);
                HY(i, j, k) = HY(i, j, k) + DB * *ps;
            }
            for (int i = NX + 1 - nx2; i <= NX - 1; ++i) {
                int ii = NX - i;
                real* ps = psiH_at(s, 1, 0, ii, j, k);
                *ps = R(p->bh[1][ii - 1]) * *ps + DIVR(s, R(p->ch[1][ii - 1]) * (EZ(i + 1, j, k) - EZ(i, j, k)), dx);
                HY(i, j, k) = HY(i, j, k) + DB * *ps;
            }
        }
    /* Hy, z faces (psi_Hyz): slot 1 */
    for (int j = 1; j <= NY - 1; ++j)
        for (int i = 1; i <= NX - 1; ++i) {
            for (int k = 1; k <= nz1 - 1; ++k) {
                real* ps = psiH_at(s, 4, 1, k, i, j);
                *ps = R(p->bh[4][k - 1]) * *ps + DIVR(s, R(p->ch[4][k - 1]) * (EX(i, j, k) - EX(i, j, k + 1)), dz);
                HY(i, j, k) = HY(i, j, k) + DB * *ps;
            }
            for (int k = NZ + 1 - nz2; k <= NZ - 1; ++k) {
                int kk = NZ - k;
                real* ps = psiH_at(s, 5, 1, kk, i, j);
                *ps = R(p->bh[5][kk - 1]) * *ps + DIVR(s, R(p->ch[5][kk - 1]) * (EX(i, j, k) - EX(i, j, k + 1)), dz);
                HY(i, j, k) = HY(i, j, k) + DB * *ps;
            }
        }
    /* Hz, x faces (psi_Hzx): slot 1 */
    for (int k = 1; k <= NZ - 1; ++k)
        for (int j = 1; j <= NY - 1; ++j) {
            for (int i = 1; i <= nx1 - 1; ++i) {
                real* ps = psiH_at(s, 0, 1, i, j, k);
                *ps = R(p->bh[0][i - 1]) * *ps + DIVR(s, R(p->ch[0][i - 1]) * (EY(i, j, k) - EY(i + 1, j, k)), dx);
                HZ(i, j, k) = HZ(i, j, k) + DB * *ps;
            }
            for (int i = NX + 1 - nx2; i <= NX - 1; ++i) {
                int ii = NX - i;
                real* ps = psiH_at(s, 1, 1, ii, j, k);
                *ps = R(p->bh[1][ii - 1]) * *ps + DIVR(s, R(p->ch[1][ii - 1]) * (EY(i, j, k) - EY(i + 1, j, k)), dx);
                HZ(i, j, k) = HZ(i, j, k) + DB * *ps;
            }
        }
    /* Hz, y faces (psi_Hzy): slot 1 */
    for (int k = 1; k <= NZ - 1; ++k)
        for (int i = 1; i <= NX - 1; ++i) {
            for (int j = 1; j <= ny1 - 1; ++j) {
                real* ps = psiH_at(s, 2, 1, j, i, k);
                *ps = R(p->bh[2][j - 1]) * *ps + DIVR(s, R(p->ch[2][j - 1]) * (EX(i, j + 1, k) - EX(i, j, k)), dy);
                HZ(i, j, k) = HZ(i, j, k) + DB * *ps;
            }
            for (int j = NY + 1 - ny2; j <= NY - 1; ++j) {
                int jj = NY - j;
                real* ps = psiH_at(s, 3, 1, jj, i, k);
                *ps = R(p->bh[3][jj - 1]) * *ps + DIVR(s, R(p->ch[3][jj - 1]) * (EX(i, j + 1, k) - EX(i, j, k)), dy);
                HZ(i, j, k) = HZ(i, j, k) + DB * *ps;
            }
        }
}

/* PBC H copies, f90:5377-5433 */
static void h_pbc(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    if (p->pbc[0])
        for (int k = 1; k <= NZ - 1; ++k)
            for (int j = 1; j <= NY - 1; ++j) { HZ(NX - 1, j, k) = HZ(1, j, k); HY(NX - 1, j, k) = HY(1, j, k); }
    if (p->pbc[1])
        for (int k = 1; k <= NZ - 1; ++k)
            for (int i = 1; i <= NX - 1; ++i) { HZ(i, NY - 1, k) = HZ(i, 1, k); HX(i, NY - 1, k) = HX(i, 1, k); }
    if (p->pbc[2])
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i) { HY(i, j, NZ - 1) = HY(i, j, 1); HX(i, j, NZ - 1) = HX(i, j, 1); }
}

/* ---- E phase ------------------------------------------------------------ */

/* bulk E, f90:5667-5710 */
static void e_bulk(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    for (int k = 2; k <= NZ - 1; ++k)
        for (int j = 2; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i) {
                int64_t o = IDX(s, i, j, k);
                s->Ex[o] = GA(s, 0, o) * s->Ex[o] + GB(s, 0, o) * ((HZ(i, j, k) - HZ(i, j - 1, k)) * R(p->den_ey[j - 1]) +
                                                                  (HY(i, j, k - 1) - HY(i, j, k)) * R(p->den_ez[k - 1]));
            }
    for (int k = 2; k <= NZ - 1; ++k)
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 2; i <= NX - 1; ++i) {
                int64_t o = IDX(s, i, j, k);
                s->Ey[o] = GA(s, 1, o) * s->Ey[o] + GB(s, 1, o) * ((HZ(i - 1, j, k) - HZ(i, j, k)) * R(p->den_ex[i - 1]) +
                                                                  (HX(i, j, k) - HX(i, j, k - 1)) * R(p->den_ez[k - 1]));
            }
    for (int k = 1; k <= NZ - 1; ++k)
        for (int j = 2; j <= NY - 1; ++j)
            for (int i = 2; i <= NX - 1; ++i) {
                int64_t o = IDX(s, i, j, k);
                s->Ez[o] = GA(s, 2, o) * s->Ez[o] + GB(s, 2, o) * ((HY(i, j, k) - HY(i - 1, j, k)) * R(p->den_ex[i - 1]) +
                                                                  (HX(i, j - 1, k) - HX(i, j, k)) * R(p->den_ey[j - 1]));
            }
}

/* sheet-normal special E, f90:5712-5780 */
static void e_sheets(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    const real dt = R(p->dt), e0 = R(8.85418782E-12f);
    for (int q = 0; q < p->n_sheet_planes[0]; ++q) {
        int i = p->sheet_pos[0][q];
        for (int k = 2; k <= NZ - 1; ++k)
            for (int j = 2; j <= NY - 1; ++j)
                if (sheet_on(s, 0, q, j, k)) {
                    real sg = sheet_sig(s, 0, q, j, k), ep = sheet_ep(s, 0, q, j, k);
                    real* e = ES(s, 0, q, j, k);
                    *e = ((R(1.0) - sg * dt / (R(2.0) * e0 * ep)) / (R(1.0) + sg * dt / (R(2.0) * e0 * ep))) * *e +
                         ((dt / (e0 * ep)) / (R(1.0) + sg * dt / (R(2.0) * e0 * ep))) *
                             ((HZ(i, j, k) - HZ(i, j - 1, k)) * R(p->den_ey[j - 1]) + (HY(i, j, k - 1) - HY(i, j, k)) * R(p->den_ez[k - 1]));
                }
    }
    for (int q = 0; q < p->n_sheet_planes[1]; ++q) {
        int j = p->sheet_pos[1][q];
        for (int k = 2; k <= NZ - 1; ++k)
            for (int i = 2; i <= NX - 1; ++i)
                if (sheet_on(s, 1, q, i, k)) {
                    real sg = sheet_sig(s, 1, q, i, k), ep = sheet_ep(s, 1, q, i, k);
                    real* e = ES(s, 1, q, i, k);
                    *e = ((R(1.0) - sg * dt / (R(2.0) * e0 * ep)) / (R(1.0) + sg * dt / (R(2.0) * e0 * ep))) * *e +
                         ((dt / (e0 * ep)) / (R(1.0) + sg * dt / (R(2.0) * e0 * ep))) *
                             ((HZ(i - 1, j, k) - HZ(i, j, k)) * R(p->den_ex[i - 1]) + (HX(i, j, k) - HX(i, j, k - 1)) * R(p->den_ez[k - 1]));
                }
    }
    for (int q = 0; q < p->n_sheet_planes[2]; ++q) {
        int k = p->sheet_pos[2][q];
        for (int j = 2; j <= NY - 1; ++j)
            for (int i = 2; i <= NX - 1; ++i)
                if (sheet_on(s, 2, q, i, j)) {
                    real sg = sheet_sig(s, 2, q, i, j), ep = sheet_ep(s, 2, q, i, j);
                    real* e = ES(s, 2, q, i, j);
                    *e = ((R(1.0) - sg * dt / (R(2.0) * e0 * ep)) / (R(1.0) + sg * dt / (R(2.0) * e0 * ep))) * *e +
                         ((dt / (e0 * ep)) / (R(1.0) + sg * dt / (R(2.0) * e0 * ep))) *
                             ((HY(i, j, k) - HY(i - 1, j, k)) * R(p->den_ex[i - 1]) + (HX(i, j - 1, k) - HX(i, j, k)) * R(p->den_ey[j - 1]));
                }
    }
}

/* Drude ADE, f90:5782-5854 */
static void e_drude(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    if (!p->has_drude) return;
    const int P = p->n_poles;
    real* Fc[3] = {s->Ex, s->Ey, s->Ez};
    for (int c = 0; c < 3; ++c) {
        const int* b = p->dbox[c];
        int64_t bx = b[1] - b[0] + 1, by = b[3] - b[2] + 1;
        int64_t nb = s->nbox[c];
        for (int k = b[4]; k <= b[5]; ++k)
            for (int j = b[2]; j <= b[3]; ++j)
                for (int i = b[0]; i <= b[1]; ++i) {
                    int64_t o = IDX(s, i, j, k);
                    int64_t bo = (i - b[0]) + bx * ((int64_t)(j - b[2]) + by * (int64_t)(k - b[4]));
                    real t = R(0.0);
                    for (int q = 0; q < P; ++q) t = t + (R(1.0) + R(p->Ja[c][q * nb + bo])) * s->J[c][q * nb + bo];
                    Fc[c][o] = Fc[c][o] - GB(s, c, o) * t / R(2.0);
                    for (int q = 0; q < P; ++q)
                        s->J[c][q * nb + bo] = R(p->Ja[c][q * nb + bo]) * s->J[c][q * nb + bo] +
                                               R(p->Jb[c][q * nb + bo]) * (Fc[c][o] + s->Eold[c][bo]);
                }
    }
}

/* TF/SF E faces, f90:5950-6033 (t = n-1/2) */
static void e_tfsf_wave(oracle_state* s, int n, int m) {
    const aem_plan_view* p = &s->p;
    const int xl = p->tf_lo[0], xh = p->tf_hi[0], yl = p->tf_lo[1], yh = p->tf_hi[1], zl = p->tf_lo[2], zh = p->tf_hi[2];
    const int px = p->pbc[0], py = p->pbc[1], pz = p->pbc[2];
    const real cf = R(n) - R(0.5);
    const real kx = R(p->kEinc[0]), ky = R(p->kEinc[1]), kz = R(p->kEinc[2]);
    const float* W = m ? p->WH_m : p->WH;
    const real WHx = R(W[0]), WHy = R(W[1]), WHz = R(W[2]);
    const real h = R(0.5);
    if (!py) {
        for (int k = zl; k <= zh - 1 + pz; ++k)
            for (int i = xl + px; i <= xh; ++i) {
                EZ(i, yl, k) = EZ(i, yl, k) + ky * inc_w(s, m, WHx, R(i), R(yl) - h, R(k) + h, cf) * R(p->wall[2]);
                EZ(i, yh, k) = EZ(i, yh, k) - ky * inc_w(s, m, WHx, R(i), R(yh) + h, R(k) + h, cf) * R(p->wall[3]);
            }
        for (int k = zl + pz; k <= zh; ++k)
            for (int i = xl; i <= xh - 1 + px; ++i) {
                EX(i, yl, k) = EX(i, yl, k) - ky * inc_w(s, m, WHz, R(i) + h, R(yl) - h, R(k), cf) * R(p->wall[2]);
                EX(i, yh, k) = EX(i, yh, k) + ky * inc_w(s, m, WHz, R(i) + h, R(yh) + h, R(k), cf) * R(p->wall[3]);
            }
    }
    if (!pz) {
        for (int j = yl; j <= yh - 1 + py; ++j)
            for (int i = xl + px; i <= xh; ++i) {
                EY(i, j, zl) = EY(i, j, zl) - kz * inc_w(s, m, WHx, R(i), R(j) + h, R(zl) - h, cf) * R(p->wall[4]);
                EY(i, j, zh) = EY(i, j, zh) + kz * inc_w(s, m, WHx, R(i), R(j) + h, R(zh) + h, cf) * R(p->wall[5]);
            }
        for (int j = yl + py; j <= yh; ++j)
            for (int i = xl; i <= xh - 1 + px; ++i) {
                EX(i, j, zl) = EX(i, j, zl) + kz * inc_w(s, m, WHy, R(i) + h, R(j), R(zl) - h, cf) * R(p->wall[4]);
                EX(i, j, zh) = EX(i, j, zh) - kz * inc_w(s, m, WHy, R(i) + h, R(j), R(zh) + h, cf) * R(p->wall[5]);
            }
    }
    if (!px) {
        for (int k = zl; k <= zh - 1 + pz; ++k)
            for (int j = yl + py; j <= yh; ++j) {
                EZ(xl, j, k) = EZ(xl, j, k) - kx * inc_w(s, m, WHy, R(xl) - h, R(j), R(k) + h, cf) * R(p->wall[0]);
                EZ(xh, j, k) = EZ(xh, j, k) + kx * inc_w(s, m, WHy, R(xh) + h, R(j), R(k) + h, cf) * R(p->wall[1]);
            }
        for (int k = zl + pz; k <= zh; ++k)
            for (int j = yl; j <= yh - 1 + py; ++j) {
                EY(xl, j, k) = EY(xl, j, k) + kx * inc_w(s, m, WHz, R(xl) - h, R(j) + h, R(k), cf) * R(p->wall[0]);
                EY(xh, j, k) = EY(xh, j, k) - kx * inc_w(s, m, WHz, R(xh) + h, R(j) + h, R(k), cf) * R(p->wall[1]);
            }
    }
}

static void e_tfsf(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    if (!p->has_tfsf) return;
    e_tfsf_wave(s, n, 0);
    if (p->mirror_type) e_tfsf_wave(s, n, 1);   /* after every face of the incident wave */
}

/* ports, E phase: 1-D line voltage, source, hard-set gap field (f90:6120-6283) */
static void e_ports(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    const int M = p->M;
    for (int q = 0; q < p->n_ports; ++q) {
        real* I = s->I + (int64_t)q * M;
        real* V = s->V + (int64_t)q * M;
        const real kV = R(p->ports[q].kV);
        for (int j = 2; j <= M - 1; ++j) V[j - 1] = V[j - 1] - kV * (I[j - 1] - I[j - 2]);
    }
    for (int q = 0; q < p->n_ports; ++q)
        if (q + 1 == p->excitation_port) {
            const aem_port_view* pt = &p->ports[q];
            real* V = s->V + (int64_t)q * M;
            V[1] = V[1] + R(pt->kS) * e_or_v(s, R(p->antenna_amp), R(n) + R(0.5), R(0.5) * R(pt->dl));
        }
    for (int q = 0; q < p->n_ports; ++q) {
        const aem_port_view* pt = &p->ports[q];
        real* V = s->V + (int64_t)q * M;
        int l1 = pt->loc[0], l2 = pt->loc[1], l3 = pt->loc[2];
        int d1 = pt->size[0] + l1 - 1, d2 = pt->size[1] + l2 - 1, d3 = pt->size[2] + l3 - 1;
        real* F = pt->dir == 0 ? s->Ex : (pt->dir == 1 ? s->Ey : s->Ez);
        real val = DIVR(s, R(1.0) * V[1], R(pt->gap_div));
        for (int ii = l1; ii <= d1; ++ii)
            for (int jj = l2; jj <= d2; ++jj)
                for (int kk = l3; kk <= d3; ++kk) F[IDX(s, ii, jj, kk)] = val;
    }
}

/* SPICE gap fields, f90:6285-6330 (basic ports): E = 1.0*V/(n*delta) on the
 * gap cells, with V = node voltage - DC bias (f90:5650-5656); the series
 * record V, and the extra nodes' raw voltages (f90:8250-8254). */
static void e_spice(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    for (int q = 0; q < p->n_spice; ++q) {
        const aem_spice_view* pt = &p->spice[q];
        int l1 = pt->loc[0], l2 = pt->loc[1], l3 = pt->loc[2];
        int d1 = pt->size[0] + l1 - 1, d2 = pt->size[1] + l2 - 1, d3 = pt->size[2] + l3 - 1;
        real* F = pt->dir == 0 ? s->Ex : (pt->dir == 1 ? s->Ey : s->Ez);
        real val = R(1.0) * s->spV[q] / R(pt->gap_div);
        for (int ii = l1; ii <= d1; ++ii)
            for (int jj = l2; jj <= d2; ++jj)
                for (int kk = l3; kk <= d3; ++kk) F[IDX(s, ii, jj, kk)] = val;
    }
    for (int q = 0; q < p->n_spice + p->n_spice_extra; ++q)
        s->spVout[(int64_t)q * p->time_steps + (n - 1)] = s->spV[q];
}

/* CPML E, f90:6522-6670 */
static void e_cpml(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    const int nx1 = p->npml[0], nx2 = p->npml[1], ny1 = p->npml[2], ny2 = p->npml[3], nz1 = p->npml[4], nz2 = p->npml[5];
    const real dx = R(p->dx), dy = R(p->dy), dz = R(p->dz);
    /* Ex, y faces */
    for (int k = 2; k <= NZ - 1; ++k)
        for (int i = 1; i <= NX - 1; ++i) {
            for (int j = 2; j <= ny1; ++j) {
                real* ps = psiE_at(s, 2, 0, j, i, k);
                *ps = R(p->be[2][j - 1]) * *ps + DIVR(s, R(p->ce[2][j - 1]) * (HZ(i, j, k) - HZ(i, j - 1, k)), dy);
                EX(i, j, k) = EX(i, j, k) + GB(s, 0, IDX(s, i, j, k)) * *ps;
            }
            for (int j = NY + 1 - ny2; j <= NY - 1; ++j) {
                int jj = NY + 1 - j;
                real* ps = psiE_at(s, 3, 0, jj, i, k);
                *ps = R(p->be[3][jj - 1]) * *ps + DIVR(s, R(p->ce[3][jj - 1]) * (HZ(i, j, k) - HZ(i, j - 1, k)), dy);
                EX(i, j, k) = EX(i, j, k) + GB(s, 0, IDX(s, i, j, k)) * *ps;
            }
        }
    /* Ex, z faces */
    for (int j = 2; j <= NY - 1; ++j)
        for (int i = 1; i <= NX - 1; ++i) {
            for (int k = 2; k <= nz1; ++k) {
                real* ps = psiE_at(s, 4, 0, k, i, j);
                *ps = R(p->be[4][k - 1]) * *ps + DIVR(s, R(p->ce[4][k - 1]) * (HY(i, j, k - 1) - HY(i, j, k)), dz);
                EX(i, j, k) = EX(i, j, k) + GB(s, 0, IDX(s, i, j, k)) * *ps;
            }
            for (int k = NZ + 1 - nz2; k <= NZ - 1; ++k) {
                int kk = NZ + 1 - k;
                real* ps = psiE_at(s, 5, 0, kk, i, j);
                *ps = R(p->be[5][kk - 1]) * *ps + DIVR(s, R(p->ce[5][kk - 1]) * (HY(i, j, k - 1) - HY(i, j, k)), dz);
                EX(i, j, k) = EX(i, j, k) + GB(s, 0, IDX(s, i, j, k)) * *ps;
            }
        }
    /* Ey, x faces */
    for (int k = 2; k <= NZ - 1; ++k)
        for (int j = 1; j <= NY - 1; ++j) {
            for (int i = 2; i <= nx1; ++i) {
                real* ps = psiE_at(s, 0, 0, i, j, k);
                *ps = R(p->be[0][i - 1]) * *ps + DIVR(s, R(p->ce[0][i - 1]) * (HZ(i - 1, j, k) - HZ(i, j, k)), dx);
                EY(i, j, k) = EY(i, j, k) + GB(s, 1, IDX(s, i, j, k)) * *ps;
            }
            for (int i = NX + 1 - nx2; i <= NX - 1; ++i) {
                int ii = NX + 1 - i;
                real* ps = psiE_at(s, 1, 0, ii, j, k);
                *ps = R(p->be[1][ii - 1]) * *ps + DIVR(s, R(p->ce[1][ii - 1]) * (HZ(i - 1, j, k) - HZ(i, j, k)), dx);
                EY(i, j, k) = EY(i, j, k) + GB(s, 1, IDX(s, i, j, k)) * *ps;
            }
        }
    /* Ey, z faces */
    for (int j = 1; j <= NY - 1; ++j)
        for (int i = 2; i <= NX - 1; ++i) {
            for (int k = 2; k <= nz1; ++k) {
                real* ps = psiE_at(s, 4, 1, k, i, j);
                *ps = R(p->be[4][k - 1]) * *ps + DIVR(s, R(p->ce[4][k - 1]) * (HX(i, j, k) - HX(i, j, k - 1)), dz);
                EY(i, j, k) = EY(i, j, k) + GB(s, 1, IDX(s, i, j, k)) * *ps;
            }
            for (int k = NZ + 1 - nz2; k <= NZ - 1; ++k) {
                int kk = NZ + 1 - k;
                real* ps = psiE_at(s, 5, 1, kk, i, j);
                *ps = R(p->be[5][kk - 1]) * *ps + DIVR(s, R(p->ce[5][kk - 1]) * (HX(i, j, k) - HX(i, j, k - 1)), dz);
                EY(i, j, k) = EY(i, j, k) + GB(s, 1, IDX(s, i, j, k)) * *ps;
            }
        }
    /* Ez, x faces */
    for (int k = 1; k <= NZ - 1; ++k)
        for (int j = 2; j <= NY - 1; ++j) {
            for (int i = 2; i <= nx1; ++i) {
                real* ps = psiE_at(s, 0, 1, i, j, k);
                *ps = R(p->be[0][i - 1]) * *ps + DIVR(s, R(p->ce[0][i - 1]) * (HY(i, j, k) - HY(i - 1, j, k)), dx);
                EZ(i, j, k) = EZ(i, j, k) + GB(s, 2, IDX(s, i, j, k)) * *ps;
            }
            for (int i = NX + 1 - nx2; i <= NX - 1; ++i) {
                int ii = NX + 1 - i;
                real* ps = psiE_at(s, 1, 1, ii, j, k);
                *ps = R(p->be[1][ii - 1]) * *ps + DIVR(s, R(p->ce[1][ii - 1]) * (HY(i, j, k) - HY(i - 1, j, k)), dx);
                EZ(i, j, k) = EZ(i, j, k) + GB(s, 2, IDX(s, i, j, k)) * *ps;
            }
        }
    /* Ez, y faces */
    for (int k = 1; k <= NZ - 1; ++k)
        for (int i = 2; i <= NX - 1; ++i) {
            for (int j = 2; j <= ny1; ++j) {
                real* ps = psiE_at(s, 2, 1, j, i, k);
                *ps = R(p->be[2][j - 1]) * *ps + DIVR(s, R(p->ce[2][j - 1]) * (HX(i, j - 1, k) - HX(i, j, k)), dy);
                EZ(i, j, k) = EZ(i, j, k) + GB(s, 2, IDX(s, i, j, k)) * *ps;
            }
            for (int j = NY + 1 - ny2; j <= NY - 1; ++j) {
                int jj = NY + 1 - j;
                real* ps = psiE_at(s, 3, 1, jj, i, k);
                *ps = R(p->be[3][jj - 1]) * *ps + DIVR(s, R(p->ce[3][jj - 1]) * (HX(i, j - 1, k) - HX(i, j, k)), dy);
                EZ(i, j, k) = EZ(i, j, k) + GB(s, 2, IDX(s, i, j, k)) * *ps;
            }
        }
}

/* PBC E and special-E copies, f90:6672-6741 */
static void e_pbc(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    if (p->pbc[0]) {
        for (int k = 1; k <= NZ - 1; ++k)
            for (int j = 1; j <= NY - 1; ++j) { EZ(1, j, k) = EZ(NX - 1, j, k); EY(1, j, k) = EY(NX - 1, j, k); }
        for (int q = 0; q < p->n_sheet_planes[2]; ++q)       /* Ez_special on z planes: (i=1,j) */
            if (p->sheet_pos[2][q] <= NZ - 1)
                for (int j = 1; j <= NY - 1; ++j) *ES(s, 2, q, 1, j) = *ES(s, 2, q, NX - 1, j);
        for (int q = 0; q < p->n_sheet_planes[1]; ++q)       /* Ey_special on y planes: (i=1,k) */
            if (p->sheet_pos[1][q] <= NY - 1)
                for (int k = 1; k <= NZ - 1; ++k) *ES(s, 1, q, 1, k) = *ES(s, 1, q, NX - 1, k);
    }
    if (p->pbc[1]) {
        for (int k = 1; k <= NZ - 1; ++k)
            for (int i = 1; i <= NX - 1; ++i) { EZ(i, 1, k) = EZ(i, NY - 1, k); EX(i, 1, k) = EX(i, NY - 1, k); }
        for (int q = 0; q < p->n_sheet_planes[2]; ++q)
            if (p->sheet_pos[2][q] <= NZ - 1)
                for (int i = 1; i <= NX - 1; ++i) *ES(s, 2, q, i, 1) = *ES(s, 2, q, i, NY - 1);
        for (int q = 0; q < p->n_sheet_planes[0]; ++q)
            if (p->sheet_pos[0][q] <= NX - 1)
                for (int k = 1; k <= NZ - 1; ++k) *ES(s, 0, q, 1, k) = *ES(s, 0, q, NY - 1, k);
    }
    if (p->pbc[2]) {
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i) { EY(i, j, 1) = EY(i, j, NZ - 1); EX(i, j, 1) = EX(i, j, NZ - 1); }
        for (int q = 0; q < p->n_sheet_planes[1]; ++q)
            if (p->sheet_pos[1][q] <= NY - 1)
                for (int i = 1; i <= NX - 1; ++i) *ES(s, 1, q, i, 1) = *ES(s, 1, q, i, NZ - 1);
        for (int q = 0; q < p->n_sheet_planes[0]; ++q)
            if (p->sheet_pos[0][q] <= NX - 1)
                for (int j = 1; j <= NY - 1; ++j) *ES(s, 0, q, j, 1) = *ES(s, 0, q, j, NZ - 1);
    }
}

/* E_oldt = E on the plasma boxes, f90:6812-6826 */
static void e_oldt(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    if (!p->has_drude) return;
    real* Fc[3] = {s->Ex, s->Ey, s->Ez};
    for (int c = 0; c < 3; ++c) {
        const int* b = p->dbox[c];
        int64_t bx = b[1] - b[0] + 1, by = b[3] - b[2] + 1;
        for (int k = b[4]; k <= b[5]; ++k)
            for (int j = b[2]; j <= b[3]; ++j)
                for (int i = b[0]; i <= b[1]; ++i)
                    s->Eold[c][(i - b[0]) + bx * ((int64_t)(j - b[2]) + by * (int64_t)(k - b[4]))] = Fc[c][IDX(s, i, j, k)];
    }
}

/* ---- per-step outputs ----------------------------------------------------- */

/* video slice, f90:6836-6990 */
static void out_video(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    if (!p->video_on) return;
    const int v1 = p->vid_size1, v2 = p->vid_size2, T = p->time_steps;
    const int64_t ps = (int64_t)v1 * v2;
    real* V[6];
    for (int q = 0; q < 6; ++q) V[q] = s->vid[q];
    for (int b = 1; b <= v2; ++b)
        for (int a = 1; a <= v1; ++a) {
            int i, j, k;
            if (p->slice == 0) { i = p->slice_location; j = a; k = b; }
            else if (p->slice == 1) { j = p->slice_location; i = a; k = b; }
            else { k = p->slice_location; i = a; j = b; }
            int64_t o = (int64_t)(a - 1) + (int64_t)v1 * (b - 1);
            V[0][(int64_t)(n - 1) * ps + o] = (EX(i, j, k) + EX(i, j + 1, k) + EX(i, j, k + 1) + EX(i, j + 1, k + 1)) / R(4.0);
            V[1][(int64_t)(n - 1) * ps + o] = (EY(i, j, k) + EY(i + 1, j, k) + EY(i, j, k + 1) + EY(i + 1, j, k + 1)) / R(4.0);
            V[2][(int64_t)(n - 1) * ps + o] = (EZ(i, j, k) + EZ(i + 1, j, k) + EZ(i, j + 1, k) + EZ(i + 1, j + 1, k)) / R(4.0);
            real hx = (HX(i, j, k) + HX(i + 1, j, k)) / R(4.0);
            real hy = (HY(i, j, k) + HY(i, j + 1, k)) / R(4.0);
            real hz = (HZ(i, j, k) + HZ(i, j, k + 1)) / R(4.0);
            V[3][(int64_t)(n - 1) * ps + o] = hx; V[4][(int64_t)(n - 1) * ps + o] = hy; V[5][(int64_t)(n - 1) * ps + o] = hz;
            if (n < T) {
                V[3][(int64_t)n * ps + o] = hx; V[4][(int64_t)n * ps + o] = hy; V[5][(int64_t)n * ps + o] = hz;
            }
        }
}

/* reflection/transmission plane sums for two periodic axes, f90:7005-7103 */
static void out_plane_sums(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    if (p->plane_sum_axis < 0) return;
    real sr = R(0.0), st = R(0.0);
    const real sp = R(p->s_pol), cp = R(p->c_pol), sph = R(p->s_ph), cph = R(p->c_ph);
    if (p->plane_sum_axis == 0) {
        int xl = p->tf_lo[0], xh = p->tf_hi[0];
        for (int k = p->ff_lo[2]; k <= p->ff_hi[2]; ++k)
            for (int j = p->ff_lo[1]; j <= p->ff_hi[1]; ++j) {
                sr = sr + EZ(xl, j, k) * sp - EY(xl, j, k) * cp * cph;
                st = st + EZ(xh, j, k) * sp - EY(xh, j, k) * cp * cph;
            }
    } else if (p->plane_sum_axis == 1) {
        int yl = p->tf_lo[1], yh = p->tf_hi[1];
        for (int k = p->ff_lo[2]; k <= p->ff_hi[2]; ++k)
            for (int i = p->ff_lo[0]; i <= p->ff_hi[0]; ++i) {
                sr = sr + EZ(i, yl, k) * sp + EX(i, yl, k) * cp * sph;
                st = st + EZ(i, yh, k) * sp + EX(i, yh, k) * cp * sph;
            }
    } else {
        int zl = p->tf_lo[2], zh = p->tf_hi[2];
        const real c1 = R(p->ps_c1), c2 = R(p->ps_c2);
        for (int j = p->ff_lo[1]; j <= p->ff_hi[1]; ++j)
            for (int i = p->ff_lo[0]; i <= p->ff_hi[0]; ++i) {
                sr = sr + EY(i, j, zl) * c1 + EX(i, j, zl) * c2;
                st = st + EY(i, j, zh) * c1 + EX(i, j, zh) * c2;
            }
    }
    s->refl[n - 1] = s->refl[n - 1] + sr;
    s->trans[n - 1] = s->trans[n - 1] + st;
}

/* face arrays: f in 0..5, slot 0/1, (u,v) 1-based */
#define FA(arr, f, q, u, v) s->arr[f][q][(int64_t)(u) + s->fdim[f][0] * (int64_t)(v)]

/* NTFF surface currents, f90:7512-7641 */
static void out_ntff_currents(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const int fxl = p->ff_lo[0], fxh = p->ff_hi[0], fyl = p->ff_lo[1], fyh = p->ff_hi[1], fzl = p->ff_lo[2], fzh = p->ff_hi[2];
    if (!p->pbc[0])
        for (int k = fzl; k <= fzh; ++k)
            for (int j = fyl; j <= fyh; ++j) {
                int i = fxl;
                FA(Mold, 0, 0, j, k) = FA(Mcur, 0, 0, j, k);
                FA(Mcur, 0, 0, j, k) = R(-1.0) * (EZ(i, j, k) + EZ(i, j + 1, k) + EZ(i + 1, j, k) + EZ(i + 1, j + 1, k)) / R(4.0) * R(p->wall[0]);
                FA(Mold, 0, 1, j, k) = FA(Mcur, 0, 1, j, k);
                FA(Mcur, 0, 1, j, k) = (EY(i, j, k) + EY(i + 1, j, k) + EY(i, j, k + 1) + EY(i + 1, j, k + 1)) / R(4.0) * R(p->wall[0]);
                FA(Jold, 0, 0, j, k) = FA(Jcur, 0, 0, j, k);
                FA(Jcur, 0, 0, j, k) = (HZ(i, j, k) + HZ(i, j, k + 1)) / R(2.0) * R(p->wall[0]);
                FA(Jold, 0, 1, j, k) = FA(Jcur, 0, 1, j, k);
                FA(Jcur, 0, 1, j, k) = R(-1.0) * (HY(i, j, k) + HY(i, j + 1, k)) / R(2.0) * R(p->wall[0]);
                i = fxh;
                FA(Mold, 1, 0, j, k) = FA(Mcur, 1, 0, j, k);
                FA(Mcur, 1, 0, j, k) = (EZ(i, j, k) + EZ(i, j + 1, k) + EZ(i + 1, j, k) + EZ(i + 1, j + 1, k)) / R(4.0) * R(p->wall[1]);
                FA(Mold, 1, 1, j, k) = FA(Mcur, 1, 1, j, k);
                FA(Mcur, 1, 1, j, k) = R(-1.0) * (EY(i, j, k) + EY(i + 1, j, k) + EY(i, j, k + 1) + EY(i + 1, j, k + 1)) / R(4.0) * R(p->wall[1]);
                FA(Jold, 1, 0, j, k) = FA(Jcur, 1, 0, j, k);
                FA(Jcur, 1, 0, j, k) = R(-1.0) * (HZ(i, j, k) + HZ(i, j, k + 1)) / R(2.0) * R(p->wall[1]);
                FA(Jold, 1, 1, j, k) = FA(Jcur, 1, 1, j, k);
                FA(Jcur, 1, 1, j, k) = (HY(i, j, k) + HY(i, j + 1, k)) / R(2.0) * R(p->wall[1]);
            }
    if (!p->pbc[1])
        for (int k = fzl; k <= fzh; ++k)
            for (int i = fxl; i <= fxh; ++i) {
                int j = fyl;
                FA(Mold, 2, 0, i, k) = FA(Mcur, 2, 0, i, k);
                FA(Mcur, 2, 0, i, k) = (EZ(i, j, k) + EZ(i + 1, j, k) + EZ(i, j + 1, k) + EZ(i + 1, j + 1, k)) / R(4.0) * R(p->wall[2]);
                FA(Mold, 2, 1, i, k) = FA(Mcur, 2, 1, i, k);
                FA(Mcur, 2, 1, i, k) = R(-1.0) * (EX(i, j, k) + EX(i, j + 1, k) + EX(i, j, k + 1) + EX(i, j + 1, k + 1)) / R(4.0) * R(p->wall[2]);
                FA(Jold, 2, 0, i, k) = FA(Jcur, 2, 0, i, k);
                FA(Jcur, 2, 0, i, k) = R(-1.0) * (HZ(i, j, k) + HZ(i, j, k + 1)) / R(2.0) * R(p->wall[2]);
                FA(Jold, 2, 1, i, k) = FA(Jcur, 2, 1, i, k);
                FA(Jcur, 2, 1, i, k) = (HX(i, j, k) + HX(i + 1, j, k)) / R(2.0) * R(p->wall[2]);
                j = fyh;
                FA(Mold, 3, 0, i, k) = FA(Mcur, 3, 0, i, k);
                FA(Mcur, 3, 0, i, k) = R(-1.0) * (EZ(i, j, k) + EZ(i + 1, j, k) + EZ(i, j + 1, k) + EZ(i + 1, j + 1, k)) / R(4.0) * R(p->wall[3]);
                FA(Mold, 3, 1, i, k) = FA(Mcur, 3, 1, i, k);
                FA(Mcur, 3, 1, i, k) = (EX(i, j, k) + EX(i, j + 1, k) + EX(i, j, k + 1) + EX(i, j + 1, k + 1)) / R(4.0) * R(p->wall[3]);
                FA(Jold, 3, 0, i, k) = FA(Jcur, 3, 0, i, k);
                FA(Jcur, 3, 0, i, k) = (HZ(i, j, k) + HZ(i, j, k + 1)) / R(2.0) * R(p->wall[3]);
                FA(Jold, 3, 1, i, k) = FA(Jcur, 3, 1, i, k);
                FA(Jcur, 3, 1, i, k) = R(-1.0) * (HX(i, j, k) + HX(i + 1, j, k)) / R(2.0) * R(p->wall[3]);
            }
    if (!p->pbc[2])
        for (int j = fyl; j <= fyh; ++j)
            for (int i = fxl; i <= fxh; ++i) {
                int k = fzl;
                FA(Mold, 4, 0, i, j) = FA(Mcur, 4, 0, i, j);
                FA(Mcur, 4, 0, i, j) = R(-1.0) * (EY(i, j, k) + EY(i + 1, j, k) + EY(i, j, k + 1) + EY(i + 1, j, k + 1)) / R(4.0) * R(p->wall[4]);
                FA(Mold, 4, 1, i, j) = FA(Mcur, 4, 1, i, j);
                FA(Mcur, 4, 1, i, j) = (EX(i, j, k) + EX(i, j + 1, k) + EX(i, j, k + 1) + EX(i, j + 1, k + 1)) / R(4.0) * R(p->wall[4]);
                FA(Jold, 4, 0, i, j) = FA(Jcur, 4, 0, i, j);
                FA(Jcur, 4, 0, i, j) = (HY(i, j, k) + HY(i, j + 1, k)) / R(2.0) * R(p->wall[4]);
                FA(Jold, 4, 1, i, j) = FA(Jcur, 4, 1, i, j);
                FA(Jcur, 4, 1, i, j) = R(-1.0) * (HX(i, j, k) + HX(i + 1, j, k)) / R(2.0) * R(p->wall[4]);
                k = fzh;
                FA(Mold, 5, 0, i, j) = FA(Mcur, 5, 0, i, j);
                FA(Mcur, 5, 0, i, j) = (EY(i, j, k) + EY(i + 1, j, k) + EY(i, j, k + 1) + EY(i + 1, j, k + 1)) / R(4.0) * R(p->wall[5]);
                FA(Mold, 5, 1, i, j) = FA(Mcur, 5, 1, i, j);
                FA(Mcur, 5, 1, i, j) = R(-1.0) * (EX(i, j, k) + EX(i, j + 1, k) + EX(i, j, k + 1) + EX(i, j + 1, k + 1)) / R(4.0) * R(p->wall[5]);
                FA(Jold, 5, 0, i, j) = FA(Jcur, 5, 0, i, j);
                FA(Jcur, 5, 0, i, j) = R(-1.0) * (HY(i, j, k) + HY(i, j + 1, k)) / R(2.0) * R(p->wall[5]);
                FA(Jold, 5, 1, i, j) = FA(Jcur, 5, 1, i, j);
                FA(Jcur, 5, 1, i, j) = (HX(i, j, k) + HX(i + 1, j, k)) / R(2.0) * R(p->wall[5]);
            }
}

/* one face cell's retarded-time scatter into U (slot pair ua,ub) and W, f90:7660-7697.
 * Index and weight arithmetic is FP32 as in the reference, whatever "real" is. */
static void bin_one(oracle_state* s, int kk, int n, int i, int j, int k, real K, real* Ua, real* Ub, real* Wa, real* Wb,
                    real dMa, real dMb, real dJa, real dJb) {
    const aem_plan_view* p = &s->p;
    const float* a = p->ang + 6 * kk;
    const float s1 = a[2], c1 = a[3], s2 = a[4], c2 = a[5];
    float tfv = (p->r_relay - ((float)(i - p->ic) * p->dx * s1 * c2 + (float)(j - p->jc) * p->dy * s1 * s2 +
                               (float)(k - p->kc) * p->dz * c1)) / (299792448.0f * p->dt);
    const int64_t L = p->len_ff + 2;
    float t = (float)n + 0.5f + tfv - 1.0f;
    int ii = (int)t;
    float w = t - (float)ii;
    real base = (R(1.0) - R(w)) * K, wk = R(w) * K;
    Ua[kk * L + ii] = Ua[kk * L + ii] + base * dMa;
    Ua[kk * L + ii + 1] = Ua[kk * L + ii + 1] + wk * dMa;
    Ub[kk * L + ii] = Ub[kk * L + ii] + base * dMb;
    Ub[kk * L + ii + 1] = Ub[kk * L + ii + 1] + wk * dMb;
    t = (float)n + tfv - 1.0f;
    ii = (int)t;
    w = t - (float)ii;
    base = (R(1.0) - R(w)) * K; wk = R(w) * K;
    Wa[kk * L + ii] = Wa[kk * L + ii] + base * dJa;
    Wa[kk * L + ii + 1] = Wa[kk * L + ii + 1] + wk * dJa;
    Wb[kk * L + ii] = Wb[kk * L + ii] + base * dJb;
    Wb[kk * L + ii + 1] = Wb[kk * L + ii + 1] + wk * dJb;
}

/* bin_one for an image current: each term times its component's sign
 * (f90:7928-8240: (1-w) K dM sign) */
static void bin_one_m(oracle_state* s, int kk, int n, int i, int j, int k, real K, real* Ua, real* Ub, real* Wa,
                      real* Wb, real dMa, real dMb, real dJa, real dJb, real sMa, real sMb, real sJa, real sJb) {
    const aem_plan_view* p = &s->p;
    const float* a = p->ang + 6 * kk;
    const float s1 = a[2], c1 = a[3], s2 = a[4], c2 = a[5];
    float tfv = (p->r_relay - ((float)(i - p->ic) * p->dx * s1 * c2 + (float)(j - p->jc) * p->dy * s1 * s2 +
                               (float)(k - p->kc) * p->dz * c1)) / (299792448.0f * p->dt);
    const int64_t L = p->len_ff + 2;
    float t = (float)n + 0.5f + tfv - 1.0f;
    int ii = (int)t;
    float w = t - (float)ii;
    real base = (R(1.0) - R(w)) * K, wk = R(w) * K;
    Ua[kk * L + ii] = Ua[kk * L + ii] + base * dMa * sMa;
    Ua[kk * L + ii + 1] = Ua[kk * L + ii + 1] + wk * dMa * sMa;
    Ub[kk * L + ii] = Ub[kk * L + ii] + base * dMb * sMb;
    Ub[kk * L + ii + 1] = Ub[kk * L + ii + 1] + wk * dMb * sMb;
    t = (float)n + tfv - 1.0f;
    ii = (int)t;
    w = t - (float)ii;
    base = (R(1.0) - R(w)) * K; wk = R(w) * K;
    Wa[kk * L + ii] = Wa[kk * L + ii] + base * dJa * sJa;
    Wa[kk * L + ii + 1] = Wa[kk * L + ii + 1] + wk * dJa * sJa;
    Wb[kk * L + ii] = Wb[kk * L + ii] + base * dJb * sJb;
    Wb[kk * L + ii + 1] = Wb[kk * L + ii + 1] + wk * dJb * sJb;
}

/* image position of a face cell across the ground plane, in the
 * reference's FP32 form i (1 - use) + (offset - i) use (f90:7957-7959) */
static int mirror_pos(const aem_plan_view* p, int a, int i) {
    return (int)((float)i * (1.0f - p->use_m[a]) + (p->off_m[a] - (float)i) * p->use_m[a]);
}

/* the image currents of a ground plane, after every face of the box, in the
 * same loop order (f90:7928-8240) */
static void out_ntff_bins_mirror(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    const int fxl = p->ff_lo[0], fxh = p->ff_hi[0], fyl = p->ff_lo[1], fyh = p->ff_hi[1], fzl = p->ff_lo[2], fzh = p->ff_hi[2];
    real **Ux = &s->U[0], **Uy = &s->U[1], **Uz = &s->U[2], **Wx = &s->W[0], **Wy = &s->W[1], **Wz = &s->W[2];
    const real mx = R(p->sgn_M[0]), my = R(p->sgn_M[1]), mz = R(p->sgn_M[2]);
    const real jx = R(p->sgn_J[0]), jy = R(p->sgn_J[1]), jz = R(p->sgn_J[2]);
    if (p->pbc[0] + p->pbc[1] + p->pbc[2] >= 2) return;
    if (!p->pbc[0]) {
        const real K = R(p->ff_K[0]);
        for (int kk = 0; kk < p->n_angles; ++kk)
            for (int k = fzl; k <= fzh; ++k)
                for (int j = fyl; j <= fyh; ++j) {
                    const int jm = mirror_pos(p, 1, j), km = mirror_pos(p, 2, k);
                    bin_one_m(s, kk, n, mirror_pos(p, 0, fxl), jm, km, K, *Uy, *Uz, *Wy, *Wz,
                              FA(Mcur, 0, 0, j, k) - FA(Mold, 0, 0, j, k), FA(Mcur, 0, 1, j, k) - FA(Mold, 0, 1, j, k),
                              FA(Jcur, 0, 0, j, k) - FA(Jold, 0, 0, j, k), FA(Jcur, 0, 1, j, k) - FA(Jold, 0, 1, j, k),
                              my, mz, jy, jz);
                    bin_one_m(s, kk, n, mirror_pos(p, 0, fxh), jm, km, K, *Uy, *Uz, *Wy, *Wz,
                              FA(Mcur, 1, 0, j, k) - FA(Mold, 1, 0, j, k), FA(Mcur, 1, 1, j, k) - FA(Mold, 1, 1, j, k),
                              FA(Jcur, 1, 0, j, k) - FA(Jold, 1, 0, j, k), FA(Jcur, 1, 1, j, k) - FA(Jold, 1, 1, j, k),
                              my, mz, jy, jz);
                }
    }
    if (!p->pbc[1]) {
        const real K = R(p->ff_K[1]);
        for (int kk = 0; kk < p->n_angles; ++kk)
            for (int k = fzl; k <= fzh; ++k)
                for (int i = fxl; i <= fxh; ++i) {
                    const int im = mirror_pos(p, 0, i), km = mirror_pos(p, 2, k);
                    bin_one_m(s, kk, n, im, mirror_pos(p, 1, fyl), km, K, *Ux, *Uz, *Wx, *Wz,
                              FA(Mcur, 2, 0, i, k) - FA(Mold, 2, 0, i, k), FA(Mcur, 2, 1, i, k) - FA(Mold, 2, 1, i, k),
                              FA(Jcur, 2, 0, i, k) - FA(Jold, 2, 0, i, k), FA(Jcur, 2, 1, i, k) - FA(Jold, 2, 1, i, k),
                              mx, mz, jx, jz);
                    bin_one_m(s, kk, n, im, mirror_pos(p, 1, fyh), km, K, *Ux, *Uz, *Wx, *Wz,
                              FA(Mcur, 3, 0, i, k) - FA(Mold, 3, 0, i, k), FA(Mcur, 3, 1, i, k) - FA(Mold, 3, 1, i, k),
                              FA(Jcur, 3, 0, i, k) - FA(Jold, 3, 0, i, k), FA(Jcur, 3, 1, i, k) - FA(Jold, 3, 1, i, k),
                              mx, mz, jx, jz);
                }
    }
    if (!p->pbc[2]) {
        const real K = R(p->ff_K[2]);
        for (int kk = 0; kk < p->n_angles; ++kk)
            for (int j = fyl; j <= fyh; ++j)
                for (int i = fxl; i <= fxh; ++i) {
                    const int im = mirror_pos(p, 0, i), jm = mirror_pos(p, 1, j);
                    bin_one_m(s, kk, n, im, jm, mirror_pos(p, 2, fzl), K, *Uy, *Ux, *Wy, *Wx,
                              FA(Mcur, 4, 1, i, j) - FA(Mold, 4, 1, i, j), FA(Mcur, 4, 0, i, j) - FA(Mold, 4, 0, i, j),
                              FA(Jcur, 4, 1, i, j) - FA(Jold, 4, 1, i, j), FA(Jcur, 4, 0, i, j) - FA(Jold, 4, 0, i, j),
                              my, mx, jy, jx);
                    bin_one_m(s, kk, n, im, jm, mirror_pos(p, 2, fzh), K, *Uy, *Ux, *Wy, *Wx,
                              FA(Mcur, 5, 1, i, j) - FA(Mold, 5, 1, i, j), FA(Mcur, 5, 0, i, j) - FA(Mold, 5, 0, i, j),
                              FA(Jcur, 5, 1, i, j) - FA(Jold, 5, 1, i, j), FA(Jcur, 5, 0, i, j) - FA(Jold, 5, 0, i, j),
                              my, mx, jy, jx);
                }
    }
}

/* NTFF retarded-time binning, f90:7643-7927 (sequential, serial-build order) */
static void out_ntff_bins(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    const int fxl = p->ff_lo[0], fxh = p->ff_hi[0], fyl = p->ff_lo[1], fyh = p->ff_hi[1], fzl = p->ff_lo[2], fzh = p->ff_hi[2];
    real **Ux = &s->U[0], **Uy = &s->U[1], **Uz = &s->U[2], **Wx = &s->W[0], **Wy = &s->W[1], **Wz = &s->W[2];
    if (!p->pbc[0]) {
        const real K = R(p->ff_K[0]);
        for (int kk = 0; kk < p->n_angles; ++kk)
            for (int k = fzl; k <= fzh; ++k)
                for (int j = fyl; j <= fyh; ++j) {
                    bin_one(s, kk, n, fxl, j, k, K, *Uy, *Uz, *Wy, *Wz,
                            FA(Mcur, 0, 0, j, k) - FA(Mold, 0, 0, j, k), FA(Mcur, 0, 1, j, k) - FA(Mold, 0, 1, j, k),
                            FA(Jcur, 0, 0, j, k) - FA(Jold, 0, 0, j, k), FA(Jcur, 0, 1, j, k) - FA(Jold, 0, 1, j, k));
                    bin_one(s, kk, n, fxh, j, k, K, *Uy, *Uz, *Wy, *Wz,
                            FA(Mcur, 1, 0, j, k) - FA(Mold, 1, 0, j, k), FA(Mcur, 1, 1, j, k) - FA(Mold, 1, 1, j, k),
                            FA(Jcur, 1, 0, j, k) - FA(Jold, 1, 0, j, k), FA(Jcur, 1, 1, j, k) - FA(Jold, 1, 1, j, k));
                }
    }
    if (!p->pbc[1]) {
        const real K = R(p->ff_K[1]);
        for (int kk = 0; kk < p->n_angles; ++kk)
            for (int k = fzl; k <= fzh; ++k)
                for (int i = fxl; i <= fxh; ++i) {
                    bin_one(s, kk, n, i, fyl, k, K, *Ux, *Uz, *Wx, *Wz,
                            FA(Mcur, 2, 0, i, k) - FA(Mold, 2, 0, i, k), FA(Mcur, 2, 1, i, k) - FA(Mold, 2, 1, i, k),
                            FA(Jcur, 2, 0, i, k) - FA(Jold, 2, 0, i, k), FA(Jcur, 2, 1, i, k) - FA(Jold, 2, 1, i, k));
                    bin_one(s, kk, n, i, fyh, k, K, *Ux, *Uz, *Wx, *Wz,
                            FA(Mcur, 3, 0, i, k) - FA(Mold, 3, 0, i, k), FA(Mcur, 3, 1, i, k) - FA(Mold, 3, 1, i, k),
                            FA(Jcur, 3, 0, i, k) - FA(Jold, 3, 0, i, k), FA(Jcur, 3, 1, i, k) - FA(Jold, 3, 1, i, k));
                }
    }
    if (!p->pbc[2]) {
        const real K = R(p->ff_K[2]);
        for (int kk = 0; kk < p->n_angles; ++kk)
            for (int j = fyl; j <= fyh; ++j)
                for (int i = fxl; i <= fxh; ++i) {
                    /* z faces accumulate y before x (f90:7861-7918): slot 0 of M is x, slot 1 is y */
                    bin_one(s, kk, n, i, j, fzl, K, *Uy, *Ux, *Wy, *Wx,
                            FA(Mcur, 4, 1, i, j) - FA(Mold, 4, 1, i, j), FA(Mcur, 4, 0, i, j) - FA(Mold, 4, 0, i, j),
                            FA(Jcur, 4, 1, i, j) - FA(Jold, 4, 1, i, j), FA(Jcur, 4, 0, i, j) - FA(Jold, 4, 0, i, j));
                    bin_one(s, kk, n, i, j, fzh, K, *Uy, *Ux, *Wy, *Wx,
                            FA(Mcur, 5, 1, i, j) - FA(Mold, 5, 1, i, j), FA(Mcur, 5, 0, i, j) - FA(Mold, 5, 0, i, j),
                            FA(Jcur, 5, 1, i, j) - FA(Jold, 5, 1, i, j), FA(Jcur, 5, 0, i, j) - FA(Jold, 5, 0, i, j));
                }
    }
}

/* phase-centre samples for plane waves, f90:8263-8290 */
static void out_phase_centre(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    if (!(p->n_angles > 0 && p->plane_wave_amp > 0)) return;
    const int i = p->ic, j = p->jc, k = p->kc, T = p->time_steps;
    s->pc[0][n - 1] = (EX(i, j, k) + EX(i, j + 1, k) + EX(i, j, k + 1) + EX(i, j + 1, k + 1)) / R(4.0);
    s->pc[1][n - 1] = (EY(i, j, k) + EY(i + 1, j, k) + EY(i, j, k + 1) + EY(i + 1, j, k + 1)) / R(4.0);
    s->pc[2][n - 1] = (EZ(i, j, k) + EZ(i + 1, j, k) + EZ(i, j + 1, k) + EZ(i + 1, j + 1, k)) / R(4.0);
    s->pc[3][n - 1] = (HX(i, j, k) + HX(i + 1, j, k)) / R(4.0);
    s->pc[4][n - 1] = (HY(i, j, k) + HY(i, j + 1, k)) / R(4.0);
    s->pc[5][n - 1] = (HZ(i, j, k) + HZ(i, j, k + 1)) / R(4.0);
    if (n < T) {
        s->pc[3][n] = s->pc[3][n - 1];
        s->pc[4][n] = s->pc[4][n - 1];
        s->pc[5][n] = s->pc[5][n - 1];
    }
}

/* ---- kmax build -------------------------------------------------------------- */

/* Inc() of the kmax build, f90:8860-8885: w exp(-g^2/2) exp(i (2 pi f_adj
 * (t + x d/c) - (k1 p1 + k2 p2))), g = (t_spread - t + x d / c) / spread,
 * t = cf dt; the complex value as (*re, *im). */
static void inc_kmax(const aem_plan_view* p, real w, real xf, real k1, real p1, real k2, real p2, real cf, real dn,
                     real* re, real* im) {
    const real g = (R(p->t_spread) - (cf * R(p->dt)) + R(p->inv_c) * (xf * dn)) / R(p->spread);
    const real a1 = w * EXP(R(-0.5) * (g * g));
    const real th = R(2.0) * R(PI_F) * R(p->f_adj) * ((cf * R(p->dt)) + R(p->inv_c) * (xf * dn)) - (k1 * p1 + k2 * p2);
    *re = a1 * COS(th);
    *im = a1 * SIN(th);
}

/* F += c0 Inc exc on both parts of a kmax pair */
#define KADD(FLD, o, c0, ir, ii, exc)                                        \
    do {                                                                     \
        s->FLD[o] = s->FLD[o] + (c0) * (ir) * (exc);                         \
        m->FLD[o] = m->FLD[o] + (c0) * (ii) * (exc);                         \
    } while (0)

/* TE source on the H plane, f90:4558-4625 (time n-1, x = 0) */
static void kmax_src_h(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    oracle_state* m = s->kim;
    if (!(p->mode_type == 0 && p->plane_wave_amp > 0)) return;
    const real cf = R(n) - R(1.0), x0 = R(0.0);
    const real kx = R(p->k_vec[0]), ky = R(p->k_vec[1]), kz = R(p->k_vec[2]);
    const real ex = R(p->k_exc[0]), ey = R(p->k_exc[1]), ez = R(p->k_exc[2]);
    const real dx = R(p->dx), dy = R(p->dy), dz = R(p->dz);
    const int kp = p->kpl_H;
    real ir, ii;
    if (p->pbc[1] + p->pbc[2] == 2) {
        const real c0 = R(p->dt) / (R(MU0_F) * dx);
        for (int k = p->tf_lo[2]; k <= p->tf_hi[2]; ++k)
            for (int j = p->tf_lo[1]; j <= p->tf_hi[1]; ++j) {
                const int64_t o = IDX(s, kp, j, k);
                inc_kmax(p, R(p->WE[1]), x0, ky, (R(j) + R(0.5)) * dy, kz, R(k) * dz, cf, dx, &ir, &ii);
                KADD(Hz, o, c0, ir, ii, ez);
                inc_kmax(p, R(p->WE[2]), x0, ky, R(j) * dy, kz, (R(k) + R(0.5)) * dz, cf, dx, &ir, &ii);
                KADD(Hy, o, c0, ir, ii, ey);
            }
    }
    if (p->pbc[0] + p->pbc[2] == 2) {
        const real c0 = R(p->dt) / (R(MU0_F) * dy);
        for (int k = p->tf_lo[2]; k <= p->tf_hi[2]; ++k)
            for (int i = p->tf_lo[0]; i <= p->tf_hi[0]; ++i) {
                const int64_t o = IDX(s, i, kp, k);
                inc_kmax(p, R(p->WE[0]), x0, kx, (R(i) + R(0.5)) * dx, kz, R(k) * dz, cf, dy, &ir, &ii);
                KADD(Hz, o, c0, ir, ii, ez);
                inc_kmax(p, R(p->WE[2]), x0, kx, R(i) * dx, kz, (R(k) + R(0.5)) * dz, cf, dy, &ir, &ii);
                KADD(Hx, o, c0, ir, ii, ex);
            }
    }
    if (p->pbc[1] + p->pbc[0] == 2) {
        const real c0 = R(p->dt) / (R(MU0_F) * dz);
        for (int j = p->tf_lo[1]; j <= p->tf_hi[1]; ++j)
            for (int i = p->tf_lo[0]; i <= p->tf_hi[0]; ++i) {
                const int64_t o = IDX(s, i, j, kp);
                inc_kmax(p, R(p->WE[1]), x0, ky, (R(j) + R(0.5)) * dy, kx, R(i) * dx, cf, dz, &ir, &ii);
                KADD(Hx, o, c0, ir, ii, ex);
                inc_kmax(p, R(p->WE[0]), x0, ky, R(j) * dy, kx, (R(i) + R(0.5)) * dx, cf, dz, &ir, &ii);
                KADD(Hy, o, c0, ir, ii, ey);
            }
    }
}

/* TM source on the E plane, f90:5857-5948 (time n-1/2, x = -1/2) */
static void kmax_src_e(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    oracle_state* m = s->kim;
    if (!(p->mode_type == 1 && p->plane_wave_amp > 0)) return;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    const real cf = R(n) - R(0.5), xh = R(-0.5);
    const real kx = R(p->k_vec[0]), ky = R(p->k_vec[1]), kz = R(p->k_vec[2]);
    const real ex = R(p->k_exc[0]), ey = R(p->k_exc[1]), ez = R(p->k_exc[2]);
    const real dx = R(p->dx), dy = R(p->dy), dz = R(p->dz);
    const int kp = p->kpl_E;
    real ir, ii;
    if (p->pbc[1] + p->pbc[2] == 2) {
        const real c0 = R(p->dt) / (R(EP0_F) * dx);
        for (int k = 1; k <= NZ - 1; ++k)
            for (int j = 2; j <= NY - 1; ++j) {
                inc_kmax(p, R(p->WH[1]), xh, ky, R(j) * dy, kz, (R(k) + R(0.5)) * dz, cf, dx, &ir, &ii);
                KADD(Ez, IDX(s, kp, j, k), c0, ir, ii, ez);
            }
        for (int k = 2; k <= NZ - 1; ++k)
            for (int j = 1; j <= NY - 1; ++j) {
                inc_kmax(p, R(p->WH[2]), xh, ky, (R(j) + R(0.5)) * dy, kz, R(k) * dz, cf, dx, &ir, &ii);
                KADD(Ey, IDX(s, kp, j, k), c0, ir, ii, ey);
            }
    }
    if (p->pbc[0] + p->pbc[2] == 2) {
        const real c0 = R(p->dt) / (R(EP0_F) * dy);
        for (int k = 1; k <= NZ - 1; ++k)
            for (int i = 2; i <= NX - 1; ++i) {
                inc_kmax(p, R(p->WH[0]), xh, kx, R(i) * dx, kz, (R(k) + R(0.5)) * dz, cf, dy, &ir, &ii);
                KADD(Ez, IDX(s, i, kp, k), c0, ir, ii, ez);
            }
        for (int k = 2; k <= NZ - 1; ++k)
            for (int i = 1; i <= NX - 1; ++i) {
                inc_kmax(p, R(p->WH[2]), xh, kx, (R(i) + R(0.5)) * dx, kz, R(k) * dz, cf, dy, &ir, &ii);
                KADD(Ex, IDX(s, i, kp, k), c0, ir, ii, ex);
            }
    }
    if (p->pbc[1] + p->pbc[0] == 2) {
        const real c0 = R(p->dt) / (R(EP0_F) * dz);
        for (int j = 2; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i) {
                inc_kmax(p, R(p->WH[1]), xh, ky, R(j) * dy, kx, (R(i) + R(0.5)) * dx, cf, dz, &ir, &ii);
                KADD(Ex, IDX(s, i, j, kp), c0, ir, ii, ex);
            }
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 2; i <= NX - 1; ++i) {
                inc_kmax(p, R(p->WH[0]), xh, ky, (R(j) + R(0.5)) * dy, kx, R(i) * dx, cf, dz, &ir, &ii);
                KADD(Ey, IDX(s, i, j, kp), c0, ir, ii, ey);
            }
    }
}

/* dst = src * (c + i sn) on both parts */
#define BLOCH(FLD, od, os, c, sn)                                            \
    do {                                                                     \
        const real a_ = s->FLD[os], b_ = m->FLD[os];                         \
        s->FLD[od] = a_ * (c) - b_ * (sn);                                   \
        m->FLD[od] = a_ * (sn) + b_ * (c);                                   \
    } while (0)

/* Bloch copies of H, f90:5387-5428: the last cell from the first times
 * exp(-i k (N-2) d) */
static void kmax_h_pbc(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    oracle_state* m = s->kim;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    if (p->pbc[0]) {
        const real th = R(p->k_vec[0]) * R(NX - 2) * R(p->dx), c = COS(-th), sn = SIN(-th);
        for (int k = 1; k <= NZ - 1; ++k)
            for (int j = 1; j <= NY - 1; ++j) {
                BLOCH(Hz, IDX(s, NX - 1, j, k), IDX(s, 1, j, k), c, sn);
                BLOCH(Hy, IDX(s, NX - 1, j, k), IDX(s, 1, j, k), c, sn);
            }
    }
    if (p->pbc[1]) {
        const real th = R(p->k_vec[1]) * R(NY - 2) * R(p->dy), c = COS(-th), sn = SIN(-th);
        for (int k = 1; k <= NZ - 1; ++k)
            for (int i = 1; i <= NX - 1; ++i) {
                BLOCH(Hz, IDX(s, i, NY - 1, k), IDX(s, i, 1, k), c, sn);
                BLOCH(Hx, IDX(s, i, NY - 1, k), IDX(s, i, 1, k), c, sn);
            }
    }
    if (p->pbc[2]) {
        const real th = R(p->k_vec[2]) * R(NZ - 2) * R(p->dz), c = COS(-th), sn = SIN(-th);
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i) {
                BLOCH(Hy, IDX(s, i, j, NZ - 1), IDX(s, i, j, 1), c, sn);
                BLOCH(Hx, IDX(s, i, j, NZ - 1), IDX(s, i, j, 1), c, sn);
            }
    }
}

/* Bloch copies of E, f90:6685-6736: the first cell from the last times
 * exp(+i k (N-2) d).  (The J_source copies of f90:6742-6806 write entries
 * that are never read, as in the ordinary build.) */
static void kmax_e_pbc(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    oracle_state* m = s->kim;
    const int NX = p->NX, NY = p->NY, NZ = p->NZ;
    if (p->pbc[0]) {
        const real th = R(p->k_vec[0]) * R(NX - 2) * R(p->dx), c = COS(th), sn = SIN(th);
        for (int k = 1; k <= NZ - 1; ++k)
            for (int j = 1; j <= NY - 1; ++j) {
                BLOCH(Ez, IDX(s, 1, j, k), IDX(s, NX - 1, j, k), c, sn);
                BLOCH(Ey, IDX(s, 1, j, k), IDX(s, NX - 1, j, k), c, sn);
            }
    }
    if (p->pbc[1]) {
        const real th = R(p->k_vec[1]) * R(NY - 2) * R(p->dy), c = COS(th), sn = SIN(th);
        for (int k = 1; k <= NZ - 1; ++k)
            for (int i = 1; i <= NX - 1; ++i) {
                BLOCH(Ez, IDX(s, i, 1, k), IDX(s, i, NY - 1, k), c, sn);
                BLOCH(Ex, IDX(s, i, 1, k), IDX(s, i, NY - 1, k), c, sn);
            }
    }
    if (p->pbc[2]) {
        const real th = R(p->k_vec[2]) * R(NZ - 2) * R(p->dz), c = COS(th), sn = SIN(th);
        for (int j = 1; j <= NY - 1; ++j)
            for (int i = 1; i <= NX - 1; ++i) {
                BLOCH(Ey, IDX(s, i, j, 1), IDX(s, i, j, NZ - 1), c, sn);
                BLOCH(Ex, IDX(s, i, j, 1), IDX(s, i, j, NZ - 1), c, sn);
            }
    }
}

/* TE/TM plane sums of one part, f90:7105-7190 */
static void kmax_plane_sums(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    const real ex = R(p->k_exc[0]), ey = R(p->k_exc[1]), ez = R(p->k_exc[2]);
    real* rte = &s->refl[n - 1];
    real* rtm = &s->refl2[n - 1];
    real* tte = &s->trans[n - 1];
    real* ttm = &s->trans2[n - 1];
    if (p->pbc[1] + p->pbc[2] == 2) {
        const int xl = p->tf_lo[0], xh = p->tf_hi[0];
        for (int k = p->ff_lo[2]; k <= p->ff_hi[2]; ++k)
            for (int j = p->ff_lo[1]; j <= p->ff_hi[1]; ++j) {
                *rte = *rte + (EZ(xl, j, k) * ey - EY(xl, j, k) * ez);
                *rtm = *rtm + (EZ(xl, j, k) * ez + EY(xl, j, k) * ey);
                *tte = *tte + (EZ(xh, j, k) * ey - EY(xh, j, k) * ez);
                *ttm = *ttm + (EZ(xh, j, k) * ez + EY(xh, j, k) * ey);
            }
    }
    if (p->pbc[0] + p->pbc[2] == 2) {
        const int yl = p->tf_lo[1], yh = p->tf_hi[1];
        for (int k = p->ff_lo[2]; k <= p->ff_hi[2]; ++k)
            for (int i = p->ff_lo[0]; i <= p->ff_hi[0]; ++i) {
                *rte = *rte + (EZ(i, yl, k) * ex - EX(i, yl, k) * ez);
                *rtm = *rtm + (EZ(i, yl, k) * ez + EX(i, yl, k) * ex);
                *tte = *tte + (EZ(i, yh, k) * ex - EX(i, yh, k) * ez);
                *ttm = *ttm + (EZ(i, yh, k) * ez + EX(i, yh, k) * ex);
            }
    }
    if (p->pbc[1] + p->pbc[0] == 2) {
        const int zl = p->tf_lo[2], zh = p->tf_hi[2];
        for (int j = p->ff_lo[1]; j <= p->ff_hi[1]; ++j)
            for (int i = p->ff_lo[0]; i <= p->ff_hi[0]; ++i) {
                *rte = *rte + (EX(i, j, zl) * ey - EY(i, j, zl) * ex);
                *rtm = *rtm + (EX(i, j, zl) * ex + EY(i, j, zl) * ey);
                *tte = *tte + (EX(i, j, zh) * ey - EY(i, j, zh) * ex);
                *ttm = *ttm + (EX(i, j, zh) * ex + EY(i, j, zh) * ey);
            }
    }
}

/* One kmax step of the pair (s = real part): each part's ordinary updates,
 * the modal source after the sheets/Drude updates, the Bloch copies in
 * place of the periodic copies. */
static void kmax_step(oracle_state* s, int n) {
    oracle_state* m = s->kim;
    oracle_state* part[2] = {s, m};
    for (int q = 0; q < 2; ++q) { h_bulk(part[q]); h_sheets(part[q]); }
    kmax_src_h(s, n);
    for (int q = 0; q < 2; ++q) { h_ports(part[q], n); h_cpml(part[q]); }
    kmax_h_pbc(s);
    for (int q = 0; q < 2; ++q) { e_bulk(part[q]); e_sheets(part[q]); e_drude(part[q]); }
    kmax_src_e(s, n);
    for (int q = 0; q < 2; ++q) { e_ports(part[q], n); e_cpml(part[q]); }
    kmax_e_pbc(s);
    for (int q = 0; q < 2; ++q) {
        oracle_state* t = part[q];
        const aem_plan_view* p = &t->p;
        e_oldt(t);
        kmax_plane_sums(t, n);
        for (int pi = 0; pi < p->n_ports; ++pi)
            t->Vout[(int64_t)pi * p->time_steps + (n - 1)] = t->V[(int64_t)pi * p->M + 1];
        t->n = n;
    }
}

/* The step splits at the SPICE exchange (f90:5435-5665): the H phase and
 * the gap currents, then the E phase with the circuit's gap voltages and
 * the per-step outputs. */
static void step_to_exchange(oracle_state* s, int n) {
    h_bulk(s);
    h_sheets(s);
    h_tfsf(s, n);
    h_ports(s, n);
    h_cpml(s);
    h_pbc(s);
    h_spice(s, n);
}

static void step_from_exchange(oracle_state* s, int n) {
    const aem_plan_view* p = &s->p;
    e_bulk(s);
    e_sheets(s);
    e_drude(s);
    e_tfsf(s, n);
    e_ports(s, n);
    e_spice(s, n);
    e_cpml(s);
    e_pbc(s);
    e_oldt(s);
    out_video(s, n);
    out_plane_sums(s, n);
    if (p->n_angles > 0) {
        out_ntff_currents(s);
        out_ntff_bins(s, n);
        if (p->mirror_type) out_ntff_bins_mirror(s, n);
    }
    for (int pi = 0; pi < p->n_ports; ++pi) s->Vout[(int64_t)pi * p->time_steps + (n - 1)] = s->V[(int64_t)pi * p->M + 1];
    out_phase_centre(s, n);
}

int oracle_step(oracle_state* s, int nsteps) {
    const aem_plan_view* p = &s->p;
    if (p->n_spice > 0) return -1;   /* stepped by oracle_port_exchange */
    for (int q = 0; q < nsteps; ++q) {
        if (s->n >= p->time_steps) return -1;
        const int n = ++s->n;
        if (s->kim) {
            kmax_step(s, n);
            continue;
        }
        step_to_exchange(s, n);
        step_from_exchange(s, n);
    }
    return 0;
}

int oracle_port_exchange(oracle_state* s, const float* v_in, float* i_out) {
    const aem_plan_view* p = &s->p;
    if (p->n_spice == 0) return -1;
    if (s->x_pending) {
        if (!v_in) return -1;
        for (int q = 0; q < p->n_spice; ++q) s->spV[q] = R(v_in[q]) - R(p->spice[q].v_bias);
        for (int q = 0; q < p->n_spice_extra; ++q) s->spV[p->n_spice + q] = R(v_in[p->n_spice + q]);
        step_from_exchange(s, s->n);
        s->x_pending = 0;
    }
    if (i_out) {
        if (s->n >= p->time_steps) return -1;
        const int n = ++s->n;
        step_to_exchange(s, n);
        for (int q = 0; q < p->n_spice; ++q) i_out[q] = (float)s->spI[q];
        s->x_pending = 1;
    }
    return 0;
}

int oracle_steps_done(const oracle_state* s) { return s->n; }

/* ---- post loop, f90:8340-8534 ------------------------------------------- */

int oracle_finalize(oracle_state* s) {
    const aem_plan_view* p = &s->p;
    const int T = p->time_steps;
    if (s->finalized) return 0;
    s->finalized = 1;
    if (s->kim) oracle_finalize(s->kim);
    /* normalisation of the plane sums (f90:8342-8361; kmax TE and TM, f90:8411-8435) */
    if (p->plane_sum_axis >= 0) {
        real den;
        if (p->plane_sum_axis == 0) den = (R(p->NY) - R(2.0)) * (R(p->NZ) - R(2.0));
        else if (p->plane_sum_axis == 1) den = (R(p->NX) - R(2.0)) * (R(p->NZ) - R(2.0));
        else den = (R(p->NY) - R(2.0)) * (R(p->NX) - R(2.0));
        for (int n = 0; n < T; ++n) { s->refl[n] = DIVR(s, s->refl[n], den); s->trans[n] = DIVR(s, s->trans[n], den); }
        if (p->kmax)
            for (int n = 0; n < T; ++n) { s->refl2[n] = cdivr(s->refl2[n], den); s->trans2[n] = cdivr(s->trans2[n], den); }
    }
    /* incident wave (f90:8366-8408) */
    if (p->plane_wave_amp > 0)
        for (int n = 1; n <= T; ++n) s->incident[n - 1] = e_or_v(s, R(p->plane_wave_amp), R(n) + R(0.0), R(0.0) * R(p->dx));
    if (p->antenna_amp > 0 && p->excitation_port > 0) {
        const aem_port_view* pt = &p->ports[p->excitation_port - 1];
        const int M = p->M, M2 = 2 * M;
        real* Vi = (real*)calloc((size_t)M2 + 1, sizeof(real));
        real* Ci = (real*)calloc((size_t)M2 + 1, sizeof(real)); /* Ci[0] stands for C_inc(0) */
        const real kI = R(pt->kI), kV = R(pt->kV), kS = R(pt->kS);
        for (int n = 1; n <= T; ++n) {
            for (int j = 1; j <= M2 - 1; ++j) Ci[j] = Ci[j] - kI * (Vi[j + 1] - Vi[j]);
            Ci[M] = Ci[M] - kI * e_or_v(s, R(p->antenna_amp), R(n) + R(0.0), R(0.0) * R(pt->dl));
            for (int j = 1; j <= M2 - 1; ++j) Vi[j] = Vi[j] - kV * (Ci[j] - Ci[j - 1]);
            Vi[M] = Vi[M] + kS * e_or_v(s, R(p->antenna_amp), R(n) + R(0.5), R(0.5) * R(pt->dl));
            s->incident[n - 1] = Vi[M];
        }
        free(Vi);
        free(Ci);
    }
    /* NTFF combine and output window (f90:8483-8534) */
    if (p->n_angles > 0) {
        const int L = p->len_ff;
        const int64_t LL = L + 2;
        const float eta = sqrtf(1.25663706E-6f / 8.85418782E-12f);
        real* Et = (real*)calloc((size_t)L + 1, sizeof(real));
        real* Ep = (real*)calloc((size_t)L + 1, sizeof(real));
        for (int kk = 0; kk < p->n_angles; ++kk) {
            const float* a = p->ang + 6 * kk;
            const real s1 = R(a[2]), c1 = R(a[3]), s2 = R(a[4]), c2 = R(a[5]);
            for (int j = 1; j <= L; ++j) {
                real wx = s->W[0][kk * LL + j], wy = s->W[1][kk * LL + j], wz = s->W[2][kk * LL + j];
                real ux = s->U[0][kk * LL + j], uy = s->U[1][kk * LL + j], uz = s->U[2][kk * LL + j];
                real wt = wx * c1 * c2 + wy * c1 * s2 - wz * s1;
                real wph = R(-1.0) * wx * s2 + wy * c2;
                real ut = ux * c1 * c2 + uy * c1 * s2 - uz * s1;
                real uph = R(-1.0) * ux * s2 + uy * c2;
                Et[j] = R(-1.0) * R(eta) * wt - uph;
                Ep[j] = R(-1.0) * R(eta) * wph + ut;
            }
            int dphi = L, dth = L;
            for (int j = 1; j <= L; ++j) if (Ep[j] != R(0.0)) { dphi = j; break; }
            for (int j = 1; j <= L; ++j) if (Et[j] != R(0.0)) { dth = j; break; }
            int dot = dphi < dth ? dphi : dth;
            if (dot > L - T + 1) dot = L - T + 1;
            for (int j = 1; j <= T; ++j) {
                s->Eth[(int64_t)kk * T + j - 1] = Et[j + dot - 1];
                s->Eph[(int64_t)kk * T + j - 1] = Ep[j + dot - 1];
            }
            s->fftime[kk] = (R(p->r_relay) / R(299792448.0f)) - R(dot - 1) * R(p->dt);
        }
        free(Et);
        free(Ep);
    }
    return 0;
}

/* ---- accessors -------------------------------------------------------------- */

/* the part of a kmax pair a component or series kind OR-ed with 0x100 selects */
#define ORACLE_IMAG 0x100
static const oracle_state* part_of(const oracle_state* s, int* sel) {
    if (!(*sel & ORACLE_IMAG)) return s;
    *sel &= ~ORACLE_IMAG;
    return s->kim;
}

int oracle_get_field(const oracle_state* s, int comp, float* out, int64_t n) {
    s = part_of(s, &comp);
    if (!s || comp < 0 || comp > 5 || n > s->nc) return -1;
    for (int64_t q = 0; q < n; ++q) out[q] = (float)s->F[comp][q];
    return 0;
}

int oracle_get_field64(const oracle_state* s, int comp, double* out, int64_t n) {
    s = part_of(s, &comp);
    if (!s || comp < 0 || comp > 5 || n > s->nc) return -1;
    for (int64_t q = 0; q < n; ++q) out[q] = (double)s->F[comp][q];
    return 0;
}

int oracle_set_field(oracle_state* s, int comp, const float* in, int64_t n) {
    if (comp & ORACLE_IMAG) {
        s = s->kim;
        comp &= ~ORACLE_IMAG;
    }
    if (!s || comp < 0 || comp > 5 || n > s->nc) return -1;
    for (int64_t q = 0; q < n; ++q) s->F[comp][q] = (real)in[q];
    return 0;
}

int oracle_get_series(const oracle_state* s, int kind, int idx, float* out, int64_t n) {
    s = part_of(s, &kind);
    if (!s) return -1;
    const aem_plan_view* p = &s->p;
    const int T = p->time_steps;
    if (n > T) return -1;
    const real* src = NULL;
    switch (kind) {
        case 0: if (idx < 0 || idx >= p->n_ports) return -1; src = s->Vout + (int64_t)idx * T; break;
        case 1: if (idx < 0 || idx >= p->n_ports) return -1; src = s->Iout + (int64_t)idx * T; break;
        case 2: src = s->incident; break;
        case 3: if (idx < 0 || idx >= p->n_angles) return -1; src = s->Eth + (int64_t)idx * T; break;
        case 4: if (idx < 0 || idx >= p->n_angles) return -1; src = s->Eph + (int64_t)idx * T; break;
        case 5: src = s->refl; break;
        case 6: src = s->trans; break;
        case 7: {
            const real c2 = R(p->ps_c2), c1 = R(p->ps_c1);
            const real c3 = R(p->s_pol) * R(p->s_th);
            for (int64_t q = 0; q < n; ++q) out[q] = (float)(s->pc[0][q] * c2 + s->pc[1][q] * c1 + s->pc[2][q] * c3);
            return 0;
        }
        case 8: case 9: case 10: case 11: case 12: case 13:
            for (int64_t q = 0; q < n; ++q) out[q] = (float)s->pc[kind - 8][q];
            return 0;
        case 20: if (idx < 0 || idx >= p->n_spice + p->n_spice_extra) return -1; src = s->spVout + (int64_t)idx * T; break;
        case 21: if (idx < 0 || idx >= p->n_spice) return -1; src = s->spIout + (int64_t)idx * T; break;
        case 22: if (!p->kmax) return -1; src = s->refl2; break;
        case 23: if (!p->kmax) return -1; src = s->trans2; break;
        default: return -1;
    }
    for (int64_t q = 0; q < n; ++q) out[q] = (float)src[q];
    return 0;
}

int oracle_get_series64(const oracle_state* s, int kind, int idx, double* out, int64_t n) {
    const aem_plan_view* p = &s->p;
    const int T = p->time_steps;
    if (n > T) return -1;
    const real* src = NULL;
    switch (kind) {
        case 0: if (idx < 0 || idx >= p->n_ports) return -1; src = s->Vout + (int64_t)idx * T; break;
        case 1: if (idx < 0 || idx >= p->n_ports) return -1; src = s->Iout + (int64_t)idx * T; break;
        case 2: src = s->incident; break;
        default: return -1;
    }
    for (int64_t q = 0; q < n; ++q) out[q] = (double)src[q];
    return 0;
}

int oracle_get_ntff_raw(const oracle_state* s, int which, double* out, int64_t n) {
    const aem_plan_view* p = &s->p;
    if (which < 0 || which > 5 || p->n_angles <= 0) return -1;
    const real* src = which < 3 ? s->U[which] : s->W[which - 3];
    const int64_t L = p->len_ff, LL = L + 2;
    if (n > (int64_t)p->n_angles * L) return -1;
    for (int64_t q = 0; q < n; ++q) {
        int64_t a = q / L, b = q % L;
        out[q] = (double)src[a * LL + b + 1];
    }
    return 0;
}

int oracle_get_ff_time(const oracle_state* s, float* out, int64_t n) {
    if (n > s->p.n_angles) return -1;
    for (int64_t q = 0; q < n; ++q) out[q] = (float)s->fftime[q];
    return 0;
}

int oracle_get_video(const oracle_state* s, int comp, int step, float* out, int64_t n) {
    const aem_plan_view* p = &s->p;
    if (!p->video_on || comp < 0 || comp > 5 || step < 1 || step > p->time_steps) return -1;
    int64_t ps = (int64_t)p->vid_size1 * p->vid_size2;
    if (n > ps) return -1;
    for (int64_t q = 0; q < n; ++q) out[q] = (float)s->vid[comp][(int64_t)(step - 1) * ps + q];
    return 0;
}

int oracle_real_bytes(void) { return (int)sizeof(real); }
