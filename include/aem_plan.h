/*
 * aem_plan.h -- read-only view of a prepared FDTD simulation ("plan").
 *
 * A plan is everything the leapfrog loop of the reference solver consumes
 * after its setup phase (main_fdtd/fdtd_solver.f90:1300-4369): grid sizes,
 * the time step, per-edge update coefficients, CPML profiles, thin-sheet
 * planes, Drude pole coefficients, TF/SF plane-wave parameters, lumped
 * ports and near-to-far-field (NTFF) bookkeeping.  The plan is built on the
 * host by aem_plan_from_inputs() (include/aem.h) and is consumed both by the
 * HIP device path and by CPU tools that need the exact same inputs.
 *
 * Index conventions
 *   All integer positions use the reference's 1-based Fortran indices.
 *   NX,NY,NZ are the reference's x_size,y_size,z_size after
 *   "x_size = x_size+1+pbc_x" (fdtd_solver.f90:1316-1318).  Every 3-D array
 *   (fields, material index) uses one common layout:
 *       offset(i,j,k) = (i-1) + PX*((j-1) + NY*(k-1)),  1<=i<=NX, ...
 *   with an x pitch PX >= NX (x fastest, as the reference's column-major
 *   arrays).  Field components that are shorter than NX/NY/NZ in the
 *   reference (e.g. Ex is (NX-1,NY,NZ), fdtd_solver.f90:3908) simply never
 *   touch the unused entries.
 *   1-D tables (den_*, CPML b/c) are stored 0-based: den_hx[i-1] is
 *   den_hx(i) of the reference.
 */
#ifndef AEM_PLAN_H
#define AEM_PLAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One basic lumped port: fdtd_solver.f90:1123-1157 (input), 4800-5013 (H
 * phase), 6120-6223 (E phase). */
typedef struct aem_port_view {
    int32_t dir;          /* E-field direction: 0=x 1=y 2=z  (ports(i,1,1)) */
    int32_t loc[3];       /* first cube, 1-based (ports(i,3,:)) */
    int32_t size[3];      /* ports(i,4,:) after the +1 on the two transverse axes */
    float   L;            /* ports(i,2,2) = Z/c   (per-unit-length inductance of the 1-D line) */
    float   C;            /* ports(i,2,3) = 1/(Z c) */
    float   Z;            /* port impedance as read */
    float   dl;           /* step_vc_out: cell size along dir */
    float   kI;           /* (1/L)*(dt/dl): current update factor  (f90:4977-4978) */
    float   kV;           /* (1/C)*(dt/dl): voltage update factor  (f90:6138-6139) */
    float   kS;           /* (1/sqrt(C*L))*(dt/dl): voltage source factor (f90:6162) */
    float   cur_num;      /* real(n_a*n_b): loop integral I = (I*cur_num)/cur_den (f90:4831/4844/4857) */
    float   cur_den;      /* real(n_dir) */
    float   gap_div;      /* ports(i,4,dir)*d_dir: gap E = V(2)/gap_div (f90:6197/6207/6217) */
    int32_t excited;      /* 1 if this is excitation_port_number */
} aem_port_view;

/* One basic SPICE port: a lumped gap whose voltage a host circuit sets
 * every step from the gap current (fdtd_solver.f90:1195-1286 input,
 * 5435-5665 current and exchange, 6285-6520 gap field). */
typedef struct aem_spice_view {
    int32_t dir;                  /* gap field component 0..2 */
    int32_t loc[3], size[3];      /* first cell; cells per axis (+1 on the transverse axes) */
    float   v_bias;               /* DC bias subtracted from the node voltage */
    float   cap;                  /* grid capacitance the reference assigns to the port's capacitor */
    float   cur_num, cur_den;     /* current scale (I*cur_num)/cur_den */
    float   gap_div;              /* gap E = V / gap_div */
} aem_spice_view;

typedef struct aem_plan_view {
    /* ---- grid ---------------------------------------------------------- */
    int32_t nx, ny, nz;       /* user cell counts as read from the deck */
    int32_t NX, NY, NZ;       /* x_size,y_size,z_size after +1+pbc */
    int32_t PX;               /* x pitch of every 3-D array, PX >= NX */
    int32_t pbc[3];           /* periodic x,y,z */
    int64_t plane;            /* PX*NY */
    int64_t ncell;            /* PX*NY*NZ */
    float   dx, dy, dz;
    float   dt;               /* del_t (f90:1312) */
    float   DB;               /* dt/mu_0 (f90:2795); DA == 1 */
    int32_t time_steps;

    /* ---- pulse ----------------------------------------------------------- */
    int32_t pulse_type;       /* 1 Gaussian, 2 differentiated Gaussian */
    float   spread, t_spread; /* f90:1711-1718 */
    float   f_center;

    /* ---- 1-D grading tables (0-based) ---------------------------------- */
    const float *den_ex, *den_ey, *den_ez;  /* lengths NX-1, NY-1, NZ-1 */
    const float *den_hx, *den_hy, *den_hz;

    /* ---- bulk E coefficients ------------------------------------------- */
    int32_t n_tuple;          /* number of distinct (gax,gbx,gay,gby,gaz,gbz) tuples */
    int32_t idx_bytes;        /* 1 or 2: width of tuple_idx8/tuple_idx16 in use */
    const float    *tuples;   /* [n_tuple][6] */
    const uint8_t  *tuple_idx8;   /* ncell entries when idx_bytes==1, else NULL */
    const uint16_t *tuple_idx16;  /* ncell entries when idx_bytes==2, else NULL */

    /* ---- CPML (fdtd_solver.f90:3293-3577) -------------------------------
     * faces: 0 x-low, 1 x-high, 2 y-low, 3 y-high, 4 z-low, 5 z-high.
     * npml[f] is nxPML_1 etc (11 surfaces = 10 cells; 0 on periodic axes).
     * be/ce have npml[f] entries, bh/ch have npml[f]-1 entries (0-based). */
    int32_t npml[6];
    const float *be[6], *ce[6], *bh[6], *ch[6];

    /* ---- thin sheets (f90:4486-4554, 5712-5780) -------------------------
     * For normal axis a, n_sheet_planes[a] unique plane positions (1-based).
     * sheet_sig/sheet_ep hold, per plane, the normal-component averaged
     * sheet properties on that plane (sheet_sig_x_x(i,:,:) etc.), laid out
     * as [plane][q][p] where (p,q) run over the two other axes in (x,y,z)
     * order with full field extents (e.g. x-sheet: p=j in 1..NY, q=k in
     * 1..NZ).  Entries outside the reference's array are vacuum (0 / 1). */
    int32_t n_sheet_planes[3];
    const int32_t *sheet_pos[3];
    const float   *sheet_sig[3];
    const float   *sheet_ep[3];
    int32_t sheet_dim[3][2];  /* (P,Q) extents of each plane array */
    float   sheet_thickness;  /* min(d)/5000 (f90:957) */

    /* ---- Drude ADE (f90:5782-5854, coefficients f90:2859-2918) ----------
     * Per E component c, the reference loops over the plasma bounding box
     * dbox[c] = {xlo,xhi,ylo,yhi,zlo,zhi} (1-based inclusive,
     * f90:3126-3245).  Ja/Jb hold the J_plasma_a and J_plasma_b arrays on that box as
     * [pole][z][y][x] (x fastest). */
    int32_t n_poles;
    int32_t has_drude;
    int32_t dbox[3][6];
    const float *Ja[3], *Jb[3];

    /* ---- TF/SF plane wave (f90:4627-4710, 5950-6033, Inc f90:8825) ----- */
    int32_t has_tfsf;
    int32_t tf_lo[3], tf_hi[3];   /* xlow..zhigh (f90:1380-1397) */
    int32_t delay[3];             /* x_delay,y_delay,z_delay (f90:1451-1494) */
    float   WE[3], WH[3];         /* WEx..WEz, WHx..WHz (f90:1499-1504) */
    float   s_th, c_th, s_ph, c_ph;   /* sin/cos of theta and phi (radians, FP32) */
    float   s_pol, c_pol;
    float   theta, phi, pol;      /* radians */
    float   kHinc[3];             /* dt/(mu_0*d_a): H-face factor for normal axis a */
    float   kEinc[3];             /* dt/(ep_0*d_a): E-face factor */
    float   inv_c;                /* 1.0/c */
    float   plane_wave_amp;

    /* ---- lumped ports --------------------------------------------------- */
    int32_t n_ports;
    int32_t M;                    /* port_array_size (f90:1339) */
    const aem_port_view *ports;
    float   antenna_amp;
    int32_t excitation_port;      /* 1-based, 0 = none */

    /* ---- NTFF (f90:7512-7927) ------------------------------------------ */
    int32_t n_angles;
    const float *ang;             /* [n_angles][6]: theta,phi, sin th, cos th, sin ph, cos ph */
    int32_t ic, jc, kc;           /* phase centre cube */
    float   r_relay;              /* r_for_time_relay (f90:1322-1329) */
    int32_t len_ff;               /* len_far_field_arrays (f90:1334) */
    int32_t ff_lo[3], ff_hi[3];   /* ff_xlow..ff_zhigh (f90:1638-1680) */
    float   ff_K[3];              /* (d_b*d_c)/(4 pi dt c) for faces normal to axis a */

    /* ---- periodic-slab reflection/transmission sums (f90:7005-7103) ---- */
    int32_t plane_sum_axis;       /* -1 none, else the non-periodic axis (0,1,2) */
    float   ps_c1, ps_c2;         /* polarisation weights of the two tangential components */

    /* ---- misc / output header ------------------------------------------ */
    int32_t excitation;           /* 0 none, 1 plane wave, 2 antenna */
    float   min_steps;            /* f90:1811 */
    int32_t video_on, slice, slice_location, vid_size1, vid_size2;

    /* ---- time-only source waveforms, evaluated once on the host --------
     * src_table[0*T + n-1] = E_or_V_1D(antenna_amp, n+0.0, 0)      current source (f90:5001)
     * src_table[1*T + n-1] = E_or_V_1D(antenna_amp, n+0.5, dl/2)   voltage source (f90:6162)
     * src_table[2*T + n-1] = E_or_V_1D(plane_wave_amp, n+0.0, 0)   incident E (f90:8368)
     * (dl of the excited port; zeros when the source is absent). */
    const float *src_table;
    float   exp_half;             /* expf(0.5) as the host evaluates it (pulse type 2) */

    /* ---- SPICE co-simulation ports (deck section f90:1195-1286) ---------- */
    int32_t n_spice;              /* basic SPICE ports */
    int32_t n_spice_extra;        /* additional circuit nodes reported beside them */
    const aem_spice_view *spice;
    const char *spice_names;      /* one name per line: the netlist file, then per port its
                                     node, current source and capacitor, then the extra nodes */

    /* ---- kmax build (plan option kmax; fdtd_solver.f90 under use_kmax_version):
     * complex fields with Bloch-periodic walls, a modal plane source in place
     * of TF/SF, TE/TM plane sums in place of the periodic-slab sums -------- */
    int32_t kmax;                 /* 1: the kmax deck grammar and setup */
    int32_t mode_type;            /* 0 TE (source on the H plane), 1 TM (source on the E plane) */
    int32_t k_direction;          /* 1 '+', 0 '-' */
    int32_t k_axis;               /* the non-periodic axis (source and sum planes), -1 without two PBC pairs */
    int32_t kpl_H, kpl_E;         /* source planes k_pl_start_H / k_pl_start_E (f90:1727-1800) */
    float   k_vec[3];             /* Bloch wavenumbers of the periodic axes, rad/m */
    float   k_exc[3];             /* k_num_{x,y,z}_exception */
    float   f_adj;                /* carrier of the kmax sources (f90:1722) */
    float   bloch_c[3], bloch_s[3]; /* cos, sin of k_a (N_a - 2) d_a: an E copy to the first cell multiplies by
                                     (c, s), an H copy to the last cell by (c, -s) (f90:5387-5428, 6685-6736) */
    const float *src_table_im;    /* imaginary parts of the complex src_table waveforms */
    const float *ksrc;            /* [4][time_steps]: the modal source's Gaussian factor and carrier phase
                                     for the H plane (time n-1), then for the E plane (time n-1/2) */

    /* ---- infinite ground plane (IGP; deck f90:746-765, setup f90:1402-1445,
     * 1571-1665): a PEC wall at a TF/SF face.  That face's TF/SF corrections
     * and far-field currents are scaled by its wall factor 0, a mirrored
     * plane wave is added after the incident one (f90:4712-4790,
     * 6035-6111), and the far field gains the image currents
     * (f90:7928-8240). */
    int32_t mirror_type;          /* 0 none; 1..6 = -x,+x,-y,+y,-z,+z */
    float   wall[6];              /* xlow_wall..zhigh_wall: 1, or 0 on the ground face */
    int32_t delay_m[3];           /* x/y/z_delay_mirror */
    float   WE_m[3], WH_m[3];     /* mirrored weights (f90:1590-1634) */
    float   s_thm, c_thm, s_phm, c_phm;   /* sin/cos of theta_mirror and phi_mirror (radians, FP32) */
    float   use_m[3];             /* use_x/y/z_mirror */
    float   off_m[3];             /* x/y/z_mirror_offset */
    float   sgn_M[3], sgn_J[3];   /* Mx..Mz_mirror, Jx..Jz_mirror */
} aem_plan_view;

/* Setup constants that the reference fixes in its source rather than reading
 * from the deck.  aem_plan_options_init fills the reference's values; a plan
 * built with them (or with no options) is the reference's bit for bit.  A
 * caller changes fields after the init call. */
typedef struct aem_plan_options {
    int32_t cpml_cells;       /* CPML thickness in cells on each non-periodic face
                                 (the reference's 11 surfaces = 10 cells, f90:34) */
    int32_t cpml_grading;     /* polynomial order m of the sigma and kappa profiles (3, f90:35) */
    float   cpml_kappa_max;   /* kappa at the outer surface (7.0, f90:1690) */
    float   cpml_alpha_max;   /* CFS alpha at the inner surface, graded linearly (0.03, f90:1687) */
    float   cpml_reflection;  /* R of sigma_max = -ln(R) (m+1) / (2 eta0 d n) (1e-6, f90:1686) */
    int32_t tfsf_inset;       /* cells between the CPML and the TF/SF box ("buffer" = 3, f90:39, 1380-1397) */
    int32_t ntff_gap;         /* far-field box faces this many cells below the TF/SF box's low
                                 faces and one fewer above its high faces (2, f90:1638-1643) */
    int32_t kmax;             /* 1: the deck is for the reference's kmax build (0) */
} aem_plan_options;

#ifdef __cplusplus
}
#endif

#endif /* AEM_PLAN_H */
