// setup.cpp -- builds the plan consumed by the leapfrog loop.
//
// Restates the setup phase of the reference solver
// (main_fdtd/fdtd_solver.f90:1300-4369) for the non-kmax, non-SPICE build.
// FP32 expressions keep the reference's left-to-right evaluation order so the
// update coefficients come out bit-identical (this file is compiled with
// -ffp-contract=off).  The full per-edge coefficient arrays Gax..Gbz of the
// reference (f90:2797-2918) are replaced by a table of distinct
// (gax,gbx,gay,gby,gaz,gbz) tuples plus one 8- or 16-bit index per cell.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

#include "plan.hpp"

namespace aem {
namespace {

inline float avg4(float a, float b, float c, float d) { return (((a + b) + c) + d) / 4.0f; }

struct Grid {
    int NX, NY, NZ;
    int pbc[3];
    size_t cidx(int i, int j, int k) const {   // cell arrays (NX-1,NY-1,NZ-1)
        return (size_t)(i - 1) + (size_t)(NX - 1) * ((size_t)(j - 1) + (size_t)(NY - 1) * (size_t)(k - 1));
    }
    // Cell (i,j,k) after the PBC replication of cell arrays, which copies
    // plane 1 onto plane N-1 in the order x, y, z (f90:2360-2508).
    size_t cell_final(int i, int j, int k) const {
        if (pbc[2] && k == NZ - 1) k = 1;
        if (pbc[1] && j == NY - 1) j = 1;
        if (pbc[0] && i == NX - 1) i = 1;
        return cidx(i, j, k);
    }
    // Edge-averaged arrays are computed for i,j,k>=2 (f90:2513-2607), then
    // plane N-1 is copied onto plane 1 in the order x, z, y (f90:2610-2757).
    // Returns false when the final value is the untouched default.
    bool edge_src(int& i, int& j, int& k) const {
        if (pbc[1] && j == 1) j = NY - 1;
        if (pbc[2] && k == 1) k = NZ - 1;
        if (pbc[0] && i == 1) i = NX - 1;
        return i >= 2 && j >= 2 && k >= 2;
    }
};

// Per-cell material lookup helpers (index 0 = vacuum / no sheet).
struct MatTab {
    const Deck* d;
    float eps(int m, int a) const { return m ? d->mats[m - 1].eps[a] : 1.0f; }
    float sig(int m, int a) const { return m ? d->mats[m - 1].sig[a] : 0.0f; }
    float wp(int m, int p, int a) const { return m ? d->mats[m - 1].wp[p][a] : 0.0f; }
    float gam(int m, int p, int a) const { return m ? d->mats[m - 1].gam[p][a] : 0.0f; }
    float sep(int m, int a) const { return m ? d->smats[m - 1].ep[a] : 1.0f; }
    float ssig(int m, int a) const { return m ? d->smats[m - 1].sig[a] : 0.0f; }
};

// Ga/Gb of one edge (f90:2837-2853, 2897-2913).
inline void ga_gb(float sig, float eps, float dt, float& ga, float& gb) {
    ga = (1.0f - sig * dt / (2.0f * kEp0 * eps)) / (1.0f + sig * dt / (2.0f * kEp0 * eps));
    gb = (dt / (kEp0 * eps)) / (1.0f + sig * dt / (2.0f * kEp0 * eps));
}

// Drude pole coefficients (f90:2870-2889): Ja in FP32, Jb through FP64.
inline float drude_a(float loss, float dt) {
    return (1.0f - loss * dt / 2.0f) / (1.0f + loss * dt / 2.0f);
}
inline float drude_b(float wp, float loss, float dt) {
    double num = (double)wp * (double)wp * (double)kEp0 * (double)dt / 2.0;
    float den = 1.0f + loss * dt / 2.0f;
    return (float)(num / (double)den);
}

// x**m for a small integer m, by repeated products as the reference's
// integer power evaluates it (m = 3: x*x*x).
float ipow(float x, int m) {
    float p = x;
    for (int q = 1; q < m; ++q) p = p * x;
    return p;
}

// CPML profile of one face (f90:3405-3577); m is the grading order.
void cpml_face(int n, int m, float sig_max, float alpha_max, float kappa_max, float dt,
               std::vector<float>& be, std::vector<float>& ce,
               std::vector<float>& bh, std::vector<float>& ch, std::vector<float>& ke,
               std::vector<float>& kh) {
    be.assign(std::max(n, 0), 0.f); ce.assign(std::max(n, 0), 0.f); ke.assign(std::max(n, 0), 1.f);
    bh.assign(std::max(n - 1, 0), 0.f); ch.assign(std::max(n - 1, 0), 0.f); kh.assign(std::max(n - 1, 0), 1.f);
    for (int i = 1; i <= n; ++i) {
        float r = (float)(n - i) / ((float)n - 1.0f);
        float r3 = ipow(r, m);
        float sg = sig_max * r3;
        float al = alpha_max * (((float)i - 1.0f) / ((float)n - 1.0f));
        float ka = 1.0f + (kappa_max - 1.0f) * r3;
        float b = std::exp(-(sg / ka + al) * dt / kEp0);
        float c;
        if (sg == 0.0f && al == 0.0f && i == n) c = 0.0f;
        else c = sg * (b - 1.0f) / (sg + ka * al) / ka;
        be[i - 1] = b; ce[i - 1] = c; ke[i - 1] = ka;
    }
    for (int i = 1; i <= n - 1; ++i) {
        float r = ((float)(n - i) - 0.5f) / ((float)n - 1.0f);
        float r3 = ipow(r, m);
        float sg = sig_max * r3;
        float al = alpha_max * (((float)i - 0.5f) / ((float)n - 1.0f));
        float ka = 1.0f + (kappa_max - 1.0f) * r3;
        float b = std::exp(-(sg / ka + al) * dt / kEp0);
        float c = sg * (b - 1.0f) / (sg + ka * al) / ka;
        bh[i - 1] = b; ch[i - 1] = c; kh[i - 1] = ka;
    }
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(2, "cannot open '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

}  // namespace

void default_options(aem_plan_options& o) {
    o.cpml_cells = 10;          // 11 surfaces (f90:34)
    o.cpml_grading = 3;         // m (f90:35)
    o.cpml_kappa_max = 7.0f;    // f90:1690
    o.cpml_alpha_max = 0.03f;   // f90:1687
    o.cpml_reflection = 1.0E-6f;// f90:1686
    o.tfsf_inset = 3;           // buffer (f90:39)
    o.ntff_gap = 2;             // ff_*low = *low - 2, ff_*high = *high + 1 (f90:1638-1643)
    o.kmax = 0;
}

void check_options(const aem_plan_options& o) {
    auto bad = [](const std::string& what) { throw Error(1, "ERROR: setup option " + what); };
    if (o.cpml_cells < 2 || o.cpml_cells > 64) bad("cpml_cells must be 2..64");
    if (o.cpml_grading < 1 || o.cpml_grading > 8) bad("cpml_grading must be 1..8");
    if (!(o.cpml_kappa_max >= 1.0f) || !std::isfinite(o.cpml_kappa_max)) bad("cpml_kappa_max must be >= 1");
    if (!(o.cpml_alpha_max >= 0.0f) || !std::isfinite(o.cpml_alpha_max)) bad("cpml_alpha_max must be >= 0");
    if (!(o.cpml_reflection > 0.0f && o.cpml_reflection < 1.0f)) bad("cpml_reflection must be in (0, 1)");
    if (o.tfsf_inset < 3 || o.tfsf_inset > 64) bad("tfsf_inset must be 3..64");
    if (o.ntff_gap < 2 || o.ntff_gap > o.tfsf_inset - 1) bad("ntff_gap must be 2..tfsf_inset-1");
    if (o.kmax != 0 && o.kmax != 1) bad("kmax must be 0 or 1");
}

Plan* build_plan(const std::string& deck_text, const std::string& base_dir, const aem_plan_options& opt) {
    check_options(opt);
    Plan* P = new Plan();
    P->opt = opt;
    try {
        P->deck = parse_deck(deck_text, opt.kmax != 0);
    } catch (...) {
        delete P;
        throw;
    }
    P->base_dir = base_dir;
    Deck& d = P->deck;
    aem_plan_view& v = P->v;
    try {
        if (d.is_mirror && d.kmax) throw Error(5, "an infinite ground plane (IGP) in a kmax deck is not supported by libaem");
        for (auto& p : d.ports)
            if (p.gridded) throw Error(5, "gridded ports are not supported by libaem");
        // SPICE ports (the reference's use_spice_version build): basic ones
        // exchange current and gap voltage with a host circuit every step
        for (auto& p : d.spice)
            if (p.gridded) throw Error(5, "gridded SPICE ports are not supported by libaem");
        if (!d.kmax && d.pulse_type != 1 && d.pulse_type != 2)
            throw Error(1, "ERROR: pulse type must be 1 (Gaussian) or 2 (differentiated Gaussian)");
        if (d.kmax) {
            // parts of the kmax build outside this library: its far field
            // (Bloch phase removal, f90:7202-7510), the E_special Bloch copies
            // of thin sheets, SPICE pairs, and the exports
            if (d.n_angles > 0) throw Error(5, "far-field angles of the kmax build are not supported by libaem");
            if (!d.sheets.empty()) throw Error(5, "thin sheets in a kmax deck are not supported by libaem");
            if (!d.spice.empty()) throw Error(5, "SPICE ports in a kmax deck are not supported by libaem");
            if (d.video_on || d.export_all == "yes")
                throw Error(5, "video and full-field exports of a kmax deck are not supported by libaem");
            // with three periodic pairs the reference launches and sums on
            // every axis into the same series (f90:4558-4625, 7105-7190)
            if (d.pbc[0] + d.pbc[1] + d.pbc[2] == 3)
                throw Error(5, "a kmax deck with three periodic pairs is not supported by libaem");
        }

        // ---- grid and derived scalars (f90:1309-1397) ----------------------
        Grid g;
        for (int a = 0; a < 3; ++a) g.pbc[a] = d.pbc[a];
        g.NX = d.nx + 1 + d.pbc[0];
        g.NY = d.ny + 1 + d.pbc[1];
        g.NZ = d.nz + 1 + d.pbc[2];
        const int NX = g.NX, NY = g.NY, NZ = g.NZ;
        v.nx = d.nx; v.ny = d.ny; v.nz = d.nz;
        v.NX = NX; v.NY = NY; v.NZ = NZ;
        v.PX = (NX + 31) & ~31;
        v.plane = (int64_t)v.PX * NY;
        v.ncell = v.plane * NZ;
        for (int a = 0; a < 3; ++a) v.pbc[a] = d.pbc[a];
        const float dx = d.dx, dy = d.dy, dz = d.dz;
        v.dx = dx; v.dy = dy; v.dz = dz;
        const float dt = 0.98f / (kC * std::sqrt(1.0f / (dx * dx) + 1.0f / (dy * dy) + 1.0f / (dz * dz))) / d.factor;
        v.dt = dt;
        v.DB = dt / kMu0;
        v.time_steps = d.time_steps;
        if (!(dt > 0) || !std::isfinite(dt)) throw Error(1, "ERROR: invalid time step");

        float r_relay = 1.0f;                                          // f90:1322-1329
        float ext = dx * (float)NX + dy * (float)NY + dz * (float)NZ;
        if (ext > r_relay) r_relay = ext * 2.0f;
        v.r_relay = r_relay;
        v.len_ff = (int)((float)d.time_steps + (r_relay + dx * (float)NX + dy * (float)NY + dz * (float)NZ) / (kC * dt));
        v.M = (int)((float)d.time_steps * dt * kC / std::min(std::min(dx, dy), dz) / 2.0f + 100.0f);  // f90:1339

        int npml[6];
        for (int a = 0; a < 3; ++a)                                    // f90:34,1342-1353
            npml[2 * a] = npml[2 * a + 1] = d.pbc[a] ? 0 : opt.cpml_cells + 1;
        for (int f = 0; f < 6; ++f) v.npml[f] = npml[f];

        // f90:1356-1367: the slice sizes follow the slice axis (0 when video
        // is off) whether or not video is on; data.dat prints them either way
        if (d.slice == 0) { v.vid_size1 = NY - 1 - d.pbc[1]; v.vid_size2 = NZ - 1 - d.pbc[2]; }
        if (d.slice == 1) { v.vid_size1 = NX - 1 - d.pbc[0]; v.vid_size2 = NZ - 1 - d.pbc[2]; }
        if (d.slice == 2) { v.vid_size1 = NX - 1 - d.pbc[0]; v.vid_size2 = NY - 1 - d.pbc[1]; }
        v.video_on = d.video_on; v.slice = d.slice; v.slice_location = d.slice_location;

        const float theta = kPi / 180.0f * d.theta;                    // f90:1370-1372
        const float phi = kPi / 180.0f * d.phi;
        const float pol = kPi / 180.0f * d.pol;
        v.theta = theta; v.phi = phi; v.pol = pol;
        v.s_th = std::sin(theta); v.c_th = std::cos(theta);
        v.s_ph = std::sin(phi); v.c_ph = std::cos(phi);
        v.s_pol = std::sin(pol); v.c_pol = std::cos(pol);

        int lo[3], hi[3];                                              // f90:1380-1397
        int N3[3] = {NX, NY, NZ};
        for (int a = 0; a < 3; ++a) {
            lo[a] = 1; hi[a] = N3[a] - 1;
            if (!d.pbc[a]) { lo[a] = npml[2 * a] + opt.tfsf_inset; hi[a] = N3[a] - npml[2 * a + 1] - opt.tfsf_inset; }
        }
        // infinite ground plane (f90:1402-1445): the TF/SF face on the ground
        // moves to the plane (low side: the cell above it) and loses its
        // corrections; that side's CPML shrinks to one surface (low) or none
        // (high), which the sigma maximum below then sees (f90:1686-1688)
        const int igp = d.is_mirror ? d.mirror_type : 0;
        v.mirror_type = igp;
        for (int f = 0; f < 6; ++f) v.wall[f] = 1.0f;
        for (int a = 0; a < 3; ++a) {
            v.use_m[a] = 0.0f; v.off_m[a] = 0.0f;
            v.sgn_M[a] = 1.0f; v.sgn_J[a] = 1.0f;
        }
        if (igp >= 1 && igp <= 6) {
            const int a = (igp - 1) / 2;
            const bool low = (igp & 1) == 1;
            if (low) {
                lo[a] = d.mirror_height + 1;
                npml[2 * a] = 1;
                v.off_m[a] = 2.0f * (float)lo[a];
            } else {
                hi[a] = d.mirror_height;
                npml[2 * a + 1] = 0;
                v.off_m[a] = 2.0f * (float)hi[a];
            }
            v.wall[2 * a + (low ? 0 : 1)] = 0.0f;
            v.use_m[a] = 1.0f;
            for (int f = 0; f < 6; ++f) v.npml[f] = npml[f];
        }
        for (int a = 0; a < 3; ++a) { v.tf_lo[a] = lo[a]; v.tf_hi[a] = hi[a]; }
        // plane-wave delays by incidence quadrant (f90:1451-1494; the mirrored
        // wave's f90:1519-1565)
        auto quadrant_delays = [&](float th, float ph, int* dl) {
            const float hp = kPi / 2.0f;
            if (th >= 0.0f && th <= hp) {
                if (ph >= 0.0f && ph <= hp) { dl[0] = lo[0]; dl[1] = lo[1]; dl[2] = lo[2]; }
                if (ph > hp && ph <= kPi) { dl[0] = hi[0]; dl[1] = lo[1]; dl[2] = lo[2]; }
                if (ph > kPi && ph <= 3.0f * kPi / 2.0f) { dl[0] = hi[0]; dl[1] = hi[1]; dl[2] = lo[2]; }
                if (ph > 3.0f * kPi / 2.0f && ph < 2.0f * kPi) { dl[0] = lo[0]; dl[1] = hi[1]; dl[2] = lo[2]; }
            }
            if (th > hp && th <= kPi) {
                if (ph >= 0.0f && ph <= hp) { dl[0] = lo[0]; dl[1] = lo[1]; dl[2] = hi[2]; }
                if (ph > hp && ph <= kPi) { dl[0] = hi[0]; dl[1] = lo[1]; dl[2] = hi[2]; }
                if (ph > kPi && ph <= 3.0f * kPi / 2.0f) { dl[0] = hi[0]; dl[1] = hi[1]; dl[2] = hi[2]; }
                if (ph > 3.0f * kPi / 2.0f && ph < 2.0f * kPi) { dl[0] = lo[0]; dl[1] = hi[1]; dl[2] = hi[2]; }
            }
        };
        int dl[3] = {0, 0, 0};
        quadrant_delays(theta, phi, dl);
        for (int a = 0; a < 3; ++a) v.delay[a] = dl[a];
        const float A = d.pw_amp;                                      // f90:1499-1504
        const float s_ep_mu = std::sqrt(kEp0 / kMu0);
        v.WH[0] = A * (std::sin(pol) * std::sin(phi) + std::cos(pol) * std::cos(theta) * std::cos(phi)) * s_ep_mu;
        v.WH[1] = A * (-1.0f * std::sin(pol) * std::cos(phi) + std::cos(pol) * std::cos(theta) * std::sin(phi)) * s_ep_mu;
        v.WH[2] = A * (-1.0f * std::cos(pol) * std::sin(theta)) * s_ep_mu;
        v.WE[0] = A * (std::cos(pol) * std::sin(phi) - std::sin(pol) * std::cos(theta) * std::cos(phi));
        v.WE[1] = A * (-1.0f * std::cos(pol) * std::cos(phi) - std::sin(pol) * std::cos(theta) * std::sin(phi));
        v.WE[2] = A * (std::sin(pol) * std::sin(theta));
        // the mirrored plane wave (f90:773-798 angles, 1375-1376, 1519-1585
        // delays, 1590-1634 weights and image-current signs)
        for (int a = 0; a < 3; ++a) { v.delay_m[a] = 0; v.WE_m[a] = 0.0f; v.WH_m[a] = 0.0f; }
        v.s_thm = v.c_thm = v.s_phm = v.c_phm = 0.0f;
        if (igp >= 1 && igp <= 6) {
            const int a = (igp - 1) / 2;
            float thm = 0.0f, phm = 0.0f;
            if (d.excitation == 1) {   // the angles are read with a plane wave only
                if (a == 0) { thm = d.theta; phm = 180.0f - d.phi; }
                else if (a == 1) { thm = d.theta; phm = 360.0f - d.phi; }
                else { thm = 180.0f - d.theta; phm = d.phi; }
            }
            thm = kPi / 180.0f * thm;
            phm = kPi / 180.0f * phm;
            v.s_thm = std::sin(thm); v.c_thm = std::cos(thm);
            v.s_phm = std::sin(phm); v.c_phm = std::cos(phm);
            int dm[3] = {0, 0, 0};
            quadrant_delays(thm, phm, dm);
            const int h = d.mirror_height;
            dm[a] = h + (h - ((igp & 1) ? hi[a] : lo[a]));
            for (int q = 0; q < 3; ++q) {
                v.delay_m[q] = dm[q];
                // E tangential to the ground flips, H normal to it flips
                v.WE_m[q] = q == a ? v.WE[q] : -1.0f * v.WE[q];
                v.WH_m[q] = q == a ? -1.0f * v.WH[q] : v.WH[q];
                if (q != a) v.sgn_J[q] = -1.0f;
            }
            v.sgn_M[a] = -1.0f;
        }
        const float dd[3] = {dx, dy, dz};
        for (int a = 0; a < 3; ++a) {
            v.kHinc[a] = dt / (kMu0 * dd[a]);
            v.kEinc[a] = dt / (kEp0 * dd[a]);
        }
        v.inv_c = 1.0f / kC;
        v.plane_wave_amp = d.pw_amp;
        v.has_tfsf = d.pw_amp > 0 ? 1 : 0;
        if (d.kmax) {                                                  // f90:1506-1517
            // both directions are launched from one plane: no TF/SF box
            v.has_tfsf = 0;
            for (int a = 0; a < 3; ++a) v.WE[a] = v.WH[a] = 0.0f;
            if (d.mode_type == 0)
                for (int a = 0; a < 3; ++a) v.WE[a] = 2.0f * d.pw_amp;
            if (d.mode_type == 1)
                for (int a = 0; a < 3; ++a) v.WH[a] = 2.0f * std::sqrt(kEp0 / kMu0) * d.pw_amp;
        }

        for (int a = 0; a < 3; ++a) {                                  // f90:1638-1680
            v.ff_lo[a] = d.pbc[a] ? 1 : lo[a] - opt.ntff_gap;
            v.ff_hi[a] = d.pbc[a] ? N3[a] - 2 : hi[a] + opt.ntff_gap - 1;
        }
        if (igp >= 1 && igp <= 6) {   // the faces stop short of the ground plane (f90:1646-1669)
            const int a = (igp - 1) / 2;
            if (!d.pbc[a]) {
                if (igp & 1) v.ff_lo[a] = lo[a] + 1;
                else v.ff_hi[a] = hi[a] - 2;
            }
        }
        {
            const float fpdc = 4.0f * kPi * dt * kC;
            v.ff_K[0] = dy * dz / fpdc;
            v.ff_K[1] = dx * dz / fpdc;
            v.ff_K[2] = dx * dy / fpdc;
        }

        // CPML maxima (f90:1686-1694)
        const float s_mu_ep = std::pow(kMu0 / kEp0 * 1.0f, 0.5f);
        float sig_max[3];
        const float lnR = std::log(opt.cpml_reflection);
        const float m1 = (float)(opt.cpml_grading + 1);
        sig_max[0] = -lnR * (m1 / (dx * s_mu_ep)) / (float)(2 * npml[0]);
        sig_max[1] = -lnR * (m1 / (dy * s_mu_ep)) / (float)(2 * npml[2]);
        sig_max[2] = -lnR * (m1 / (dz * s_mu_ep)) / (float)(2 * npml[4]);
        float alpha_max = opt.cpml_alpha_max;
        const float kappa_max = opt.cpml_kappa_max;
        const float kmag = std::sqrt(d.k_count[1] * d.k_count[1] + d.k_count[2] * d.k_count[2] +
                                     d.k_count[0] * d.k_count[0]);
        if (d.kmax) {                                                  // f90:1696-1705
            for (int a = 0; a < 3; ++a) sig_max[a] = sig_max[a] * 1.25f;
            alpha_max = 0.03f + 0.9f * kmag / 377.0f;
        }

        if (d.kmax) {                                                  // f90:1720-1724
            v.spread = 4.0f / (2.0f * kPi * d.f_center);
            v.t_spread = 5.0f * v.spread + d.pw_delay + d.ant_delay;
            v.f_adj = (kmag * kC) / (2.0f * kPi) + d.f_center;
        } else if (d.pulse_type == 1) {                                // f90:1711-1718
            v.spread = 1.0f / (2.0f * kPi * d.f_center);
            v.t_spread = 5.0f * v.spread + d.pw_delay + d.ant_delay;
        } else {
            v.spread = 1.0f / (2.0f * kPi * d.f_center);
            v.t_spread = 6.0f * v.spread + d.pw_delay + d.ant_delay;
        }
        v.pulse_type = d.pulse_type;
        v.f_center = d.f_center;
        {
            float sx = std::sin(theta) * std::cos(phi) * (float)NX * dx;   // f90:1811-1812
            float sy = std::sin(theta) * std::sin(phi) * (float)NY * dy;
            float sz = std::cos(theta) * (float)NZ * dz;
            v.min_steps = 2.0f * v.t_spread / dt + std::sqrt(sx * sx + sy * sy + sz * sz) / (kC * dt);
        }
        v.excitation = d.excitation;

        // ---- rasterise cells (f90:1968-2352) -------------------------------
        MatTab mt{&d};
        const size_t ncells = (size_t)(NX - 1) * (NY - 1) * (NZ - 1);
        std::vector<uint16_t> cellm(ncells, 0);
        if (d.mats.size() > 65535) throw Error(1, "ERROR: too many materials");
        auto find_mat = [&](float id) -> int {
            int r = 0;
            for (size_t rr = 0; rr < d.mats.size(); ++rr)
                if (d.mats[rr].id == id) r = (int)rr + 1;
            if (r == 0) throw Error(1, "ERROR: object refers to undefined material " + std::to_string((int)id));
            return r;
        };
        auto in_cells = [&](int i, int j, int k) {
            return i >= 1 && i <= NX - 1 && j >= 1 && j <= NY - 1 && k >= 1 && k <= NZ - 1;
        };
        const int gnx = d.nx, gny = d.ny, gnz = d.nz;
        std::vector<float>& geom = P->geom;
        geom.assign((size_t)4 * gnx * gny * gnz, 0.0f);
        auto gidx = [&](int c, int i, int j, int k) {
            return (size_t)c + 4 * ((size_t)(i - 1) + (size_t)gnx * ((size_t)(j - 1) + (size_t)gny * (size_t)(k - 1)));
        };
        auto geom_set = [&](int c, int i, int j, int k, float val) {
            if (i <= gnx && j <= gny && k <= gnz) geom[gidx(c, i, j, k)] = val;
        };

        if (d.use_opt_geom) {                                          // f90:2051-2095
            std::string path = d.opt_geom_file;
            if (!path.empty() && path[0] != '/') path = base_dir + "/" + path;
            std::string raw = read_file(path);
            size_t need = (size_t)gnx * gny * gnz * 4;
            if (raw.size() < need) throw Error(2, "optional geometry file '" + path + "' is too short");
            const float* og = reinterpret_cast<const float*>(raw.data());
            int maxid = 0;
            for (auto& m : d.mats) maxid = std::max(maxid, (int)m.id);
            std::vector<int> tmp(maxid + 1, 0);
            for (size_t rr = 0; rr < d.mats.size(); ++rr) tmp[(int)d.mats[rr].id] = (int)rr + 1;
            for (int k = 1; k <= gnz; ++k)
                for (int j = 1; j <= gny; ++j)
                    for (int i = 1; i <= gnx; ++i) {
                        float val = og[(size_t)(i - 1) + (size_t)gnx * ((size_t)(j - 1) + (size_t)gny * (k - 1))];
                        int iv = (int)val;
                        int m = (iv >= 0 && iv <= maxid) ? tmp[iv] : 0;
                        if (m != 0) cellm[g.cidx(i, j, k)] = (uint16_t)m;
                        geom[gidx(0, i, j, k)] = val;
                    }
        }

        for (auto& o : d.objects) {
            int m = find_mat((float)o.mat_id);
            auto put = [&](int i, int j, int k) {
                if (!in_cells(i, j, k))
                    throw Error(1, "ERROR: object extends outside the grid at cell (" + std::to_string(i) + "," +
                                       std::to_string(j) + "," + std::to_string(k) + ")");
                cellm[g.cidx(i, j, k)] = (uint16_t)m;
                geom_set(0, i, j, k, (float)o.mat_id);
            };
            if (o.kind == 0) {                                         // block f90:2104-2139
                for (int k = o.p[2]; k <= o.p[2] + o.s[2] - 1; ++k)
                    for (int j = o.p[1]; j <= o.p[1] + o.s[1] - 1; ++j)
                        for (int i = o.p[0]; i <= o.p[0] + o.s[0] - 1; ++i) put(i, j, k);
            } else if (o.kind == 1) {                                  // sphere f90:2141-2181
                const int r = o.s[0];
                const float r2 = (float)r * (float)r;
                for (int k = o.p[2] - r; k <= o.p[2] + r - 1; ++k)
                    for (int j = o.p[1] - r; j <= o.p[1] + r - 1; ++j)
                        for (int i = o.p[0] - r; i <= o.p[0] + r - 1; ++i) {
                            float a = (float)i - (float)o.p[0] + 0.5f;
                            float b = (float)j - (float)o.p[1] + 0.5f;
                            float c = (float)k - (float)o.p[2] + 0.5f;
                            if (a * a + b * b + c * c <= r2) put(i, j, k);
                        }
            } else {                                                   // cylinder f90:2183-2285
                const int len = o.s[0], r = o.s[1];
                const float r2 = (float)r * (float)r;
                if (o.axis == 0) {
                    for (int k = o.p[2] - r; k <= o.p[2] + r - 1; ++k)
                        for (int j = o.p[1] - r; j <= o.p[1] + r - 1; ++j)
                            for (int i = o.p[0]; i <= o.p[0] + len - 1; ++i) {
                                float b = (float)j - (float)o.p[1] + 0.5f, c = (float)k - (float)o.p[2] + 0.5f;
                                if (b * b + c * c <= r2) put(i, j, k);
                            }
                } else if (o.axis == 1) {
                    for (int k = o.p[2] - r; k <= o.p[2] + r - 1; ++k)
                        for (int j = o.p[1]; j <= o.p[1] + len - 1; ++j)
                            for (int i = o.p[0] - r; i <= o.p[0] + r - 1; ++i) {
                                float a = (float)i - (float)o.p[0] + 0.5f, c = (float)k - (float)o.p[2] + 0.5f;
                                if (a * a + c * c <= r2) put(i, j, k);
                            }
                } else {
                    for (int k = o.p[2]; k <= o.p[2] + len - 1; ++k)
                        for (int j = o.p[1] - r; j <= o.p[1] + r - 1; ++j)
                            for (int i = o.p[0] - r; i <= o.p[0] + r - 1; ++i) {
                                float a = (float)i - (float)o.p[0] + 0.5f, b = (float)j - (float)o.p[1] + 0.5f;
                                if (a * a + b * b <= r2) put(i, j, k);
                            }
                }
            }
        }

        // sheet cells per normal axis (f90:2290-2352)
        int nsheets[3] = {0, 0, 0};
        for (auto& s : d.sheets) nsheets[s.axis]++;
        std::vector<uint16_t> sheetm[3];
        for (int a = 0; a < 3; ++a)
            if (nsheets[a]) sheetm[a].assign(ncells, 0);
        for (auto& s : d.sheets) {
            int sm = 0;
            for (size_t rr = 0; rr < d.smats.size(); ++rr)
                if (d.smats[rr].id == (float)s.mat_id) sm = (int)rr + 1;
            if (sm == 0) throw Error(1, "ERROR: sheet refers to undefined sheet material " + std::to_string(s.mat_id));
            for (int q = s.start[1]; q <= s.start[1] + s.size[1] - 1; ++q)
                for (int p = s.start[0]; p <= s.start[0] + s.size[0] - 1; ++p) {
                    int i, j, k;
                    if (s.axis == 0) { i = s.pos; j = p; k = q; }
                    else if (s.axis == 1) { j = s.pos; i = p; k = q; }
                    else { k = s.pos; i = p; j = q; }
                    if (!in_cells(i, j, k)) throw Error(1, "ERROR: sheet extends outside the grid");
                    sheetm[s.axis][g.cidx(i, j, k)] = (uint16_t)sm;
                    geom_set(1 + s.axis, i, j, k, (float)s.mat_id);
                }
        }
        const bool any_sheet = nsheets[0] + nsheets[1] + nsheets[2] > 0;
        const int NP = d.num_poles;
        const bool plasma_in = d.is_plasma == 1 && NP > 0;
        const float tsh = d.sheet_thickness;
        const float tdx = tsh / dx, tdy = tsh / dy, tdz = tsh / dz;

        // helpers reading final (post-replication) cell materials
        auto cm = [&](int i, int j, int k) -> int { return cellm[g.cell_final(i, j, k)]; };
        auto sm = [&](int a, int i, int j, int k) -> int { return sheetm[a][g.cell_final(i, j, k)]; };

        // The four cells averaged onto each edge (f90:2516-2527): x-edges mix
        // (j,j-1)x(k,k-1), y-edges (i,i-1)x(k,k-1), z-edges (i,i-1)x(j,j-1).
        auto edge_cells = [&](int c, int i, int j, int k, int out[4][3]) {
            int o[4][3];
            if (c == 0) { int t[4][3] = {{i, j, k}, {i, j - 1, k}, {i, j, k - 1}, {i, j - 1, k - 1}}; std::memcpy(o, t, sizeof o); }
            else if (c == 1) { int t[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j, k - 1}, {i - 1, j, k - 1}}; std::memcpy(o, t, sizeof o); }
            else { int t[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j - 1, k}, {i - 1, j - 1, k}}; std::memcpy(o, t, sizeof o); }
            std::memcpy(out, o, sizeof o);
        };

        // averaged bulk eps/sigma of component c at edge (i,j,k), final values
        auto bulk_avg = [&](int c, int i, int j, int k, float& eps, float& sig) {
            int ii = i, jj = j, kk = k;
            if (!g.edge_src(ii, jj, kk)) { eps = 1.0f; sig = 0.0f; return; }
            int e[4][3];
            edge_cells(c, ii, jj, kk, e);
            int m0 = cm(e[0][0], e[0][1], e[0][2]), m1 = cm(e[1][0], e[1][1], e[1][2]);
            int m2 = cm(e[2][0], e[2][1], e[2][2]), m3 = cm(e[3][0], e[3][1], e[3][2]);
            eps = avg4(mt.eps(m0, c), mt.eps(m1, c), mt.eps(m2, c), mt.eps(m3, c));
            sig = avg4(mt.sig(m0, c), mt.sig(m1, c), mt.sig(m2, c), mt.sig(m3, c));
        };
        auto plasma_avg = [&](int c, int p, int i, int j, int k, float& wp, float& loss) {
            int ii = i, jj = j, kk = k;
            if (!g.edge_src(ii, jj, kk)) { wp = 0.0f; loss = 0.0f; return; }
            int e[4][3];
            edge_cells(c, ii, jj, kk, e);
            int m[4];
            for (int q = 0; q < 4; ++q) m[q] = cm(e[q][0], e[q][1], e[q][2]);
            wp = avg4(mt.wp(m[0], p, c), mt.wp(m[1], p, c), mt.wp(m[2], p, c), mt.wp(m[3], p, c));
            loss = avg4(mt.gam(m[0], p, c), mt.gam(m[1], p, c), mt.gam(m[2], p, c), mt.gam(m[3], p, c));
        };
        // Averaged sheet property (f90:2532-2582), including the reference's
        // double-counted cells on the in-plane components.  a = sheet normal,
        // c = field component, which = 0 ep, 1 sig.
        auto sheet_avg = [&](int a, int c, int which, int i, int j, int k) -> float {
            int ii = i, jj = j, kk = k;
            float dflt = which ? 0.0f : 1.0f;
            if (!g.edge_src(ii, jj, kk)) return dflt;
            // Offsets (di,dj,dk) of the four summed cells, in the reference's
            // summation order (f90:2537-2578).  Tangential components sample
            // one cell pair twice; the y-sheet normal component sums in a
            // different order from the bulk y-edge average.
            static const int T[3][3][4][3] = {
                // x-sheet: c=x (regular), c=y, c=z
                {{{0, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, -1, -1}},
                 {{0, 0, 0}, {0, 0, -1}, {0, 0, 0}, {0, 0, -1}},
                 {{0, 0, 0}, {0, -1, 0}, {0, 0, 0}, {0, -1, 0}}},
                // y-sheet: c=x, c=y (k before i), c=z
                {{{0, 0, 0}, {0, 0, 0}, {0, 0, -1}, {0, 0, -1}},
                 {{0, 0, 0}, {0, 0, -1}, {-1, 0, 0}, {-1, 0, -1}},
                 {{0, 0, 0}, {-1, 0, 0}, {0, 0, 0}, {-1, 0, 0}}},
                // z-sheet: c=x, c=y, c=z (regular)
                {{{0, 0, 0}, {0, -1, 0}, {0, 0, 0}, {0, -1, 0}},
                 {{0, 0, 0}, {-1, 0, 0}, {0, 0, 0}, {-1, 0, 0}},
                 {{0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {-1, -1, 0}}}};
            int e[4][3];
            for (int q = 0; q < 4; ++q) {
                e[q][0] = ii + T[a][c][q][0];
                e[q][1] = jj + T[a][c][q][1];
                e[q][2] = kk + T[a][c][q][2];
            }
            float vals[4];
            for (int q = 0; q < 4; ++q) {
                int mm = sm(a, e[q][0], e[q][1], e[q][2]);
                vals[q] = which ? mt.ssig(mm, c) : mt.sep(mm, c);
            }
            return avg4(vals[0], vals[1], vals[2], vals[3]);
        };

        // ---- Ga/Gb tuples, Drude coefficients (f90:2832-3120) --------------
        // uniform-neighbourhood fast path: coefficient of material m when all
        // four averaged cells carry m
        const int nm = (int)d.mats.size();
        std::vector<float> uni_ga((nm + 1) * 3), uni_gb((nm + 1) * 3);
        for (int m = 0; m <= nm; ++m)
            for (int c = 0; c < 3; ++c) {
                float e = mt.eps(m, c), s = mt.sig(m, c);
                float ea = avg4(e, e, e, e), sa = avg4(s, s, s, s);
                ga_gb(sa, ea, dt, uni_ga[m * 3 + c], uni_gb[m * 3 + c]);
            }

        struct Key {
            uint32_t w[6];
            bool operator==(const Key& o) const { return std::memcmp(w, o.w, sizeof w) == 0; }
        };
        struct KeyHash {
            size_t operator()(const Key& k) const {
                uint64_t h = 1469598103934665603ull;
                for (int i = 0; i < 6; ++i) { h ^= k.w[i]; h *= 1099511628211ull; }
                return (size_t)h;
            }
        };
        std::unordered_map<Key, uint32_t, KeyHash> tmap;
        std::vector<uint32_t> cell_tuple((size_t)v.ncell, 0);
        auto add_tuple = [&](const float t6[6]) -> uint32_t {
            Key k;
            std::memcpy(k.w, t6, sizeof k.w);
            auto it = tmap.find(k);
            if (it != tmap.end()) return it->second;
            uint32_t id = (uint32_t)tmap.size();
            tmap.emplace(k, id);
            P->tuples.insert(P->tuples.end(), t6, t6 + 6);
            return id;
        };
        // vacuum tuple first so that index 0 is vacuum (unused entries)
        {
            float t6[6];
            for (int c = 0; c < 3; ++c) { t6[2 * c] = uni_ga[c]; t6[2 * c + 1] = uni_gb[c]; }
            // vacuum edges at index 1 are not averaged: eps=1, sig=0 exactly
            float ga0, gb0;
            ga_gb(0.0f, 1.0f, dt, ga0, gb0);
            for (int c = 0; c < 3; ++c) { t6[2 * c] = ga0; t6[2 * c + 1] = gb0; }
            add_tuple(t6);
        }

        // Drude: full-grid Ja/Jb are only materialised on the plasma bounding
        // box.  First pass finds the box (f90:3126-3245), second fills it.
        const bool need_plasma = plasma_in;
        int box[3][6];
        bool any_plasma = false;
        if (need_plasma) {
            // seeds of the reference (f90:3130-3148)
            int seeds[3][6] = {
                {NX - 1, 1, NY - 1, 2, NZ - 1, 2},   // x-fields: min_x,max_x,min_y,max_y,min_z,max_z
                {NX - 1, 2, NY - 1, 1, NZ - 1, 2},
                {NX - 1, 2, NY - 1, 2, NZ - 1, 1}};
            std::memcpy(box, seeds, sizeof box);
            int rng[3][3] = {{1, 2, 2}, {2, 1, 2}, {2, 2, 1}};  // lower loop bounds i,j,k per component
            for (int c = 0; c < 3; ++c)
                for (int k = rng[c][2]; k <= NZ - 1; ++k)
                    for (int j = rng[c][1]; j <= NY - 1; ++j)
                        for (int i = rng[c][0]; i <= NX - 1; ++i)
                            for (int p = 0; p < NP; ++p) {
                                float wp, loss;
                                plasma_avg(c, p, i, j, k, wp, loss);
                                float ja = drude_a(loss, dt), jb = drude_b(wp, loss, dt);
                                if (ja > 1.0f || jb > 0.0f) {
                                    any_plasma = true;
                                    box[c][0] = std::min(box[c][0], i); box[c][1] = std::max(box[c][1], i);
                                    box[c][2] = std::min(box[c][2], j); box[c][3] = std::max(box[c][3], j);
                                    box[c][4] = std::min(box[c][4], k); box[c][5] = std::max(box[c][5], k);
                                    break;
                                }
                            }
            // The ANY() test of f90:3126 also covers index-1 edges outside the
            // detection loops; those only matter if no interior edge is plasma.
            if (!any_plasma) {
                for (int c = 0; c < 3 && !any_plasma; ++c)
                    for (int k = 1; k <= NZ - 1 && !any_plasma; ++k)
                        for (int j = 1; j <= NY - 1 && !any_plasma; ++j)
                            for (int i = 1; i <= NX - 1 && !any_plasma; ++i)
                                for (int p = 0; p < NP; ++p) {
                                    float wp, loss;
                                    plasma_avg(c, p, i, j, k, wp, loss);
                                    if (drude_a(loss, dt) > 1.0f || drude_b(wp, loss, dt) > 0.0f) { any_plasma = true; break; }
                                }
            }
        }
        v.n_poles = NP;
        v.has_drude = (need_plasma && any_plasma) ? 1 : 0;
        if (v.has_drude) {
            for (int c = 0; c < 3; ++c) {
                for (int q = 0; q < 6; ++q) v.dbox[c][q] = box[c][q];
                int bx = std::max(0, box[c][1] - box[c][0] + 1), by = std::max(0, box[c][3] - box[c][2] + 1),
                    bz = std::max(0, box[c][5] - box[c][4] + 1);
                size_t nb = (size_t)bx * by * bz;
                P->Ja[c].assign(nb * NP, 1.0f);
                P->Jb[c].assign(nb * NP, 0.0f);
                for (int k = box[c][4]; k <= box[c][5]; ++k)
                    for (int j = box[c][2]; j <= box[c][3]; ++j)
                        for (int i = box[c][0]; i <= box[c][1]; ++i)
                            for (int p = 0; p < NP; ++p) {
                                float wp, loss;
                                plasma_avg(c, p, i, j, k, wp, loss);
                                size_t o = (size_t)p * nb + (size_t)(i - box[c][0]) +
                                           (size_t)bx * ((size_t)(j - box[c][2]) + (size_t)by * (k - box[c][4]));
                                P->Ja[c][o] = drude_a(loss, dt);
                                P->Jb[c][o] = drude_b(wp, loss, dt);
                            }
            }
        }

        // per-edge coefficients
        for (int k = 1; k <= NZ - 1; ++k)
            for (int j = 1; j <= NY - 1; ++j)
                for (int i = 1; i <= NX - 1; ++i) {
                    float t6[6];
                    bool simple = !any_sheet && !plasma_in;
                    if (simple) {
                        int ii = i, jj = j, kk = k;
                        bool interior = g.edge_src(ii, jj, kk);
                        for (int c = 0; c < 3 && simple; ++c) {
                            if (!interior) {
                                ga_gb(0.0f, 1.0f, dt, t6[2 * c], t6[2 * c + 1]);
                                continue;
                            }
                            int e[4][3];
                            edge_cells(c, ii, jj, kk, e);
                            int m0 = cm(e[0][0], e[0][1], e[0][2]);
                            bool same = true;
                            for (int q = 1; q < 4; ++q) same &= cm(e[q][0], e[q][1], e[q][2]) == m0;
                            if (same) {
                                t6[2 * c] = uni_ga[m0 * 3 + c];
                                t6[2 * c + 1] = uni_gb[m0 * 3 + c];
                            } else {
                                float eps, sig;
                                bulk_avg(c, i, j, k, eps, sig);
                                ga_gb(sig, eps, dt, t6[2 * c], t6[2 * c + 1]);
                            }
                        }
                    } else {
                        float eps[3], sig[3], tpl[3] = {0, 0, 0};
                        for (int c = 0; c < 3; ++c) bulk_avg(c, i, j, k, eps[c], sig[c]);
                        if (plasma_in) {
                            for (int c = 0; c < 3; ++c) {
                                float t = 0.0f;
                                for (int p = 0; p < NP; ++p) {
                                    float wp, loss;
                                    plasma_avg(c, p, i, j, k, wp, loss);
                                    t = t + drude_b(wp, loss, dt);
                                }
                                tpl[c] = t;
                                ga_gb(sig[c] + t, eps[c], dt, t6[2 * c], t6[2 * c + 1]);
                            }
                        } else {
                            for (int c = 0; c < 3; ++c) ga_gb(sig[c], eps[c], dt, t6[2 * c], t6[2 * c + 1]);
                        }
                        if (any_sheet) {                               // f90:2922-3120
                            float sea[3] = {eps[0], eps[1], eps[2]}, ssa[3] = {sig[0], sig[1], sig[2]};
                            int cnt[3] = {0, 0, 0};
                            const float tda[3] = {tdx, tdy, tdz};
                            for (int a = 0; a < 3; ++a) {
                                if (!nsheets[a]) continue;
                                for (int c = 0; c < 3; ++c) {
                                    if (c == a) continue;
                                    float se = sheet_avg(a, c, 0, i, j, k), ss = sheet_avg(a, c, 1, i, j, k);
                                    if (ss > 0 || se > 1) {
                                        sea[c] = sea[c] + tda[a] * (se - eps[c]);
                                        ssa[c] = ssa[c] + tda[a] * (ss - sig[c]);
                                        cnt[c] = 1;
                                    }
                                }
                            }
                            for (int c = 2; c >= 0; --c)
                                if (cnt[c]) ga_gb(ssa[c] + (plasma_in ? tpl[c] : 0.0f), sea[c], dt, t6[2 * c], t6[2 * c + 1]);
                        }
                    }
                    size_t off = (size_t)(i - 1) + (size_t)v.PX * ((size_t)(j - 1) + (size_t)NY * (k - 1));
                    cell_tuple[off] = add_tuple(t6);
                }
        P->cell_mat = cellm;
        const size_t ntup = tmap.size();
        if (ntup > 65536) throw Error(5, "more than 65536 distinct update-coefficient tuples");
        v.n_tuple = (int32_t)ntup;
        if (ntup <= 256) {
            P->idx8.resize((size_t)v.ncell);
            for (size_t q = 0; q < (size_t)v.ncell; ++q) P->idx8[q] = (uint8_t)cell_tuple[q];
            v.idx_bytes = 1;
        } else {
            P->idx16.resize((size_t)v.ncell);
            for (size_t q = 0; q < (size_t)v.ncell; ++q) P->idx16[q] = (uint16_t)cell_tuple[q];
            v.idx_bytes = 2;
        }
        cell_tuple.clear();
        cell_tuple.shrink_to_fit();

        // ---- sheets: unique planes and normal-component arrays -------------
        v.sheet_thickness = tsh;
        for (int a = 0; a < 3; ++a) {
            P->sheet_pos[a].clear();
            if (!nsheets[a]) continue;
            for (int q = 1; q <= N3[a] - 1; ++q)                       // f90:3685-3716
                for (auto& s : d.sheets)
                    if (s.axis == a && s.pos == q) { P->sheet_pos[a].push_back(q); break; }
            int Pd, Qd;
            if (a == 0) { Pd = NY; Qd = NZ; } else if (a == 1) { Pd = NX; Qd = NZ; } else { Pd = NX; Qd = NY; }
            v.sheet_dim[a][0] = Pd; v.sheet_dim[a][1] = Qd;
            size_t np = P->sheet_pos[a].size();
            P->sheet_sig[a].assign(np * Pd * Qd, 0.0f);
            P->sheet_ep[a].assign(np * Pd * Qd, 1.0f);
            for (size_t s = 0; s < np; ++s) {
                int pos = P->sheet_pos[a][s];
                for (int q = 1; q <= Qd - 1; ++q)
                    for (int p = 1; p <= Pd - 1; ++p) {
                        int i, j, k;
                        if (a == 0) { i = pos; j = p; k = q; } else if (a == 1) { j = pos; i = p; k = q; } else { k = pos; i = p; j = q; }
                        size_t o = s * Pd * Qd + (size_t)(p - 1) + (size_t)Pd * (q - 1);
                        P->sheet_sig[a][o] = sheet_avg(a, a, 1, i, j, k);
                        P->sheet_ep[a][o] = sheet_avg(a, a, 0, i, j, k);
                    }
            }
        }

        // ---- CPML (f90:3405-3577) and den_* tables (f90:3586-3666) --------
        std::vector<float> ke[6], kh[6];
        for (int f = 0; f < 6; ++f) {
            int ax = f / 2;
            cpml_face(npml[f], opt.cpml_grading, sig_max[ax], alpha_max, kappa_max, dt, P->cpml[f][0], P->cpml[f][1],
                      P->cpml[f][2], P->cpml[f][3], ke[f], kh[f]);
        }
        for (int ax = 0; ax < 3; ++ax) {
            const int N = N3[ax];
            const int n1 = npml[2 * ax], n2 = npml[2 * ax + 1];
            const float del = dd[ax];
            std::vector<float>& dh = P->den[3 + ax];
            std::vector<float>& de = P->den[ax];
            dh.assign(N - 1 > 0 ? N - 1 : 1, 0.f);
            de.assign(N - 1 > 0 ? N - 1 : 1, 0.f);
            int ii = n2 - 1;
            for (int i = 1; i <= N - 1; ++i) {
                if (i <= n1 - 1) dh[i - 1] = 1.0f / (kh[2 * ax][i - 1] * del);
                else if (i >= N + 1 - n2) { dh[i - 1] = 1.0f / (kh[2 * ax + 1][ii - 1] * del); ii = ii - 1; }
                else dh[i - 1] = 1.0f / del;
            }
            ii = n2;
            for (int i = 1; i <= N - 1; ++i) {
                if (i <= n1) de[i - 1] = 1.0f / (ke[2 * ax][i - 1] * del);
                else if (i >= N + 1 - n2) { de[i - 1] = 1.0f / (ke[2 * ax + 1][ii - 1] * del); ii = ii - 1; }
                else de[i - 1] = 1.0f / del;
            }
        }

        // ---- ports ---------------------------------------------------------
        v.antenna_amp = d.ant_amp;
        v.excitation_port = d.exc_port;
        if (d.exc_port > (int)d.ports.size()) throw Error(1, "ERROR: excitation port number exceeds the number of ports");
        for (size_t pi = 0; pi < d.ports.size(); ++pi) {
            const PortIn& in = d.ports[pi];
            aem_port_view pv{};
            pv.dir = in.dir;
            for (int a = 0; a < 3; ++a) { pv.loc[a] = in.loc[a]; pv.size[a] = in.size[a]; }
            pv.Z = in.Z;
            pv.L = in.Z / kC;                                          // f90:1136-1139
            pv.C = 1.0f / (in.Z * kC);
            pv.dl = dd[in.dir];
            pv.kI = (1.0f / pv.L) * (dt / pv.dl);
            pv.kV = (1.0f / pv.C) * (dt / pv.dl);
            pv.kS = (1.0f / std::sqrt(pv.C * pv.L)) * (dt / pv.dl);
            float s0 = (float)in.size[0], s1 = (float)in.size[1], s2 = (float)in.size[2];
            if (in.dir == 0) { pv.cur_num = s1 * s2; pv.cur_den = s0; }
            else if (in.dir == 1) { pv.cur_num = s0 * s2; pv.cur_den = s1; }
            else { pv.cur_num = s0 * s1; pv.cur_den = s2; }
            pv.gap_div = (float)in.size[in.dir] * dd[in.dir];
            pv.excited = ((int)pi + 1 == d.exc_port) ? 1 : 0;
            // every cell the port touches must exist in the field arrays
            for (int a = 0; a < 3; ++a) {
                int hiq = in.loc[a] + in.size[a] - 1;
                if (in.loc[a] < 2 || hiq > N3[a] - 1)
                    throw Error(1, "ERROR: lumped port " + std::to_string(pi + 1) + " lies outside the grid");
            }
            P->ports.push_back(pv);
        }
        v.n_ports = (int32_t)P->ports.size();

        // ---- SPICE ports (f90:1195-1286, grid capacitance f90:3800-3890) ----
        P->spice_names = d.netlist + "\n";
        for (size_t si = 0; si < d.spice.size(); ++si) {
            const SpiceIn& in = d.spice[si];
            aem_spice_view sv{};
            sv.dir = in.dir;
            for (int a = 0; a < 3; ++a) { sv.loc[a] = in.loc[a]; sv.size[a] = in.size[a]; }
            for (int a = 0; a < 3; ++a) {
                int hiq = in.loc[a] + in.size[a] - 1;
                if (in.loc[a] < 2 || hiq > N3[a] - 1)
                    throw Error(1, "ERROR: SPICE port " + std::to_string(si + 1) + " lies outside the grid");
            }
            sv.v_bias = in.v_bias;
            const float n0 = (float)in.size[0], n1 = (float)in.size[1], n2 = (float)in.size[2];
            if (in.dir == 0) { sv.cur_num = n1 * n2; sv.cur_den = n0; }
            else if (in.dir == 1) { sv.cur_num = n0 * n2; sv.cur_den = n1; }
            else { sv.cur_num = n0 * n1; sv.cur_den = n2; }
            sv.gap_div = (float)in.size[in.dir] * dd[in.dir];
            // parallel-plate capacitance of the gap from the average edge
            // permittivity, normalised to the gap length as the current is
            float cap = 0.0f;
            for (int ii = in.loc[0]; ii <= in.loc[0] + in.size[0] - 1; ++ii)
                for (int jj = in.loc[1]; jj <= in.loc[1] + in.size[1] - 1; ++jj)
                    for (int kk = in.loc[2]; kk <= in.loc[2] + in.size[2] - 1; ++kk) {
                        float eps, sig;
                        bulk_avg(in.dir, ii, jj, kk, eps, sig);
                        cap = cap + eps * kEp0;
                    }
            cap = cap / (n0 * n1 * n2);
            if (in.dir == 0) cap = cap * (n1 * dy * n2 * dz) / (n0 * dx);
            else if (in.dir == 1) cap = cap * (n0 * dx * n2 * dz) / (n1 * dy);
            else cap = cap * (n0 * dx * n1 * dy) / (n2 * dz);
            cap = cap / (float)in.size[in.dir];
            sv.cap = cap;
            P->spice.push_back(sv);
            P->spice_names += in.node + "\n" + in.isrc + "\n" + in.cap + "\n";
        }
        for (auto& e : d.spice_extra) P->spice_names += e + "\n";
        v.n_spice = (int32_t)P->spice.size();
        v.n_spice_extra = v.n_spice > 0 ? (int32_t)d.spice_extra.size() : 0;

        // ---- far field -----------------------------------------------------
        v.n_angles = d.n_angles;
        v.ic = d.ic; v.jc = d.jc; v.kc = d.kc;
        for (int a = 0; a < d.n_angles; ++a) {
            float t = d.ang_theta[a], p = d.ang_phi[a];
            float vals[6] = {t, p, std::sin(t), std::cos(t), std::sin(p), std::cos(p)};
            P->ang.insert(P->ang.end(), vals, vals + 6);
        }
        if (d.n_angles > 0 && d.pw_amp > 0 &&
            (d.ic < 1 || d.ic > NX - 1 || d.jc < 1 || d.jc > NY - 1 || d.kc < 1 || d.kc > NZ - 1))
            throw Error(1, "ERROR: far-field phase centre lies outside the grid");
        if (d.n_angles > 0)
            for (int a = 0; a < 3; ++a)
                if (!d.pbc[a] && (v.ff_lo[a] < 1 || v.ff_hi[a] + 1 > N3[a] || v.ff_lo[a] > v.ff_hi[a]))
                    throw Error(1, "ERROR: grid too small for the far-field box (needs more than 28 cells per axis)");

        // ---- periodic-slab plane sums (f90:7009-7102) ----------------------
        v.plane_sum_axis = -1;
        if (d.pbc[1] + d.pbc[2] == 2) v.plane_sum_axis = 0;
        else if (d.pbc[0] + d.pbc[2] == 2) v.plane_sum_axis = 1;
        else if (d.pbc[0] + d.pbc[1] == 2) v.plane_sum_axis = 2;
        v.ps_c1 = -1.0f * std::cos(pol) * std::cos(phi) - std::sin(pol) * std::cos(theta) * std::sin(phi);
        v.ps_c2 = std::cos(pol) * std::sin(phi) - std::sin(pol) * std::cos(theta) * std::cos(phi);

        // ---- kmax: source planes, mode weights, Bloch phases ----------------
        v.kmax = d.kmax;
        v.k_axis = -1;
        v.mode_type = d.mode_type;
        v.k_direction = d.k_direction;
        for (int a = 0; a < 3; ++a) v.k_vec[a] = d.k_count[a];
        if (d.kmax) {
            int a = -1;
            if (d.pbc[1] + d.pbc[2] == 2) a = 0;
            else if (d.pbc[0] + d.pbc[2] == 2) a = 1;
            else if (d.pbc[0] + d.pbc[1] == 2) a = 2;
            v.k_axis = a;
            if (a >= 0) {                                              // f90:1727-1806
                if (d.k_direction == 1) { v.kpl_H = lo[a] - 1; v.kpl_E = lo[a]; }
                else { v.kpl_H = hi[a]; v.kpl_E = hi[a]; }
                // the periodic pair (p, q) in the reference's order: (y,z), (x,z), (x,y)
                const int p = a == 0 ? 1 : 0, q = a == 2 ? 1 : 2;
                const float kp = d.k_count[p], kq = d.k_count[q];
                const float nrm = std::sqrt(kq * kq + kp * kp);
                if (kp == 0.0f && kq != 0.0f) { v.k_exc[p] = 0.0f; v.k_exc[q] = kq / nrm; }
                if (kp != 0.0f && kq == 0.0f) { v.k_exc[p] = kp / nrm; v.k_exc[q] = 0.0f; }
                if (kp == 0.0f && kq == 0.0f) { v.k_exc[p] = 0.0f; v.k_exc[q] = 1.0f; }
                if (kp != 0.0f && kq != 0.0f) { v.k_exc[p] = kp / nrm; v.k_exc[q] = kq / nrm; }
            }
            // E copies to the first cell carry exp(i k (N-2) d), H copies to
            // the last cell exp(-i k (N-2) d) (f90:5387-5428, 6685-6736)
            for (int b = 0; b < 3; ++b) {
                const float th = d.k_count[b] * (float)(N3[b] - 2) * dd[b];
                v.bloch_c[b] = std::cos(th);
                v.bloch_s[b] = std::sin(th);
            }
            // modal source factors per step (Inc, f90:8860-8885): the Gaussian
            // exp(-((t_spread - t + x d/c) / spread)^2 / 2) and the carrier
            // phase 2 pi f_adj (t + x d / c), with t = (n-1) dt, x = 0 on the
            // H plane and t = (n-1/2) dt, x = -1/2 on the E plane, d the cell
            // size along the source axis
            const int T = d.time_steps;
            P->ksrc.assign((size_t)4 * T, 0.0f);
            const float dn = a >= 0 ? dd[a] : 0.0f;
            const float tpf = 2.0f * kPi * v.f_adj;
            for (int n = 1; n <= T; ++n) {
                for (int ph = 0; ph < 2; ++ph) {
                    const float ctr = ph == 0 ? (float)n - 1.0f : (float)n - 0.5f;
                    const float x = ph == 0 ? 0.0f : -0.5f;
                    const float q = ((v.t_spread - (ctr * dt)) + 1.0f / kC * (x * dn)) / v.spread;
                    P->ksrc[(size_t)(2 * ph) * T + n - 1] = std::exp(-0.5f * (q * q));
                    P->ksrc[(size_t)(2 * ph + 1) * T + n - 1] = tpf * ((ctr * dt) + 1.0f / kC * (x * dn));
                }
            }
        }

        // ---- time-only source waveforms (f90:5001, 6162, 8368) -------------
        // Evaluated here once so that the device uses the same FP32 values
        // (and the same libm exp) as the reference's serial loop.
        {
            const int T = d.time_steps;
            P->src.assign((size_t)3 * T, 0.0f);
            P->src_im.assign((size_t)3 * T, 0.0f);
            if (d.kmax) {
                // complex E_or_V_1D of the kmax build (f90:8888-8900):
                // amp exp(-((t_spread - (tc-1) dt - pos/c) / spread)^2 / 2)
                // exp(i 2 pi f_adj ((tc-1) dt - pos/c))
                auto eov = [&](float amp, float tc, float pos, size_t at) {
                    const float g = ((v.t_spread - ((tc - 1.0f) * dt)) - 1.0f / kC * pos) / v.spread;
                    const float m = amp * std::exp(-0.5f * (g * g));
                    const float th = 2.0f * kPi * v.f_adj * (((tc - 1.0f) * dt) - 1.0f / kC * pos);
                    P->src[at] = m * std::cos(th);
                    P->src_im[at] = m * std::sin(th);
                };
                if (d.exc_port > 0 && d.ant_amp > 0) {
                    const float dl = P->ports[d.exc_port - 1].dl;
                    for (int n = 1; n <= T; ++n) {
                        eov(d.ant_amp, (float)n + 0.0f, 0.0f * dl, (size_t)n - 1);
                        eov(d.ant_amp, (float)n + 0.5f, 0.5f * dl, (size_t)T + n - 1);
                    }
                }
                if (d.pw_amp > 0)
                    for (int n = 1; n <= T; ++n) eov(d.pw_amp, (float)n + 0.0f, 0.0f, (size_t)2 * T + n - 1);
            } else if (d.exc_port > 0 && d.ant_amp > 0) {
                const float dl = P->ports[d.exc_port - 1].dl;
                for (int n = 1; n <= T; ++n) {
                    P->src[n - 1] = e_or_v_1d(d.ant_amp, (float)n + 0.0f, 0.0f * dl, v.spread, v.t_spread,
                                              d.pulse_type, dt);
                    P->src[(size_t)T + n - 1] = e_or_v_1d(d.ant_amp, (float)n + 0.5f, 0.5f * dl, v.spread,
                                                          v.t_spread, d.pulse_type, dt);
                }
            }
            if (d.pw_amp > 0 && !d.kmax)
                for (int n = 1; n <= T; ++n)
                    P->src[(size_t)2 * T + n - 1] = e_or_v_1d(d.pw_amp, (float)n + 0.0f, 0.0f * dx, v.spread,
                                                              v.t_spread, d.pulse_type, dt);
            v.exp_half = std::exp(0.5f);
        }

        // ---- wire up the view pointers ---------------------------------------
        v.den_ex = P->den[0].data(); v.den_ey = P->den[1].data(); v.den_ez = P->den[2].data();
        v.den_hx = P->den[3].data(); v.den_hy = P->den[4].data(); v.den_hz = P->den[5].data();
        v.tuples = P->tuples.data();
        v.tuple_idx8 = P->idx8.empty() ? nullptr : P->idx8.data();
        v.tuple_idx16 = P->idx16.empty() ? nullptr : P->idx16.data();
        for (int f = 0; f < 6; ++f) {
            v.be[f] = P->cpml[f][0].data(); v.ce[f] = P->cpml[f][1].data();
            v.bh[f] = P->cpml[f][2].data(); v.ch[f] = P->cpml[f][3].data();
        }
        for (int a = 0; a < 3; ++a) {
            v.n_sheet_planes[a] = (int32_t)P->sheet_pos[a].size();
            v.sheet_pos[a] = P->sheet_pos[a].empty() ? nullptr : P->sheet_pos[a].data();
            v.sheet_sig[a] = P->sheet_sig[a].empty() ? nullptr : P->sheet_sig[a].data();
            v.sheet_ep[a] = P->sheet_ep[a].empty() ? nullptr : P->sheet_ep[a].data();
            v.Ja[a] = P->Ja[a].empty() ? nullptr : P->Ja[a].data();
            v.Jb[a] = P->Jb[a].empty() ? nullptr : P->Jb[a].data();
        }
        v.ports = P->ports.empty() ? nullptr : P->ports.data();
        v.spice = P->spice.empty() ? nullptr : P->spice.data();
        v.spice_names = P->spice_names.c_str();
        v.ang = P->ang.empty() ? nullptr : P->ang.data();
        v.src_table = P->src.data();
        v.src_table_im = P->src_im.data();
        v.ksrc = P->ksrc.empty() ? nullptr : P->ksrc.data();
    } catch (...) {
        delete P;
        throw;
    }
    return P;
}

// E_or_V_1D of f90:8844-8857 (FP32, left-to-right).
float e_or_v_1d(float amp, float tc, float pos, float spread, float t_spread, int ptype, float dt) {
    const float x = ((t_spread - ((tc - 1.0f) * dt)) - 1.0f / kC * pos) / spread;
    if (ptype == 1) return amp * std::exp(-0.5f * (x * x));
    return amp * -1.0f * x * std::exp(0.5f) * std::exp(-0.5f * (x * x));
}

void design_coefficients(const Plan& P, const int lo[3], const int hi[3], const float* eps3, float* coef) {
    const aem_plan_view& v = P.v;
    const int n[3] = {hi[0] - lo[0] + 2, hi[1] - lo[1] + 2, hi[2] - lo[2] + 2};
    const size_t nedge = (size_t)n[0] * n[1] * n[2];
    const size_t ncb = (size_t)(n[0] - 1) * (n[1] - 1) * (n[2] - 1);
    const float dt = v.dt;
    auto cell_eps_sig = [&](int i, int j, int k, int c, float& e, float& sg) {
        const int m = P.cell_mat[(size_t)(i - 1) + (size_t)(v.NX - 1) * ((size_t)(j - 1) + (size_t)(v.NY - 1) * (k - 1))];
        sg = m ? P.deck.mats[m - 1].sig[c] : 0.0f;
        if (i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2])
            e = eps3[(size_t)c * ncb + (size_t)(i - lo[0]) +
                     (size_t)(n[0] - 1) * ((size_t)(j - lo[1]) + (size_t)(n[1] - 1) * (k - lo[2]))];
        else
            e = m ? P.deck.mats[m - 1].eps[c] : 1.0f;
    };
    for (int k = lo[2]; k <= hi[2] + 1; ++k)
        for (int j = lo[1]; j <= hi[1] + 1; ++j)
            for (int i = lo[0]; i <= hi[0] + 1; ++i) {
                const size_t e = (size_t)(i - lo[0]) + (size_t)n[0] * ((size_t)(j - lo[1]) + (size_t)n[1] * (k - lo[2]));
                for (int c = 0; c < 3; ++c) {
                    int q[4][3];
                    if (c == 0) { int t[4][3] = {{i, j, k}, {i, j - 1, k}, {i, j, k - 1}, {i, j - 1, k - 1}}; std::memcpy(q, t, sizeof q); }
                    else if (c == 1) { int t[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j, k - 1}, {i - 1, j, k - 1}}; std::memcpy(q, t, sizeof q); }
                    else { int t[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j - 1, k}, {i - 1, j - 1, k}}; std::memcpy(q, t, sizeof q); }
                    float ev[4], sv[4];
                    for (int r = 0; r < 4; ++r) cell_eps_sig(q[r][0], q[r][1], q[r][2], c, ev[r], sv[r]);
                    ga_gb(avg4(sv[0], sv[1], sv[2], sv[3]), avg4(ev[0], ev[1], ev[2], ev[3]), dt,
                          coef[(2 * c) * nedge + e], coef[(2 * c + 1) * nedge + e]);
                }
            }
}

void write_geometry(const Plan& p, const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(2, "cannot write '" + path + "'");
    size_t n = p.geom.size();
    size_t w = std::fwrite(p.geom.data(), sizeof(float), n, f);
    std::fclose(f);
    if (w != n) throw Error(2, "short write to '" + path + "'");
}

}  // namespace aem
