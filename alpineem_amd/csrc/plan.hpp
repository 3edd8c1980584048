// plan.hpp -- host-side deck and plan structures of libaem.
//
// Deck mirrors the positional input grammar of the reference solver
// (main_fdtd/fdtd_solver.f90:726-1298, documented in main_fdtd/master.py).
// Plan holds the storage behind aem_plan_view (include/aem_plan.h).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/aem_plan.h"

namespace aem {

// Deck / STOP errors carry the status they map to.
struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

// Physical constants exactly as the reference declares them (FP32 parameters,
// fdtd_solver.f90:24-26, 43).
constexpr float kC    = 299792458.0f;           // rounds to 299792448 in FP32
constexpr float kMu0  = 1.25663706E-6f;
constexpr float kEp0  = 8.85418782E-12f;
constexpr float kPi   = 3.14159265358979323846f;

struct Material {            // materials_properties(i,:,:) (f90:922-946)
    float id = 0;
    bool plasma = false;
    float eps[3] = {1, 1, 1};
    float sig[3] = {0, 0, 0};
    int npoles = 0;
    float wp[6][3] = {};     // plasma angular frequency per pole, per axis
    float gam[6][3] = {};    // collision frequency per pole, per axis
};

struct SheetMat {            // sheet_properties(j,:,:) (f90:948-972)
    float id = 0;
    float ep[3] = {1, 1, 1};
    float sig[3] = {0, 0, 0};
};

struct Object {              // blocks / spheres / cylinders (f90:995-1036)
    int kind = 0;            // 0 block, 1 sphere, 2 cylinder
    int mat_id = 0;
    int p[3] = {0, 0, 0};    // start (block, cylinder) or centre (sphere)
    int s[3] = {0, 0, 0};    // block size; sphere radius in s[0]; cylinder (length, radius) in s[0], s[1]
    int axis = 0;            // cylinder major axis 0/1/2
};

struct SheetIn {             // sheets_x/y/z (f90:1059-1093)
    int axis = 0;
    int mat_id = 0;
    int pos = 0;
    int start[2] = {0, 0};
    int size[2] = {0, 0};
};

struct PortIn {              // ports(i,:,:) (f90:1116-1193)
    int gridded = 0;
    int dir = 2;
    float Z = 50;
    int loc[3] = {0, 0, 0};
    int size[3] = {0, 0, 0}; // after the +1 on the transverse axes
};

struct SpiceIn {             // ports_spice(i,:,:) and names (f90:1195-1286)
    int gridded = 0;
    int dir = 2;
    int loc[3] = {0, 0, 0};
    int size[3] = {0, 0, 0}; // after the +1 on the transverse axes
    float v_bias = 0;        // ports_spice(i,4,1): subtracted from the node voltage
    std::string node, isrc, cap;
};

struct Deck {
    int nx = 0, ny = 0, nz = 0;
    float dx = 0, dy = 0, dz = 0;
    int time_steps = 0;
    float factor = 1;
    float f_center = 0;
    int pulse_type = 0;
    int bnd[6] = {0, 0, 0, 0, 0, 0};
    int pbc[3] = {0, 0, 0};
    int is_mirror = 0, mirror_type = 0, mirror_height = 0;
    int kmax = 0;                        // deck of the kmax build (complex fields, Bloch-periodic walls)
    int mode_type = 0;                   // kmax plane wave: 0 TE (E source), 1 TM (H source)
    int k_direction = 0;                 // kmax plane wave: 1 '+', 0 '-' (antenna decks keep 0, f90:189)
    float k_count[3] = {0, 0, 0};        // kmax Bloch wavenumbers of the periodic axes (rad/m)
    int excitation = 0;                  // 0 none, 1 plane wave, 2 antenna
    float pw_amp = 0, pw_delay = 0;
    float theta = 0, phi = 0, pol = 0;   // degrees as read
    float ant_amp = 0, ant_delay = 0;
    int exc_port = 0;
    int n_angles = 0;
    int ic = 0, jc = 0, kc = 0;
    std::vector<float> ang_theta, ang_phi;  // radians, FP32 (f90:895-896)
    int video_on = 0, slice = 0, slice_location = 0;
    std::string export_all = "no";
    std::string filename = "data.dat";
    std::vector<Material> mats;
    std::vector<SheetMat> smats;
    float sheet_thickness = 0;
    int num_poles = 0;
    int is_plasma = 0;
    int use_opt_geom = 0;
    std::string opt_geom_file;
    std::vector<Object> objects;
    std::vector<SheetIn> sheets;
    std::vector<PortIn> ports;
    int num_spice_ports = 0;
    std::string netlist;                 // SPICE section, read by the SPICE build only
    std::vector<SpiceIn> spice;
    std::vector<std::string> spice_extra;   // additional node names
};

// kmax: the grammar of the reference's kmax build (fdtd_solver.f90:832-882
// in place of 769-830).
Deck parse_deck(const std::string& text, bool kmax = false);

struct Plan {
    Deck deck;
    std::string base_dir;
    aem_plan_options opt{};                       // setup constants the plan was built with
    aem_plan_view v{};

    std::vector<float> den[6];                    // ex ey ez hx hy hz
    std::vector<float> tuples;
    std::vector<uint8_t> idx8;
    std::vector<uint16_t> idx16;
    std::vector<float> cpml[6][4];                // be ce bh ch per face
    std::vector<int32_t> sheet_pos[3];
    std::vector<float> sheet_sig[3], sheet_ep[3];
    std::vector<float> Ja[3], Jb[3];
    std::vector<aem_port_view> ports;
    std::vector<aem_spice_view> spice;
    std::string spice_names;
    std::vector<float> ang;
    std::vector<float> src;                       // [3][time_steps] source waveforms
    std::vector<float> src_im;                    // kmax: their imaginary parts
    std::vector<float> ksrc;                      // kmax: [4][time_steps] modal source factors

    // rasterised geometry kept for the _geometry.bin writer (f90:2355-2357)
    std::vector<float> geom;                      // (4, nx, ny, nz) Fortran order

    // bulk material index of every cell (0 = background, m = deck.mats[m-1]),
    // (i-1) + (NX-1)*((j-1) + (NY-1)*(k-1)); kept for permittivity gradients
    std::vector<uint16_t> cell_mat;
};

// The reference's setup constants (f90:34-39, 1686-1694).
void default_options(aem_plan_options& o);
// Throws Error(1, ...) for values outside the accepted ranges.
void check_options(const aem_plan_options& o);

Plan* build_plan(const std::string& deck_text, const std::string& base_dir, const aem_plan_options& opt);

// Ga/Gb of every edge of the box [lo, hi+1] when the cells of [lo, hi] have
// the per-axis relative permittivity eps3 ([3][cells of the box], x fastest)
// and every other cell its deck material; the setup's four-cell averaging and
// formulas for non-dispersive media without sheets (f90:2516-2527,
// 2837-2853).  coef is [6][edges]: ga_x, gb_x, ga_y, gb_y, ga_z, gb_z.
void design_coefficients(const Plan& P, const int lo[3], const int hi[3], const float* eps3, float* coef);
void write_geometry(const Plan& p, const std::string& path);

// FP32 pulse functions of the reference (fdtd_solver.f90:8825-8857).
float e_or_v_1d(float amp, float tc, float pos, float spread, float t_spread, int ptype, float dt);

}  // namespace aem
