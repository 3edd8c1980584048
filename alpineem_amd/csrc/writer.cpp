// writer.cpp -- data.dat writer of libaem (aem_write_outputs, include/aem.h).
//
// Reproduces the record order and labels of the reference writer
// (main_fdtd/fdtd_solver.f90:8536-8645) so that
// utility_scripts/post_processor.py:43-160, which reads the file by line and
// whitespace-separated column, consumes it unchanged.  Reals are printed in
// the shortest form that round-trips FP32 (the reference's list-directed
// output does the same), with an upper-case exponent.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/aem.h"
#include "api_internal.hpp"
#include "plan.hpp"

namespace {

std::string fmt(float x) {
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, x);
    std::string s(buf, r.ptr);
    for (auto& ch : s)
        if (ch == 'e') ch = 'E';
    return s;
}

struct File {
    std::FILE* f;
    explicit File(const std::string& p) : f(std::fopen(p.c_str(), "w")) {
        if (!f) throw aem::Error(AEM_ERR_IO, "cannot write '" + p + "'");
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void line(const std::string& s) { std::fprintf(f, " %s\n", s.c_str()); }
};

void get(aem_sim* s, int kind, int idx, std::vector<float>& out) {
    aem_status st = aem_get_series(s, kind, idx, out.data(), (int64_t)out.size());
    if (st != AEM_OK) throw aem::Error(st, aem_last_error());
}

// T lines "re im" of a kmax simulation's complex series
void complex_lines(File& f, aem_sim* s, int kind, int idx, int T) {
    std::vector<float> re(T), im(T);
    get(s, kind, idx, re);
    get(s, kind | AEM_IMAG, idx, im);
    for (int n = 0; n < T; ++n) f.line(fmt(re[n]) + " " + fmt(im[n]));
}

// data.dat of the kmax build (f90:8648-8785): the same header with "N/A"
// for the minimum steps and pulse type, the k vector and mode records,
// complex series as "re im" lines, and the TE/TM plane sums, whose 'Refl'
// records hold the sums of the plane on the launch side for k_direction 1
// and of the far plane for k_direction 0.
void write_kmax(File& f, aem_sim* s, const aem_plan_view& v, const aem::Deck& d) {
    const int T = v.time_steps;
    f.line("x steps =  " + std::to_string(v.nx));
    f.line("y steps =  " + std::to_string(v.ny));
    f.line("z steps =  " + std::to_string(v.nz));
    f.line("delta x,y,z =  " + fmt(v.dx) + " " + fmt(v.dy) + " " + fmt(v.dz));
    f.line("time step =  " + fmt(v.dt));
    f.line("number of time steps =  " + std::to_string(T));
    f.line("minimum time steps needed N/A");
    f.line("freq center =  " + fmt(v.f_center));
    f.line("pulse type = N/A");
    f.line("slice =  " + std::to_string(v.slice));
    f.line("slice location =  " + std::to_string(v.slice_location));
    f.line("video size_a =  " + std::to_string(v.vid_size1));
    f.line("video size_b =  " + std::to_string(v.vid_size2));
    f.line("number of far field angles =  " + std::to_string(v.n_angles));
    f.line("number of internal ports =  " + std::to_string(v.n_ports));
    f.line("number of spice ports =  0");
    f.line("simulation type " + std::to_string(v.pbc[0]) + " " + std::to_string(v.pbc[1]) + " " +
           std::to_string(v.pbc[2]) + " " + std::to_string(v.k_direction));
    f.line("simulation is k vector method w/ k1,k2:");
    if (v.k_axis >= 0) {
        const int p = v.k_axis == 0 ? 1 : 0, q = v.k_axis == 2 ? 1 : 2;
        f.line(fmt(v.k_vec[p]) + " " + fmt(v.k_vec[q]));
    }
    f.line("mode type is 0 for TE and 1 for TM: " + std::to_string(v.mode_type));
    if (d.excitation == 1) f.line("incident E-field wave: (not used in post processor bc r(k))");
    if (d.excitation == 2) f.line("incident voltage wave at port  " + std::to_string(d.exc_port));
    complex_lines(f, s, AEM_SERIES_INCIDENT, 0, T);
    const bool fwd = v.k_direction == 1;
    const int sec[4] = {fwd ? AEM_SERIES_REFLECTED : AEM_SERIES_TRANSMITTED,
                        fwd ? AEM_SERIES_REFLECTED_TM : AEM_SERIES_TRANSMITTED_TM,
                        fwd ? AEM_SERIES_TRANSMITTED : AEM_SERIES_REFLECTED,
                        fwd ? AEM_SERIES_TRANSMITTED_TM : AEM_SERIES_REFLECTED_TM};
    const char* lab[4] = {"TE Refl Wave:", "TM Refl Wave:", "TE Trans Wave:", "TM Trans Wave:"};
    for (int q = 0; q < 4; ++q) {
        f.line(lab[q]);
        if (v.k_axis >= 0) {
            complex_lines(f, s, sec[q], 0, T);
        } else {
            for (int n = 0; n < T; ++n) f.line("0 0");   // no periodic pair: the sums stay zero
        }
    }
    for (int q = 0; q < v.n_ports; ++q) {
        f.line("recieved voltage at port  " + std::to_string(q + 1) + " w/ Z");
        f.line(fmt(std::sqrt(v.ports[q].L / v.ports[q].C)));
        complex_lines(f, s, AEM_SERIES_PORT_V, q, T);
    }
}

// <stem>_video_{E,H}{x,y,z}.bin: the whole (size_a, size_b, steps) array of
// a component as ONE Fortran unformatted sequential record -- a 4-byte byte
// count, the floats in Fortran order, the count again (f90:8788-8807, in the
// reference's file order Ex, Hx, Ey, Hy, Ez, Hz).  Frames are copied from
// the device in chunks.
void write_video(aem_sim* s, const aem_plan_view& v, const std::string& stem) {
    static const int order[6] = {0, 3, 1, 4, 2, 5};
    static const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    const int64_t ps = (int64_t)v.vid_size1 * v.vid_size2, T = v.time_steps;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(T, (64ll << 20) / std::max<int64_t>(ps * 4, 1)));
    std::vector<float> buf((size_t)(chunk * ps));
    for (int c : order) {
        const std::string p = stem + "_video_" + names[c] + ".bin";
        std::FILE* f = std::fopen(p.c_str(), "wb");
        if (!f) throw aem::Error(AEM_ERR_IO, "cannot write '" + p + "'");
        const int32_t len = (int32_t)(ps * T * 4);
        bool ok = std::fwrite(&len, 4, 1, f) == 1;
        for (int64_t n0 = 1; ok && n0 <= T; n0 += chunk) {
            const int64_t m = std::min(chunk, T - n0 + 1);
            aem_status st = aem_get_video(s, c, n0, m, buf.data(), m * ps);
            if (st != AEM_OK) {
                std::fclose(f);
                throw aem::Error(st, aem_last_error());
            }
            ok = std::fwrite(buf.data(), 4, (size_t)(m * ps), f) == (size_t)(m * ps);
        }
        ok = ok && std::fwrite(&len, 4, 1, f) == 1;
        if (std::fclose(f) != 0 || !ok) throw aem::Error(AEM_ERR_IO, "writing '" + p + "' failed");
    }
}

}  // namespace

extern "C" aem_status aem_write_outputs(aem_sim* s, const char* path) {
    const aem::Plan* plan = s ? aem::sim_plan(s) : nullptr;
    if (!plan) {
        aem::set_error("aem_write_outputs: null handle");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        const aem_plan_view& v = plan->v;
        const std::string name = path ? std::string(path) : plan->deck.filename;
        const int T = v.time_steps;
        aem_status fs = aem_finalize(s);
        if (fs != AEM_OK) throw aem::Error(fs, aem_last_error());
        File f(name);
        const aem::Deck& d = plan->deck;
        if (v.kmax) {
            write_kmax(f, s, v, d);
            aem::write_geometry(*plan, (name.size() > 4 ? name.substr(0, name.size() - 4) : name) + "_geometry.bin");
            return;
        }
        f.line("x steps =  " + std::to_string(v.nx));
        f.line("y steps =  " + std::to_string(v.ny));
        f.line("z steps =  " + std::to_string(v.nz));
        f.line("delta x,y,z =  " + fmt(v.dx) + " " + fmt(v.dy) + " " + fmt(v.dz));
        f.line("time step =  " + fmt(v.dt));
        f.line("number of time steps =  " + std::to_string(T));
        f.line("minimum time steps needed (but more recommended) =  " + fmt(v.min_steps));
        f.line("freq parameter =  " + fmt(v.f_center));
        f.line("pulse type =  " + std::to_string(v.pulse_type));
        f.line("slice =  " + std::to_string(v.slice));
        f.line("slice location =  " + std::to_string(v.slice_location));
        f.line("video size_a =  " + std::to_string(v.vid_size1));
        f.line("video size_b =  " + std::to_string(v.vid_size2));
        f.line("number of far field angles =  " + std::to_string(v.n_angles));
        f.line("number of internal ports =  " + std::to_string(v.n_ports));
        f.line("number of spice ports =  " + std::to_string(v.n_spice + v.n_spice_extra));
        f.line("simulation type " + std::to_string(v.pbc[0]) + " " + std::to_string(v.pbc[1]) + " " +
               std::to_string(v.pbc[2]));
        std::vector<float> buf(T);
        if (d.excitation == 1) f.line("incident E-field wave:");
        if (d.excitation == 2) f.line("incident voltage wave at port  " + std::to_string(d.exc_port));
        get(s, AEM_SERIES_INCIDENT, 0, buf);
        for (int n = 0; n < T; ++n) f.line(fmt(buf[n]));
        if (v.pbc[0] + v.pbc[1] + v.pbc[2] == 2 && d.excitation == 1) {       // f90:8573-8596
            const float pi = aem::kPi;
            bool straight = (v.pbc[0] == 0 && v.phi == 0) ||
                            (v.pbc[1] == 0 && (v.phi > 89 * pi / 180 && v.phi < 91 * pi / 180)) ||
                            (v.pbc[2] == 0 && v.theta == 0);
            std::vector<float> r(T), t(T);
            get(s, AEM_SERIES_REFLECTED, 0, r);
            get(s, AEM_SERIES_TRANSMITTED, 0, t);
            f.line("reflected wave for PBC:");
            for (int n = 0; n < T; ++n) f.line(fmt(straight ? r[n] : t[n]));
            f.line("transmitted wave for PBC:");
            for (int n = 0; n < T; ++n) f.line(fmt(straight ? t[n] : r[n]));
        }
        for (int a = 0; a < v.n_angles; ++a) {
            const float th = 180.0f / aem::kPi * v.ang[6 * a], ph = 180.0f / aem::kPi * v.ang[6 * a + 1];
            f.line("far field rE theta polarization at theta,phi:");
            f.line(fmt(th) + " " + fmt(ph));
            get(s, AEM_SERIES_E_THETA, a, buf);
            for (int n = 0; n < T; ++n) f.line(fmt(buf[n]));
            f.line("far field rE phi polarization at theta,phi:");
            f.line(fmt(th) + " " + fmt(ph));
            get(s, AEM_SERIES_E_PHI, a, buf);
            for (int n = 0; n < T; ++n) f.line(fmt(buf[n]));
        }
        for (int q = 0; q < v.n_ports; ++q) {
            f.line("recieved voltage at port  " + std::to_string(q + 1) + " w/ Z");
            f.line(fmt(std::sqrt(v.ports[q].L / v.ports[q].C)));
            get(s, AEM_SERIES_PORT_V, q, buf);
            for (int n = 0; n < T; ++n) f.line(fmt(buf[n]));
        }
        for (int q = 0; q < v.n_spice + v.n_spice_extra; ++q) {              // f90:8619-8627 (SPICE build)
            f.line("recieved voltage at spice port  " + std::to_string(q + 1) + " w/ Z");
            f.line("1");
            get(s, AEM_SERIES_SPICE_V, q, buf);
            for (int n = 0; n < T; ++n) f.line(fmt(buf[n]));
        }
        if (v.n_angles > 0) {
            std::vector<float> ft(v.n_angles);
            aem_status st = aem_get_ff_time(s, ft.data(), v.n_angles);
            if (st != AEM_OK) throw aem::Error(st, aem_last_error());
            f.line("far field phase correction time information:");
            for (float x : ft) f.line(fmt(x));
        }
        if (v.n_angles > 0 && v.plane_wave_amp > 0) {
            f.line("incident for far field phase centering:");
            get(s, AEM_SERIES_FF_PC, 0, buf);
            for (int n = 0; n < T; ++n) f.line(fmt(buf[n]));
        }
        const std::string stem = name.size() > 4 ? name.substr(0, name.size() - 4) : name;
        aem::write_geometry(*plan, stem + "_geometry.bin");
        if (v.video_on) write_video(s, v, stem);
    });
}
