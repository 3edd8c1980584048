// deck.cpp -- parser of the reference's positional inputs.txt deck.
//
// Grammar follows the read sequence of main_fdtd/fdtd_solver.f90:726-1298
// (non-kmax, non-SPICE build) and the comments of main_fdtd/master.py.  Each
// Fortran "read(1,*)" is list-directed: it starts on a new record, takes
// comma/blank separated values (continuing onto following records when a
// record runs short) and discards what is left of its last record.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <sstream>

#include "plan.hpp"

namespace aem {
namespace {

class Reader {
  public:
    explicit Reader(const std::string& text) {
        std::string line;
        std::istringstream is(text);
        while (std::getline(is, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            lines_.push_back(line);
        }
    }

    // One list-directed read of n items.
    std::vector<std::string> items(size_t n, const char* what) {
        std::vector<std::string> out;
        while (out.size() < n) {
            if (cur_ >= lines_.size())
                throw Error(1, std::string("ERROR: end of deck while reading ") + what);
            split(lines_[cur_++], out, n);
        }
        return out;
    }

    // read(1,'(a)'): the whole next record.
    std::string line(const char* what) {
        if (cur_ >= lines_.size())
            throw Error(1, std::string("ERROR: end of deck while reading ") + what);
        return lines_[cur_++];
    }

    int geti(const char* what) { return to_int(items(1, what)[0], what); }
    float getf(const char* what) { return to_float(items(1, what)[0], what); }
    std::string gets(const char* what) { return items(1, what)[0]; }
    void get3f(float* v, const char* what) {
        auto t = items(3, what);
        for (int i = 0; i < 3; ++i) v[i] = to_float(t[i], what);
    }
    void getni(int* v, int n, const char* what) {
        auto t = items(n, what);
        for (int i = 0; i < n; ++i) v[i] = to_int(t[i], what);
    }

    static int to_int(const std::string& s, const char* what) {
        char* end = nullptr;
        long v = std::strtol(s.c_str(), &end, 10);
        if (s.empty() || *end != '\0')
            throw Error(1, std::string("ERROR: bad integer '") + s + "' for " + what);
        return (int)v;
    }
    static float to_float(std::string s, const char* what) {
        for (auto& ch : s)
            if (ch == 'd' || ch == 'D') ch = 'E';
        char* end = nullptr;
        float v = std::strtof(s.c_str(), &end);
        if (s.empty() || *end != '\0')
            throw Error(1, std::string("ERROR: bad real '") + s + "' for " + what);
        return v;
    }

  private:
    static void split(const std::string& l, std::vector<std::string>& out, size_t n) {
        size_t i = 0;
        while (i < l.size() && out.size() < n) {
            while (i < l.size() && (l[i] == ',' || std::isspace((unsigned char)l[i]))) {
                // a comma directly after another separator would be a null
                // value in Fortran; decks never use it
                ++i;
            }
            if (i >= l.size()) break;
            if (l[i] == '/') return;  // list-directed terminator
            if (l[i] == '\'' || l[i] == '"') {
                char q = l[i++];
                size_t j = l.find(q, i);
                if (j == std::string::npos) j = l.size();
                out.push_back(l.substr(i, j - i));
                i = j + 1;
                continue;
            }
            size_t j = i;
            while (j < l.size() && l[j] != ',' && !std::isspace((unsigned char)l[j])) ++j;
            std::string tok = l.substr(i, j - i);
            // repeat count r*c
            size_t star = tok.find('*');
            if (star != std::string::npos && star > 0 &&
                std::all_of(tok.begin(), tok.begin() + star, ::isdigit)) {
                int r = std::atoi(tok.substr(0, star).c_str());
                for (int k = 0; k < r && out.size() < n; ++k) out.push_back(tok.substr(star + 1));
            } else {
                out.push_back(tok);
            }
            i = j;
        }
    }

    std::vector<std::string> lines_;
    size_t cur_ = 0;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

int axis_of(const std::string& s, const char* what) {
    std::string t = trim(s);
    if (t == "x") return 0;
    if (t == "y") return 1;
    if (t == "z") return 2;
    throw Error(1, std::string("ERROR: bad axis '") + t + "' for " + what);
}

}  // namespace

Deck parse_deck(const std::string& text, bool kmax) {
    Reader r(text);
    Deck d;
    d.kmax = kmax ? 1 : 0;
    int n3[3];
    r.getni(n3, 3, "grid size");                       // f90:728
    d.nx = n3[0]; d.ny = n3[1]; d.nz = n3[2];
    float s3[3];
    r.get3f(s3, "cell size");                          // f90:729
    d.dx = s3[0]; d.dy = s3[1]; d.dz = s3[2];
    d.time_steps = r.geti("time steps");               // f90:730
    d.factor = r.getf("time step reduction factor");   // f90:731
    d.f_center = r.getf("frequency parameter");        // f90:732
    d.pulse_type = r.geti("pulse type");               // f90:733
    r.getni(d.bnd, 6, "boundary types");               // f90:735
    for (int a = 0; a < 3; ++a) d.pbc[a] = (d.bnd[2 * a] == 1 && d.bnd[2 * a + 1] == 1) ? 1 : 0;
    if (d.nx < 1 || d.ny < 1 || d.nz < 1 || d.time_steps < 1)
        throw Error(1, "ERROR: grid size and time steps must be positive");
    if (!(d.dx > 0 && d.dy > 0 && d.dz > 0)) throw Error(1, "ERROR: cell sizes must be positive");

    std::string mirror = trim(r.gets("IGP flag"));     // f90:746-765
    if (mirror == "IGP") {
        d.is_mirror = 1;
        std::string mt = trim(r.gets("IGP wall"));
        const char* names[6] = {"-x", "+x", "-y", "+y", "-z", "+z"};
        for (int i = 0; i < 6; ++i)
            if (mt == names[i]) d.mirror_type = i + 1;
        d.mirror_height = r.geti("IGP height");
    }

    std::string exc = trim(r.line("excitation type"));  // f90:770 read(1,'(a)')
    if (exc.size() > 15) exc = trim(exc.substr(0, 15));
    if (kmax) {                                           // f90:832-882
        if (exc == "plane wave") {
            d.excitation = 1;
            auto t = r.items(2, "plane wave amplitude, delay");
            d.pw_amp = Reader::to_float(t[0], "amplitude");
            d.pw_delay = Reader::to_float(t[1], "delay");
            const std::string mt = trim(r.gets("mode type"));
            if (mt == "TE") d.mode_type = 0;
            else if (mt == "TM") d.mode_type = 1;
            const std::string kd = trim(r.gets("propagation direction"));
            if (kd == "-") d.k_direction = 0;
            else if (kd == "+") d.k_direction = 1;
            else throw Error(1, "ERROR: INVALID PROPOGATION DIRECTION DETECTED");
        } else if (exc == "antenna") {
            d.excitation = 2;
            auto t = r.items(2, "antenna amplitude, delay");
            d.ant_amp = Reader::to_float(t[0], "amplitude");
            d.ant_delay = Reader::to_float(t[1], "delay");
            d.exc_port = r.geti("excitation port");
            if (d.exc_port == 0) { d.ant_amp = 0; d.ant_delay = 0; }
        }
        int a0 = -1, a1 = -1;
        if (d.pbc[1] + d.pbc[2] == 2) { a0 = 1; a1 = 2; }
        else if (d.pbc[0] + d.pbc[2] == 2) { a0 = 0; a1 = 2; }
        else if (d.pbc[0] + d.pbc[1] == 2) { a0 = 0; a1 = 1; }
        if (a0 >= 0) {
            auto t = r.items(2, "k vector");
            d.k_count[a0] = Reader::to_float(t[0], "k vector");
            d.k_count[a1] = Reader::to_float(t[1], "k vector");
        }
        const int np = d.pbc[0] + d.pbc[1] + d.pbc[2];
        if (np == 2 && d.is_mirror == 1)
            throw Error(1, "ERROR: USER ADDED AN IGP CONDITION TO A UNIT CELL W/ 2 SETS OF PBC - THIS IS NOT NEEDED AND INVALID");
        if (np == 1) throw Error(1, "ERROR: USER ONLY ADDED 1 PBC W/ KMAX - THIS IS INVALID");
    } else if (exc == "plane wave") {
        d.excitation = 1;
        auto t = r.items(2, "plane wave amplitude, delay");   // f90:773
        d.pw_amp = Reader::to_float(t[0], "amplitude");
        d.pw_delay = Reader::to_float(t[1], "delay");
        float a3[3];
        r.get3f(a3, "incident angles");                        // f90:774
        d.theta = a3[0]; d.phi = a3[1]; d.pol = a3[2];
    } else if (exc == "antenna") {
        d.excitation = 2;
        auto t = r.items(2, "antenna amplitude, delay");      // f90:800
        d.ant_amp = Reader::to_float(t[0], "amplitude");
        d.ant_delay = Reader::to_float(t[1], "delay");
        d.exc_port = r.geti("excitation port");               // f90:801
        if (d.exc_port == 0) { d.ant_amp = 0; d.ant_delay = 0; }
    }

    // input checks of f90:809-829 (the pbc==3 case is only a warning there)
    int npbc = d.pbc[0] + d.pbc[1] + d.pbc[2];
    if (!kmax && npbc == 2 && d.is_mirror == 1)
        throw Error(1, "ERROR: USER ADDED AN IGP CONDITION TO A UNIT CELL W/ 2 SETS OF PBC - THIS IS NOT NEEDED AND INVALID");
    if (!kmax && d.pbc[0] == 1 && d.pw_amp > 0)
        if (std::fmod(d.theta, 180.0f) != 0.0f && std::fmod(d.phi, 180.0f) != 90.0f)
            throw Error(1, "ERROR: INCIDENT PLANE WAVE W/ OBLIQUE ANGLE W/ RESPECT TO PBC DETECTED - USE KMAX");
    if (!kmax && d.pbc[1] == 1 && d.pw_amp > 0)
        if (std::fmod(d.theta, 180.0f) != 0.0f && std::fmod(d.phi, 180.0f) != 0.0f)
            throw Error(1, "ERROR: INCIDENT PLANE WAVE W/ OBLIQUE ANGLE W/ RESPECT TO PBC DETECTED - USE KMAX");
    if (!kmax && d.pbc[2] == 1 && d.pw_amp > 0)
        if (std::fmod(d.theta, 180.0f) != 90.0f)
            throw Error(1, "ERROR: INCIDENT PLANE WAVE W/ OBLIQUE ANGLE W/ RESPECT TO PBC DETECTED - USE KMAX");

    d.n_angles = r.geti("number of far field angles");         // f90:884
    if (d.n_angles < 0) throw Error(1, "ERROR: negative number of far field angles");
    if (d.n_angles > 0) {
        int c3[3];
        r.getni(c3, 3, "far field phase centre");               // f90:891
        d.ic = c3[0]; d.jc = c3[1]; d.kc = c3[2];
    }
    for (int i = 0; i < d.n_angles; ++i) {
        auto t = r.items(2, "far field angle");                  // f90:894-896
        float th = Reader::to_float(t[0], "theta"), ph = Reader::to_float(t[1], "phi");
        d.ang_theta.push_back(kPi / 180.0f * th);
        d.ang_phi.push_back(kPi / 180.0f * ph);
    }

    std::string vid = trim(r.gets("video flag"));              // f90:899-912
    if (vid == "yes") {
        d.video_on = 1;
        d.slice = axis_of(r.gets("slice axis"), "slice axis");
        d.slice_location = r.geti("slice location");
    }
    d.export_all = trim(r.gets("full field export flag"));     // f90:913
    d.filename = trim(r.gets("output file name"));             // f90:915
    if (d.filename.size() > 30) d.filename = d.filename.substr(0, 30);

    int ntot = r.geti("number of materials");                  // f90:917
    for (int ii = 0; ii < ntot; ++ii) {
        std::string mtype = trim(r.gets("material type"));     // f90:927
        if (mtype == "volume") {
            Material m;
            std::string kind = trim(r.gets("material kind"));  // f90:931
            m.id = r.getf("material id");
            r.get3f(m.eps, "relative epsilon");
            r.get3f(m.sig, "sigma");
            if (kind == "plasma") {
                m.plasma = true;
                d.is_plasma = 1;
                m.npoles = r.geti("number of poles");
                if (m.npoles < 0 || m.npoles > 6) throw Error(1, "ERROR: at most 6 Drude poles are supported");
                for (int p = 0; p < m.npoles; ++p) {
                    r.get3f(m.wp[p], "plasma frequency");
                    r.get3f(m.gam[p], "collision frequency");
                }
                d.num_poles = std::max(d.num_poles, m.npoles);
            }
            d.mats.push_back(m);
        } else if (mtype == "sheet") {
            SheetMat s;
            s.id = r.getf("sheet material id");                // f90:955
            float imped = r.getf("sheet impedance");           // f90:956
            d.sheet_thickness = std::min(std::min(d.dx, d.dy), d.dz) / 5000.0f;   // f90:957
            for (int a = 0; a < 3; ++a) {
                s.ep[a] = 1.0f;
                s.sig[a] = (imped == 0) ? 1E20f : 1.0f / (imped * d.sheet_thickness);  // f90:961-971
            }
            d.smats.push_back(s);
        } else {
            throw Error(1, "ERROR: unknown material type '" + mtype + "'");
        }
    }

    std::string og = trim(r.gets("optional geometry flag"));   // f90:990-993
    if (og == "yes") {
        d.use_opt_geom = 1;
        d.opt_geom_file = trim(r.gets("optional geometry file"));
    }

    int nobj = r.geti("number of objects");                    // f90:995
    for (int ii = 0; ii < nobj; ++ii) {
        std::string ot = trim(r.gets("object type"));
        Object o;
        if (ot == "block") {
            o.kind = 0;
            o.mat_id = r.geti("block material");
            r.getni(o.p, 3, "block start");
            r.getni(o.s, 3, "block size");
        } else if (ot == "sphere") {
            o.kind = 1;
            o.mat_id = r.geti("sphere material");
            r.getni(o.p, 3, "sphere centre");
            o.s[0] = r.geti("sphere radius");
        } else if (ot == "cylinder") {
            o.kind = 2;
            o.mat_id = r.geti("cylinder material");
            o.axis = axis_of(r.gets("cylinder axis"), "cylinder axis");
            r.getni(o.p, 3, "cylinder start");
            r.getni(o.s, 2, "cylinder length, radius");
        } else {
            throw Error(1, "ERROR: unknown object type '" + ot + "'");
        }
        d.objects.push_back(o);
    }

    int nsh = r.geti("number of sheets");                      // f90:1059
    for (int ii = 0; ii < nsh; ++ii) {
        SheetIn s;
        s.axis = axis_of(r.gets("sheet normal"), "sheet normal");
        s.mat_id = r.geti("sheet material");
        s.pos = r.geti("sheet location");
        r.getni(s.start, 2, "sheet start");
        r.getni(s.size, 2, "sheet size");
        d.sheets.push_back(s);
    }

    int np = r.geti("number of ports");                        // f90:1116
    for (int i = 0; i < np; ++i) {
        PortIn p;
        std::string pt = trim(r.gets("port type"));
        if (pt == "basic") {
            p.dir = axis_of(r.gets("port direction"), "port direction");
            p.Z = r.getf("port impedance");
            r.getni(p.loc, 3, "port location");
            if (p.loc[0] == 1 || p.loc[1] == 1 || p.loc[2] == 1)
                throw Error(1, "ERROR: FDTD SOFTWARE DOES NOT CURRENTLY SUPPORT PLACING AN INTERNAL LUMPED PORT AT A FIRST CUBE");
            r.getni(p.size, 3, "port size");
            for (int a = 0; a < 3; ++a)                          // f90:1145-1155
                if (a != p.dir) p.size[a] += 1;
        } else if (pt == "gridded") {
            p.gridded = 1;
            std::string dir = trim(r.gets("gridded port direction"));
            p.Z = r.getf("port impedance");
            r.getni(p.loc, 3, "port location");
            if (p.loc[0] == 1 || p.loc[1] == 1 || p.loc[2] == 1)
                throw Error(1, "ERROR: FDTD SOFTWARE DOES NOT CURRENTLY SUPPORT PLACING AN INTERNAL LUMPED PORT AT A FIRST CUBE");
            r.getni(p.size, 2, "gridded port size");
            r.gets("gridded feed file");
        } else {
            throw Error(1, "ERROR: unknown port type '" + pt + "'");
        }
        d.ports.push_back(p);
    }
    d.num_spice_ports = r.geti("number of spice ports");      // f90:1195
    // The SPICE section (f90:1196-1286, compiled with use_spice_version):
    // netlist name, then per port its type, direction, location, size, DC
    // bias and names, then the additional nodes.  The reference reads the
    // location, size and bias into a real array.
    if (d.num_spice_ports > 0) {
        d.netlist = trim(r.gets("SPICE netlist"));
        for (int i = 0; i < d.num_spice_ports; ++i) {
            SpiceIn p;
            std::string pt = trim(r.gets("SPICE port type"));
            auto geti_real = [&](int* v, int n, const char* what) {
                auto t = r.items((size_t)n, what);
                for (int q = 0; q < n; ++q) v[q] = (int)Reader::to_float(t[q], what);
            };
            if (pt == "basic") {
                p.dir = axis_of(r.gets("SPICE port direction"), "SPICE port direction");
                geti_real(p.loc, 3, "SPICE port location");
                if (p.loc[0] == 1 || p.loc[1] == 1 || p.loc[2] == 1)
                    throw Error(1, "ERROR: FDTD SOFTWARE DOES NOT CURRENTLY SUPPORT PLACING A SPICE LUMPED PORT AT A FIRST CUBE");
                geti_real(p.size, 3, "SPICE port size");
                p.v_bias = r.getf("SPICE port DC bias");
                p.node = trim(r.gets("SPICE port node"));
                p.isrc = trim(r.gets("SPICE port current source"));
                p.cap = trim(r.gets("SPICE port capacitance"));
                for (int a = 0; a < 3; ++a)                          // f90:1225-1235
                    if (a != p.dir) p.size[a] += 1;
            } else if (pt == "gridded") {
                p.gridded = 1;
                r.gets("SPICE gridded port direction");
                geti_real(p.loc, 3, "SPICE port location");
                if (p.loc[0] == 1 || p.loc[1] == 1 || p.loc[2] == 1)
                    throw Error(1, "ERROR: FDTD SOFTWARE DOES NOT CURRENTLY SUPPORT PLACING A SPICE LUMPED PORT AT A FIRST CUBE");
                geti_real(p.size, 2, "SPICE gridded port size");
                r.gets("SPICE gridded feed file");
                p.v_bias = r.getf("SPICE port DC bias");
                p.node = trim(r.gets("SPICE port node"));
                p.isrc = trim(r.gets("SPICE port current source"));
            } else {
                throw Error(1, "ERROR: unknown SPICE port type '" + pt + "'");
            }
            d.spice.push_back(p);
        }
        const int extra = r.geti("number of additional SPICE nodes");
        for (int i = 0; i < extra; ++i) d.spice_extra.push_back(trim(r.gets("additional SPICE node")));
    }
    return d;
}

}  // namespace aem
