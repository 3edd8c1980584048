// sim.hip -- device state and step orchestration of libaem (include/aem.h).
//
// aem_sim_create uploads a plan (the reference's "!$acc data" region,
// fdtd_solver.f90:4380-4430); aem_run enqueues the leapfrog step
// (fdtd_solver.f90:4440-8299) on the sim's stream, optionally replaying a
// captured hipGraph per step (one per buffer set when the single-pass step
// alternates two sets of fields); aem_finalize runs the post-loop products
// (fdtd_solver.f90:8340-8534) on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/aem.h"
#include "api_internal.hpp"
#include "kernels.hpp"
#include "kernels_kmax.hpp"
#include "kernels_out.hpp"
#include "plan.hpp"

using namespace aemk;

namespace {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw aem::Error(AEM_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <class T>
T* dalloc(size_t n, std::vector<void*>& owned) {
    if (n == 0) n = 1;
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    owned.push_back(p);
    return static_cast<T*>(p);
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// Replay loops keep at most two chunks of steps queued ahead of the device:
// after every kRunAheadChunk launches the host records an event and waits
// for the one it recorded a chunk earlier.  The device always has a chunk
// of work queued, while the runtime is never handed a whole run at once (a
// 3000-step, 64-member population is about 10^6 kernel dispatches).
constexpr int kRunAheadChunk = 32;

// A run checks every field for NaN or infinity after every kWatchSteps
// steps and after its last step (one read of the six fields: about 0.3 ms
// at 400^3, 0.1% of 256 steps).
constexpr int kWatchSteps = 256;
struct RunAhead {
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool live[2] = {false, false};
    void create() {
        for (auto& e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    void destroy() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        ev[0] = ev[1] = nullptr;
    }
    // after launch q (0-based) of a replay loop on stream s
    void pace(hipStream_t s, int64_t q) {
        if ((q + 1) % kRunAheadChunk) return;
        const int slot = (int)(((q + 1) / kRunAheadChunk) & 1);
        HIPCHK(hipEventRecord(ev[slot], s));
        live[slot] = true;
        if (live[slot ^ 1]) HIPCHK(hipEventSynchronize(ev[slot ^ 1]));
    }
};

// The CPML correction can be folded into the bulk update when no cell of a
// slab also receives a sheet, TF/SF, port or Drude update: the reference
// applies those between the bulk update and the CPML correction
// (fdtd_solver.f90:4486-5375, 5712-6670), so only then is the per-cell
// order unchanged.
bool cpml_fusable(const aem_plan_view& v) {
    const int N[3] = {v.NX, v.NY, v.NZ};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = v.npml[2 * a] > 0 ? v.npml[2 * a] + 1 : 1;
        hi[a] = v.npml[2 * a + 1] > 0 ? N[a] - v.npml[2 * a + 1] : N[a];
    }
    auto inside = [&](int a, int x) { return x >= lo[a] && x <= hi[a]; };
    auto box_inside = [&](const int b[6]) {   // {xlo,xhi,ylo,yhi,zlo,zhi}
        for (int a = 0; a < 3; ++a)
            if (b[2 * a] <= b[2 * a + 1] && !(inside(a, b[2 * a]) && inside(a, b[2 * a + 1]))) return false;
        return true;
    };
    for (int q = 0; q < v.n_ports; ++q) {
        const aem_port_view& p = v.ports[q];
        int b[6];
        for (int a = 0; a < 3; ++a) { b[2 * a] = p.loc[a] - 1; b[2 * a + 1] = p.loc[a] + p.size[a]; }
        if (!box_inside(b)) return false;
    }
    for (int q = 0; q < v.n_spice; ++q) {
        const aem_spice_view& p = v.spice[q];
        int b[6];
        for (int a = 0; a < 3; ++a) { b[2 * a] = p.loc[a] - 1; b[2 * a + 1] = p.loc[a] + p.size[a]; }
        if (!box_inside(b)) return false;
    }
    if (v.has_tfsf) {
        int b[6];
        for (int a = 0; a < 3; ++a) { b[2 * a] = v.tf_lo[a] - 1; b[2 * a + 1] = v.tf_hi[a] + 1; }
        for (int a = 0; a < 3; ++a)
            if (v.pbc[a]) { b[2 * a] = lo[a]; b[2 * a + 1] = hi[a]; }   // periodic axes carry no slab
        if (!box_inside(b)) return false;
    }
    if (v.kmax && v.k_axis >= 0) {   // the modal source planes, across the periodic axes
        int b[6];
        for (int a = 0; a < 3; ++a) { b[2 * a] = lo[a]; b[2 * a + 1] = hi[a]; }
        b[2 * v.k_axis] = std::min(v.kpl_H, v.kpl_E) - 1;
        b[2 * v.k_axis + 1] = std::max(v.kpl_H, v.kpl_E) + 1;
        if (!box_inside(b)) return false;
    }
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < v.n_sheet_planes[a]; ++s) {
            const int P = v.sheet_dim[a][0], Q = v.sheet_dim[a][1];
            const float* sg = v.sheet_sig[a] + (size_t)s * P * Q;
            const float* ep = v.sheet_ep[a] + (size_t)s * P * Q;
            const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
            for (int q = 1; q <= Q; ++q)
                for (int p = 1; p <= P; ++p) {
                    const size_t k = (size_t)(p - 1) + (size_t)P * (q - 1);
                    if (!(sg[k] > 0 || ep[k] > 1)) continue;
                    if (!inside(a, v.sheet_pos[a][s]) || !inside(o1, p) || !inside(o1, p + 1) || !inside(o2, q) ||
                        !inside(o2, q + 1))
                        return false;
                }
        }
    if (v.has_drude)
        for (int c = 0; c < 3; ++c)
            if (!box_inside(v.dbox[c])) return false;
    return true;
}

// The single-pass step (k_step_sp) needs the CPML folded into the bulk and
// nothing applied between the bulk H and bulk E updates of a core cell: no
// sheets, TF/SF, PBC or Drude media.  Ports qualify: their H phase only reads
// H (fdtd_solver.f90:4917-4957) and their E phase follows the bulk E.
bool single_pass_ok(const aem_plan_view& v) {
    if (v.has_tfsf || v.has_drude || v.n_spice > 0 || v.kmax) return false;
    for (int a = 0; a < 3; ++a)
        if (v.pbc[a] || v.n_sheet_planes[a] > 0) return false;
    return true;
}

// Image position of a far-field face cell across an infinite ground plane
// along axis a, in the reference's FP32 form i (1 - use) + (offset - i) use
// (f90:7957-7959); i itself off the plane's axis.
int mirror_pos(const aem_plan_view& v, int a, int i) {
    return (int)((float)i * (1.0f - v.use_m[a]) + (v.off_m[a] - (float)i) * v.use_m[a]);
}

// One bulk launch: a set of boxes and the slab tests compiled into its kernel.
struct Launch {
    BoxSet bs;
    int c;
    unsigned blocks() const { return (unsigned)bs.first[bs.n]; }
    long long cells() const {
        long long n = 0;
        for (int q = 0; q < bs.n; ++q) n += (long long)bs.b[q].ni * bs.b[q].nj * bs.b[q].nk;
        return n;
    }
};

constexpr int kPlanes = 4;     // default z planes marched by one thread of a bulk kernel
constexpr int kSpPlanes = 10;  // target z planes marched by one column of the single-pass kernel

// Parts of a step, in the order the reference applies them.
enum : unsigned {
    kPartHBulk = 1u << 0,    // bulk H (+ fused CPML)
    kPartHPre = 1u << 1,     // H sheets, TF/SF
    kPartHPort = 1u << 2,    // ports, H phase
    kPartHPost = 1u << 3,    // separate CPML H, PBC
    kPartEBulk = 1u << 4,    // bulk E (+ fused CPML)
    kPartEPre = 1u << 5,     // E sheets, Drude, TF/SF
    kPartEPort = 1u << 6,    // ports, E phase
    kPartEPost = 1u << 7,    // separate CPML E, PBC
    kPartOutPre = 1u << 8,   // plane sums
    kPartNtff = 1u << 9,     // far-field currents and binning
    kPartOutPost = 1u << 10, // phase-centre samples
    kPartStep = 1u << 11,    // step counter
    kPartRecord = 1u << 12,  // design-box record for the adjoint (between H and E)
    kPartObjective = 1u << 13, // far-field objective accumulation
    kPartSpiceH = 1u << 14,  // SPICE gap currents (after the H phase)
    kPartSpiceE = 1u << 15,  // SPICE gap fields (after the lumped ports' E phase)
    kPartEOld = 1u << 16,    // E_old of the Drude boxes (after the PBC copies)
    kPartAll = (1u << 17) - 1,
    // a step split at the SPICE exchange: everything up to the gap currents,
    // then the rest
    kPartToExchange = kPartHBulk | kPartHPre | kPartHPort | kPartHPost | kPartRecord | kPartSpiceH,
    kPartFromExchange = kPartAll & ~kPartToExchange
};

Launch make_launch(const std::vector<Box>& boxes, int c) {
    Launch L{};
    L.c = c;
    L.bs.n = (int)boxes.size();
    long long acc = 0;
    for (int q = 0; q < L.bs.n; ++q) {
        const Box& b = boxes[q];
        L.bs.b[q] = b;
        L.bs.first[q] = (int)acc;
        acc += (long long)cdiv(b.ni, kTX) * cdiv(b.nj, kTY) * cdiv(b.nk, b.kc);
    }
    if (acc > 0x7fffffffLL) throw aem::Error(AEM_ERR_STATE, "grid too large for one bulk launch");
    L.bs.first[L.bs.n] = (int)acc;
    return L;
}

// Launches covering the updated cells [1,NX-1]x[1,NY-1]x[1,NZ-1] with
// disjoint boxes.  Without fusion: one box.  With fusion: the core (all x,
// y and z outside the y/z slabs) with the x-slab test, then one launch over
// the y-slab and z-slab boxes.  H slab layers are 1..n-1 and N+1-n..N-1, E
// layers 2..n and N+1-n..N-1, so one split serves both phases.
std::vector<Launch> bulk_launches(const aem_plan_view& v, bool fuse, int planes) {
    const int N[3] = {v.NX, v.NY, v.NZ};
    auto box = [&](int a0, int a1, int b0, int b1, int c0, int c1) {
        // z chunks of equal length, at most `planes` each; a box of at most
        // 16 planes (a z slab) in chunks of 2, for more work-groups (at 400^3
        // 2.5% faster slab launches than 4 planes; 1 plane the same, the whole
        // box 23% slower)
        const int nk = c1 - c0 + 1;
        const int kc = nk > 0 ? (nk <= 16 ? 2 : (int)cdiv(nk, cdiv(nk, planes))) : planes;
        return Box{a0, b0, c0, a1 - a0 + 1, b1 - b0 + 1, nk, kc};
    };
    std::vector<Launch> out;
    if (!fuse) {
        out.push_back(make_launch({box(1, N[0] - 1, 1, N[1] - 1, 1, N[2] - 1)}, 0));
        return out;
    }
    int lo[3], hi[3];
    std::vector<std::pair<int, int>> edge[3];
    for (int a = 1; a < 3; ++a) {
        lo[a] = v.npml[2 * a] > 0 ? v.npml[2 * a] + 1 : 1;
        hi[a] = v.npml[2 * a + 1] > 0 ? N[a] - v.npml[2 * a + 1] : N[a] - 1;
        if (lo[a] > hi[a]) {
            edge[a].push_back({1, N[a] - 1});
            continue;
        }
        if (lo[a] > 1) edge[a].push_back({1, lo[a] - 1});
        if (hi[a] < N[a] - 1) edge[a].push_back({hi[a] + 1, N[a] - 1});
    }
    const bool cy = lo[1] <= hi[1], cz = lo[2] <= hi[2];
    if (cy && cz) out.push_back(make_launch({box(1, N[0] - 1, lo[1], hi[1], lo[2], hi[2])}, kCx));
    std::vector<Box> slabs;
    if (cz)
        for (auto e : edge[1]) slabs.push_back(box(1, N[0] - 1, e.first, e.second, lo[2], hi[2]));
    for (auto e : edge[2]) slabs.push_back(box(1, N[0] - 1, 1, N[1] - 1, e.first, e.second));
    if (!slabs.empty()) out.push_back(make_launch(slabs, kCx | kCy | kCz));
    return out;
}

}  // namespace

struct aem_sim {
    const aem::Plan* plan = nullptr;
    aem::Plan* owned_plan = nullptr;
    hipStream_t st = nullptr;
    std::vector<void*> owned;
    Grid g{};
    int* d_step = nullptr;
    int64_t steps_done = 0;
    float* F[6] = {};              // fields now (Ex, Ey, Ez, Hx, Hy, Hz)
    float* F2[6] = {};             // second field set of the single-pass step
    void* d_idx = nullptr;
    uint32_t* d_idx4 = nullptr;    // uint8_t indices packed four z planes per word (single pass)
    int n_tup = 0;                 // coefficient tuples on the device (a design box adds one)
    float* d_tup = nullptr;
    float* d_den[6] = {};
    Cpml cpH{}, cpE{};
    Sheets sh[3]{};
    DrudeBox db[3]{};
    PortDev* d_ports = nullptr;
    float *d_V = nullptr, *d_I = nullptr, *d_Vout = nullptr, *d_Iout = nullptr;
    NtffFaces nf{};
    NtffBins nb{};
    NtffExact nx_acc{};
    int win_max = 0;
    NtffList nl{};                 // ordered binning by contributor lists (nl.n_waves > 0)
    bool ntff_lists = true;        // ordered binning from contributor lists (else the row scan)
    void build_ntff_lists(const std::vector<int>& win, unsigned stride);
    double *d_refl = nullptr, *d_trans = nullptr;
    double *d_refl2 = nullptr, *d_trans2 = nullptr;     // kmax: the TM sums (d_refl, d_trans hold TE)
    float *d_reflf2 = nullptr, *d_transf2 = nullptr;
    // kmax (complex fields): this simulation holds the real part and kim the
    // imaginary part, a second simulation of the same plan stepped with it on
    // this stream; the parts meet only in the Bloch copies of the periodic
    // walls and in the modal plane source
    aem_sim* kim = nullptr;
    bool kim_part = false;         // this simulation is the imaginary part of a kmax pair
    KmaxSrc kmH{}, kmE{};
    KmaxTab ktab{};                // the modal source as host-built increments (ktab.ncell > 0)
    void kmax_setup();
    void kmax_table();
    void enqueue_kmax_step(hipStream_t s);
    float* d_pc = nullptr;
    float *d_incident = nullptr, *d_Eth = nullptr, *d_Eph = nullptr, *d_fftime = nullptr;
    float *d_reflf = nullptr, *d_transf = nullptr, *d_ffpc = nullptr, *d_scratch = nullptr;
    Tfsf tf{};
    Tfsf tfm{};                    // an infinite ground plane's mirrored plane wave (v.mirror_type > 0)
    Pulse pu{};
    bool fuse_cpml = false;        // CPML corrections folded into the bulk kernels
    RDiv rdv[3] = {};              // the CPML kernels' divisions by dx, dy, dz
    std::vector<Launch> bulk;      // bulk launches of each phase; [0] is the core when fused
    HArgs ha{};                    // bulk kernel arguments (fields, tables, CPML state)
    EArgs ea{};
    PortArgs pa{};                 // port kernel arguments
    // SPICE co-simulation: gap currents out and node voltages in through
    // pinned staging, once per step (aem_port_exchange)
    SpiceArgs spa{};
    SpiceDev* d_spice = nullptr;
    float *d_spI = nullptr, *d_spV = nullptr, *d_spIout = nullptr, *d_spVout = nullptr;
    float *h_spI = nullptr, *h_spV = nullptr;
    bool x_pending = false;        // a step waits at the exchange for its node voltages
    NtffArgs na{};                 // far-field kernel arguments (batched form)
    int ntff_mode = AEM_NTFF_ORDERED;
    int ntff_mu = 0, ntff_mv = 0;  // largest face extents of the far-field box
    int planes = kPlanes;
    bool planes_set = false;       // planes given by the caller (also used by the single-pass kernel)
    bool off32 = false;            // every field array under 4 GiB: 32-bit byte offsets in the kernels
    bool allow_fuse = true;
    // single-pass step: fields and x-face psi alternate between two buffer
    // sets, so a step reads one set and writes the other
    bool allow_sp = true;
    bool sp_ok = false;            // the deck admits it (set at creation)
    bool batched = false;          // a batch steps this simulation: in-place kernels only
    Box sp_box{};                  // core box of the single-pass kernel
    std::vector<std::pair<Box, int>> sp_parts;   // its launches (box, slab tests compiled in)
    int par = 0;                   // which buffer set holds the fields now
    // k_adj_record reads E before its update, which the single pass has
    // already written; a design box records inside k_e_design, after it
    bool sp_active() const { return sp_ok && !batched && !(rec_on && !des_on); }
    void flip();
    void bind();
    // differentiation
    Objective obj{};
    bool obj_on = false;
    FfDft ff{};                    // running-DFT far field
    bool ff_on = false;
    AdjRec rec{};
    bool rec_on = false;
    int rec_lo[3] = {0, 0, 0}, rec_hi[3] = {0, 0, 0};   // design cell box
    bool des_on = false;           // per-cell permittivity in the design box
    DesignDev des{};
    std::vector<float> des_eps;    // [3][cells of the box], x fastest
    float* d_des_coef = nullptr;
    void upload_design();
    float* lam[9] = {};            // adjoint fields and mu = gb lambda_E (allocated at the first backward step)
    Cpml lpE{}, lpH{};             // adjoint CPML states
    DrudeBox adb[3]{};             // adjoint Drude states: J holds lambda_J, Eold lambda_Eold
    float *d_lV = nullptr, *d_lI = nullptr;   // adjoint lumped-port lines [np][M]
    // adjoint checkpoints: with ck_K > 0 the design record keeps ck_K steps,
    // the state is saved before every ck_K-th step, and the backward sweep
    // recomputes each interval from its snapshot before transposing it
    int ck_K = 0;
    bool ck_last_ok = false;       // the record holds the last interval (true after the forward run)
    size_t ck_bytes = 0;
    std::vector<char*> ck_snap;    // one per interval, then the state at the end of the run
    std::vector<std::pair<void*, size_t>> state_regions();
    void snapshot(int q);
    void restore(int q);
    int* d_bstep = nullptr;
    float* d_srcD = nullptr;       // adjoint source: [6][4][mv][mu] increments of a step
    float* d_desm = nullptr;       // design transpose: [3][nedge] gb lambda_E
    void ensure_adjoint();
    template <class IdxT>
    void enqueue_backward_t(hipStream_t s, bool objective, double Fr, double Fi, const double* wV, const double* wI,
                            double* dga, double* dgb);
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph2 = nullptr;   // single-pass steps: the step from the other buffer set
    hipGraphExec_t gexec2 = nullptr;
    int gpar = 0;                  // buffer set the first graph starts from
    bool gsp = false;              // the graphs are single-pass steps
    void drop_graphs();
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    RunAhead ra;                   // bounds the steps a replay loop queues
    float last_ms = 0.f;
    bool finalized = false;
    int dev = 0;                   // HIP device current at creation
    double* d_dft = nullptr;   // aem_get_dft staging: frequencies, then (re,im) pairs
    int dft_cap = 0;
    // exports (f90:6833-7003): the video slice stays on the device for the
    // run; the full-field records go to their files after every step
    Video vid{};
    bool vid_on = false;
    bool full_on = false;
    int full_n[3] = {0, 0, 0};     // cells of a full-field record per axis
    float* d_full[6] = {};
    float** d_fullp = nullptr;     // device copy of d_full for k_full
    float* h_full = nullptr;       // pinned staging of one record set
    std::FILE* full_f[6] = {};
    void export_step();
    // numeric watchdog: set by k_watch, read after every run
    int* d_watch = nullptr;
    void enqueue_watch(hipStream_t s);
    void check_watch();

    ~aem_sim() {
        if (st) (void)hipStreamSynchronize(st);
        delete kim;
        for (auto& f : full_f)
            if (f) std::fclose(f);
        if (h_full) (void)hipHostFree(h_full);
        if (h_spI) (void)hipHostFree(h_spI);
        if (h_spV) (void)hipHostFree(h_spV);
        drop_graphs();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        ra.destroy();
        if (d_dft) (void)hipFree(d_dft);
        for (void* p : owned) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
        delete owned_plan;
    }

    template <class T>
    T* up(const T* host, size_t n) {
        T* d = dalloc<T>(n, owned);
        if (n && host) HIPCHK(hipMemcpyAsync(d, host, n * sizeof(T), hipMemcpyHostToDevice, st));
        else HIPCHK(hipMemsetAsync(d, 0, (n ? n : 1) * sizeof(T), st));
        return d;
    }
    template <class T>
    T* zeros(size_t n) {
        T* d = dalloc<T>(n, owned);
        HIPCHK(hipMemsetAsync(d, 0, (n ? n : 1) * sizeof(T), st));
        return d;
    }

    void create(const aem_sim_options* opt);
    // one full step; a single-pass step leaves the fields in the other buffer
    // set and flips the host's pointers to it
    void enqueue_step(hipEvent_t* mark = nullptr) {
        if (kim) enqueue_kmax_step(st);
        else enqueue_parts(mark, st, kPartAll);
    }
    void enqueue_parts(hipEvent_t* mark, hipStream_t s, unsigned parts);
    template <class IdxT>
    void enqueue_step_t(hipEvent_t* mark, hipStream_t s, unsigned parts);
};

void aem_sim::create(const aem_sim_options* opt) {
    const aem_plan_view& v = plan->v;
    if (opt) {
        ntff_mode = opt->ntff_binning;
        if (opt->planes > 0) {
            planes = opt->planes;
            planes_set = true;
        }
        allow_fuse = opt->no_cpml_fusion == 0;
        allow_sp = opt->no_single_pass == 0;
        ntff_lists = opt->ntff_row_scan == 0;
    }
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    ra.create();
    g.NX = v.NX; g.NY = v.NY; g.NZ = v.NZ; g.PX = v.PX; g.plane = v.plane;
    const size_t nc = (size_t)v.ncell;
    for (int q = 0; q < 6; ++q) F[q] = zeros<float>(nc + v.plane);  // one spare plane keeps halo reads in bounds
    d_step = zeros<int>(1);
    {
        int one = 1;
        HIPCHK(hipMemcpyAsync(d_step, &one, sizeof(int), hipMemcpyHostToDevice, st));
    }
    if (v.idx_bytes == 1) d_idx = up<uint8_t>(v.tuple_idx8, nc);
    else d_idx = up<uint16_t>(v.tuple_idx16, nc);
    d_tup = up<float>(v.tuples, (size_t)v.n_tuple * 6);
    n_tup = v.n_tuple;
    const int dl[6] = {v.NX - 1, v.NY - 1, v.NZ - 1, v.NX - 1, v.NY - 1, v.NZ - 1};
    const float* dsrc[6] = {v.den_ex, v.den_ey, v.den_ez, v.den_hx, v.den_hy, v.den_hz};
    for (int q = 0; q < 6; ++q) d_den[q] = up<float>(dsrc[q], std::max(dl[q], 1));

    // CPML: coefficients and psi slabs (the two x faces of a slot in one
    // allocation, face 0 first: the single-pass kernel reaches both from
    // one base)
    for (int f = 0; f < 6; ++f) {
        const int n = v.npml[f];
        const int a = f / 2;
        cpH.n[f] = cpE.n[f] = n;
        cpH.b[f] = up<float>(v.bh[f], std::max(n - 1, 0));
        cpH.c[f] = up<float>(v.ch[f], std::max(n - 1, 0));
        cpE.b[f] = up<float>(v.be[f], std::max(n, 0));
        cpE.c[f] = up<float>(v.ce[f], std::max(n, 0));
        const size_t nt = (size_t)std::max(n - 1, 0);
        size_t trans;
        if (a == 0) trans = (size_t)v.NY * v.NZ;
        else if (a == 1) trans = (size_t)v.PX * v.NZ;
        else trans = (size_t)v.PX * v.NY;
        for (int s = 0; s < 2; ++s) {
            if (f == 1) continue;
            if (f == 0) {
                const size_t nt1 = (size_t)std::max(v.npml[1] - 1, 0);
                cpH.psi[0][s] = zeros<float>((nt + nt1) * trans);
                cpE.psi[0][s] = zeros<float>((nt + nt1) * trans);
                cpH.psi[1][s] = cpH.psi[0][s] + nt * trans;
                cpE.psi[1][s] = cpE.psi[0][s] + nt * trans;
                continue;
            }
            cpH.psi[f][s] = zeros<float>(nt * trans);
            cpE.psi[f][s] = zeros<float>(nt * trans);
        }
    }

    // sheets
    for (int a = 0; a < 3; ++a) {
        Sheets& s = sh[a];
        s.np = v.n_sheet_planes[a];
        s.P = v.sheet_dim[a][0];
        s.Q = v.sheet_dim[a][1];
        const size_t n = (size_t)s.np * s.P * s.Q;
        s.pos = up<int>(v.sheet_pos[a], s.np);
        s.sig = up<float>(v.sheet_sig[a], n);
        s.ep = up<float>(v.sheet_ep[a], n);
        s.es = zeros<float>(n);
    }

    // Drude boxes
    if (v.has_drude) {
        for (int c = 0; c < 3; ++c) {
            DrudeBox& b = db[c];
            for (int q = 0; q < 3; ++q) { b.lo[q] = v.dbox[c][2 * q]; b.hi[q] = v.dbox[c][2 * q + 1]; }
            long long bx = std::max(0, b.hi[0] - b.lo[0] + 1), by = std::max(0, b.hi[1] - b.lo[1] + 1),
                      bz = std::max(0, b.hi[2] - b.lo[2] + 1);
            b.nb = bx * by * bz;
            b.Ja = up<float>(v.Ja[c], (size_t)b.nb * v.n_poles);
            b.Jb = up<float>(v.Jb[c], (size_t)b.nb * v.n_poles);
            b.J = zeros<float>((size_t)b.nb * v.n_poles);
            b.Eold = zeros<float>((size_t)b.nb);
        }
    }

    // ports
    const int T = v.time_steps;
    const int np = std::max(v.n_ports, 1);
    std::vector<PortDev> pd(v.n_ports);
    for (int q = 0; q < v.n_ports; ++q) {
        const aem_port_view& p = v.ports[q];
        PortDev& d = pd[q];
        d.dir = p.dir;
        for (int a = 0; a < 3; ++a) { d.loc[a] = p.loc[a]; d.size[a] = p.size[a]; }
        d.kI = p.kI; d.kV = p.kV; d.kS = p.kS; d.dl = p.dl;
        d.cur_num = p.cur_num; d.cur_den = p.cur_den; d.gap_div = p.gap_div;
        d.cdv = make_rdiv(p.cur_den, v.kmax != 0);
        d.gdv = make_rdiv(p.gap_div, v.kmax != 0);
        d.excited = p.excited;
    }
    d_ports = up<PortDev>(pd.data(), pd.size());
    d_V = zeros<float>((size_t)np * v.M);
    d_I = zeros<float>((size_t)np * v.M);
    d_Vout = zeros<float>((size_t)np * T);
    d_Iout = zeros<float>((size_t)np * T);
    if (v.n_spice > 0) {
        std::vector<SpiceDev> sd(v.n_spice);
        for (int q = 0; q < v.n_spice; ++q) {
            const aem_spice_view& p = v.spice[q];
            sd[q].dir = p.dir;
            for (int a = 0; a < 3; ++a) { sd[q].loc[a] = p.loc[a]; sd[q].size[a] = p.size[a]; }
            sd[q].cur_num = p.cur_num; sd[q].cur_den = p.cur_den; sd[q].gap_div = p.gap_div; sd[q].bias = p.v_bias;
        }
        const size_t nv = (size_t)v.n_spice + v.n_spice_extra;
        d_spice = up<SpiceDev>(sd.data(), sd.size());
        d_spI = zeros<float>((size_t)v.n_spice);
        d_spV = zeros<float>(nv);
        d_spIout = zeros<float>((size_t)v.n_spice * T);
        d_spVout = zeros<float>(nv * T);
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h_spI), (size_t)v.n_spice * sizeof(float)));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h_spV), nv * sizeof(float)));
    }
    pu.src = up<float>(kim_part ? v.src_table_im : v.src_table, (size_t)3 * T);
    pu.T = T;

    // TF/SF
    for (int a = 0; a < 3; ++a) {
        tf.lo[a] = v.tf_lo[a]; tf.hi[a] = v.tf_hi[a]; tf.delay[a] = v.delay[a]; tf.pbc[a] = v.pbc[a];
        tf.WE[a] = v.WE[a]; tf.WH[a] = v.WH[a]; tf.kH[a] = v.kHinc[a]; tf.kE[a] = v.kEinc[a];
    }
    tf.s_th = v.s_th; tf.c_th = v.c_th; tf.s_ph = v.s_ph; tf.c_ph = v.c_ph;
    tf.dx = v.dx; tf.dy = v.dy; tf.dz = v.dz; tf.dt = v.dt; tf.inv_c = v.inv_c;
    tf.spread = v.spread; tf.t_spread = v.t_spread; tf.pulse_type = v.pulse_type; tf.exp_half = v.exp_half;
    for (int f = 0; f < 6; ++f) tf.wall[f] = v.wall[f];
    tfm = tf;   // the mirrored wave: its weights, angles and delays, the same faces
    for (int a = 0; a < 3; ++a) { tfm.WE[a] = v.WE_m[a]; tfm.WH[a] = v.WH_m[a]; tfm.delay[a] = v.delay_m[a]; }
    tfm.s_th = v.s_thm; tfm.c_th = v.c_thm; tfm.s_ph = v.s_phm; tfm.c_ph = v.c_phm;

    // NTFF
    if (v.n_angles > 0) {
        for (int a = 0; a < 3; ++a) { nf.lo[a] = v.ff_lo[a]; nf.hi[a] = v.ff_hi[a]; nf.pbc[a] = v.pbc[a]; }
        for (int f = 0; f < 6; ++f) nf.wall[f] = v.wall[f];
        // the face arrays of every (face, slot) in one allocation, each at a
        // power-of-two stride, in blocks M, Mo, J, Jo (the list kernel reaches
        // any of them from one base and an element index)
        unsigned stride = 1;
        for (int f = 0; f < 6; ++f) {
            const int a = f / 2, b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
            const size_t n = (size_t)(v.ff_hi[b] + 2) * (v.ff_hi[c] + 2);
            while (stride < n) stride <<= 1;
        }
        float* fb = zeros<float>((size_t)48 * stride);
        for (int f = 0; f < 6; ++f) {
            const int a = f / 2, b = a == 0 ? 1 : 0;
            nf.fd0[f] = v.ff_hi[b] + 2;
            for (int s = 0; s < 2; ++s) {
                const size_t o = (size_t)(2 * f + s) * stride;
                nf.M[f][s] = fb + o;
                nf.Mo[f][s] = fb + (size_t)12 * stride + o;
                nf.J[f][s] = fb + (size_t)24 * stride + o;
                nf.Jo[f][s] = fb + (size_t)36 * stride + o;
            }
        }
        nl.Mbase = fb;
        nl.Jbase = fb + (size_t)24 * stride;
        nl.off_old = 12u * stride;
        {
            int sh = 1;
            while ((1u << (sh - 1)) < stride) ++sh;
            nl.face_shift = sh;           // (2 face + slot) * stride >> (log2 stride + 1) = face
        }
        nb.ang = up<float>(v.ang, (size_t)v.n_angles * 6);
        nb.n_angles = v.n_angles;
        nb.ic = v.ic; nb.jc = v.jc; nb.kc = v.kc;
        nb.r_relay = v.r_relay; nb.dx = v.dx; nb.dy = v.dy; nb.dz = v.dz;
        nb.c_dt = aem::kC * v.dt;
        for (int a = 0; a < 3; ++a) nb.K[a] = v.ff_K[a];
        nb.LL = (long long)v.len_ff + 2;
        nb.exact = ntff_mode == AEM_NTFF_ORDERED ? 1 : 0;
        {
            // retarded-time offsets of every face cell, per angle, bound the
            // window of bins one step can touch
            std::vector<int> win(2 * (size_t)v.n_angles);
            win_max = 0;
            for (int kk = 0; kk < v.n_angles; ++kk) {
                const float* an = v.ang + 6 * kk;
                const float s1 = an[2], c1 = an[3], s2 = an[4], c2 = an[5];
                float tmin = 1e30f, tmax = -1e30f;
                for (int m = 0; m < (v.mirror_type > 0 ? 2 : 1); ++m)   // with a ground plane, the image cells too
                for (int a = 0; a < 3; ++a) {
                    if (v.pbc[a]) continue;
                    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
                    for (int side = 0; side < 2; ++side) {
                        const int pos = side ? v.ff_hi[a] : v.ff_lo[a];
                        for (int vv = v.ff_lo[c]; vv <= v.ff_hi[c]; ++vv)
                            for (int uu = v.ff_lo[b]; uu <= v.ff_hi[b]; ++uu) {
                                int i, j, k;
                                if (a == 0) { i = pos; j = uu; k = vv; }
                                else if (a == 1) { i = uu; j = pos; k = vv; }
                                else { i = uu; j = vv; k = pos; }
                                if (m) { i = mirror_pos(v, 0, i); j = mirror_pos(v, 1, j); k = mirror_pos(v, 2, k); }
                                const float tfv = (v.r_relay - ((float)(i - v.ic) * v.dx * s1 * c2 +
                                                                (float)(j - v.jc) * v.dy * s1 * s2 +
                                                                (float)(k - v.kc) * v.dz * c1)) / (aem::kC * v.dt);
                                tmin = std::min(tmin, tfv);
                                tmax = std::max(tmax, tfv);
                            }
                    }
                }
                const int lo = (int)std::floor(tmin) - 3, hi = (int)std::floor(tmax) + 3;
                win[2 * kk] = lo;
                win[2 * kk + 1] = std::max(hi - lo + 1, 0);
                win_max = std::max(win_max, win[2 * kk + 1]);
            }
            nx_acc.win = up<int>(win.data(), win.size());
            if (nb.exact) {
                const size_t row = (size_t)v.n_angles * nb.LL;
                nl.acc = zeros<float>(6 * row);
                for (int q = 0; q < 3; ++q) {
                    nb.Uf[q] = nl.acc + q * row;
                    nb.Wf[q] = nl.acc + (3 + q) * row;
                    nx_acc.acc[q] = nb.Uf[q];
                    nx_acc.acc[3 + q] = nb.Wf[q];
                }
                if (ntff_lists) build_ntff_lists(win, stride);
            }
        }
        if (!nb.exact) {
            for (int q = 0; q < 3; ++q) {
                nb.U[q] = zeros<double>((size_t)v.n_angles * nb.LL);
                nb.W[q] = zeros<double>((size_t)v.n_angles * nb.LL);
            }
        }
        // the histogram keeps one window of bins in LDS; larger windows bin
        // through global atomics
        if (ntff_mode == AEM_NTFF_HISTOGRAM && win_max > kHistMax) ntff_mode = AEM_NTFF_ATOMIC;
        // an infinite ground plane's image currents are binned by the
        // contributor lists only
        if (v.mirror_type > 0 && (ntff_mode != AEM_NTFF_ORDERED || nl.n_waves == 0))
            throw aem::Error(AEM_ERR_UNSUPPORTED,
                             "the far field of an infinite ground plane deck needs the ordered binning from "
                             "contributor lists (AEM_NTFF_ORDERED, ntff_row_scan 0)");
        if (ntff_mode == AEM_NTFF_ORDERED && nl.n_waves == 0)   // the row-scan kernel's limit
            for (int a = 0; a < 2; ++a)
                if (!v.pbc[a] && 2 * (v.ff_hi[a] - v.ff_lo[a] + 1) > kExactCells)
                    throw aem::Error(AEM_ERR_UNSUPPORTED,
                                     "ordered far-field binning stages rows of at most " +
                                         std::to_string(kExactCells / 2) +
                                         " cells; use AEM_NTFF_ATOMIC or AEM_NTFF_HISTOGRAM for this grid");
        d_Eth = zeros<float>((size_t)v.n_angles * T);
        d_Eph = zeros<float>((size_t)v.n_angles * T);
        d_fftime = zeros<float>(v.n_angles);
        d_scratch = zeros<float>((size_t)v.n_angles * 2 * (v.len_ff + 1));
    }
    d_refl = zeros<double>(T);
    d_trans = zeros<double>(T);
    d_reflf = zeros<float>(T);
    d_transf = zeros<float>(T);
    if (v.kmax) {
        d_refl2 = zeros<double>(T);
        d_trans2 = zeros<double>(T);
        d_reflf2 = zeros<float>(T);
        d_transf2 = zeros<float>(T);
    }
    d_pc = zeros<float>((size_t)6 * (T + 1));
    d_ffpc = zeros<float>(T);
    d_incident = zeros<float>(T);
    d_watch = zeros<int>(1);

    // exports: a component's video (one Fortran record) and a full-field
    // record must each fit the 4-byte record length of the files
    const unsigned long long rec_max = 0x7fffffffull;
    if (v.video_on) {
        vid.slice = v.slice;
        vid.loc = v.slice_location;
        vid.v1 = v.vid_size1;
        vid.v2 = v.vid_size2;
        vid.T = T;
        const unsigned long long n = (unsigned long long)vid.v1 * vid.v2 * (unsigned long long)T;
        if (n * 4 > rec_max)
            throw aem::Error(AEM_ERR_UNSUPPORTED, "video slice of " + std::to_string(n * 4) +
                                                      " bytes per component exceeds one 2 GiB file record");
        for (int q = 0; q < 6; ++q) vid.V[q] = zeros<float>((size_t)n);
        vid_on = true;
    }
    if (plan->deck.export_all == "yes") {
        full_n[0] = v.NX - 1 - v.pbc[0];
        full_n[1] = v.NY - 1 - v.pbc[1];
        full_n[2] = v.NZ - 1 - v.pbc[2];
        const unsigned long long n = (unsigned long long)full_n[0] * full_n[1] * full_n[2];
        if (n * 4 > rec_max)
            throw aem::Error(AEM_ERR_UNSUPPORTED, "full-field record of " + std::to_string(n * 4) +
                                                      " bytes exceeds one 2 GiB file record");
        for (int q = 0; q < 6; ++q) d_full[q] = zeros<float>((size_t)n);
        d_fullp = up<float*>(d_full, 6);
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h_full), (size_t)n * 6 * sizeof(float)));
        full_on = true;
    }

    // (a kmax pair runs the separate CPML kernels: they carry the complex
    // division of its psi updates)
    fuse_cpml = allow_fuse && !v.kmax && cpml_fusable(v);
    rdv[0] = make_rdiv(v.dx, v.kmax != 0);
    rdv[1] = make_rdiv(v.dy, v.kmax != 0);
    rdv[2] = make_rdiv(v.dz, v.kmax != 0);
    bulk = bulk_launches(v, fuse_cpml, planes);
    for (int f = 0; f < 6; ++f)
        for (int q = 0; q < 2; ++q) {
            cpH.pso[f][q] = cpH.psi[f][q];
            cpE.pso[f][q] = cpE.psi[f][q];
        }
    // the single-pass kernel takes the core launch's box (every x, y and z
    // outside the y/z slabs); the slab launch stays as it is
    // k_step_sp addresses the fields by 32-bit byte offsets, as the bulk
    // kernels do where they can
    const bool small = ((unsigned long long)v.ncell + (unsigned long long)v.plane) * 4ull < (1ull << 32);
    off32 = small;
    sp_ok = allow_sp && fuse_cpml && small && single_pass_ok(v) && !bulk.empty() && bulk[0].c == kCx &&
            bulk[0].bs.n == 1;
    if (sp_ok) {
        sp_box = bulk[0].bs.b[0];
        sp_box.kc = planes_set ? planes : kSpPlanes;
        // small grids: fewer planes per column until the launch has enough
        // work-groups to fill the chip (16 per CU); then chunks of equal
        // length (at 400^3: 380 planes in 38 chunks of 10, 2% faster than 16)
        auto groups = [&](int kc) {
            return (long long)cdiv(sp_box.ni, 64) * cdiv(sp_box.nj, kSY * kSpRows) * cdiv(sp_box.nk, kc);
        };
        if (!planes_set) {
            while (sp_box.kc > 2 && groups(sp_box.kc) < 4096) sp_box.kc /= 2;
            sp_box.kc = (int)cdiv(sp_box.nk, cdiv(sp_box.nk, sp_box.kc));
        }
        // one launch with the x-slab test compiled in: measured at 400^3,
        // separate launches for the slab-free x tiles (no slab test, fewer
        // registers) and the two x-slab tiles of each row took 5-12% longer,
        // in sequence or on a second stream beside them
        sp_parts.push_back({sp_box, (int)kCx});
        if (v.idx_bytes == 1) {   // word (cell of plane 1, q) = bytes of planes 4q+1 .. 4q+4
            const size_t pl = (size_t)v.plane, nq = (size_t)(v.NZ + 3) / 4;
            std::vector<uint32_t> w(pl * nq, 0u);
            for (size_t k = 0; k < (size_t)v.NZ; ++k)
                for (size_t c = 0; c < pl; ++c)
                    w[(k >> 2) * pl + c] |= (uint32_t)v.tuple_idx8[k * pl + c] << (8 * (k & 3));
            d_idx4 = up<uint32_t>(w.data(), w.size());
        }
        for (int q = 0; q < 6; ++q) F2[q] = zeros<float>(nc + v.plane);
        {
            const size_t nt0 = (size_t)std::max(v.npml[0] - 1, 0), nt1 = (size_t)std::max(v.npml[1] - 1, 0);
            const size_t trans = (size_t)v.NY * v.NZ;
            for (int q = 0; q < 2; ++q) {   // same layout as psi: face 0, then face 1
                cpH.pso[0][q] = zeros<float>((nt0 + nt1) * trans);
                cpE.pso[0][q] = zeros<float>((nt0 + nt1) * trans);
                cpH.pso[1][q] = cpH.pso[0][q] + nt0 * trans;
                cpE.pso[1][q] = cpE.pso[0][q] + nt0 * trans;
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        ntff_mu = std::max(ntff_mu, v.ff_hi[b] - v.ff_lo[b] + 1);
        ntff_mv = std::max(ntff_mv, v.ff_hi[c] - v.ff_lo[c] + 1);
    }
    bind();
    HIPCHK(hipStreamSynchronize(st));
    if (v.kmax && !kim_part) {
        kim = new aem_sim();
        kim->plan = plan;
        kim->kim_part = true;
        kim->create(opt);
        kmax_setup();
    }
}

// Kernel argument blocks from the current field and psi pointers.  In place,
// the bulk kernels read and write F and psi; ping-pong launches write F2 and
// the x-face pso.
void aem_sim::bind() {
    const aem_plan_view& v = plan->v;
    float* const* O = sp_ok ? F2 : F;
    ha = HArgs{g, F[3], F[4], F[5], O[3], O[4], O[5], F[0], F[1], F[2], d_den[3], d_den[4], d_den[5], v.DB, cpH,
               v.dx, v.dy, v.dz};
    ea = EArgs{g, F[0], F[1], F[2], O[0], O[1], O[2], F[3], F[4], F[5], d_idx, d_tup, d_den[0], d_den[1], d_den[2],
               cpE, v.dx, v.dy, v.dz};
    pa = PortArgs{g, d_ports, v.n_ports, v.M, v.time_steps, d_V, d_I, d_Vout, d_Iout, F[0], F[1], F[2], F[3], F[4],
                  F[5], v.dx, v.dy, v.dz, pu, d_step};
    na = NtffArgs{g, nf, nb, nx_acc.win, F[0], F[1], F[2], F[3], F[4], F[5], d_step};
    spa = SpiceArgs{g, d_spice, v.n_spice, v.n_spice_extra, v.time_steps, d_spI, d_spV, d_spIout, d_spVout,
                    F[0], F[1], F[2], F[3], F[4], F[5], v.dx, v.dy, v.dz, d_step};
}

// After a single-pass step: the other buffer set now holds the fields.
// The device state one step depends on, as (address, bytes): the six fields
// and the CPML psi of the current buffer set, the port lines, the Drude
// currents and E_old, the step counter.  (Thin sheets and SPICE ports, which
// the adjoint does not take, are not part of it.)
std::vector<std::pair<void*, size_t>> aem_sim::state_regions() {
    const aem_plan_view& v = plan->v;
    std::vector<std::pair<void*, size_t>> r;
    for (int q = 0; q < 6; ++q) r.push_back({F[q], ((size_t)v.ncell + v.plane) * 4});
    for (int f = 0; f < 6; ++f) {
        if (f == 1) continue;   // the two x faces share one allocation, face 0 first
        const int a = f / 2;
        size_t nt = (size_t)std::max(v.npml[f] - 1, 0);
        if (f == 0) nt += (size_t)std::max(v.npml[1] - 1, 0);
        const size_t trans = a == 0 ? (size_t)v.NY * v.NZ : (a == 1 ? (size_t)v.PX * v.NZ : (size_t)v.PX * v.NY);
        for (int q = 0; q < 2; ++q)
            if (nt * trans) {
                r.push_back({cpH.psi[f][q], nt * trans * 4});
                r.push_back({cpE.psi[f][q], nt * trans * 4});
            }
    }
    if (v.n_ports > 0) {
        r.push_back({d_V, (size_t)v.n_ports * v.M * 4});
        r.push_back({d_I, (size_t)v.n_ports * v.M * 4});
    }
    if (v.has_drude)
        for (int c = 0; c < 3; ++c)
            if (db[c].nb > 0) {
                r.push_back({db[c].J, (size_t)db[c].nb * v.n_poles * 4});
                r.push_back({db[c].Eold, (size_t)db[c].nb * 4});
            }
    r.push_back({d_step, sizeof(int)});
    return r;
}

void aem_sim::snapshot(int q) {
    size_t o = 0;
    for (const auto& reg : state_regions()) {
        HIPCHK(hipMemcpyAsync(ck_snap[q] + o, reg.first, reg.second, hipMemcpyDeviceToDevice, st));
        o += (reg.second + 255) & ~(size_t)255;
    }
}

void aem_sim::restore(int q) {
    size_t o = 0;
    for (const auto& reg : state_regions()) {
        HIPCHK(hipMemcpyAsync(reg.first, ck_snap[q] + o, reg.second, hipMemcpyDeviceToDevice, st));
        o += (reg.second + 255) & ~(size_t)255;
    }
}

void aem_sim::flip() {
    for (int q = 0; q < 6; ++q) std::swap(F[q], F2[q]);
    for (int f = 0; f < 2; ++f)
        for (int q = 0; q < 2; ++q) {
            std::swap(cpH.psi[f][q], cpH.pso[f][q]);
            std::swap(cpE.psi[f][q], cpE.pso[f][q]);
        }
    par ^= 1;
    bind();
}

// Modal source descriptors of a kmax pair (the H plane for TE, the E plane
// for TM), set once both parts exist.  The per-step Gaussian factor and
// carrier phase of each plane come from the plan's tables.
void aem_sim::kmax_setup() {
    const aem_plan_view& v = plan->v;
    if (v.k_axis < 0) return;
    const int T = v.time_steps;
    const float* ks = up<float>(v.ksrc, (size_t)4 * T);
    const float dd[3] = {v.dx, v.dy, v.dz};
    for (int e = 0; e < 2; ++e) {
        KmaxSrc& k = e ? kmE : kmH;
        k.g = g;
        k.plane = e ? v.kpl_E : v.kpl_H;
        for (int a = 0; a < 3; ++a) {
            k.lo[a] = v.tf_lo[a];
            k.hi[a] = v.tf_hi[a];
            k.w[a] = e ? v.WH[a] : v.WE[a];
            k.kexc[a] = v.k_exc[a];
            k.kv[a] = v.k_vec[a];
            k.d[a] = dd[a];
        }
        k.c0 = e ? v.dt / (aem::kEp0 * dd[v.k_axis]) : v.dt / (aem::kMu0 * dd[v.k_axis]);
        k.gss = ks + (size_t)(2 * e) * T;
        k.ph = ks + (size_t)(2 * e + 1) * T;
        const int o = e ? 0 : 3;
        k.xr = F[o]; k.yr = F[o + 1]; k.zr = F[o + 2];
        k.xi = kim->F[o]; k.yi = kim->F[o + 1]; k.zi = kim->F[o + 2];
        k.step = d_step;
    }
    kmax_table();
    HIPCHK(hipStreamSynchronize(st));
}

// The modal source's increments c0 Inc k_exception of every step, built on
// the host with the formulas of k_kmax_src_h / k_kmax_src_e and the host
// libm's sincosf (what the reference's CEXP evaluates), when the table fits
// kKmaxTabBytes; larger decks evaluate Inc on the device instead, whose
// cosine and sine are correctly rounded and so differ from the reference's
// in the last bit for about 1% of the arguments.
constexpr size_t kKmaxTabBytes = size_t(2) << 30;

void aem_sim::kmax_table() {
    const aem_plan_view& v = plan->v;
    const int A = v.k_axis;
    const bool te = A >= 0 && v.mode_type == 0 && v.plane_wave_amp > 0;
    const bool tm = A >= 0 && v.mode_type == 1 && v.plane_wave_amp > 0;
    if (!te && !tm) return;
    struct Cell { long long o; float w, kd, exc; };
    std::vector<Cell> grp[2];
    int fld[2];
    const float dx = v.dx, dy = v.dy, dz = v.dz;
    const float kx = v.k_vec[0], ky = v.k_vec[1], kz = v.k_vec[2];
    const float* W = te ? v.WE : v.WH;
    auto at = [&](int i, int j, int k) {
        return (long long)(i - 1) + (long long)v.PX * ((long long)(j - 1) + (long long)v.NY * (long long)(k - 1));
    };
    auto h = [](int q) { return (float)q + 0.5f; };
    if (te) {
        const int pl = v.kpl_H;
        const int* lo = v.tf_lo;
        const int* hi = v.tf_hi;
        if (A == 0) {
            fld[0] = 5; fld[1] = 4;   // Hz, Hy
            for (int k = lo[2]; k <= hi[2]; ++k)
                for (int j = lo[1]; j <= hi[1]; ++j) {
                    grp[0].push_back({at(pl, j, k), W[1], ky * (h(j) * dy) + kz * ((float)k * dz), v.k_exc[2]});
                    grp[1].push_back({at(pl, j, k), W[2], ky * ((float)j * dy) + kz * (h(k) * dz), v.k_exc[1]});
                }
        } else if (A == 1) {
            fld[0] = 5; fld[1] = 3;   // Hz, Hx
            for (int k = lo[2]; k <= hi[2]; ++k)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    grp[0].push_back({at(i, pl, k), W[0], kx * (h(i) * dx) + kz * ((float)k * dz), v.k_exc[2]});
                    grp[1].push_back({at(i, pl, k), W[2], kx * ((float)i * dx) + kz * (h(k) * dz), v.k_exc[0]});
                }
        } else {
            fld[0] = 3; fld[1] = 4;   // Hx, Hy
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    grp[0].push_back({at(i, j, pl), W[1], ky * (h(j) * dy) + kx * ((float)i * dx), v.k_exc[0]});
                    grp[1].push_back({at(i, j, pl), W[0], ky * ((float)j * dy) + kx * (h(i) * dx), v.k_exc[1]});
                }
        }
    } else {
        const int pl = v.kpl_E;
        if (A == 0) {
            fld[0] = 2; fld[1] = 1;   // Ez, Ey
            for (int k = 1; k <= v.NZ - 1; ++k)
                for (int j = 2; j <= v.NY - 1; ++j)
                    grp[0].push_back({at(pl, j, k), W[1], ky * ((float)j * dy) + kz * (h(k) * dz), v.k_exc[2]});
            for (int k = 2; k <= v.NZ - 1; ++k)
                for (int j = 1; j <= v.NY - 1; ++j)
                    grp[1].push_back({at(pl, j, k), W[2], ky * (h(j) * dy) + kz * ((float)k * dz), v.k_exc[1]});
        } else if (A == 1) {
            fld[0] = 2; fld[1] = 0;   // Ez, Ex
            for (int k = 1; k <= v.NZ - 1; ++k)
                for (int i = 2; i <= v.NX - 1; ++i)
                    grp[0].push_back({at(i, pl, k), W[0], kx * ((float)i * dx) + kz * (h(k) * dz), v.k_exc[2]});
            for (int k = 2; k <= v.NZ - 1; ++k)
                for (int i = 1; i <= v.NX - 1; ++i)
                    grp[1].push_back({at(i, pl, k), W[2], kx * (h(i) * dx) + kz * ((float)k * dz), v.k_exc[0]});
        } else {
            fld[0] = 0; fld[1] = 1;   // Ex, Ey
            for (int j = 2; j <= v.NY - 1; ++j)
                for (int i = 1; i <= v.NX - 1; ++i)
                    grp[0].push_back({at(i, j, pl), W[1], ky * ((float)j * dy) + kx * (h(i) * dx), v.k_exc[0]});
            for (int j = 1; j <= v.NY - 1; ++j)
                for (int i = 2; i <= v.NX - 1; ++i)
                    grp[1].push_back({at(i, j, pl), W[0], ky * (h(j) * dy) + kx * ((float)i * dx), v.k_exc[1]});
        }
    }
    const int T = v.time_steps;
    const int n0 = (int)grp[0].size(), nc = n0 + (int)grp[1].size();
    if (nc == 0 || (size_t)T * 2 * nc * sizeof(float) > kKmaxTabBytes) return;
    std::vector<Cell> cells(grp[0]);
    cells.insert(cells.end(), grp[1].begin(), grp[1].end());
    const float dd = A == 0 ? dx : (A == 1 ? dy : dz);
    const float c0 = te ? v.dt / (aem::kMu0 * dd) : v.dt / (aem::kEp0 * dd);
    const float* gss = v.ksrc + (size_t)(te ? 0 : 2) * T;
    const float* ph = v.ksrc + (size_t)(te ? 1 : 3) * T;
    std::vector<long long> offs(nc);
    for (int q = 0; q < nc; ++q) offs[q] = cells[q].o;
    ktab.off = up<long long>(offs.data(), offs.size());
    float* dinc = dalloc<float>((size_t)T * 2 * nc, owned);
    // chunks of steps, rows split over host threads (each writes its own rows)
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)T, (size_t(64) << 20) / (2 * nc * sizeof(float))));
    std::vector<float> buf((size_t)chunk * 2 * nc);
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (int n0c = 1; n0c <= T; n0c += chunk) {
        const int m = std::min(chunk, T - n0c + 1);
        auto work = [&](int t0) {
            for (int r = t0; r < m; r += nth) {
                const int n = n0c + r;
                const float gs = gss[n - 1], tht = ph[n - 1];
                float* row = buf.data() + (size_t)r * 2 * nc;
                for (int q = 0; q < nc; ++q) {
                    const Cell& cl = cells[q];
                    float sn, cs;
                    sincosf(tht - cl.kd, &sn, &cs);
                    const float amp = cl.w * gs;
                    row[q] = c0 * (amp * cs) * cl.exc;
                    row[nc + q] = c0 * (amp * sn) * cl.exc;
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
        HIPCHK(hipMemcpy(dinc + (size_t)(n0c - 1) * 2 * nc, buf.data(), (size_t)m * 2 * nc * sizeof(float),
                         hipMemcpyHostToDevice));
    }
    ktab.inc = dinc;
    ktab.n0 = n0;
    ktab.ncell = nc;
    ktab.r0 = F[fld[0]]; ktab.i0 = kim->F[fld[0]];
    ktab.r1 = F[fld[1]]; ktab.i1 = kim->F[fld[1]];
    ktab.step = d_step;
}

// One step of a kmax pair on stream s: both parts run each phase of the
// ordinary step; the modal source follows the H (TE) or E (TM) sheets/Drude
// updates where the TF/SF source would be (f90:4558-4625, 5857-5948), and
// the Bloch copies replace the periodic copies after the CPML corrections
// (f90:5377-5433, 6672-6741).
void aem_sim::enqueue_kmax_step(hipStream_t s) {
    const aem_plan_view& v = plan->v;
    aem_sim* im = kim;
    const dim3 bb(32, 8);
    const int N[3] = {v.NX, v.NY, v.NZ};
    const int a = v.k_axis;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;   // the source plane's axes
    const bool te = a >= 0 && v.mode_type == 0 && v.plane_wave_amp > 0;
    const bool tm = a >= 0 && v.mode_type == 1 && v.plane_wave_amp > 0;
    auto both = [&](unsigned parts) {
        enqueue_parts(nullptr, s, parts);
        im->enqueue_parts(nullptr, s, parts);
    };
    // Bloch copies of one phase, x then y then z; the H copy to the last
    // cell multiplies by exp(-i th), the E copy to the first by exp(+i th)
    auto bloch = [&](bool e) {
        for (int q = 0; q < 3; ++q) {
            if (!v.pbc[q]) continue;
            const int u = q == 0 ? 1 : 0, w = q == 2 ? 1 : 2;
            const dim3 gp(cdiv(N[u] - 1, 32), cdiv(N[w] - 1, 8));
            // (Hz, Hy) / (Ez, Ey) on x, (z, x) on y, (y, x) on z
            const int ca = q == 2 ? 1 : 2, cb = q == 0 ? 1 : 0;
            const int o = e ? 0 : 3;
            if (e)
                k_kmax_pbc<true><<<gp, bb, 0, s>>>(g, q, v.bloch_c[q], v.bloch_s[q], F[o + ca], im->F[o + ca],
                                                    F[o + cb], im->F[o + cb]);
            else
                k_kmax_pbc<false><<<gp, bb, 0, s>>>(g, q, v.bloch_c[q], -v.bloch_s[q], F[o + ca], im->F[o + ca],
                                                     F[o + cb], im->F[o + cb]);
        }
    };
    both(kPartHBulk | kPartHPre);
    if (te && ktab.ncell > 0) {
        k_kmax_src_tab<<<cdiv(ktab.ncell, 256), 256, 0, s>>>(ktab);
    } else if (te) {
        const dim3 gs(cdiv(kmH.hi[b] - kmH.lo[b] + 1, 32), cdiv(kmH.hi[c] - kmH.lo[c] + 1, 8));
        if (a == 0) k_kmax_src_h<0><<<gs, bb, 0, s>>>(kmH);
        else if (a == 1) k_kmax_src_h<1><<<gs, bb, 0, s>>>(kmH);
        else k_kmax_src_h<2><<<gs, bb, 0, s>>>(kmH);
    }
    both(kPartHPort | kPartHPost);
    bloch(false);
    both(kPartEBulk | kPartEPre);
    if (tm && ktab.ncell > 0) {
        k_kmax_src_tab<<<cdiv(ktab.ncell, 256), 256, 0, s>>>(ktab);
    } else if (tm) {
        const dim3 gs(cdiv(N[b] - 1, 32), cdiv(N[c] - 1, 8));
        if (a == 0) k_kmax_src_e<0><<<gs, bb, 0, s>>>(kmE);
        else if (a == 1) k_kmax_src_e<1><<<gs, bb, 0, s>>>(kmE);
        else k_kmax_src_e<2><<<gs, bb, 0, s>>>(kmE);
    }
    both(kPartEPort | kPartEPost);
    bloch(true);
    both(kPartEOld | kPartOutPre | kPartNtff | kPartOutPost | kPartStep);
}

void aem_sim::drop_graphs() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (gexec2) (void)hipGraphExecDestroy(gexec2);
    if (graph2) (void)hipGraphDestroy(graph2);
    gexec = gexec2 = nullptr;
    graph = graph2 = nullptr;
}

// Enqueue the selected parts of one step on stream s, in the reference's
// order.  A simulation run alone enqueues every part; a batch replaces the
// bulk, port, far-field and step-counter parts by batched launches and
// enqueues the remaining parts per member between them.
template <class IdxT>
void aem_sim::enqueue_step_t(hipEvent_t* mark, hipStream_t s, unsigned parts) {
    const aem_plan_view& v = plan->v;
    const IdxT* idx = static_cast<const IdxT*>(d_idx);
    float *Ex = F[0], *Ey = F[1], *Ez = F[2], *Hx = F[3], *Hy = F[4], *Hz = F[5];
    const dim3 bb(32, 8, 1);
    const dim3 bt(kTX, kTY, 1);   // bulk tile
    const int T = v.time_steps;
    // 32-bit byte offsets where the arrays stay under 4 GiB
    auto h_bulk = [&](const Launch& L) {
        if (off32) switch (L.c) {
            case 0: k_h_bulk<0, false, uint32_t><<<L.blocks(), bt, 0, s>>>(ha, L.bs); break;
            case kCx: k_h_bulk<kCx, false, uint32_t><<<L.blocks(), bt, 0, s>>>(ha, L.bs); break;
            default: k_h_bulk<kCx | kCy | kCz, false, uint32_t><<<L.blocks(), bt, 0, s>>>(ha, L.bs); break;
        }
        else switch (L.c) {
            case 0: k_h_bulk<0><<<L.blocks(), bt, 0, s>>>(ha, L.bs); break;
            case kCx: k_h_bulk<kCx><<<L.blocks(), bt, 0, s>>>(ha, L.bs); break;
            default: k_h_bulk<kCx | kCy | kCz><<<L.blocks(), bt, 0, s>>>(ha, L.bs); break;
        }
    };
    auto e_bulk = [&](const Launch& L) {
        if (off32) switch (L.c) {
            case 0: k_e_bulk<IdxT, 0, false, uint32_t><<<L.blocks(), bt, 0, s>>>(ea, L.bs); break;
            case kCx: k_e_bulk<IdxT, kCx, false, uint32_t><<<L.blocks(), bt, 0, s>>>(ea, L.bs); break;
            default: k_e_bulk<IdxT, kCx | kCy | kCz, false, uint32_t><<<L.blocks(), bt, 0, s>>>(ea, L.bs); break;
        }
        else switch (L.c) {
            case 0: k_e_bulk<IdxT, 0><<<L.blocks(), bt, 0, s>>>(ea, L.bs); break;
            case kCx: k_e_bulk<IdxT, kCx><<<L.blocks(), bt, 0, s>>>(ea, L.bs); break;
            default: k_e_bulk<IdxT, kCx | kCy | kCz><<<L.blocks(), bt, 0, s>>>(ea, L.bs); break;
        }
    };

    // single pass: the core box's H and E in one launch, then the y/z slab
    // boxes' H and E as ping-pong bulk launches; the rest of the step then
    // runs on the new buffer set
    const bool sp = (parts & kPartHBulk) && (parts & kPartEBulk) && sp_active();

    // ---- H phase ----  (marks: [0,1] the core bulk launch -- the single-pass
    // launch when sp --, [1,2] the rest of H)
    if (mark) HIPCHK(hipEventRecord(mark[0], st));
    if (sp) {
        const SpArgs sa{g, Ex, Ey, Ez, Hx, Hy, Hz, F2[0], F2[1], F2[2], F2[3], F2[4], F2[5], d_idx, d_idx4, d_tup,
                        n_tup, d_den[0], d_den[1], d_den[2], d_den[3], d_den[4], d_den[5], v.DB, cpH, cpE, v.dx, v.dy, v.dz};
        for (const auto& part : sp_parts) {
            const Box& b = part.first;
            const unsigned nblk = cdiv(b.ni, 64) * cdiv(b.nj, kSY * kSpRows) * cdiv(b.nk, b.kc);
            if (part.second == kCx) k_step_sp<IdxT, kCx><<<nblk, dim3(64, kSY), 0, s>>>(sa, b);
            else k_step_sp<IdxT, 0><<<nblk, dim3(64, kSY), 0, s>>>(sa, b);
        }
    } else if (parts & kPartHBulk) {
        h_bulk(bulk[0]);
    }
    if (mark) HIPCHK(hipEventRecord(mark[1], st));
    if (sp) {
        EArgs es = ea;   // E of the slabs reads the new H
        es.Hx = F2[3];
        es.Hy = F2[4];
        es.Hz = F2[5];
        // (single-pass decks always have 32-bit offsets)
        for (size_t q = 1; q < bulk.size(); ++q)
            k_h_bulk<kCx | kCy | kCz, true, uint32_t><<<bulk[q].blocks(), bt, 0, s>>>(ha, bulk[q].bs);
        for (size_t q = 1; q < bulk.size(); ++q)
            k_e_bulk<IdxT, kCx | kCy | kCz, true, uint32_t><<<bulk[q].blocks(), bt, 0, s>>>(es, bulk[q].bs);
        flip();
        Ex = F[0]; Ey = F[1]; Ez = F[2]; Hx = F[3]; Hy = F[4]; Hz = F[5];
    } else if (parts & kPartHBulk) {
        for (size_t q = 1; q < bulk.size(); ++q) h_bulk(bulk[q]);
    }
    const float dd[3] = {v.dx, v.dy, v.dz};
    for (int a = 0; a < 3; ++a)
        if ((parts & kPartHPre) && sh[a].np > 0) {
            const float f = v.DB * (v.sheet_thickness / dd[a]);
            const dim3 gs(cdiv(sh[a].P, 32), cdiv(sh[a].Q, 8));
            k_h_sheets<<<gs, bb, 0, s>>>(g, sh[a], a, Hx, Hy, Hz, Ex, Ey, Ez, d_den[3], d_den[4], d_den[5], f);
        }
    // the incident wave's faces, then (infinite ground plane) the mirrored
    // wave's, as the reference orders them (f90:4627-4790, 5950-6111)
    const int ntf = v.mirror_type > 0 ? 2 : 1;
    if ((parts & kPartHPre) && v.has_tfsf) {
        const int nx = v.tf_hi[0] - v.tf_lo[0] + 2, ny = v.tf_hi[1] - v.tf_lo[1] + 2, nz = v.tf_hi[2] - v.tf_lo[2] + 2;
        for (int w = 0; w < ntf; ++w) {
            const Tfsf& t = w ? tfm : tf;
            if (fuse_cpml) {   // fused path: the three face pairs in one launch
                if (nx > 0 && ny > 0 && nz > 0)
                    k_h_tfsf3<<<dim3(cdiv(std::max(nx, ny), 32), cdiv(std::max(ny, nz), 8), 3), bb, 0, s>>>(g, t, Hx, Hy,
                                                                                                            Hz, d_step);
            } else {           // one launch per face pair, in the reference's order
                if (!v.pbc[1] && nx > 0 && nz > 0) k_h_tfsf<<<dim3(cdiv(nx, 32), cdiv(nz, 8)), bb, 0, s>>>(g, t, 0, Hx, Hy, Hz, d_step);
                if (!v.pbc[2] && nx > 0 && ny > 0) k_h_tfsf<<<dim3(cdiv(nx, 32), cdiv(ny, 8)), bb, 0, s>>>(g, t, 1, Hx, Hy, Hz, d_step);
                if (!v.pbc[0] && ny > 0 && nz > 0) k_h_tfsf<<<dim3(cdiv(ny, 32), cdiv(nz, 8)), bb, 0, s>>>(g, t, 2, Hx, Hy, Hz, d_step);
            }
        }
    }
    if ((parts & kPartHPort) && v.n_ports > 0) k_ports_h<<<1, 256, 0, s>>>(pa);
    if (parts & kPartHPost) {
        const long long tx = (long long)(std::max(v.npml[0] - 1, 0) + std::max(v.npml[1] - 1, 0)) * (v.NY - 1) * (v.NZ - 1);
        const long long ty = (long long)(v.NX - 1) * (std::max(v.npml[2] - 1, 0) + std::max(v.npml[3] - 1, 0)) * (v.NZ - 1);
        const long long tz = (long long)(v.NX - 1) * (v.NY - 1) * (std::max(v.npml[4] - 1, 0) + std::max(v.npml[5] - 1, 0));
        if (!fuse_cpml && tx > 0) k_h_cpml_x<<<cdiv(tx, 256), 256, 0, s>>>(g, cpH, Hy, Hz, Ey, Ez, v.DB, rdv[0]);
        if (!fuse_cpml && ty > 0) k_h_cpml_y<<<cdiv(ty, 256), 256, 0, s>>>(g, cpH, Hx, Hz, Ex, Ez, v.DB, rdv[1]);
        if (!fuse_cpml && tz > 0) k_h_cpml_z<<<cdiv(tz, 256), 256, 0, s>>>(g, cpH, Hx, Hy, Ex, Ey, v.DB, rdv[2]);
        // (a kmax pair makes its Bloch copies across both parts)
        if (v.pbc[0] && !v.kmax) k_h_pbc<<<dim3(cdiv(v.NY - 1, 32), cdiv(v.NZ - 1, 8)), bb, 0, s>>>(g, 0, Hz, Hy);
        if (v.pbc[1] && !v.kmax) k_h_pbc<<<dim3(cdiv(v.NX - 1, 32), cdiv(v.NZ - 1, 8)), bb, 0, s>>>(g, 1, Hz, Hx);
        if (v.pbc[2] && !v.kmax) k_h_pbc<<<dim3(cdiv(v.NX - 1, 32), cdiv(v.NY - 1, 8)), bb, 0, s>>>(g, 2, Hy, Hx);
    }

    if ((parts & kPartSpiceH) && v.n_spice > 0) k_spice_h<<<1, 64, 0, s>>>(spa);

    if ((parts & kPartRecord) && rec_on && !des_on)   // a design box records in k_e_design
        k_adj_record<<<cdiv(rec.nedge, 256), 256, 0, s>>>(g, rec, Ex, Ey, Ez, Hx, Hy, Hz, d_den[0], d_den[1], d_den[2],
                                                          d_step);

    // ---- E phase ----  (marks: [2,3] the core bulk launch, [3,4] the rest of E)
    if (mark) HIPCHK(hipEventRecord(mark[2], st));
    if ((parts & kPartEBulk) && !sp) e_bulk(bulk[0]);
    if (mark) HIPCHK(hipEventRecord(mark[3], st));
    if ((parts & kPartEBulk) && !sp)
        for (size_t q = 1; q < bulk.size(); ++q) e_bulk(bulk[q]);
    if ((parts & kPartEPre) && des_on)
        k_e_design<<<cdiv(des.nedge, 256), 256, 0, s>>>(g, des, Ex, Ey, Ez, Hx, Hy, Hz, d_den[0], d_den[1], d_den[2],
                                                         rec, rec_on ? 1 : 0, d_step);
    for (int a = 0; a < 3; ++a)
        if ((parts & kPartEPre) && sh[a].np > 0) {
            const dim3 gs(cdiv(sh[a].P, 32), cdiv(sh[a].Q, 8));
            k_e_sheets<<<gs, bb, 0, s>>>(g, sh[a], a, Hx, Hy, Hz, d_den[0], d_den[1], d_den[2], v.dt, aem::kEp0);
        }
    if ((parts & kPartEPre) && v.has_drude)
        for (int c = 0; c < 3; ++c)
            if (db[c].nb > 0) k_drude<IdxT><<<cdiv(db[c].nb, 256), 256, 0, s>>>(g, db[c], c, v.n_poles, F[c], idx, d_tup);
    if ((parts & kPartEPre) && v.has_tfsf) {
        const int nx = v.tf_hi[0] - v.tf_lo[0] + 2, ny = v.tf_hi[1] - v.tf_lo[1] + 2, nz = v.tf_hi[2] - v.tf_lo[2] + 2;
        for (int w = 0; w < ntf; ++w) {
            const Tfsf& t = w ? tfm : tf;
            if (fuse_cpml) {
                if (nx > 0 && ny > 0 && nz > 0)
                    k_e_tfsf3<<<dim3(cdiv(std::max(nx, ny), 32), cdiv(std::max(ny, nz), 8), 3), bb, 0, s>>>(g, t, Ex, Ey,
                                                                                                            Ez, d_step);
            } else {
                if (!v.pbc[1] && nx > 0 && nz > 0) k_e_tfsf<<<dim3(cdiv(nx, 32), cdiv(nz, 8)), bb, 0, s>>>(g, t, 0, Ex, Ey, Ez, d_step);
                if (!v.pbc[2] && nx > 0 && ny > 0) k_e_tfsf<<<dim3(cdiv(nx, 32), cdiv(ny, 8)), bb, 0, s>>>(g, t, 1, Ex, Ey, Ez, d_step);
                if (!v.pbc[0] && ny > 0 && nz > 0) k_e_tfsf<<<dim3(cdiv(ny, 32), cdiv(nz, 8)), bb, 0, s>>>(g, t, 2, Ex, Ey, Ez, d_step);
            }
        }
    }
    if ((parts & kPartEPort) && v.n_ports > 0) k_ports_e<<<1, 256, 0, s>>>(pa);
    if ((parts & kPartSpiceE) && v.n_spice > 0) k_spice_e<<<1, 64, 0, s>>>(spa);
    if (parts & kPartEPost) {
        const long long tx = (long long)(std::max(v.npml[0] - 1, 0) + std::max(v.npml[1] - 1, 0)) * (v.NY - 1) * (v.NZ - 1);
        const long long ty = (long long)(v.NX - 1) * (std::max(v.npml[2] - 1, 0) + std::max(v.npml[3] - 1, 0)) * (v.NZ - 1);
        const long long tz = (long long)(v.NX - 1) * (v.NY - 1) * (std::max(v.npml[4] - 1, 0) + std::max(v.npml[5] - 1, 0));
        if (!fuse_cpml && tx > 0) k_e_cpml_x<IdxT><<<cdiv(tx, 256), 256, 0, s>>>(g, cpE, Ey, Ez, Hy, Hz, idx, d_tup, rdv[0]);
        if (!fuse_cpml && ty > 0) k_e_cpml_y<IdxT><<<cdiv(ty, 256), 256, 0, s>>>(g, cpE, Ex, Ez, Hx, Hz, idx, d_tup, rdv[1]);
        if (!fuse_cpml && tz > 0) k_e_cpml_z<IdxT><<<cdiv(tz, 256), 256, 0, s>>>(g, cpE, Ex, Ey, Hx, Hy, idx, d_tup, rdv[2]);
        if (v.pbc[0] && !v.kmax) {
            k_e_pbc<<<dim3(cdiv(v.NY - 1, 32), cdiv(v.NZ - 1, 8)), bb, 0, s>>>(g, 0, Ez, Ey);
            if (sh[2].np) k_sheet_pbc<<<dim3(cdiv(v.NY - 1, 256), sh[2].np), 256, 0, s>>>(sh[2], 0, v.NX - 1, v.NY - 1);
            if (sh[1].np) k_sheet_pbc<<<dim3(cdiv(v.NZ - 1, 256), sh[1].np), 256, 0, s>>>(sh[1], 0, v.NX - 1, v.NZ - 1);
        }
        if (v.pbc[1] && !v.kmax) {
            k_e_pbc<<<dim3(cdiv(v.NX - 1, 32), cdiv(v.NZ - 1, 8)), bb, 0, s>>>(g, 1, Ez, Ex);
            if (sh[2].np) k_sheet_pbc<<<dim3(cdiv(v.NX - 1, 256), sh[2].np), 256, 0, s>>>(sh[2], 1, v.NY - 1, v.NX - 1);
            if (sh[0].np) k_sheet_pbc<<<dim3(cdiv(v.NZ - 1, 256), sh[0].np), 256, 0, s>>>(sh[0], 0, v.NY - 1, v.NZ - 1);
        }
        if (v.pbc[2] && !v.kmax) {
            k_e_pbc<<<dim3(cdiv(v.NX - 1, 32), cdiv(v.NY - 1, 8)), bb, 0, s>>>(g, 2, Ey, Ex);
            if (sh[1].np) k_sheet_pbc<<<dim3(cdiv(v.NX - 1, 256), sh[1].np), 256, 0, s>>>(sh[1], 1, v.NZ - 1, v.NX - 1);
            if (sh[0].np) k_sheet_pbc<<<dim3(cdiv(v.NY - 1, 256), sh[0].np), 256, 0, s>>>(sh[0], 1, v.NZ - 1, v.NY - 1);
        }
    }
    if ((parts & kPartEOld) && v.has_drude)
        for (int c = 0; c < 3; ++c)
            if (db[c].nb > 0) k_save_eold<<<cdiv(db[c].nb, 256), 256, 0, s>>>(g, db[c], F[c]);

    // ---- per-step outputs ----
    if (mark) HIPCHK(hipEventRecord(mark[4], st));
    if ((parts & kPartOutPre) && v.kmax && v.k_axis >= 0) {
        // TE and TM sums of the kmax build (f90:7105-7190): TE = Ea k1 - Eb k2, TM = Ea k2 + Eb k1
        const int ax = v.k_axis;
        const int lo = v.tf_lo[ax], hi = v.tf_hi[ax];
        const int p = ax == 0 ? 1 : 0, q = ax == 2 ? 1 : 2;
        const float* Ea = ax == 2 ? Ex : Ez;
        const float* Eb = ax == 0 ? Ey : (ax == 1 ? Ex : Ey);
        // (x normal: Ez, Ey with k_y, k_z; y normal: Ez, Ex with k_x, k_z; z normal: Ex, Ey with k_y, k_x)
        const float k1 = ax == 2 ? v.k_exc[1] : v.k_exc[p], k2 = ax == 2 ? v.k_exc[0] : v.k_exc[q];
        k_plane_sums<<<1, 256, 0, s>>>(g, ax, lo, hi, v.ff_lo[p], v.ff_hi[p], v.ff_lo[q], v.ff_hi[q], Ea, Eb, k1, k2,
                                        1.0f, 1, d_refl, d_trans, d_step);
        k_plane_sums<<<1, 256, 0, s>>>(g, ax, lo, hi, v.ff_lo[p], v.ff_hi[p], v.ff_lo[q], v.ff_hi[q], Ea, Eb, k2, k1,
                                        1.0f, 0, d_refl2, d_trans2, d_step);
    } else if ((parts & kPartOutPre) && v.plane_sum_axis >= 0) {
        const int ax = v.plane_sum_axis;
        const int lo = v.tf_lo[ax], hi = v.tf_hi[ax];
        if (ax == 0)
            k_plane_sums<<<1, 256, 0, s>>>(g, 0, lo, hi, v.ff_lo[1], v.ff_hi[1], v.ff_lo[2], v.ff_hi[2], Ez, Ey, v.s_pol,
                                            v.c_pol, v.c_ph, 1, d_refl, d_trans, d_step);
        else if (ax == 1)
            k_plane_sums<<<1, 256, 0, s>>>(g, 1, lo, hi, v.ff_lo[0], v.ff_hi[0], v.ff_lo[2], v.ff_hi[2], Ez, Ex, v.s_pol,
                                            v.c_pol, v.s_ph, 0, d_refl, d_trans, d_step);
        else
            k_plane_sums<<<1, 256, 0, s>>>(g, 2, lo, hi, v.ff_lo[0], v.ff_hi[0], v.ff_lo[1], v.ff_hi[1], Ey, Ex, v.ps_c1,
                                            v.ps_c2, 1.0f, 0, d_refl, d_trans, d_step);
    }
    if ((parts & kPartNtff) && v.n_angles > 0 && ntff_mode != AEM_NTFF_OFF) {
        const dim3 gf(cdiv(ntff_mu, 32), cdiv(ntff_mv, 8), 6);
        k_ntff_currents<<<gf, bb, 0, s>>>(g, nf, Ex, Ey, Ez, Hx, Hy, Hz);
        if (ntff_mode == AEM_NTFF_ORDERED) {
            if (nl.n_waves > 0)
                k_ntff_list<<<cdiv(nl.n_waves, 4), 256, 0, s>>>(nl, nb, d_step);
            else if (win_max > 0)
                k_ntff_exact<<<dim3(cdiv(win_max, 256), 6, v.n_angles), 256, 0, s>>>(nf, nb, nx_acc, d_step);
        } else if (ntff_mode == AEM_NTFF_HISTOGRAM) {
            k_ntff_hist<<<dim3(v.n_angles, 6), 256, 0, s>>>(nf, nb, nx_acc.win, d_step);
        } else {
            k_ntff_bins<<<dim3(cdiv(ntff_mu, 32), cdiv(ntff_mv, 8), 6 * v.n_angles), bb, 0, s>>>(nf, nb, d_step);
        }
    }
    if ((parts & kPartObjective) && obj_on) {
        const dim3 go(cdiv(ntff_mu, 32), cdiv(ntff_mv, 8), 6);
        k_objective<<<go, 256, 0, s>>>(g, obj, Ex, Ey, Ez, Hx, Hy, Hz, d_step);
    }
    if ((parts & kPartObjective) && ff_on) {
        const dim3 gf(cdiv(ntff_mu, 32), cdiv(ntff_mv, 8), 6);
        k_ffdft<<<gf, 256, 0, s>>>(g, ff, Ex, Ey, Ez, Hx, Hy, Hz);
        k_ffdft_sum<<<ff.n_ang * ff.n_freq, 256, 0, s>>>(ff, (int)(gf.x * gf.y * gf.z), d_step);
    }
    if ((parts & kPartOutPost) && vid_on)
        k_video<<<dim3(cdiv(vid.v1, 32), cdiv(vid.v2, 8)), bb, 0, s>>>(g, vid, Ex, Ey, Ez, Hx, Hy, Hz, d_step);
    const bool pc_on = (parts & kPartOutPost) && v.n_angles > 0 && v.plane_wave_amp > 0;
    if (parts & kPartStep)
        k_step_end<<<1, 1, 0, s>>>(g, v.ic, v.jc, v.kc, T, Ex, Ey, Ez, Hx, Hy, Hz, d_pc, d_step, pc_on ? 1 : 0);
    else if (pc_on)
        k_phase_centre<<<1, 1, 0, s>>>(g, v.ic, v.jc, v.kc, T, Ex, Ey, Ez, Hx, Hy, Hz, d_pc, d_step);
    if (mark) HIPCHK(hipEventRecord(mark[5], st));
}

void aem_sim::enqueue_parts(hipEvent_t* mark, hipStream_t s, unsigned parts) {
    if (plan->v.idx_bytes == 1) enqueue_step_t<uint8_t>(mark, s, parts);
    else enqueue_step_t<uint16_t>(mark, s, parts);
}

// One record of each <name>_full_*.bin file after the step just enqueued
// (f90:6992-7003; files opened at the first step as the reference opens
// them at setup, f90:4348-4355).  Each record is a Fortran unformatted
// sequential record: a 4-byte length, the cell-centred floats, the length.
void aem_sim::export_step() {
    if (!full_f[0]) {
        const std::string& fn = plan->deck.filename;
        const std::string stem = fn.size() > 4 ? fn.substr(0, fn.size() - 4) : fn;
        static const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
        for (int q = 0; q < 6; ++q) {
            const std::string p = stem + "_full_" + names[q] + ".bin";
            full_f[q] = std::fopen(p.c_str(), "wb");
            if (!full_f[q]) throw aem::Error(AEM_ERR_IO, "cannot write '" + p + "'");
        }
    }
    const size_t n = (size_t)full_n[0] * full_n[1] * full_n[2];
    k_full<<<dim3(cdiv(full_n[0], 32), cdiv(full_n[1], 8), full_n[2]), dim3(32, 8, 1), 0, st>>>(
        g, full_n[0], full_n[1], full_n[2], d_fullp, F[0], F[1], F[2], F[3], F[4], F[5]);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 6; ++q)
        HIPCHK(hipMemcpyAsync(h_full + q * n, d_full[q], n * sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int32_t len = (int32_t)(n * sizeof(float));
    for (int q = 0; q < 6; ++q)
        if (std::fwrite(&len, 4, 1, full_f[q]) != 1 || std::fwrite(h_full + q * n, 4, n, full_f[q]) != n ||
            std::fwrite(&len, 4, 1, full_f[q]) != 1)
            throw aem::Error(AEM_ERR_IO, "writing a full-field record failed");
}

// Contributor lists of the ordered far-field binning (k_ntff_list).  For
// angle kk and accumulator arr (Ux,Uy,Uz,Wx,Wy,Wz) a face cell with retarded
// time tfv lands at t = n + tfv - 0.5 (W: n + tfv - 1), evaluated in FP32 as
// the reference does.  Its bin offset ii - n is floor of the exact value,
// except that the FP32 rounding of t (at most one ulp of the largest t, here
// doubled) can move it across a bin edge; the cell goes into the list of
// every offset it can reach with either weight.  Lists keep the reference's
// loop order (the order k_ntff_exact walks).  Larger than kListBytes in
// total, the builder gives up and the row-scan kernel stays in use.  A
// wave's descriptor: row, first bin offset, first entry, and the entries per
// lane (64 short lists) or minus the length of one long list.
void aem_sim::build_ntff_lists(const std::vector<int>& win, unsigned stride) {
    constexpr size_t kListBytes = size_t(4) << 30;
    constexpr size_t kLong = 192;     // entries from which a bin's list gets a wave of its own
    const aem_plan_view& v = plan->v;
    const float cdt = aem::kC * v.dt;
    const double tmax = (double)v.time_steps + (double)v.len_ff + 8.0;
    const double eps = 2.0 * std::ldexp(1.0, (int)std::ceil(std::log2(tmax)) - 23);
    std::vector<int4> waves;
    std::vector<unsigned long long> skip;
    std::vector<uint2> ent;
    std::vector<std::vector<uint2>> lists;
    for (int kk = 0; kk < v.n_angles; ++kk) {
        const float* an = v.ang + 6 * kk;
        const float s1 = an[2], c1 = an[3], s2 = an[4], c2 = an[5];
        const int wlo = win[2 * kk], wlen = win[2 * kk + 1];
        for (int arr = 0; arr < 6; ++arr) {
            const int comp = arr % 3;
            const double shift = arr >= 3 ? -1.0 : -0.5;
            lists.assign((size_t)wlen, {});
            // with a ground plane the image currents follow every face of the
            // box, in the same order, their entries flagged with the top bit
            // where the component's image sign is -1 (f90:7928-8240)
            const int passes = v.mirror_type > 0 && v.pbc[0] + v.pbc[1] + v.pbc[2] < 2 ? 2 : 1;
            for (int m = 0; m < passes; ++m)
            for (int a = 0; a < 3; ++a) {                     // face pairs in the reference's order
                if (a == comp || v.pbc[a]) continue;
                const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
                const int slot = comp == b ? 0 : 1;
                for (int vv = v.ff_lo[c]; vv <= v.ff_hi[c]; ++vv)
                    for (int uu = v.ff_lo[b]; uu <= v.ff_hi[b]; ++uu)
                        for (int side = 0; side < 2; ++side) {
                            const int pos = side ? v.ff_hi[a] : v.ff_lo[a];
                            int i, j, k;
                            if (a == 0) { i = pos; j = uu; k = vv; }
                            else if (a == 1) { i = uu; j = pos; k = vv; }
                            else { i = uu; j = vv; k = pos; }
                            if (m) { i = mirror_pos(v, 0, i); j = mirror_pos(v, 1, j); k = mirror_pos(v, 2, k); }
                            const float tfv = (v.r_relay - ((float)(i - v.ic) * v.dx * s1 * c2 +
                                                            (float)(j - v.jc) * v.dy * s1 * s2 +
                                                            (float)(k - v.kc) * v.dz * c1)) / cdt;
                            const double x = (double)tfv + shift;
                            const int olo = (int)std::floor(x - eps), ohi = (int)std::floor(x + eps);
                            const int face = 2 * a + side;
                            const float sg = m ? (arr >= 3 ? v.sgn_J[comp] : v.sgn_M[comp]) : 1.0f;
                            const unsigned idx = ((unsigned)(2 * face + slot) * stride +
                                                  (unsigned)(uu + nf.fd0[face] * vv)) |
                                                 (sg < 0.0f ? 0x80000000u : 0u);
                            uint32_t tb;
                            std::memcpy(&tb, &tfv, 4);
                            for (int o = olo; o <= ohi + 1; ++o) {
                                const int q = o - wlo;
                                if (q < 0 || q >= wlen)
                                    throw aem::Error(AEM_ERR_STATE, "far-field list outside its bin window");
                                lists[q].push_back(make_uint2(idx, tb));
                            }
                        }
            }
            const int row = arr * v.n_angles + kk;
            for (int w0 = 0; w0 < wlen; w0 += 64) {
                // lists longer than kLong get a wave of their own; the rest
                // of the 64 bins share one wave, one lane per bin
                size_t m = 0;
                unsigned long long sk = 0;
                for (int l = 0; l < 64 && w0 + l < wlen; ++l) {
                    const std::vector<uint2>& li = lists[w0 + l];
                    if (li.size() > kLong) {
                        sk |= 1ull << l;
                        waves.push_back(make_int4(row, wlo + w0 + l, (int)ent.size(), -(int)li.size()));
                        skip.push_back(0ull);
                        ent.insert(ent.end(), li.begin(), li.end());
                    } else {
                        m = std::max(m, li.size());
                    }
                }
                if ((ent.size() + 64 * m) * sizeof(uint2) > kListBytes) return;
                waves.push_back(make_int4(row, wlo + w0, (int)ent.size(), (int)m));
                skip.push_back(sk);
                for (size_t p = 0; p < m; ++p)
                    for (int l = 0; l < 64; ++l) {
                        const int q = w0 + l;
                        const bool in = q < wlen && !((sk >> l) & 1ull) && p < lists[q].size();
                        ent.push_back(in ? lists[q][p] : make_uint2(0xffffffffu, 0u));
                    }
            }
        }
    }
    if (ent.size() > 0x7fffffffull) return;
    nl.ent = up<uint2>(ent.data(), ent.size());
    nl.waves = up<int4>(waves.data(), waves.size());
    nl.skip = up<unsigned long long>(skip.data(), skip.size());
    nl.n_waves = (int)waves.size();
}

void aem_sim::enqueue_watch(hipStream_t s) {
    const WatchSet ws{{F[0], F[1], F[2], F[3], F[4], F[5]}};
    k_watch<<<1024, 256, 0, s>>>(ws, (long long)(plan->v.ncell / 4), d_watch);
    if (kim) kim->enqueue_watch(s);
}

void aem_sim::check_watch() {
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, d_watch, sizeof bad, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (bad)
        throw aem::Error(AEM_ERR_NUMERIC, "non-finite field value (NaN or Inf) by step " + std::to_string(steps_done) +
                                              ": the update is unstable (check the deck's Courant factor)");
    if (kim) kim->check_watch();
}

namespace aem {
const Plan* sim_plan(const aem_sim* s) { return s ? s->plan : nullptr; }
}  // namespace aem

__global__ void k_ffpc(int T, const float* pc, float c2, float c1, float c3, float* out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= T) return;
    out[n] = pc[n] * c2 + pc[(T + 1) + n] * c1 + pc[2 * (T + 1) + n] * c3;
}

extern "C" {

aem_status aem_sim_create(const aem_plan* plan, aem_sim** out) { return aem_sim_create_ex(plan, nullptr, out); }

aem_status aem_sim_create_ex(const aem_plan* plan, const aem_sim_options* opt, aem_sim** out) {
    if (!plan || !out) {
        aem::set_error("aem_sim_create: null argument");
        return AEM_ERR_STATE;
    }
    *out = nullptr;
    if (opt && (opt->ntff_binning < AEM_NTFF_ORDERED || opt->ntff_binning > AEM_NTFF_OFF ||
                opt->planes < 0 || opt->planes > 64)) {
        aem::set_error("aem_sim_create_ex: ntff_binning must be an AEM_NTFF_* value and planes 0..64");
        return AEM_ERR_STATE;
    }
    aem_sim* s = new (std::nothrow) aem_sim();
    if (!s) {
        aem::set_error("host allocation failed");
        return AEM_ERR_STATE;
    }
    s->plan = reinterpret_cast<const aem::Plan*>(plan);
    aem_status stt = aem::guard([&]() { s->create(opt); });
    if (stt != AEM_OK) {
        delete s;
        return stt;
    }
    *out = s;
    return AEM_OK;
}

aem_status aem_create_from_inputs(const char* inputs_path, aem_sim** out) {
    aem_plan* p = nullptr;
    aem_status stt = aem_plan_from_inputs(inputs_path, &p);
    if (stt != AEM_OK) return stt;
    stt = aem_sim_create(p, out);
    if (stt != AEM_OK) {
        aem_plan_destroy(p);
        return stt;
    }
    (*out)->owned_plan = reinterpret_cast<aem::Plan*>(p);
    return AEM_OK;
}

void aem_destroy(aem_sim* sim) { delete sim; }

aem_status aem_run(aem_sim* s, int64_t n_steps, int32_t use_graph) {
    if (!s) {
        aem::set_error("aem_run: null handle");
        return AEM_ERR_STATE;
    }
    if (n_steps < 0 || s->steps_done + n_steps > s->plan->v.time_steps) {
        aem::set_error("aem_run: step count exceeds the deck's time_steps");
        return AEM_ERR_STATE;
    }
    if (s->finalized) {
        aem::set_error("aem_run: simulation already finalized");
        return AEM_ERR_STATE;
    }
    if (s->plan->v.n_spice > 0) {
        aem::set_error("aem_run: the deck has SPICE ports; step it with aem_port_exchange");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        if (n_steps == 0) return;
        if (use_graph && !s->gexec) {
            // a single-pass step alternates buffer sets: one graph per set
            const int p0 = s->par;
            HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
            s->enqueue_step();
            HIPCHK(hipStreamEndCapture(s->st, &s->graph));
            HIPCHK(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
            s->gsp = s->par != p0;
            s->gpar = p0;
            if (s->gsp) {
                HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
                s->enqueue_step();
                HIPCHK(hipStreamEndCapture(s->st, &s->graph2));
                HIPCHK(hipGraphInstantiate(&s->gexec2, s->graph2, nullptr, nullptr, 0));
            }
        }
        HIPCHK(hipEventRecord(s->ev0, s->st));
        for (int64_t q = 0; q < n_steps; ++q) {
            if (s->ck_K > 0 && (s->steps_done + q) % s->ck_K == 0) s->snapshot((int)((s->steps_done + q) / s->ck_K));
            if (use_graph) {
                HIPCHK(hipGraphLaunch(s->par == s->gpar ? s->gexec : s->gexec2, s->st));
                if (s->gsp) s->flip();
            } else {
                s->enqueue_step();
            }
            if (s->full_on) s->export_step();
            if ((q + 1) % kWatchSteps == 0 && q + 1 < n_steps) s->enqueue_watch(s->st);
            s->ra.pace(s->st, q);
        }
        s->enqueue_watch(s->st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(s->ev1, s->st));
        HIPCHK(hipEventSynchronize(s->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        s->last_ms = ms / (float)n_steps;
        s->steps_done += n_steps;
        if (s->kim) s->kim->steps_done = s->steps_done;
        s->ck_last_ok = true;
        s->check_watch();
    });
}

int64_t aem_steps_done(const aem_sim* s) { return s ? s->steps_done : -1; }

int64_t aem_core_cells(const aem_sim* s) {
    if (!s || s->bulk.empty()) return -1;
    return (int64_t)s->bulk[0].cells();
}

int32_t aem_single_pass(const aem_sim* s) { return s ? (s->sp_active() ? 1 : 0) : -1; }

float aem_last_run_ms(const aem_sim* s) { return s ? s->last_ms : 0.f; }

aem_status aem_sync(aem_sim* s) {
    if (!s) return AEM_ERR_STATE;
    return aem::guard([&]() { HIPCHK(hipStreamSynchronize(s->st)); });
}

aem_status aem_port_exchange(aem_sim* s, const float* v_in, float* i_out) {
    if (!s) {
        aem::set_error("aem_port_exchange: null handle");
        return AEM_ERR_STATE;
    }
    const aem_plan_view& v = s->plan->v;
    if (v.n_spice == 0) {
        aem::set_error("aem_port_exchange: the deck has no SPICE ports");
        return AEM_ERR_STATE;
    }
    if (s->finalized) {
        aem::set_error("aem_port_exchange: simulation already finalized");
        return AEM_ERR_STATE;
    }
    if (s->x_pending && !v_in) {
        aem::set_error("aem_port_exchange: a step waits for its node voltages (v_in)");
        return AEM_ERR_STATE;
    }
    if (i_out && s->steps_done + (s->x_pending ? 1 : 0) >= v.time_steps) {
        aem::set_error("aem_port_exchange: no step left (the deck's time_steps are done)");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        const size_t nv = (size_t)v.n_spice + v.n_spice_extra;
        if (s->x_pending) {
            // rest of the waiting step, from the gap fields on
            std::memcpy(s->h_spV, v_in, nv * sizeof(float));
            HIPCHK(hipMemcpyAsync(s->d_spV, s->h_spV, nv * sizeof(float), hipMemcpyHostToDevice, s->st));
            s->enqueue_parts(nullptr, s->st, kPartFromExchange);
            HIPCHK(hipGetLastError());
            s->x_pending = false;
            s->steps_done += 1;
            if (s->full_on) s->export_step();
            if (s->steps_done % kWatchSteps == 0 || s->steps_done == v.time_steps) {
                s->enqueue_watch(s->st);
                s->check_watch();
            }
        }
        if (i_out) {
            // next step up to its gap currents
            s->enqueue_parts(nullptr, s->st, kPartToExchange);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(s->h_spI, s->d_spI, (size_t)v.n_spice * sizeof(float), hipMemcpyDeviceToHost, s->st));
            HIPCHK(hipStreamSynchronize(s->st));
            std::memcpy(i_out, s->h_spI, (size_t)v.n_spice * sizeof(float));
            s->x_pending = true;
        } else {
            HIPCHK(hipStreamSynchronize(s->st));
        }
    });
}

aem_status aem_finalize(aem_sim* s) {
    if (!s) {
        aem::set_error("aem_finalize: null handle");
        return AEM_ERR_STATE;
    }
    if (s->finalized) return AEM_OK;
    if (s->x_pending) {
        aem::set_error("aem_finalize: a step waits at the SPICE exchange; finish it with aem_port_exchange");
        return AEM_ERR_STATE;
    }
    if (s->kim) {   // the imaginary part of a kmax pair: its steps ran on this simulation's stream
        const aem_status st = aem::guard([&]() { HIPCHK(hipStreamSynchronize(s->st)); });
        if (st != AEM_OK) return st;
        const aem_status sk = aem_finalize(s->kim);
        if (sk != AEM_OK) return sk;
    }
    return aem::guard([&]() {
        const aem_plan_view& v = s->plan->v;
        const int T = v.time_steps;
        if (v.plane_sum_axis >= 0) {
            float den;
            if (v.plane_sum_axis == 0) den = ((float)v.NY - 2.0f) * ((float)v.NZ - 2.0f);
            else if (v.plane_sum_axis == 1) den = ((float)v.NX - 2.0f) * ((float)v.NZ - 2.0f);
            else den = ((float)v.NY - 2.0f) * ((float)v.NX - 2.0f);
            const RDiv dv = make_rdiv(den, v.kmax != 0);
            k_plane_norm<<<cdiv(T, 256), 256, 0, s->st>>>(T, s->d_refl, s->d_trans, dv, s->d_reflf, s->d_transf);
            if (v.kmax)   // the TM sums (f90:8411-8435)
                k_plane_norm<<<cdiv(T, 256), 256, 0, s->st>>>(T, s->d_refl2, s->d_trans2, dv, s->d_reflf2,
                                                                s->d_transf2);
        }
        if (v.plane_wave_amp > 0)
            k_plane_incident<<<cdiv(T, 256), 256, 0, s->st>>>(T, s->pu, s->d_incident);
        if (v.antenna_amp > 0 && v.excitation_port > 0) {
            const aem_port_view& pt = v.ports[v.excitation_port - 1];
            float* Vi = s->zeros<float>((size_t)2 * v.M + 2);
            float* Ci = s->zeros<float>((size_t)2 * v.M + 2);
            k_clear_line<<<1, 1024, 0, s->st>>>(v.M, T, pt.kI, pt.kV, pt.kS, s->pu, Vi, Ci, s->d_incident);
        }
        if (v.n_angles > 0) {
            const float eta = std::sqrt(aem::kMu0 / aem::kEp0);
            k_ntff_combine<<<v.n_angles, 256, 0, s->st>>>(s->nb, v.len_ff, T, eta, v.dt, s->d_Eth, s->d_Eph, s->d_fftime,
                                                          s->d_scratch);
        }
        if (v.n_angles > 0 && v.plane_wave_amp > 0)
            k_ffpc<<<cdiv(T, 256), 256, 0, s->st>>>(T, s->d_pc, v.ps_c2, v.ps_c1, v.s_pol * v.s_th, s->d_ffpc);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->st));
        s->finalized = true;
    });
}

}  // extern "C"

namespace {
// Device address of an output series (T floats), or null with the error set.
// Series recorded step by step (valid for the steps run so far), as opposed
// to the post-loop products.
bool per_step_series(int32_t kind) {
    return kind == AEM_SERIES_PORT_V || kind == AEM_SERIES_PORT_I || kind == AEM_SERIES_SPICE_V ||
           kind == AEM_SERIES_SPICE_I;
}

// Frequencies (Hz) to the device buffer buf, the DFT of src (n samples) into
// buf + n_freq, the (re, im) pairs back to re_im.  buf holds 3 n_freq doubles.
void dft_into(aem_sim* s, const float* src, int n, const double* freqs_hz, int n_freq, double* buf, double* re_im) {
    HIPCHK(hipMemcpyAsync(buf, freqs_hz, (size_t)n_freq * sizeof(double), hipMemcpyHostToDevice, s->st));
    k_dft<<<n_freq, 256, 0, s->st>>>(src, n, buf, (double)s->plan->v.dt, buf + n_freq);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(re_im, buf + n_freq, (size_t)n_freq * 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
}

// Static scratch of aem_get_dft_noalloc, for callers that cannot allocate
// device memory: every device the library runs on holds one copy
// (kDftScratchFreqs frequencies x 3 doubles = 384 KiB) until the process
// ends.  In host code the array's name is not a device address; each
// device's address comes from hipGetSymbolAddress, looked up once with that
// device current.  All calls on one device share its copy, so concurrent
// calls on one device race.
constexpr int kDftScratchFreqs = 16384;
static __device__ __align__(256) double dft_scratch[kDftScratchFreqs * 3];

double* dft_scratch_ptr(int dev) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* ptr[kMaxDevices] = {};
    if (dev < 0 || dev >= kMaxDevices) throw aem::Error(AEM_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != dev) HIPCHK(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(dft_scratch));
        if (cur != dev) HIPCHK(hipSetDevice(cur));
        HIPCHK(e);
        ptr[dev] = static_cast<double*>(p);
    }
    return ptr[dev];
}

const float* series_src(const aem_sim* s, int32_t kind, int32_t idx) {
    if (kind & AEM_IMAG) {
        if (!s->kim) { aem::set_error("AEM_IMAG selects the imaginary part of a kmax simulation"); return nullptr; }
        return series_src(s->kim, kind & ~AEM_IMAG, idx);
    }
    const aem_plan_view& v = s->plan->v;
    const int T = v.time_steps;
    switch (kind) {
        case AEM_SERIES_REFLECTED_TM:
        case AEM_SERIES_TRANSMITTED_TM:
            if (!v.kmax) { aem::set_error("the TM sums belong to kmax simulations"); return nullptr; }
            if (!s->finalized) { aem::set_error("call aem_finalize first"); return nullptr; }
            return kind == AEM_SERIES_REFLECTED_TM ? s->d_reflf2 : s->d_transf2;
        case AEM_SERIES_PORT_V:
        case AEM_SERIES_PORT_I:
            if (idx < 0 || idx >= v.n_ports) { aem::set_error("port index out of range"); return nullptr; }
            return (kind == AEM_SERIES_PORT_V ? s->d_Vout : s->d_Iout) + (size_t)idx * T;
        case AEM_SERIES_E_THETA:
        case AEM_SERIES_E_PHI:
            if (idx < 0 || idx >= v.n_angles) { aem::set_error("angle index out of range"); return nullptr; }
            if (s->ntff_mode == AEM_NTFF_OFF) { aem::set_error("the time-domain far field is off (AEM_NTFF_OFF)"); return nullptr; }
            if (!s->finalized) { aem::set_error("call aem_finalize first"); return nullptr; }
            return (kind == AEM_SERIES_E_THETA ? s->d_Eth : s->d_Eph) + (size_t)idx * T;
        case AEM_SERIES_SPICE_V:
            if (idx < 0 || idx >= v.n_spice + v.n_spice_extra) { aem::set_error("SPICE port index out of range"); return nullptr; }
            return s->d_spVout + (size_t)idx * T;
        case AEM_SERIES_SPICE_I:
            if (idx < 0 || idx >= v.n_spice) { aem::set_error("SPICE port index out of range"); return nullptr; }
            return s->d_spIout + (size_t)idx * T;
        case AEM_SERIES_INCIDENT:
        case AEM_SERIES_REFLECTED:
        case AEM_SERIES_TRANSMITTED:
        case AEM_SERIES_FF_PC:
            if (!s->finalized) { aem::set_error("call aem_finalize first"); return nullptr; }
            return kind == AEM_SERIES_INCIDENT ? s->d_incident
                   : kind == AEM_SERIES_REFLECTED ? s->d_reflf
                   : kind == AEM_SERIES_TRANSMITTED ? s->d_transf
                                                    : s->d_ffpc;
        default: aem::set_error("unknown series kind"); return nullptr;
    }
}
}  // namespace

extern "C" {

aem_status aem_get_series(aem_sim* s, int32_t kind, int32_t idx, float* out, int64_t n) {
    if (!s || !out) {
        aem::set_error("aem_get_series: null argument");
        return AEM_ERR_STATE;
    }
    const int T = s->plan->v.time_steps;
    if (n < 0 || n > T) {
        aem::set_error("aem_get_series: length exceeds time_steps");
        return AEM_ERR_STATE;
    }
    const float* src = series_src(s, kind, idx);
    if (!src) return AEM_ERR_STATE;
    return aem::guard([&]() {
        HIPCHK(hipMemcpyAsync(out, src, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_get_dft(aem_sim* s, int32_t kind, int32_t idx, const double* freqs_hz, int32_t n_freq,
                       double* re_im) {
    if (!s || !freqs_hz || !re_im || n_freq < 0) {
        aem::set_error("aem_get_dft: bad argument");
        return AEM_ERR_STATE;
    }
    const float* src = series_src(s, kind, idx);
    if (!src) return AEM_ERR_STATE;
    if (n_freq == 0) return AEM_OK;
    const aem_plan_view& v = s->plan->v;
    // port series hold the steps run so far; the post-loop series all T
    const int n = per_step_series(kind & ~AEM_IMAG) ? (int)s->steps_done : v.time_steps;
    return aem::guard([&]() {
        if (n_freq > s->dft_cap) {
            if (s->d_dft) HIPCHK(hipFree(s->d_dft));
            s->d_dft = nullptr;
            s->dft_cap = 0;
            HIPCHK(hipMalloc(&s->d_dft, (size_t)n_freq * 3 * sizeof(double)));
            s->dft_cap = n_freq;
        }
        dft_into(s, src, n, freqs_hz, n_freq, s->d_dft, re_im);
    });
}

aem_status aem_get_dft_noalloc(aem_sim* s, int32_t kind, int32_t idx, const double* freqs_hz, int32_t n_freq,
                               double* re_im) {
    if (!s || !freqs_hz || !re_im || n_freq < 0) {
        aem::set_error("aem_get_dft_noalloc: bad argument");
        return AEM_ERR_STATE;
    }
    if (n_freq > kDftScratchFreqs) {
        aem::set_error("aem_get_dft_noalloc: " + std::to_string(n_freq) + " frequencies exceed the static scratch of " +
                       std::to_string(kDftScratchFreqs) + "; use aem_get_dft");
        return AEM_ERR_STATE;
    }
    const float* src = series_src(s, kind, idx);
    if (!src) return AEM_ERR_STATE;
    if (n_freq == 0) return AEM_OK;
    const int n = per_step_series(kind & ~AEM_IMAG) ? (int)s->steps_done : s->plan->v.time_steps;
    return aem::guard([&]() { dft_into(s, src, n, freqs_hz, n_freq, dft_scratch_ptr(s->dev), re_im); });
}

// The part of a kmax pair that holds a field component OR-ed with AEM_IMAG.
static aem_sim* field_part(aem_sim* s, int32_t& comp) {
    if (!s || !(comp & AEM_IMAG)) return s;
    comp &= ~AEM_IMAG;
    return s->kim;
}

aem_status aem_get_field(aem_sim* s, int32_t comp, float* out, int64_t n) {
    s = field_part(s, comp);
    if (!s || !out || comp < 0 || comp > 5 || n < 0 || n > s->plan->v.ncell) {
        aem::set_error("aem_get_field: bad argument");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        HIPCHK(hipMemcpyAsync(out, s->F[comp], (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_get_video(aem_sim* s, int32_t comp, int64_t first_step, int64_t n_steps, float* out, int64_t n) {
    if (!s || !out || comp < 0 || comp > 5) {
        aem::set_error("aem_get_video: bad argument");
        return AEM_ERR_STATE;
    }
    if (!s->vid_on) {
        aem::set_error("aem_get_video: the deck exports no video slice");
        return AEM_ERR_STATE;
    }
    const int64_t ps = (int64_t)s->vid.v1 * s->vid.v2;
    if (first_step < 1 || n_steps < 0 || first_step + n_steps - 1 > s->vid.T || n != ps * n_steps) {
        aem::set_error("aem_get_video: frames outside [1, time_steps] or n is not frames x slice size");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        HIPCHK(hipMemcpyAsync(out, s->vid.V[comp] + (first_step - 1) * ps, (size_t)n * sizeof(float),
                              hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_set_field(aem_sim* s, int32_t comp, const float* in, int64_t n) {
    s = field_part(s, comp);
    if (!s || !in || comp < 0 || comp > 5 || n < 0 || n > s->plan->v.ncell) {
        aem::set_error("aem_set_field: bad argument");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        HIPCHK(hipMemcpyAsync(s->F[comp], in, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s->st));
        // cells no step writes keep their value in both buffer sets
        if (s->F2[comp]) HIPCHK(hipMemcpyAsync(s->F2[comp], in, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_get_ntff_raw(aem_sim* s, int32_t which, double* out, int64_t n) {
    if (!s || !out || which < 0 || which > 5) {
        aem::set_error("aem_get_ntff_raw: bad argument");
        return AEM_ERR_STATE;
    }
    const aem_plan_view& v = s->plan->v;
    if (v.n_angles <= 0 || s->ntff_mode == AEM_NTFF_OFF || n > (int64_t)v.n_angles * v.len_ff) {
        aem::set_error("aem_get_ntff_raw: no time-domain far field or length too large");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        std::vector<double> all((size_t)v.n_angles * s->nb.LL);
        if (s->nb.exact) {
            const float* src = which < 3 ? s->nb.Uf[which] : s->nb.Wf[which - 3];
            std::vector<float> f32(all.size());
            HIPCHK(hipMemcpyAsync(f32.data(), src, f32.size() * sizeof(float), hipMemcpyDeviceToHost, s->st));
            HIPCHK(hipStreamSynchronize(s->st));
            for (size_t q = 0; q < all.size(); ++q) all[q] = f32[q];
        } else {
            const double* src = which < 3 ? s->nb.U[which] : s->nb.W[which - 3];
            HIPCHK(hipMemcpyAsync(all.data(), src, all.size() * sizeof(double), hipMemcpyDeviceToHost, s->st));
            HIPCHK(hipStreamSynchronize(s->st));
        }
        for (int64_t q = 0; q < n; ++q) {
            int64_t a = q / v.len_ff, b = q % v.len_ff;
            out[q] = all[(size_t)(a * s->nb.LL + b + 1)];
        }
    });
}

aem_status aem_get_ff_time(aem_sim* s, float* out, int64_t n) {
    if (!s || !out || n > s->plan->v.n_angles || !s->finalized || s->ntff_mode == AEM_NTFF_OFF) {
        aem::set_error("aem_get_ff_time: bad argument, not finalized or no time-domain far field");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        HIPCHK(hipMemcpyAsync(out, s->d_fftime, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_time_phases(aem_sim* s, int64_t n_steps, float* ms, int32_t n_ms) {
    if (!s || !ms || n_ms < 5 || n_steps <= 0 || s->steps_done + n_steps > s->plan->v.time_steps || s->finalized) {
        aem::set_error("aem_time_phases: bad argument or not enough steps left");
        return AEM_ERR_STATE;
    }
    if (s->kim) {
        aem::set_error("aem_time_phases: not available for a kmax simulation");
        return AEM_ERR_UNSUPPORTED;
    }
    if (s->ck_K > 0) {
        aem::set_error("aem_time_phases: not available for a simulation with adjoint checkpoints");
        return AEM_ERR_UNSUPPORTED;
    }
    return aem::guard([&]() {
        hipEvent_t ev[6];
        for (int q = 0; q < 6; ++q) HIPCHK(hipEventCreate(&ev[q]));
        double acc[5] = {0, 0, 0, 0, 0};
        for (int64_t n = 0; n < n_steps; ++n) {
            s->enqueue_step(ev);
            HIPCHK(hipEventSynchronize(ev[5]));
            for (int q = 0; q < 5; ++q) {
                float t = 0.f;
                HIPCHK(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
                acc[q] += t;
            }
        }
        for (int q = 0; q < 6; ++q) (void)hipEventDestroy(ev[q]);
        for (int q = 0; q < 5; ++q) ms[q] = (float)(acc[q] / (double)n_steps);
        s->steps_done += n_steps;
    });
}

}  // extern "C"

// ===========================================================================
// Batched stepping of independent simulations (a GA population, an angle or
// parameter sweep).  The reference runs such populations as separate solver
// processes (optimization_scripts/GA.py:11-65, main_ga.py:34-72).  Here one
// graph-captured step advances every member: the bulk updates of all members
// run as one launch per kernel variant, the ports, far-field currents and
// histogram binning and the step counters as one launch each, and any other
// part a member needs (sheets, TF/SF, Drude, separate CPML, PBC, plane sums,
// ordered or atomic far-field binning) is enqueued for that member between
// them in the reference's order.  Members keep their own device state, so
// series, fields and finalisation are read per member afterwards.
// ===========================================================================

struct aem_batch {
    std::vector<aem_sim*> m;
    hipStream_t st = nullptr;
    std::vector<void*> owned;
    struct Group {
        int c = 0, idx_bytes = 1;
        BatchEntry* ents = nullptr;
        int n_ents = 0;
        unsigned blocks = 0;
    };
    std::vector<Group> hg, eg;
    bool off32 = true;             // every member's arrays under 4 GiB
    HArgs* d_hargs = nullptr;
    EArgs* d_eargs = nullptr;
    PortArgs* d_pargs = nullptr;
    int n_port = 0;
    // batched far fields: the face currents of every member binned by the
    // LDS histogram or by contributor lists, then each binning's own launch
    NtffArgs* d_nargs = nullptr;
    NtffArgs* d_hist_args = nullptr;
    NtffListArgs* d_list_args = nullptr;
    int n_ntff = 0, n_hist = 0, n_list = 0, max_angles = 0, max_list_blocks = 0, ntff_mu = 0, ntff_mv = 0;
    std::vector<int> ntff_alone;   // members whose far field bins outside the batch
    int** d_steps = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    RunAhead ra;
    float last_ms = 0.f;

    ~aem_batch() {
        if (st) (void)hipStreamSynchronize(st);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        ra.destroy();
        for (void* p : owned) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }

    template <class T>
    T* upload(const std::vector<T>& h) {
        T* d = dalloc<T>(h.size(), owned);
        if (!h.empty()) HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
        return d;
    }

    void create() {
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&ev0));
        HIPCHK(hipEventCreate(&ev1));
        ra.create();
        const int n = (int)m.size();
        std::vector<HArgs> ha(n);
        std::vector<EArgs> ea(n);
        for (int q = 0; q < n; ++q) {
            ha[q] = m[q]->ha;
            ea[q] = m[q]->ea;
        }
        d_hargs = upload(ha);
        d_eargs = upload(ea);
        for (int q = 0; q < n; ++q) off32 = off32 && m[q]->off32;
        // bulk groups: one launch per kernel variant (H: by slab tests; E: by
        // slab tests and index width), entries in member order
        auto build = [&](bool e_phase, std::vector<Group>& out) {
            for (int c : {0, (int)kCx, (int)(kCx | kCy | kCz)})
                for (int ib : {1, 2}) {
                    if (!e_phase && ib == 2) continue;
                    std::vector<BatchEntry> ents;
                    long long blocks = 0;
                    for (int q = 0; q < n; ++q) {
                        if (e_phase && m[q]->plan->v.idx_bytes != ib) continue;
                        for (const Launch& L : m[q]->bulk) {
                            if (L.c != c) continue;
                            for (int k = 0; k < L.bs.n; ++k) {
                                const Box& b = L.bs.b[k];
                                ents.push_back(BatchEntry{(int)blocks, q, b});
                                blocks += (long long)cdiv(b.ni, kTX) * cdiv(b.nj, kTY) * cdiv(b.nk, b.kc);
                            }
                        }
                    }
                    if (ents.empty()) continue;
                    if (blocks > 0x7fffffffLL) throw aem::Error(AEM_ERR_STATE, "batch too large for one bulk launch");
                    Group gr;
                    gr.c = c;
                    gr.idx_bytes = ib;
                    gr.ents = upload(ents);
                    gr.n_ents = (int)ents.size();
                    gr.blocks = (unsigned)blocks;
                    out.push_back(gr);
                }
        };
        build(false, hg);
        build(true, eg);
        std::vector<PortArgs> pa;
        std::vector<NtffArgs> na, nh;
        std::vector<NtffListArgs> nla;
        std::vector<int*> steps(n);
        for (int q = 0; q < n; ++q) {
            aem_sim* s = m[q];
            const aem_plan_view& v = s->plan->v;
            if (v.n_ports > 0) pa.push_back(s->pa);
            if (v.n_angles > 0 && s->ntff_mode != AEM_NTFF_OFF) {
                const bool hist = s->ntff_mode == AEM_NTFF_HISTOGRAM;
                const bool list = s->ntff_mode == AEM_NTFF_ORDERED && s->nl.n_waves > 0;
                if (hist || list) {
                    na.push_back(s->na);
                    ntff_mu = std::max(ntff_mu, s->ntff_mu);
                    ntff_mv = std::max(ntff_mv, s->ntff_mv);
                }
                if (hist) {
                    nh.push_back(s->na);
                    max_angles = std::max(max_angles, v.n_angles);
                } else if (list) {
                    nla.push_back(NtffListArgs{s->nl, s->nb, s->d_step});
                    max_list_blocks = std::max(max_list_blocks, (int)cdiv(s->nl.n_waves, 4));
                } else {
                    ntff_alone.push_back(q);
                }
            }
            steps[q] = s->d_step;
        }
        n_port = (int)pa.size();
        n_ntff = (int)na.size();
        n_hist = (int)nh.size();
        n_list = (int)nla.size();
        d_pargs = upload(pa);
        d_nargs = upload(na);
        d_hist_args = upload(nh);
        d_list_args = upload(nla);
        d_steps = upload(steps);
        HIPCHK(hipStreamSynchronize(st));
    }

    // batched bulk launches; 32-bit byte offsets when every member's arrays
    // stay under 4 GiB
    template <class OffT>
    void launch_h_t(const Group& gr, hipStream_t s) {
        const dim3 bt(kTX, kTY, 1);
        if (gr.c == 0) k_h_bulk_batch<0, OffT><<<gr.blocks, bt, 0, s>>>(d_hargs, gr.ents, gr.n_ents);
        else if (gr.c == kCx) k_h_bulk_batch<kCx, OffT><<<gr.blocks, bt, 0, s>>>(d_hargs, gr.ents, gr.n_ents);
        else k_h_bulk_batch<kCx | kCy | kCz, OffT><<<gr.blocks, bt, 0, s>>>(d_hargs, gr.ents, gr.n_ents);
    }
    void launch_h(const Group& gr, hipStream_t s) {
        if (off32) launch_h_t<uint32_t>(gr, s);
        else launch_h_t<uint64_t>(gr, s);
    }
    template <class IdxT, class OffT>
    void launch_e_t(const Group& gr, hipStream_t s) {
        const dim3 bt(kTX, kTY, 1);
        if (gr.c == 0) k_e_bulk_batch<IdxT, 0, OffT><<<gr.blocks, bt, 0, s>>>(d_eargs, gr.ents, gr.n_ents);
        else if (gr.c == kCx) k_e_bulk_batch<IdxT, kCx, OffT><<<gr.blocks, bt, 0, s>>>(d_eargs, gr.ents, gr.n_ents);
        else k_e_bulk_batch<IdxT, kCx | kCy | kCz, OffT><<<gr.blocks, bt, 0, s>>>(d_eargs, gr.ents, gr.n_ents);
    }
    template <class IdxT>
    void launch_e(const Group& gr, hipStream_t s) {
        if (off32) launch_e_t<IdxT, uint32_t>(gr, s);
        else launch_e_t<IdxT, uint64_t>(gr, s);
    }

    void enqueue(hipStream_t s) {
        const dim3 bb(32, 8, 1);
        const int n = (int)m.size();
        auto each = [&](unsigned part) {
            for (aem_sim* x : m) x->enqueue_parts(nullptr, s, part);
        };
        for (const Group& gr : hg)
            launch_h(gr, s);
        each(kPartHPre);
        if (n_port) k_ports_h_batch<<<n_port, 256, 0, s>>>(d_pargs);
        each(kPartHPost | kPartRecord);
        for (const Group& gr : eg) {
            if (gr.idx_bytes == 1) launch_e<uint8_t>(gr, s);
            else launch_e<uint16_t>(gr, s);
        }
        each(kPartEPre);
        if (n_port) k_ports_e_batch<<<n_port, 256, 0, s>>>(d_pargs);
        each(kPartEPost | kPartEOld);
        each(kPartOutPre);
        if (n_ntff)
            k_ntff_currents_batch<<<dim3(cdiv(ntff_mu, 32), cdiv(ntff_mv, 8), 6 * n_ntff), bb, 0, s>>>(d_nargs);
        if (n_hist) k_ntff_hist_batch<<<dim3(max_angles, 6, n_hist), 256, 0, s>>>(d_hist_args);
        if (n_list) k_ntff_list_batch<<<dim3(max_list_blocks, n_list), 256, 0, s>>>(d_list_args);
        for (int q : ntff_alone) m[q]->enqueue_parts(nullptr, s, kPartNtff);
        each(kPartObjective | kPartOutPost);
        k_step_inc_batch<<<cdiv(n, 256), 256, 0, s>>>(d_steps, n);
    }
};

extern "C" {

aem_status aem_batch_create(aem_sim* const* sims, int32_t n, aem_batch** out) {
    if (!sims || !out || n <= 0) {
        aem::set_error("aem_batch_create: need at least one simulation");
        return AEM_ERR_STATE;
    }
    *out = nullptr;
    for (int q = 0; q < n; ++q) {
        if (!sims[q] || sims[q]->finalized) {
            aem::set_error("aem_batch_create: null or finalized member");
            return AEM_ERR_STATE;
        }
        for (int r = 0; r < q; ++r)
            if (sims[r] == sims[q]) {
                aem::set_error("aem_batch_create: a simulation appears twice");
                return AEM_ERR_STATE;
            }
        if (sims[q]->full_on) {
            aem::set_error("aem_batch_create: a member exports the full field every step; run it alone");
            return AEM_ERR_UNSUPPORTED;
        }
        if (sims[q]->plan->v.n_spice > 0) {
            aem::set_error("aem_batch_create: a member has SPICE ports; step it alone with aem_port_exchange");
            return AEM_ERR_UNSUPPORTED;
        }
        if (sims[q]->kim) {
            aem::set_error("aem_batch_create: a member is a kmax (complex-field) simulation; run it alone");
            return AEM_ERR_UNSUPPORTED;
        }
        if (sims[q]->ck_K > 0) {
            aem::set_error("aem_batch_create: a member keeps adjoint checkpoints; run it alone");
            return AEM_ERR_UNSUPPORTED;
        }
        if (sims[q]->plan->v.mirror_type > 0 && sims[q]->plan->v.n_angles > 0) {
            aem::set_error("aem_batch_create: a member's far field has ground-plane images; run it alone");
            return AEM_ERR_UNSUPPORTED;
        }
    }
    aem_batch* b = new (std::nothrow) aem_batch();
    if (!b) {
        aem::set_error("host allocation failed");
        return AEM_ERR_STATE;
    }
    b->m.assign(sims, sims + n);
    aem_status stt = aem::guard([&]() {
        for (aem_sim* s : b->m) {
            HIPCHK(hipStreamSynchronize(s->st));
            // the batch steps its members with in-place kernels from here on
            s->batched = true;
            s->drop_graphs();
        }
        b->create();
    });
    if (stt != AEM_OK) {
        delete b;
        return stt;
    }
    *out = b;
    return AEM_OK;
}

aem_status aem_batch_run(aem_batch* b, int64_t n_steps, int32_t use_graph) {
    if (!b || n_steps < 0) {
        aem::set_error("aem_batch_run: bad argument");
        return AEM_ERR_STATE;
    }
    for (aem_sim* s : b->m)
        if (s->finalized || s->steps_done + n_steps > s->plan->v.time_steps) {
            aem::set_error("aem_batch_run: a member would exceed its time_steps or is finalized");
            return AEM_ERR_STATE;
        }
    return aem::guard([&]() {
        if (n_steps == 0) return;
        for (aem_sim* s : b->m) HIPCHK(hipStreamSynchronize(s->st));
        if (use_graph && !b->gexec) {
            HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            b->enqueue(b->st);
            HIPCHK(hipStreamEndCapture(b->st, &b->graph));
            HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        HIPCHK(hipEventRecord(b->ev0, b->st));
        for (int64_t q = 0; q < n_steps; ++q) {
            if (use_graph) HIPCHK(hipGraphLaunch(b->gexec, b->st));
            else b->enqueue(b->st);
            if ((q + 1) % kWatchSteps == 0 && q + 1 < n_steps)
                for (aem_sim* s : b->m) s->enqueue_watch(b->st);
            b->ra.pace(b->st, q);
        }
        for (aem_sim* s : b->m) s->enqueue_watch(b->st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b->ev1, b->st));
        HIPCHK(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->last_ms = ms / (float)n_steps;
        for (aem_sim* s : b->m) s->steps_done += n_steps;
        for (aem_sim* s : b->m) s->check_watch();
    });
}

float aem_batch_last_run_ms(const aem_batch* b) { return b ? b->last_ms : 0.f; }

int32_t aem_batch_size(const aem_batch* b) { return b ? (int32_t)b->m.size() : -1; }

void aem_batch_destroy(aem_batch* b) { delete b; }

}  // extern "C"

// ===========================================================================
// Differentiation (no counterpart in the reference solver; its adjoint
// workflow, optimization_scripts/adjoint_optimization/adjoint_optimization.py
// :76-134, runs a second forward simulation with an adjoint source and
// correlates FFTs of field slices).  Here the objective is the far-field
// amplitude of the frequency-domain near-to-far-field transform, and its
// gradient is the exact discrete adjoint of this solver's update: the field
// update is linear, so the adjoint state runs the transposed H/E/CPML update
// backwards from the last step, and the only forward data it needs are the
// values multiplying the design coefficients (E and the curl term S of every
// design edge, recorded each step).  The design is the permittivity of the
// deck's materials inside a box of cells.
// ===========================================================================

void aem_sim::ensure_adjoint() {
    if (lam[0]) return;
    const aem_plan_view& v = plan->v;
    for (int q = 0; q < 9; ++q) lam[q] = zeros<float>((size_t)v.ncell + v.plane);
    lpE = cpE;
    lpH = cpH;
    for (int f = 0; f < 6; ++f) {
        const int a = f / 2;
        const size_t nt = (size_t)std::max(v.npml[f] - 1, 0);
        size_t trans;
        if (a == 0) trans = (size_t)v.NY * v.NZ;
        else if (a == 1) trans = (size_t)v.PX * v.NZ;
        else trans = (size_t)v.PX * v.NY;
        for (int s = 0; s < 2; ++s) {
            lpE.psi[f][s] = zeros<float>(nt * trans);
            lpH.psi[f][s] = zeros<float>(nt * trans);
        }
    }
    if (v.has_drude)
        for (int c = 0; c < 3; ++c) {
            adb[c] = db[c];
            adb[c].J = zeros<float>((size_t)db[c].nb * v.n_poles);
            adb[c].Eold = zeros<float>((size_t)db[c].nb);
        }
    if (v.n_ports > 0) {
        d_lV = zeros<float>((size_t)v.n_ports * v.M);
        d_lI = zeros<float>((size_t)v.n_ports * v.M);
    }
    d_bstep = zeros<int>(1);
    d_srcD = zeros<float>((size_t)24 * std::max(ntff_mu, 1) * std::max(ntff_mv, 1));
    d_desm = zeros<float>((size_t)3 * std::max(des.nedge, 1LL));
    HIPCHK(hipStreamSynchronize(st));
}

// One step of the adjoint sweep (the adjoint state after step n -> after
// step n-1).  With `objective`, first adds the source of J = |F|^2 at step n
// of the far-field objective; wV, wI (device [np][T], or null) add dJ/dV and
// dJ/dI of the port series at step n; with dga, dgb it accumulates the
// design-edge derivatives.
template <class IdxT>
void aem_sim::enqueue_backward_t(hipStream_t s, bool objective, double Fr, double Fi, const double* wV,
                                 const double* wI, double* dga, double* dgb) {
    const aem_plan_view& v = plan->v;
    const IdxT* idx = static_cast<const IdxT*>(d_idx);
    float *lEx = lam[0], *lEy = lam[1], *lEz = lam[2], *lHx = lam[3], *lHy = lam[4], *lHz = lam[5];
    float *mEx = lam[6], *mEy = lam[7], *mEz = lam[8];
    // the source and the design transpose in a fixed summation order, so a
    // gradient is bit-identical from run to run
    if (objective) {
        k_adj_source_terms<<<dim3(cdiv(ntff_mu, 32), cdiv(ntff_mv, 8), 6), 256, 0, s>>>(g, obj, Fr, Fi, d_srcD,
                                                                                        d_bstep);
        k_adj_source_gather<IdxT><<<dim3(cdiv(ntff_mu + 1, 32), cdiv(ntff_mv + 1, 8), 12), 256, 0, s>>>(
            g, obj, d_srcD, lEx, lEy, lEz, lHx, lHy, lHz, mEx, mEy, mEz, idx, d_tup);
    }
    const AdjPorts ap{d_ports, v.n_ports, v.M, v.time_steps, d_lV, d_lI, wV, wI};
    // lambda_E of the step's end -> lambda_E of the bulk update's output
    if (v.n_ports > 0) k_adj_ports_e<<<1, 64, 0, s>>>(g, ap, lEx, lEy, lEz, mEx, mEy, mEz, d_bstep);
    if (v.has_drude) {
        float* lE[3] = {lEx, lEy, lEz};
        float* mE[3] = {mEx, mEy, mEz};
        for (int c = 0; c < 3; ++c)
            if (adb[c].nb > 0)
                k_adj_drude<IdxT><<<cdiv(adb[c].nb, 256), 256, 0, s>>>(g, adb[c], c, v.n_poles, lE[c], mE[c], idx, d_tup);
    }
    if (dga && rec_on) k_adj_grad<<<cdiv(rec.nedge, 256), 256, 0, s>>>(g, rec, lEx, lEy, lEz, dga, dgb, d_bstep);
    if (des_on) {
        k_adj_design_m<<<cdiv(des.nedge, 256), 256, 0, s>>>(g, des, lEx, lEy, lEz, d_desm);
        const long long nh = (long long)(des.n[0] + 1) * (des.n[1] + 1) * (des.n[2] + 1);
        k_adj_design_gather<<<cdiv(nh, 256), 256, 0, s>>>(g, des, d_desm, lHx, lHy, lHz, d_den[0], d_den[1],
                                                          d_den[2]);
    }
    const dim3 bt(64, 4, 1);
    k_adj_h<<<dim3(cdiv(v.NX - 1, 64), cdiv(v.NY - 1, 4), v.NZ - 1), bt, 0, s>>>(
        g, lHx, lHy, lHz, mEx, mEy, mEz, d_den[0], d_den[1], d_den[2], cpE, lpE, cpH, lpH, v.DB, v.dx, v.dy, v.dz);
    if (v.n_ports > 0) k_adj_ports_h<<<1, 64, 0, s>>>(g, ap, lHx, lHy, lHz, v.dx, v.dy, v.dz, d_bstep);
    k_adj_e<IdxT><<<dim3(cdiv(v.NX, 64), cdiv(v.NY, 4), v.NZ), bt, 0, s>>>(
        g, lEx, lEy, lEz, mEx, mEy, mEz, lHx, lHy, lHz, idx, d_tup, d_den[3], d_den[4], d_den[5], cpE, lpE, cpH, lpH,
        v.DB, v.dx, v.dy, v.dz);
    k_step_dec<<<1, 1, 0, s>>>(d_bstep);
}

namespace {

// The Drude box of component c as cells [lo, hi] per axis; false when empty.
bool drude_box(const aem_plan_view& v, int c, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = v.dbox[c][2 * a];
        hi[a] = v.dbox[c][2 * a + 1];
        if (lo[a] > hi[a]) return false;
    }
    return true;
}

// The transposed Drude step covers boxes clear of the CPML slabs and the
// outer boundary by a cell: every component in them is updated by the bulk
// loop and by no psi term.
bool drude_interior(const aem_plan_view& v) {
    const int N[3] = {v.NX, v.NY, v.NZ};
    for (int c = 0; c < 3; ++c) {
        int lo[3], hi[3];
        if (!drude_box(v, c, lo, hi)) continue;
        for (int a = 0; a < 3; ++a)
            if (lo[a] < std::max(v.npml[2 * a], 0) + 2 || hi[a] > N[a] - std::max(v.npml[2 * a + 1], 0) - 2)
                return false;
    }
    return true;
}

// The cells a lumped port touches: its gap [loc, loc+size-1] and the H of
// the loop around it, within [loc-1, loc+size-1] per axis.
void port_reach(const aem_port_view& p, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = p.loc[a] - 1;
        hi[a] = p.loc[a] + p.size[a] - 1;
    }
}

// The transposed port step covers ports clear of the CPML slabs and the
// outer boundary by a cell (no psi term and no fixed component in reach)
// and of the Drude boxes (the gap's overwrite follows the Drude update).
const char* ports_unsupported(const aem_plan_view& v) {
    const int N[3] = {v.NX, v.NY, v.NZ};
    for (int q = 0; q < v.n_ports; ++q) {
        int lo[3], hi[3];
        port_reach(v.ports[q], lo, hi);
        for (int a = 0; a < 3; ++a)
            if (lo[a] < std::max(v.npml[2 * a], 0) + 2 || hi[a] > N[a] - std::max(v.npml[2 * a + 1], 0) - 2)
                return "lumped ports within a cell of the CPML or the outer boundary";
        if (v.has_drude)
            for (int c = 0; c < 3; ++c) {
                int dl[3], dh[3];
                if (!drude_box(v, c, dl, dh)) continue;
                bool meet = true;
                for (int a = 0; a < 3; ++a) meet = meet && lo[a] <= dh[a] && hi[a] >= dl[a];
                if (meet) return "lumped ports inside the Drude media's bounding box";
            }
    }
    return nullptr;
}

// Features whose update is not part of the transposed step.
const char* adjoint_unsupported(const aem_plan_view& v) {
    if (v.mirror_type > 0) return "an infinite ground plane";
    if (v.n_spice > 0) return "SPICE ports";
    if (const char* why = ports_unsupported(v)) return why;
    if (v.n_sheet_planes[0] + v.n_sheet_planes[1] + v.n_sheet_planes[2] > 0) return "thin sheets";
    if (v.has_drude && !drude_interior(v)) return "Drude media within a cell of the CPML or the outer boundary";
    if (v.pbc[0] || v.pbc[1] || v.pbc[2]) return "periodic boundaries";
    if (v.kmax) return "the kmax build";
    return nullptr;
}

// dJ/d eps_r of every cell of the design box [lo, hi] (x fastest, the three
// axes together) from the per-edge derivatives dJ/dga, dJ/dgb of the record
// box [lo, hi+1].  Each edge's eps and sigma are the four-cell averages of the
// setup (fdtd_solver.f90:2516-2527) and its Ga/Gb the formulas of
// f90:2837-2853, so d/d eps_cell = (dGa/d eps dJ/dGa + dGb/d eps dJ/dGb) / 4
// for each of the edge's cells in the box.  eps3 is the box's per-cell
// permittivity ([3][cells]) or null for the deck's materials.
void cell_gradient(const aem::Plan& P, const AdjRec& r, const int lo[3], const int hi[3], const float* eps3,
                   const std::vector<double>& dga, const std::vector<double>& dgb, std::vector<double>& out) {
    const aem_plan_view& v = P.v;
    const double dt = v.dt, e0 = aem::kEp0;
    const int nb[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    const size_t ncb = (size_t)nb[0] * nb[1] * nb[2];
    auto cidx = [&](int i, int j, int k) {
        return (size_t)(i - 1) + (size_t)(v.NX - 1) * ((size_t)(j - 1) + (size_t)(v.NY - 1) * (size_t)(k - 1));
    };
    auto bidx = [&](int i, int j, int k) -> long long {
        if (i < lo[0] || i > hi[0] || j < lo[1] || j > hi[1] || k < lo[2] || k > hi[2]) return -1;
        return (long long)(i - lo[0]) + (long long)nb[0] * ((long long)(j - lo[1]) + (long long)nb[1] * (k - lo[2]));
    };
    out.assign(ncb, 0.0);
    for (long long e = 0; e < r.nedge; ++e) {
        const int i = r.lo[0] + (int)(e % r.n[0]);
        const long long rr = e / r.n[0];
        const int j = r.lo[1] + (int)(rr % r.n[1]), k = r.lo[2] + (int)(rr / r.n[1]);
        for (int c = 0; c < 3; ++c) {
            int cell[4][3];
            if (c == 0) { int t[4][3] = {{i, j, k}, {i, j - 1, k}, {i, j, k - 1}, {i, j - 1, k - 1}}; std::memcpy(cell, t, sizeof t); }
            else if (c == 1) { int t[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j, k - 1}, {i - 1, j, k - 1}}; std::memcpy(cell, t, sizeof t); }
            else { int t[4][3] = {{i, j, k}, {i - 1, j, k}, {i, j - 1, k}, {i - 1, j - 1, k}}; std::memcpy(cell, t, sizeof t); }
            long long b[4];
            float ev[4], sv[4];
            for (int q = 0; q < 4; ++q) {
                const int m = P.cell_mat[cidx(cell[q][0], cell[q][1], cell[q][2])];
                b[q] = bidx(cell[q][0], cell[q][1], cell[q][2]);
                sv[q] = m ? P.deck.mats[m - 1].sig[c] : 0.0f;
                ev[q] = (eps3 && b[q] >= 0) ? eps3[(size_t)c * ncb + (size_t)b[q]] : (m ? P.deck.mats[m - 1].eps[c] : 1.0f);
            }
            const double eps = (((ev[0] + ev[1]) + ev[2]) + ev[3]) / 4.0f;
            const double sig = (((sv[0] + sv[1]) + sv[2]) + sv[3]) / 4.0f;
            const double al = sig * dt / (2.0 * e0 * eps), A = dt / (e0 * eps);
            const double dgb_de = -A / (eps * (1.0 + al) * (1.0 + al));
            const double dga_de = 2.0 * al / (eps * (1.0 + al) * (1.0 + al));
            const double dj = dga[c * r.nedge + e] * dga_de + dgb[c * r.nedge + e] * dgb_de;
            for (int q = 0; q < 4; ++q)
                if (b[q] >= 0) out[(size_t)b[q]] += dj / 4.0;
        }
    }
}

bool fresh(const aem_sim* s) { return s->steps_done == 0 && !s->gexec; }

// A batch captured its members' kernel arguments (coefficient table, record
// and objective state) when it was created; a member cannot gain a design
// box, a record or an objective afterwards.
bool refuse_batched(const aem_sim* s, const char* who) {
    if (!s->batched) return false;
    aem::set_error(std::string(who) + ": a batch steps this simulation; set it up before aem_batch_create");
    return true;
}

}  // namespace

extern "C" {

aem_status aem_objective_set(aem_sim* s, float theta_deg, float phi_deg, double freq_hz, int32_t pol) {
    if (!s || freq_hz <= 0 || (pol != 0 && pol != 1)) {
        aem::set_error("aem_objective_set: bad argument");
        return AEM_ERR_STATE;
    }
    if (refuse_batched(s, "aem_objective_set")) return AEM_ERR_STATE;
    const aem_plan_view& v = s->plan->v;
    if (!fresh(s) || v.n_angles <= 0) {
        aem::set_error("aem_objective_set: needs a deck with a far-field box and a sim that has not stepped");
        return AEM_ERR_STATE;
    }
    if (v.mirror_type > 0) {
        aem::set_error("aem_objective_set: the frequency-domain far field has no image currents of a ground plane");
        return AEM_ERR_UNSUPPORTED;
    }
    return aem::guard([&]() {
        Objective& o = s->obj;
        for (int a = 0; a < 3; ++a) { o.lo[a] = v.ff_lo[a]; o.hi[a] = v.ff_hi[a]; o.pbc[a] = v.pbc[a]; }
        const double d[3] = {v.dx, v.dy, v.dz};
        const double omega = 2.0 * M_PI * freq_hz, k = omega / 299792458.0, dt = v.dt;
        const double th = theta_deg * M_PI / 180.0, ph = phi_deg * M_PI / 180.0;
        const double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
        const double u[3] = {st * cp, st * sp, ct};
        const double eta = std::sqrt((double)aem::kMu0 / (double)aem::kEp0);
        const double th_v[3] = {ct * cp, ct * sp, -st}, ph_v[3] = {-sp, cp, 0.0};
        for (int a = 0; a < 3; ++a) {
            if (pol == 0) { o.gM[a] = ph_v[a]; o.gJ[a] = eta * th_v[a]; }
            else { o.gM[a] = th_v[a]; o.gJ[a] = -eta * ph_v[a]; }
        }
        const double pre_im = (pol == 0 ? -1.0 : 1.0) * k / (4.0 * M_PI);   // pre = j pre_im
        // face-cell table: pre dS e^{jk r^.r'}, r' = ((i - ic) dx, (j - jc) dy, (k - kc) dz)
        const int mu = s->ntff_mu, mv = s->ntff_mv;
        std::vector<double2> gt((size_t)6 * mv * mu, make_double2(0.0, 0.0));
        for (int face = 0; face < 6; ++face) {
            const int a = face / 2, b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
            if (v.pbc[a]) continue;
            const double dS = d[b] * d[c];
            const int pos = (face & 1) ? o.hi[a] : o.lo[a];
            for (int vv = 0; vv < mv; ++vv)
                for (int uu = 0; uu < mu; ++uu) {
                    const int uc = o.lo[b] + uu, vc = o.lo[c] + vv;
                    if (uc > o.hi[b] || vc > o.hi[c]) continue;
                    int ijk[3];
                    ijk[a] = pos; ijk[b] = uc; ijk[c] = vc;
                    const double phs = k * (u[0] * (ijk[0] - v.ic) * d[0] + u[1] * (ijk[1] - v.jc) * d[1] +
                                            u[2] * (ijk[2] - v.kc) * d[2]);
                    gt[((size_t)face * mv + vv) * mu + uu] =
                        make_double2(-pre_im * dS * std::sin(phs), pre_im * dS * std::cos(phs));
                }
        }
        const int T = v.time_steps;
        std::vector<double2> te(T), thh(T);
        for (int n = 1; n <= T; ++n) {
            te[n - 1] = make_double2(std::cos(omega * n * dt), -std::sin(omega * n * dt));
            thh[n - 1] = make_double2(std::cos(omega * (n - 0.5) * dt), -std::sin(omega * (n - 0.5) * dt));
        }
        o.mu = mu;
        o.mv = mv;
        o.g = s->up(gt.data(), gt.size());
        o.tE = s->up(te.data(), te.size());
        o.tH = s->up(thh.data(), thh.size());
        o.part = s->zeros<double>((size_t)2 * cdiv(mu, 32) * cdiv(mv, 8) * 6);
        o.done = s->zeros<unsigned>(1);
        if (!o.F) o.F = s->zeros<double>(2);
        s->obj_on = true;
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_ffdft_set(aem_sim* s, int32_t n_ang, const float* theta_deg, const float* phi_deg, double f0_hz,
                         double df_hz, int32_t n_freq) {
    if (!s || n_ang <= 0 || n_freq <= 0 || !theta_deg || !phi_deg || f0_hz < 0 || df_hz < 0) {
        aem::set_error("aem_ffdft_set: bad argument");
        return AEM_ERR_STATE;
    }
    if (refuse_batched(s, "aem_ffdft_set")) return AEM_ERR_STATE;
    const aem_plan_view& v = s->plan->v;
    if (!fresh(s) || v.n_angles <= 0 || s->ff_on) {
        aem::set_error("aem_ffdft_set: needs a deck with a far-field box and a sim that has not stepped (once)");
        return AEM_ERR_STATE;
    }
    if (v.mirror_type > 0) {
        aem::set_error("aem_ffdft_set: the frequency-domain far field has no image currents of a ground plane");
        return AEM_ERR_UNSUPPORTED;
    }
    return aem::guard([&]() {
        FfDft& p = s->ff;
        for (int a = 0; a < 3; ++a) { p.lo[a] = v.ff_lo[a]; p.hi[a] = v.ff_hi[a]; p.pbc[a] = v.pbc[a]; }
        p.mu = s->ntff_mu; p.mv = s->ntff_mv;
        p.n_ang = n_ang; p.n_freq = n_freq;
        p.dt = v.dt;
        p.w0 = 2.0 * M_PI * f0_hz; p.dw = 2.0 * M_PI * df_hz;
        p.eta = std::sqrt((double)aem::kMu0 / (double)aem::kEp0);
        const double d[3] = {v.dx, v.dy, v.dz};
        const int ctr[3] = {v.ic, v.jc, v.kc};
        const int npair = n_ang * n_freq;
        std::vector<double> dir(9 * (size_t)n_ang);
        for (int a = 0; a < n_ang; ++a) {
            const double th = theta_deg[a] * M_PI / 180.0, ph = phi_deg[a] * M_PI / 180.0;
            const double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
            const double d9[9] = {st * cp, st * sp, ct, ct * cp, ct * sp, -st, -sp, cp, 0.0};
            std::copy(d9, d9 + 9, dir.begin() + 9 * a);
        }
        auto kq = [&](int pr) { return (p.w0 + (pr % n_freq) * p.dw) / 299792458.0; };
        // per face: the spatial factor's step along u (axis b) and v (axis c)
        std::vector<double> stp((size_t)6 * npair * 4);
        for (int face = 0; face < 6; ++face) {
            const int A = face / 2, b = A == 0 ? 1 : 0, c = A == 2 ? 1 : 2;
            for (int pr = 0; pr < npair; ++pr) {
                const double* rh = &dir[9 * (pr / n_freq)];
                const double k = kq(pr), xu = k * rh[b] * d[b], xv = k * rh[c] * d[c];
                double* o = &stp[((size_t)face * npair + pr) * 4];
                o[0] = std::cos(xu); o[1] = std::sin(xu); o[2] = std::cos(xv); o[3] = std::sin(xv);
            }
        }
        // per tile: dS e^{jk r^.r'} at its first cell
        const int gx = cdiv(p.mu, 32), gy = cdiv(p.mv, 8);
        const size_t ntile = (size_t)gx * gy * 6;
        std::vector<double> z0(ntile * npair * 2);
        for (int face = 0; face < 6; ++face) {
            const int A = face / 2, b = A == 0 ? 1 : 0, c = A == 2 ? 1 : 2;
            const double dS = d[b] * d[c];
            for (int by = 0; by < gy; ++by)
                for (int bx = 0; bx < gx; ++bx) {
                    int ijk[3];
                    ijk[A] = (face & 1) ? p.hi[A] : p.lo[A];
                    ijk[b] = p.lo[b] + bx * 32;
                    ijk[c] = p.lo[c] + by * 8;
                    const size_t tile = (size_t)bx + (size_t)gx * ((size_t)by + (size_t)gy * face);
                    for (int pr = 0; pr < npair; ++pr) {
                        const double* rh = &dir[9 * (pr / n_freq)];
                        double x = 0.0;
                        for (int q = 0; q < 3; ++q) x += rh[q] * (ijk[q] - ctr[q]) * d[q];
                        const double phs = kq(pr) * x;
                        z0[(tile * npair + pr) * 2] = dS * std::cos(phs);
                        z0[(tile * npair + pr) * 2 + 1] = dS * std::sin(phs);
                    }
                }
        }
        p.dir = s->up(dir.data(), dir.size());
        p.step = s->up(stp.data(), stp.size());
        p.z0 = s->up(z0.data(), z0.size());
        p.part = s->zeros<double>(ntile * npair * 8);
        p.tot = s->zeros<double>((size_t)npair * 4);
        p.comp = s->zeros<double>((size_t)npair * 4);
        s->ff_on = true;
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_ffdft_get(aem_sim* s, double* out, int64_t n) {
    if (!s || !out || !s->ff_on || n != (int64_t)s->ff.n_ang * s->ff.n_freq * 4) {
        aem::set_error("aem_ffdft_get: no running-DFT far field, or n is not n_ang * n_freq * 4");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        std::vector<double> S((size_t)n);
        HIPCHK(hipMemcpyAsync(S.data(), s->ff.tot, S.size() * sizeof(double), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        const FfDft& p = s->ff;
        for (int a = 0; a < p.n_ang; ++a)
            for (int q = 0; q < p.n_freq; ++q) {
                const size_t o = ((size_t)a * p.n_freq + q) * 4;
                const double kq = (p.w0 + q * p.dw) / 299792458.0 / (4.0 * M_PI);
                // r E_theta = -j k/(4 pi) S_theta,  r E_phi = j k/(4 pi) S_phi
                out[o] = kq * S[o + 1];
                out[o + 1] = -kq * S[o];
                out[o + 2] = -kq * S[o + 3];
                out[o + 3] = kq * S[o + 2];
            }
    });
}

aem_status aem_objective_get(aem_sim* s, double* re_im) {
    if (!s || !re_im || !s->obj_on) {
        aem::set_error("aem_objective_get: no objective set");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        HIPCHK(hipMemcpyAsync(re_im, s->obj.F, 2 * sizeof(double), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_adjoint_record(aem_sim* s, const int32_t* lo, const int32_t* hi) {
    if (!s || !lo || !hi) {
        aem::set_error("aem_adjoint_record: null argument");
        return AEM_ERR_STATE;
    }
    if (refuse_batched(s, "aem_adjoint_record")) return AEM_ERR_STATE;
    const aem_plan_view& v = s->plan->v;
    if (!fresh(s)) {
        aem::set_error("aem_adjoint_record: call before the first step");
        return AEM_ERR_STATE;
    }
    if (const char* why = adjoint_unsupported(v)) {
        aem::set_error(std::string("aem_adjoint_record: the adjoint step does not include ") + why);
        return AEM_ERR_UNSUPPORTED;
    }
    const int N[3] = {v.NX, v.NY, v.NZ};
    for (int a = 0; a < 3; ++a) {
        // the design edges [lo, hi+1] must lie in the interior, outside every CPML slab
        const int lo_ok = std::max(2, v.npml[2 * a] + 1), hi_ok = N[a] - 1 - std::max(v.npml[2 * a + 1], 0);
        if (lo[a] > hi[a] || lo[a] < lo_ok || hi[a] + 1 > hi_ok) {
            aem::set_error("aem_adjoint_record: design box must lie inside the grid, outside the CPML slabs");
            return AEM_ERR_STATE;
        }
    }
    if (v.has_drude)
        for (int c = 0; c < 3; ++c) {
            // the record and design updates assume the plain E update on their edges [lo, hi+1]
            int dl[3], dh[3];
            if (!drude_box(v, c, dl, dh)) continue;
            bool meet = true;
            for (int a = 0; a < 3; ++a) meet = meet && lo[a] <= dh[a] && hi[a] + 1 >= dl[a];
            if (meet) {
                aem::set_error("aem_adjoint_record: design box must not meet the Drude media's bounding box");
                return AEM_ERR_UNSUPPORTED;
            }
        }
    for (int q = 0; q < v.n_ports; ++q) {
        // the gap's overwrite and the loop integral stay outside the design edges
        int pl[3], ph[3];
        port_reach(v.ports[q], pl, ph);
        bool meet = true;
        for (int a = 0; a < 3; ++a) meet = meet && lo[a] <= ph[a] && hi[a] + 1 >= pl[a];
        if (meet) {
            aem::set_error("aem_adjoint_record: design box must not meet a lumped port's gap or loop");
            return AEM_ERR_UNSUPPORTED;
        }
    }
    return aem::guard([&]() {
        AdjRec& r = s->rec;
        for (int a = 0; a < 3; ++a) {
            s->rec_lo[a] = lo[a];
            s->rec_hi[a] = hi[a];
            r.lo[a] = lo[a];
            r.n[a] = hi[a] - lo[a] + 2;
        }
        r.nedge = (long long)r.n[0] * r.n[1] * r.n[2];
        r.slots = s->ck_K > 0 ? std::min(s->ck_K, v.time_steps) : v.time_steps;
        r.hist = s->zeros<float>((size_t)r.slots * 6 * r.nedge);
        s->rec_on = true;
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

void aem_sim::upload_design() {
    const long long ne = des.nedge;
    std::vector<float> coef(6 * (size_t)ne);
    aem::design_coefficients(*plan, rec_lo, rec_hi, des_eps.data(), coef.data());
    HIPCHK(hipMemcpyAsync(d_des_coef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
}

aem_status aem_design_box(aem_sim* s, const int32_t* lo, const int32_t* hi) {
    if (s && refuse_batched(s, "aem_design_box")) return AEM_ERR_STATE;
    if (s && (s->rec_on || s->des_on)) {
        aem::set_error("aem_design_box: the simulation already records a box");
        return AEM_ERR_STATE;
    }
    const aem_status st = aem_adjoint_record(s, lo, hi);
    if (st != AEM_OK) return st;
    const aem_plan_view& v = s->plan->v;
    return aem::guard([&]() {
        const int max_tuples = v.idx_bytes == 1 ? 256 : 65536;
        if (v.n_tuple + 1 > max_tuples)
            throw aem::Error(AEM_ERR_UNSUPPORTED, "aem_design_box: the coefficient table has no room for the design marker");
        DesignDev& d = s->des;
        for (int a = 0; a < 3; ++a) {
            d.lo[a] = s->rec.lo[a];
            d.n[a] = s->rec.n[a];
        }
        d.nedge = s->rec.nedge;
        // initial design: the deck's materials
        const int nb[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
        const size_t ncb = (size_t)nb[0] * nb[1] * nb[2];
        s->des_eps.assign(3 * ncb, 1.0f);
        size_t b = 0;
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i, ++b) {
                    const int m = s->plan->cell_mat[(size_t)(i - 1) +
                                                    (size_t)(v.NX - 1) * ((size_t)(j - 1) + (size_t)(v.NY - 1) * (k - 1))];
                    for (int c = 0; c < 3; ++c) s->des_eps[c * ncb + b] = m ? s->plan->deck.mats[m - 1].eps[c] : 1.0f;
                }
        s->d_des_coef = s->zeros<float>(6 * (size_t)d.nedge);
        d.coef = s->d_des_coef;
        // the coefficient table gains the neutral tuple (Ga = 1, Gb = 0), which
        // every edge of the box now selects
        const int nt = v.n_tuple;
        float* tup = s->zeros<float>((size_t)(nt + 1) * 6);
        HIPCHK(hipMemcpyAsync(tup, s->d_tup, (size_t)nt * 6 * sizeof(float), hipMemcpyDeviceToDevice, s->st));
        const float neutral[6] = {1.0f, 0.0f, 1.0f, 0.0f, 1.0f, 0.0f};
        HIPCHK(hipMemcpyAsync(tup + (size_t)nt * 6, neutral, sizeof neutral, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        s->d_tup = tup;
        s->ea.tup = tup;
        s->n_tup = nt + 1;
        if (v.idx_bytes == 1) {
            k_set_idx<uint8_t><<<cdiv(d.nedge, 256), 256, 0, s->st>>>(s->g, d, static_cast<uint8_t*>(s->d_idx), (uint8_t)nt);
            if (s->d_idx4) {   // the single pass reads the packed copy
                const long long nw = (long long)v.plane * ((v.NZ + 3) / 4);
                k_pack_idx4<<<cdiv(nw, 256), 256, 0, s->st>>>(static_cast<const uint8_t*>(s->d_idx), s->d_idx4,
                                                                (long long)v.plane, v.NZ);
            }
        } else
            k_set_idx<uint16_t><<<cdiv(d.nedge, 256), 256, 0, s->st>>>(s->g, d, static_cast<uint16_t*>(s->d_idx),
                                                                       (uint16_t)nt);
        HIPCHK(hipGetLastError());
        s->upload_design();
        s->des_on = true;
    });
}

aem_status aem_set_design_eps(aem_sim* s, const float* eps, int64_t n) {
    if (!s || !eps || !s->des_on) {
        aem::set_error("aem_set_design_eps: needs a design box (aem_design_box)");
        return AEM_ERR_STATE;
    }
    if (!fresh(s)) {
        aem::set_error("aem_set_design_eps: call before the first step");
        return AEM_ERR_STATE;
    }
    const int64_t ncb = (int64_t)s->des_eps.size() / 3;
    if (n != ncb) {
        aem::set_error("aem_set_design_eps: n must equal the design box's cell count " + std::to_string(ncb));
        return AEM_ERR_STATE;
    }
    for (int64_t q = 0; q < n; ++q)
        if (!(eps[q] >= 1.0f) || !std::isfinite(eps[q])) {
            aem::set_error("aem_set_design_eps: permittivities must be finite and >= 1");
            return AEM_ERR_INPUT;
        }
    return aem::guard([&]() {
        for (int c = 0; c < 3; ++c) std::memcpy(s->des_eps.data() + c * ncb, eps, (size_t)n * sizeof(float));
        s->upload_design();
    });
}

namespace {

// The full backward sweep of J = |F|^2: dJ/dGa and dJ/dGb of every record
// edge ([3][edges] each).  Clears the adjoint state, captures one backward
// step as a graph and replays it time_steps times.
// The adjoint source weights a step's terms by 2 Re(conj(F) w), the
// derivative of J = |F|^2.  A general real J(F) with dJ/dF = dJ/dRe F +
// i dJ/dIm F (dj, two doubles) has dJ = Re(conj(dj) dF): the same source
// with F replaced by dj / 2.  With port series sources (wV, wI: host
// [np][T] or null) the far-field source is off unless dj is given.
void backward_sweep(aem_sim* s, std::vector<double>& ha, std::vector<double>& hb, const double* dj = nullptr,
                    const double* wV = nullptr, const double* wI = nullptr) {
    const aem_plan_view& v = s->plan->v;
    s->ensure_adjoint();
    const bool ports_src = wV || wI;
    const bool ff_src = s->obj_on && (dj || !ports_src);
    double F[2] = {0.0, 0.0};
    if (dj) {
        F[0] = 0.5 * dj[0];
        F[1] = 0.5 * dj[1];
    } else if (ff_src) {
        HIPCHK(hipMemcpy(F, s->obj.F, sizeof F, hipMemcpyDeviceToHost));
    }
    const long long ne = s->rec.nedge;
    std::vector<void*> tmp;
    const size_t nw = (size_t)v.n_ports * v.time_steps;
    double* d_wV = wV ? dalloc<double>(nw, tmp) : nullptr;
    double* d_wI = wI ? dalloc<double>(nw, tmp) : nullptr;
    if (wV) HIPCHK(hipMemcpyAsync(d_wV, wV, nw * sizeof(double), hipMemcpyHostToDevice, s->st));
    if (wI) HIPCHK(hipMemcpyAsync(d_wI, wI, nw * sizeof(double), hipMemcpyHostToDevice, s->st));
    double* dga = dalloc<double>(3 * ne, tmp);
    double* dgb = dalloc<double>(3 * ne, tmp);
    HIPCHK(hipMemsetAsync(dga, 0, 3 * ne * sizeof(double), s->st));
    HIPCHK(hipMemsetAsync(dgb, 0, 3 * ne * sizeof(double), s->st));
    for (int q = 0; q < 9; ++q) HIPCHK(hipMemsetAsync(s->lam[q], 0, ((size_t)v.ncell + v.plane) * 4, s->st));
    for (int f = 0; f < 6; ++f) {
        const int a = f / 2;
        const size_t nt = (size_t)std::max(v.npml[f] - 1, 0);
        const size_t trans = a == 0 ? (size_t)v.NY * v.NZ : (a == 1 ? (size_t)v.PX * v.NZ : (size_t)v.PX * v.NY);
        for (int q = 0; q < 2; ++q) {
            if (nt * trans) HIPCHK(hipMemsetAsync(s->lpE.psi[f][q], 0, nt * trans * 4, s->st));
            if (nt * trans) HIPCHK(hipMemsetAsync(s->lpH.psi[f][q], 0, nt * trans * 4, s->st));
        }
    }
    if (v.has_drude)
        for (int c = 0; c < 3; ++c)
            if (s->adb[c].nb > 0) {
                HIPCHK(hipMemsetAsync(s->adb[c].J, 0, (size_t)s->adb[c].nb * v.n_poles * 4, s->st));
                HIPCHK(hipMemsetAsync(s->adb[c].Eold, 0, (size_t)s->adb[c].nb * 4, s->st));
            }
    if (v.n_ports > 0) {
        HIPCHK(hipMemsetAsync(s->d_lV, 0, (size_t)v.n_ports * v.M * 4, s->st));
        HIPCHK(hipMemsetAsync(s->d_lI, 0, (size_t)v.n_ports * v.M * 4, s->st));
    }
    const int T = v.time_steps;
    HIPCHK(hipMemcpyAsync(s->d_bstep, &T, sizeof(int), hipMemcpyHostToDevice, s->st));
    hipGraph_t gr = nullptr;
    hipGraphExec_t gx = nullptr;
    HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    if (v.idx_bytes == 1) s->enqueue_backward_t<uint8_t>(s->st, ff_src, F[0], F[1], d_wV, d_wI, dga, dgb);
    else s->enqueue_backward_t<uint16_t>(s->st, ff_src, F[0], F[1], d_wV, d_wI, dga, dgb);
    HIPCHK(hipStreamEndCapture(s->st, &gr));
    HIPCHK(hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0));
    HIPCHK(hipEventRecord(s->ev0, s->st));
    if (s->ck_K > 0) {
        // intervals from the last: the record of the last one is still in
        // place right after the forward run; every other one is recomputed
        // from its snapshot (field update and record only, no outputs) and
        // then transposed.  The state at the end of the run comes back at the
        // end.
        const int K = s->ck_K, S = (T + K - 1) / K;
        constexpr unsigned kRecompute = kPartAll & ~(kPartOutPre | kPartNtff | kPartOutPost | kPartObjective);
        s->snapshot(S);
        int launched = 0;
        for (int seg = S - 1; seg >= 0; --seg) {
            const int n0 = seg * K, n1 = std::min(T, n0 + K);
            if (seg < S - 1 || !s->ck_last_ok) {
                s->restore(seg);
                for (int n = n0; n < n1; ++n) {
                    s->enqueue_parts(nullptr, s->st, kRecompute);
                    s->ra.pace(s->st, launched++);
                }
            }
            for (int n = n1; n > n0; --n) {
                HIPCHK(hipGraphLaunch(gx, s->st));
                s->ra.pace(s->st, launched++);
            }
        }
        s->restore(S);
        s->ck_last_ok = false;   // the record now holds the first interval
    } else {
        for (int n = 0; n < T; ++n) {
            HIPCHK(hipGraphLaunch(gx, s->st));
            s->ra.pace(s->st, n);
        }
    }
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms / (float)T;
    (void)hipGraphExecDestroy(gx);
    (void)hipGraphDestroy(gr);
    ha.resize(3 * ne);
    hb.resize(3 * ne);
    HIPCHK(hipMemcpy(ha.data(), dga, ha.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hb.data(), dgb, hb.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (void* p : tmp) (void)hipFree(p);
}

bool gradient_ready(const aem_sim* s, const char* who) {
    if (!s->obj_on || !s->rec_on || s->steps_done != s->plan->v.time_steps) {
        aem::set_error(std::string(who) + ": needs an objective, a design record and a completed run");
        return false;
    }
    return true;
}

}  // namespace

aem_status aem_adjoint_eps_gradient(aem_sim* s, int32_t n_mat, double* grad) {
    if (!s || !grad || n_mat < 0) {
        aem::set_error("aem_adjoint_eps_gradient: bad argument");
        return AEM_ERR_STATE;
    }
    if (!gradient_ready(s, "aem_adjoint_eps_gradient")) return AEM_ERR_STATE;
    if (s->des_on) {
        aem::set_error("aem_adjoint_eps_gradient: the design box sets its own permittivity; use aem_adjoint_design_gradient");
        return AEM_ERR_STATE;
    }
    if (n_mat > (int)s->plan->deck.mats.size()) {
        aem::set_error("aem_adjoint_eps_gradient: more materials requested than the deck defines");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        std::vector<double> ha, hb, cells;
        backward_sweep(s, ha, hb);
        cell_gradient(*s->plan, s->rec, s->rec_lo, s->rec_hi, nullptr, ha, hb, cells);
        const aem_plan_view& v = s->plan->v;
        for (int m = 0; m < n_mat; ++m) grad[m] = 0.0;
        size_t b = 0;
        for (int k = s->rec_lo[2]; k <= s->rec_hi[2]; ++k)
            for (int j = s->rec_lo[1]; j <= s->rec_hi[1]; ++j)
                for (int i = s->rec_lo[0]; i <= s->rec_hi[0]; ++i, ++b) {
                    const int m = s->plan->cell_mat[(size_t)(i - 1) +
                                                    (size_t)(v.NX - 1) * ((size_t)(j - 1) + (size_t)(v.NY - 1) * (k - 1))];
                    if (m >= 1 && m <= n_mat) grad[m - 1] += cells[b];
                }
    });
}

aem_status aem_adjoint_design_gradient(aem_sim* s, double* grad, int64_t n) {
    if (!s || !grad) {
        aem::set_error("aem_adjoint_design_gradient: bad argument");
        return AEM_ERR_STATE;
    }
    if (!s->des_on) {
        aem::set_error("aem_adjoint_design_gradient: no design box (aem_design_box)");
        return AEM_ERR_STATE;
    }
    if (!gradient_ready(s, "aem_adjoint_design_gradient")) return AEM_ERR_STATE;
    const int64_t ncb = (int64_t)s->des_eps.size() / 3;
    if (n != ncb) {
        aem::set_error("aem_adjoint_design_gradient: n must equal the design box's cell count " + std::to_string(ncb));
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        std::vector<double> ha, hb, cells;
        backward_sweep(s, ha, hb);
        cell_gradient(*s->plan, s->rec, s->rec_lo, s->rec_hi, s->des_eps.data(), ha, hb, cells);
        std::memcpy(grad, cells.data(), cells.size() * sizeof(double));
    });
}

aem_status aem_adjoint(aem_sim* s, const double* dJ_dF, double* grad, int64_t n) {
    if (!s || !dJ_dF || !grad || n < 0) {
        aem::set_error("aem_adjoint: bad argument");
        return AEM_ERR_STATE;
    }
    if (!gradient_ready(s, "aem_adjoint")) return AEM_ERR_STATE;
    const aem_plan_view& v = s->plan->v;
    const int64_t want = s->des_on ? (int64_t)s->des_eps.size() / 3 : (int64_t)s->plan->deck.mats.size();
    if (s->des_on ? n != want : n > want) {
        aem::set_error("aem_adjoint: n must be the design box's cell count " + std::string(s->des_on ? "" : "or at most the deck's material count ") +
                       "(" + std::to_string(want) + ")");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        std::vector<double> ha, hb, cells;
        backward_sweep(s, ha, hb, dJ_dF);
        cell_gradient(*s->plan, s->rec, s->rec_lo, s->rec_hi, s->des_on ? s->des_eps.data() : nullptr, ha, hb, cells);
        if (s->des_on) {
            std::memcpy(grad, cells.data(), cells.size() * sizeof(double));
            return;
        }
        for (int64_t m = 0; m < n; ++m) grad[m] = 0.0;
        size_t b = 0;
        for (int k = s->rec_lo[2]; k <= s->rec_hi[2]; ++k)
            for (int j = s->rec_lo[1]; j <= s->rec_hi[1]; ++j)
                for (int i = s->rec_lo[0]; i <= s->rec_hi[0]; ++i, ++b) {
                    const int m = s->plan->cell_mat[(size_t)(i - 1) +
                                                    (size_t)(v.NX - 1) * ((size_t)(j - 1) + (size_t)(v.NY - 1) * (k - 1))];
                    if (m >= 1 && m <= n) grad[m - 1] += cells[b];
                }
    });
}

aem_status aem_adjoint_checkpoint(aem_sim* s, int32_t every) {
    if (!s || every < 1) {
        aem::set_error("aem_adjoint_checkpoint: bad argument");
        return AEM_ERR_STATE;
    }
    if (refuse_batched(s, "aem_adjoint_checkpoint")) return AEM_ERR_STATE;
    if (!fresh(s) || s->rec_on || s->ck_K > 0) {
        aem::set_error("aem_adjoint_checkpoint: call once, before aem_adjoint_record / aem_design_box and the first step");
        return AEM_ERR_STATE;
    }
    if (const char* why = adjoint_unsupported(s->plan->v)) {
        aem::set_error(std::string("aem_adjoint_checkpoint: the adjoint step does not include ") + why);
        return AEM_ERR_UNSUPPORTED;
    }
    return aem::guard([&]() {
        const int T = s->plan->v.time_steps;
        const int K = std::min((int)every, T);
        size_t bytes = 0;
        for (const auto& reg : s->state_regions()) bytes += (reg.second + 255) & ~(size_t)255;
        const int S = (T + K - 1) / K;
        s->ck_bytes = bytes;
        s->ck_snap.assign((size_t)S + 1, nullptr);
        for (int q = 0; q <= S; ++q) s->ck_snap[q] = dalloc<char>(bytes, s->owned);
        s->ck_K = K;
    });
}

aem_status aem_adjoint_series(aem_sim* s, const double* dJ_dV, const double* dJ_dI, double* grad, int64_t n) {
    if (!s || (!dJ_dV && !dJ_dI) || !grad || n < 0) {
        aem::set_error("aem_adjoint_series: bad argument");
        return AEM_ERR_STATE;
    }
    const aem_plan_view& v = s->plan->v;
    if (v.n_ports == 0) {
        aem::set_error("aem_adjoint_series: the deck has no lumped port");
        return AEM_ERR_STATE;
    }
    if (!s->rec_on || s->steps_done != v.time_steps) {
        aem::set_error("aem_adjoint_series: needs a design record and a completed run");
        return AEM_ERR_STATE;
    }
    const int64_t want = s->des_on ? (int64_t)s->des_eps.size() / 3 : (int64_t)s->plan->deck.mats.size();
    if (s->des_on ? n != want : n > want) {
        aem::set_error("aem_adjoint_series: n must be the design box's cell count " +
                       std::string(s->des_on ? "" : "or at most the deck's material count ") + "(" +
                       std::to_string(want) + ")");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        std::vector<double> ha, hb, cells;
        backward_sweep(s, ha, hb, nullptr, dJ_dV, dJ_dI);
        cell_gradient(*s->plan, s->rec, s->rec_lo, s->rec_hi, s->des_on ? s->des_eps.data() : nullptr, ha, hb, cells);
        if (s->des_on) {
            std::memcpy(grad, cells.data(), cells.size() * sizeof(double));
            return;
        }
        for (int64_t m = 0; m < n; ++m) grad[m] = 0.0;
        size_t b = 0;
        for (int k = s->rec_lo[2]; k <= s->rec_hi[2]; ++k)
            for (int j = s->rec_lo[1]; j <= s->rec_hi[1]; ++j)
                for (int i = s->rec_lo[0]; i <= s->rec_hi[0]; ++i, ++b) {
                    const int m = s->plan->cell_mat[(size_t)(i - 1) +
                                                    (size_t)(v.NX - 1) * ((size_t)(j - 1) + (size_t)(v.NY - 1) * (k - 1))];
                    if (m >= 1 && m <= n) grad[m - 1] += cells[b];
                }
    });
}

aem_status aem_adjoint_step(aem_sim* s, int64_t n_steps) {
    if (!s || n_steps < 0) {
        aem::set_error("aem_adjoint_step: bad argument");
        return AEM_ERR_STATE;
    }
    if (const char* why = adjoint_unsupported(s->plan->v)) {
        aem::set_error(std::string("aem_adjoint_step: the adjoint step does not include ") + why);
        return AEM_ERR_UNSUPPORTED;
    }
    return aem::guard([&]() {
        s->ensure_adjoint();
        for (int64_t n = 0; n < n_steps; ++n) {
            if (s->plan->v.idx_bytes == 1)
                s->enqueue_backward_t<uint8_t>(s->st, false, 0, 0, nullptr, nullptr, nullptr, nullptr);
            else s->enqueue_backward_t<uint16_t>(s->st, false, 0, 0, nullptr, nullptr, nullptr, nullptr);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_adjoint_get_field(aem_sim* s, int32_t comp, float* out, int64_t n) {
    if (!s || !out || comp < 0 || comp > 5 || n < 0 || n > s->plan->v.ncell) {
        aem::set_error("aem_adjoint_get_field: bad argument");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        s->ensure_adjoint();
        HIPCHK(hipMemcpyAsync(out, s->lam[comp], (size_t)n * 4, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

aem_status aem_adjoint_set_field(aem_sim* s, int32_t comp, const float* in, int64_t n) {
    if (!s || !in || comp < 0 || comp > 5 || n < 0 || n > s->plan->v.ncell) {
        aem::set_error("aem_adjoint_set_field: bad argument");
        return AEM_ERR_STATE;
    }
    return aem::guard([&]() {
        s->ensure_adjoint();
        const aem_plan_view& v = s->plan->v;
        // the adjoint CPML state carries gb lambda_E of the E psi cells: swap
        // the old field's share for the new one's
        auto adjust = [&](float sign) {
            if (comp >= 3) return;
            const dim3 gr(cdiv(v.NX, 64), cdiv(v.NY, 4), v.NZ), bt(64, 4, 1);
            if (v.idx_bytes == 1)
                k_adj_nu_adjust<uint8_t><<<gr, bt, 0, s->st>>>(s->g, s->lam[comp], comp, sign,
                                                               static_cast<const uint8_t*>(s->d_idx), s->d_tup, s->cpE,
                                                               s->lpE);
            else
                k_adj_nu_adjust<uint16_t><<<gr, bt, 0, s->st>>>(s->g, s->lam[comp], comp, sign,
                                                                static_cast<const uint16_t*>(s->d_idx), s->d_tup,
                                                                s->cpE, s->lpE);
            HIPCHK(hipGetLastError());
        };
        adjust(-1.0f);
        HIPCHK(hipMemcpyAsync(s->lam[comp], in, (size_t)n * 4, hipMemcpyHostToDevice, s->st));
        adjust(1.0f);
        if (comp < 3) {   // rebuild mu of this component
            const dim3 gr(cdiv(v.NX, 64), cdiv(v.NY, 4), v.NZ), bt(64, 4, 1);
            if (v.idx_bytes == 1)
                k_adj_mu<uint8_t><<<gr, bt, 0, s->st>>>(s->g, s->lam[comp], s->lam[6 + comp], comp,
                                                        static_cast<const uint8_t*>(s->d_idx), s->d_tup);
            else
                k_adj_mu<uint16_t><<<gr, bt, 0, s->st>>>(s->g, s->lam[comp], s->lam[6 + comp], comp,
                                                         static_cast<const uint16_t*>(s->d_idx), s->d_tup);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(s->st));
    });
}

}  // extern "C"
