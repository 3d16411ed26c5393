// api.hip -- C ABI of libfcfd_hip.so (declared in include/fcfd.h).
#include <dlfcn.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fcfd.h"
#include "fcfd_common.hpp"
#include "fluid.hpp"
#include "chem.hpp"
#include "surrogate_impl.hpp"
#include "tran.hpp"
#include "aux.hpp"

namespace fcfd {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

}  // namespace fcfd

using namespace fcfd;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
            return FCFD_E_HIP;                                                                 \
        }                                                                                      \
    } while (0)

namespace {

constexpr int NTIMERS = 10;  // 0 flux, 1 stage end, 2 CFL, 3 chemistry SDIRK, 4 source Qc refresh, 5 surrogate, 6 explicit screen, 7 transport flux, 8 explicit retry, 9 auxiliary sources

int lvir_of(const fcfd_opts &o) {                                  // SRC/params.f90:161-173
    if (o.flag_acu == 0) return 1;
    if (o.flag_acu == 1) return 2;
    if (o.flag_weno == 1) return 2;
    if (o.flag_weno == 4) return 4;
    return 3;
}

int check_opts(const fcfd_opts &o, int kk) {
    if (o.is_spherical != 0 && o.is_spherical != 1) { set_error("is_spherical must be 0 (Cartesian) or 1 (spherical)"); return FCFD_E_ARG; }
    if (o.is_spherical == 1 && o.select_time == 2) { set_error("is_spherical = 1 is supported with select_time 0 or 1 (the reference's RK4 has no spherical update)"); return FCFD_E_UNSUPPORTED; }
    if ((o.flag_lbound != 0 && o.flag_lbound != 1) || (o.flag_rbound != 0 && o.flag_rbound != 1)) { set_error("flag_lbound / flag_rbound must be 0 (wall) or 1 (inflow / outflow)"); return FCFD_E_ARG; }
    if ((o.flag_lbound == 1 || o.flag_rbound == 1) && o.select_time == 2) { set_error("inflow / outflow boundaries are supported with select_time 0 or 1"); return FCFD_E_UNSUPPORTED; }
    if (o.select_time < 0 || o.select_time > 2) { set_error("select_time must be 0 (Euler), 1 (TVD-RK3) or 2 (RK4)"); return FCFD_E_UNSUPPORTED; }
    if (o.select_riemann < 0 || o.select_riemann > 2) { set_error("select_riemann must be 0, 1 or 2"); return FCFD_E_ARG; }
    if (o.flag_acu < 0 || o.flag_acu > 2) { set_error("flag_acu must be 0, 1 or 2"); return FCFD_E_ARG; }
    if (o.flag_acu == 1 && (o.flag_muscl < 1 || o.flag_muscl > 4)) { set_error("flag_muscl must be 1..4"); return FCFD_E_ARG; }
    if (o.flag_acu == 2 && (o.flag_weno < 1 || o.flag_weno > 4)) { set_error("flag_weno must be 1..4"); return FCFD_E_ARG; }
    if (o.do_reaction && !(o.rtol > 0.0 && o.atol > 0.0)) { set_error("do_reaction needs atol > 0 and rtol > 0"); return FCFD_E_ARG; }
    bool ok = false;
#define FCFD_CASE(N) ok = ok || (kk == N);
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    if (!ok) { set_error("species count " + std::to_string(kk) + " has no kernel instantiation"); return FCFD_E_UNSUPPORTED; }
    return FCFD_OK;
}

long round_up(long n, long m) { return (n + m - 1) / m * m; }

}  // namespace

struct fcfd_ctx {
    const fcfd_mech *mech = nullptr;
    fcfd_opts o{};
    int nblk = 0, ncell = 0, lv = 0, S = 0, kk = 0, nequ = 0;
    long P = 0, PF = 0;
    double dx = 0.0;
    int chem_wg = 0;                     // integrator grid (0: chem_workgroups of the mechanism)
    hipStream_t stream = nullptr;
    DevThermo *thermo = nullptr;
    double *Q[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    double *qf = nullptr, *F = nullptr, *god = nullptr, *cells = nullptr;
    double *d_dt = nullptr, *d_time = nullptr;
    unsigned long long *d_cvel = nullptr;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    double graph_cfl = -1.0;
    int graph_tune_gen = -1;             // tuning table generation the graphs were captured with
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pending;   // (timer id, event index of the start event)
    size_t ev_used = 0;
    double ms[NTIMERS] = {};
    long long cnt[NTIMERS] = {};
    // chemistry
    unsigned int *d_counter = nullptr;
    double *d_hrr = nullptr, *d_hrr_dot = nullptr;
    ChemCellStats *d_stats = nullptr;
    fcfd_surrogate *sur = nullptr;
    int *d_accepted = nullptr;
    int *d_list = nullptr;               // cells the explicit screen sends to the implicit integrator
    int *d_order = nullptr;              // that list (or the surrogate's fallback list) reordered, slowest first
    double *d_hmem = nullptr;            // per interior cell: the integrator's last step-size proposal
    int *d_list2 = nullptr;              // the stiff list after the lane-per-cell retry of a long list
    unsigned *d_list2_n = nullptr;
    int *d_retry = nullptr;              // 1 when the cells of d_list2 already failed that retry
    double *d_f0 = nullptr;              // [kk + 1][cells]: f at the start state of the screen's rejects
    unsigned *d_list_n = nullptr;
    double *acc = nullptr;               // RK4 stage accumulator
    int graph_next[3] = {0, 0, 0};       // state buffer after the captured step from buffer i
    // step composition of fcfd_advance (fcfd_set_step)
    int strang = 0;                      // 0: Lie (the reference driver's order), 1: Strang
    int step_hrr = 0;                    // 1: the source calls of fcfd_advance compute HRR / HRR_dot
    double *d_dth = nullptr;             // dt / 2 on the device (Strang)
    // transport
    DevTran *d_tran = nullptr;
    double *props = nullptr;
    long PP = 0;
    // open boundaries and auxiliary sources
    FlowBC *d_bc = nullptr;              // inflow / outflow state (fcfd_set_flow_boundary)
    bool bc_set = false;
    AuxParams aux{};                     // host copy of d_aux
    AuxParams *d_aux = nullptr;
    double *d_xm = nullptr;              // [nblk * ncell] cell centres, m
    double *d_qrad = nullptr, *d_qw = nullptr;
    // stretched mesh (fcfd_create_smr): per-block cell size, coarse block counts at each end, the
    // left block of every pair of unequal blocks
    double *d_dxb = nullptr;             // nullptr: dx in every block
    std::vector<double> h_dxb;
    int smr_left = 0, smr_right = 0;
    int idx0 = 1;                        // the reference's index of the first cell (nx_all_0)
    int *d_smr_ifb = nullptr;
    int n_smr_if = 0;
};

namespace fcfd {

// Device memory released when it goes out of scope (host-array entry points, error paths).
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

// Device copies of a mechanism's tables on the current device (created once).
int ensure_dev_tables(const fcfd_mech *cm) {
    fcfd_mech *m = const_cast<fcfd_mech *>(cm);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (m->d_thermo && m->device == dev) return FCFD_OK;
    if (m->d_thermo) { set_error("mechanism tables already live on another device"); return FCFD_E_ARG; }
    // the tables go into the mechanism only when every step succeeded; a failure frees what was made
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    DevBuf th, kin, counter;
    HIP_TRY(hipMalloc(&th.p, sizeof(DevThermo)));
    HIP_TRY(hipMemcpy(th.p, &m->host.thermo, sizeof(DevThermo), hipMemcpyHostToDevice));
    auto k = std::make_unique<DevKin>();
    build_dev_kin(m->host, *k);
    HIP_TRY(hipMalloc(&kin.p, sizeof(DevKin)));
    HIP_TRY(hipMemcpy(kin.p, k.get(), sizeof(DevKin), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&counter.p, sizeof(unsigned int)));
    m->ncu = ncu;
    m->d_thermo = static_cast<DevThermo *>(th.p);
    m->d_kin = kin.p;
    m->d_counter = static_cast<unsigned int *>(counter.p);
    th.p = kin.p = counter.p = nullptr;
    m->device = dev;
    return FCFD_OK;
}

// ---- launch shapes chosen by the tuner (fortrancfd_amd/tune.py) ------------------------------
// A text table of "<arch> <species count> <key> <value>" lines: by default data/tuned_launch.txt
// next to the library's directory (FCFD_TUNE_FILE overrides), read once per process; entries for
// other GPU architectures are ignored.  Keys: flux_variant, stage_variant (fluid.hpp),
// chem_wg_per_cu (integrator workgroups per CU), chem_four_wave_factor (four-wave mode for lists of
// at most factor x workgroups cells), chem_lu (the four-wave integrator's LU: 0 step by step through
// LDS, 1 blocked with matrix-core updates, 2 with the matrix in registers).
struct TuneTable {
    std::mutex mu;
    bool loaded = false;
    std::map<std::string, int> v;   // "<nsp>:<key>"
    int gen = 0;                    // bumped by every change, so contexts re-capture their graphs
    std::string source;
};

TuneTable &tune_table() {
    static TuneTable t;
    return t;
}

std::string device_arch() {
    int dev = 0;
    hipDeviceProp_t p{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return "none";
    std::string a(p.gcnArchName);
    return a.substr(0, a.find(':'));
}

std::string default_tune_path() {
    if (const char *e = std::getenv("FCFD_TUNE_FILE")) return e;
    Dl_info info{};
    if (!dladdr(reinterpret_cast<void *>(&chem_coop_max), &info) || !info.dli_fname) return "";
    std::string lib(info.dli_fname);
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
    return dir + "/../data/tuned_launch.txt";
}

// Replaces the table with the entries of path for this device's architecture; false if unreadable.
bool load_tuning_locked(TuneTable &t, const std::string &path) {
    t.loaded = true;
    t.v.clear();
    t.source = path;
    ++t.gen;
    std::ifstream f(path);
    if (!f) return false;
    const std::string arch = device_arch();
    std::string line;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream in(line);
        std::string a, key;
        int nsp = 0, val = 0;
        if (!(in >> a >> nsp >> key >> val)) continue;
        if (a == arch) t.v[std::to_string(nsp) + ":" + key] = val;
    }
    return true;
}

int tuned(int nsp, const char *key, int dflt) {
    TuneTable &t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    if (!t.loaded) load_tuning_locked(t, default_tune_path());
    auto it = t.v.find(std::to_string(nsp) + ":" + key);
    return it == t.v.end() ? dflt : it->second;
}

int tune_gen() {
    TuneTable &t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    return t.gen;
}

// Value of a tuned entry that the table does not hold.
static int launch_default(const char *key) {
    if (std::strcmp(key, "chem_wg_per_cu") == 0) return 5;
    if (std::strcmp(key, "chem_four_wave_factor") == 0) return 8;
    if (std::strcmp(key, "chem_lu") == 0) return 2;
    return 0;
}

int chem_coop_max(int n_workgroups, int nsp) {
    return tuned(nsp, "chem_four_wave_factor", launch_default("chem_four_wave_factor")) * n_workgroups;
}

int chem_lu_variant(int nsp) { return tuned(nsp, "chem_lu", launch_default("chem_lu")); }

// Persistent grid size: one-wave workgroups per CU (default five, as many as the LDS of a CU holds)
// times the CUs.
int chem_workgroups(const fcfd_mech *m) {
    return tuned(m->host.kk, "chem_wg_per_cu", launch_default("chem_wg_per_cu")) * (m->ncu > 0 ? m->ncu : 256);
}

}  // namespace fcfd

namespace {

hipEvent_t next_event(fcfd_ctx *c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// Runs `launch` bracketed by timing events when timing is on.
template <class Fn>
hipError_t timed(fcfd_ctx *c, int id, Fn launch) {
    if (!c->timing) return launch();
    const size_t i0 = c->ev_used;
    hipEvent_t e0 = next_event(c), e1 = next_event(c);
    if (!e0 || !e1) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(e0, c->stream);
    if (e != hipSuccess) return e;
    e = launch();
    if (e != hipSuccess) return e;
    e = hipEventRecord(e1, c->stream);
    c->ev_pending.push_back({id, int(i0)});
    return e;
}

hipError_t collect_timing(fcfd_ctx *c) {
    if (c->ev_pending.empty()) { c->ev_used = 0; return hipSuccess; }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    for (auto &p : c->ev_pending) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, c->ev_pool[p.second], c->ev_pool[p.second + 1]);
        if (e != hipSuccess) return e;
        c->ms[p.first] += ms;
        c->cnt[p.first] += 1;
    }
    c->ev_pending.clear();
    c->ev_used = 0;
    return hipSuccess;
}

FluxArgs flux_args(fcfd_ctx *c, bool god) {
    FluxArgs a{};
    a.thermo = c->thermo; a.qf = c->qf; a.F = c->F; a.godunov = god ? c->god : nullptr;
    a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P; a.PF = c->PF;
    a.muscl = c->o.flag_muscl; a.riemann = c->o.select_riemann;
    a.spherical = c->o.is_spherical;
    a.variant = tuned(c->kk, "flux_variant", 0);
    return a;
}

StageArgs stage_args(fcfd_ctx *c, int mode, const double *Q0, const double *Qin, double *Qout,
                     double dt, const double *dt_dev, double h1, double h2) {
    StageArgs a{};
    a.thermo = c->thermo; a.Q0 = Q0; a.Qin = Qin; a.F = c->F; a.Qout = Qout; a.qf = c->qf;
    a.dt_dev = dt_dev; a.dt = dt; a.dx = c->dx; a.h1 = h1; a.h2 = h2; a.mode = mode; a.ghosts_only = 0;
    a.spherical = c->o.is_spherical; a.xm = c->d_xm; a.god = c->god; a.dxb = c->d_dxb;
    a.variant = tuned(c->kk, "stage_variant", 0);
    a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P; a.PF = c->PF;
    return a;
}

// The halos between blocks of unequal cell size, over the plain copies a stage or refresh has
// just written.
hipError_t smr_halo(fcfd_ctx *c, double *Q) {
    if (c->n_smr_if == 0) return hipSuccess;
    SmrArgs a{};
    a.thermo = c->thermo; a.Q = Q; a.qf = c->qf; a.ifb = c->d_smr_ifb; a.dxb = c->d_dxb;
    a.nif = c->n_smr_if; a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.nequ = c->nequ;
    a.P = c->P;
    return launch_smr_halo(a, c->kk, c->stream);
}

// The inflow / outflow ghost cells over the wall ghosts a stage or refresh has just written.
hipError_t open_bc(fcfd_ctx *c, double *Q) {
    if (c->o.flag_lbound != 1 && c->o.flag_rbound != 1) return hipSuccess;
    OpenBCArgs a{};
    a.thermo = c->thermo; a.bc = c->d_bc; a.Q = Q; a.qf = c->qf;
    a.lbound = c->o.flag_lbound; a.rbound = c->o.flag_rbound;
    a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P;
    return launch_open_bc(a, c->kk, c->stream);
}

// One solve_fluid on the stream; dt from the host (dt_dev == nullptr) or from device memory.
hipError_t enqueue_fluid(fcfd_ctx *c, double dt, const double *dt_dev) {
    const int nsp = c->kk;
    FluxArgs fa = flux_args(c, c->o.is_spherical == 1);   // the spherical update reads the Roe pressures
    fa.accumulate = c->o.do_tran == 1 ? 1 : 0;
    // Fplus of one stage (SRC/timeintegration.f90:138-146): transport from the conserved state
    // Qin, then advection from the primitives on top
    auto flux = [&](const double *Qin) {
        if (c->o.do_tran == 1) {
            TranArgs ta{};
            ta.thermo = c->thermo; ta.tran = c->d_tran; ta.Qc = Qin; ta.props = c->props; ta.F = c->F;
            ta.nblk = c->nblk; ta.ncell = c->ncell; ta.lv = c->lv; ta.S = c->S; ta.P = c->P; ta.PF = c->PF;
            ta.PP = c->PP; ta.dx = c->dx; ta.dxb = c->d_dxb;
            hipError_t e2 = timed(c, 7, [&]() { return launch_flux_tran(ta, nsp, c->stream); });
            if (e2 != hipSuccess) return e2;
        }
        if (c->o.do_hydro != 1) return hipSuccess;
        if (!c->d_dxb) return timed(c, 0, [&]() { return launch_flux(fa, nsp, c->o.flag_acu, c->o.flag_weno, c->stream); });
        // refined mesh: the blocks at dx_comp with the configured reconstruction, the coarser
        // blocks at first order (SRC/flux.f90:240)
        return timed(c, 0, [&]() {
            const int nl = c->smr_left, nr = c->smr_right, nf = c->nblk - nl - nr;
            FluxArgs f = fa;
            f.b0 = nl; f.nblk = nf;
            hipError_t e2 = nf > 0 ? launch_flux(f, nsp, c->o.flag_acu, c->o.flag_weno, c->stream) : hipSuccess;
            if (e2 == hipSuccess && nl > 0) {
                f.b0 = 0; f.nblk = nl;
                e2 = launch_flux(f, nsp, 0, c->o.flag_weno, c->stream);
            }
            if (e2 == hipSuccess && nr > 0) {
                f.b0 = c->nblk - nr; f.nblk = nr;
                e2 = launch_flux(f, nsp, 0, c->o.flag_weno, c->stream);
            }
            return e2;
        });
    };
    auto stage = [&](const StageArgs &s) {
        return timed(c, 1, [&]() {
            hipError_t e2 = launch_stage_end(s, nsp, c->stream);
            if (e2 == hipSuccess) e2 = smr_halo(c, s.Qout);
            return e2 != hipSuccess ? e2 : open_bc(c, s.Qout);
        });
    };
    hipError_t e;
    const int a = c->cur, b = (c->cur + 1) % 3, d = (c->cur + 2) % 3;
    if (c->o.select_time == 0) {                                   // explicit_euler
        if ((e = flux(c->Q[a])) != hipSuccess) return e;
        if ((e = stage(stage_args(c, 0, nullptr, c->Q[a], c->Q[b], dt, dt_dev, 1.0, 1.0))) != hipSuccess) return e;
        c->cur = b;
        return hipSuccess;
    }
    if (c->o.select_time == 2) {                                   // RK4 (:222-356)
        auto rk4 = [&](int st, const double *Qin, double *Qout) {
            StageArgs s = stage_args(c, st, c->Q[a], Qin, Qout, dt, dt_dev, 1.0, 1.0);
            s.acc = c->acc;
            return timed(c, 1, [&]() { return launch_rk4_stage(s, nsp, c->stream); });
        };
        if ((e = flux(c->Q[a])) != hipSuccess) return e;
        if ((e = rk4(0, c->Q[a], c->Q[b])) != hipSuccess) return e;
        if ((e = flux(c->Q[b])) != hipSuccess) return e;
        if ((e = rk4(1, c->Q[b], c->Q[d])) != hipSuccess) return e;
        if ((e = flux(c->Q[d])) != hipSuccess) return e;
        if ((e = rk4(2, c->Q[d], c->Q[b])) != hipSuccess) return e;
        if ((e = flux(c->Q[b])) != hipSuccess) return e;
        if ((e = rk4(3, c->Q[b], c->Q[d])) != hipSuccess) return e;
        c->cur = d;
        return hipSuccess;
    }
    // TVDRK3 (SRC/timeintegration.f90:132-217)
    if ((e = flux(c->Q[a])) != hipSuccess) return e;
    if ((e = stage(stage_args(c, 0, nullptr, c->Q[a], c->Q[b], dt, dt_dev, 1.0, 1.0))) != hipSuccess) return e;
    if ((e = flux(c->Q[b])) != hipSuccess) return e;
    if ((e = stage(stage_args(c, 1, c->Q[a], c->Q[b], c->Q[d], dt, dt_dev, 3.0 / 4.0, 1.0 / 4.0))) != hipSuccess) return e;
    if ((e = flux(c->Q[d])) != hipSuccess) return e;
    if ((e = stage(stage_args(c, 1, c->Q[a], c->Q[d], c->Q[b], dt, dt_dev, 1.0 / 3.0, 2.0 / 3.0))) != hipSuccess) return e;
    c->cur = b;
    return hipSuccess;
}

hipError_t enqueue_dt(fcfd_ctx *c, double cfl, double *time_dev) {
    DtArgs a{};
    a.thermo = c->thermo; a.qf = c->qf; a.cvel_bits = c->d_cvel; a.dt_out = c->d_dt; a.time = time_dev;
    a.cfl = cfl; a.dx = c->dx; a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P;
    return timed(c, 2, [&]() { return launch_dt(a, c->kk, c->stream); });
}

hipError_t enqueue_refresh(fcfd_ctx *c) {
    StageArgs s = stage_args(c, 2, nullptr, c->Q[c->cur], c->Q[c->cur], 0.0, nullptr, 1.0, 1.0);
    return timed(c, 1, [&]() {
        hipError_t e = launch_stage_end(s, c->kk, c->stream);
        if (e == hipSuccess) e = smr_halo(c, c->Q[c->cur]);
        return e != hipSuccess ? e : open_bc(c, c->Q[c->cur]);
    });
}

bool aux_on(const fcfd_ctx *c) { return c->aux.nrad > 0 || c->aux.wall_heat || c->aux.ig_type != 0; }

// The auxiliary sources of solve_source (SRC/update.f90:127-147) on the interior cells.
hipError_t enqueue_aux(fcfd_ctx *c, double dt, const double *dt_dev, double time, const double *time_dev) {
    AuxArgs a{};
    a.thermo = c->thermo; a.tran = c->d_tran; a.par = c->d_aux; a.Q = c->Q[c->cur]; a.qf = c->qf;
    a.xm = c->d_xm; a.qrad = c->d_qrad; a.qw = c->d_qw;
    a.dt_dev = dt_dev; a.time_dev = time_dev; a.dt = dt; a.time = time;
    a.rad = c->aux.nrad > 0; a.wall = c->aux.wall_heat; a.ig = c->aux.ig_type != 0;
    a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P;
    return timed(c, 9, [&]() { return launch_aux(a, c->kk, c->stream); });
}

// reaction_macks (SRC/source.f90:17-84): chemistry on every interior cell, then Qc = qfToQc(qf).
// The heat-release diagnostics cost one more rate evaluation per cell and are only computed when
// the caller asks for them.  With a surrogate attached, its accepted cells skip the integrator.
hipError_t enqueue_source(fcfd_ctx *c, double dt, const double *dt_dev, bool want_hrr) {
    ChemArgs a{};
    a.thermo = c->thermo; a.kin = static_cast<const DevKin *>(c->mech->d_kin); a.counter = c->d_counter;
    a.ncells = c->nblk * c->ncell; a.rtol = c->o.rtol; a.atol = c->o.atol;
    a.qf = c->qf; a.P = c->P; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S;
    a.dt = dt; a.dt_dev = dt_dev; a.stats = c->d_stats; a.order = c->d_order; a.hmem = c->d_hmem;
    if (want_hrr) { a.hrr = c->d_hrr; a.hrr_dot = c->d_hrr_dot; }
    a.layout = 0;
    const int chem_wg = c->chem_wg > 0 ? c->chem_wg : chem_workgroups(c->mech);
    a.route = c->d_accepted;
    a.f0 = c->d_f0;
    const int n = c->nblk * c->ncell;
    // explicit screen, one cell per lane: quiescent cells are final, the rest form the stiff list
    hipError_t e = timed(c, 6, [&]() {
        hipError_t e2 = launch_zero_u32(c->d_list_n, c->stream);
        if (e2 != hipSuccess) return e2;
        a.stiff = c->d_list;
        a.stiff_n = c->d_list_n;
        return launch_chem_screen(a, c->kk, c->stream);
    });
    if (e != hipSuccess) return e;
    // a long stiff list (the cold region failing the screen at once) first gets the explicit
    // half-step retry lane-per-cell; what remains, or a short list unchanged, is d_list2
    e = timed(c, 8, [&]() {
        hipError_t e2 = launch_zero_u32(c->d_list2_n, c->stream);
        if (e2 != hipSuccess) return e2;
        ChemArgs r = a;
        r.list = c->d_list;
        r.list_n = c->d_list_n;
        r.stiff = c->d_list2;
        r.stiff_n = c->d_list2_n;
        r.retry_done = c->d_retry;
        r.coop_max = chem_coop_max(chem_wg, c->kk);
        return launch_chem_retry(r, c->kk, c->stream);
    });
    if (e != hipSuccess) return e;
    a.retry_done = c->d_retry;
    if (c->sur) {
        fcfd_surrogate *s = c->sur;
        e = timed(c, 5, [&]() {
            FeatArgs fa{};
            fa.layout = 0; fa.n = n; fa.qf = c->qf; fa.P = c->P; fa.ncell = c->ncell; fa.lv = c->lv; fa.S = c->S;
            fa.dt_dev = dt_dev; fa.dt = dt;
            hipError_t e2 = s->network(fa, c->d_list2, c->d_list2_n, n, c->stream);
            if (e2 != hipSuccess) return e2;
            ApplyArgs ap{};
            ap.layout = 0; ap.qf = c->qf; ap.P = c->P; ap.ncell = c->ncell; ap.lv = c->lv; ap.S = c->S;
            ap.route = c->d_accepted;
            return s->accept(ap, a, c->d_list2, c->d_list2_n, n, c->stream);
        });
        if (e != hipSuccess) return e;
        e = timed(c, 3, [&]() { return s->fallback(a, chem_wg, c->stream); });
    } else {
        a.list = c->d_list2;
        a.list_n = c->d_list2_n;
        a.skip_explicit = 1;
        e = timed(c, 3, [&]() { return launch_chem(a, c->kk, chem_wg, c->stream); });
    }
    if (e != hipSuccess) return e;
    return timed(c, 4, [&]() {
        return launch_qf_to_qc_interior(c->thermo, c->kk, c->nblk, c->ncell, c->lv, c->S, c->qf,
                                        c->Q[c->cur], c->P, c->stream);
    });
}

// One fcfd_advance step: calc_dt, then the fluid and chemistry in the context's splitting
// (SRC/main.f90:134-163 for Lie; Strang puts half a chemistry step on either side of the fluid step,
// with a refresh after the first so that the fluid stage reads updated ghost cells).
hipError_t enqueue_step(fcfd_ctx *c, double cfl) {
    const bool react = c->o.do_reaction == 1;
    const bool fluid = c->o.do_hydro == 1 || c->o.do_tran == 1;
    const bool hrr = c->step_hrr == 1;
    hipError_t e = enqueue_dt(c, cfl, c->d_time);
    if (e != hipSuccess) return e;
    if (react && c->strang) {
        if ((e = launch_scale_scalar(c->d_dt, c->d_dth, 0.5, c->stream)) != hipSuccess) return e;
        if ((e = enqueue_source(c, 0.0, c->d_dth, false)) != hipSuccess) return e;
        if ((e = enqueue_refresh(c)) != hipSuccess) return e;
    }
    if (fluid && (e = enqueue_fluid(c, 0.0, c->d_dt)) != hipSuccess) return e;
    if (react && (e = enqueue_source(c, 0.0, c->strang ? c->d_dth : c->d_dt, hrr)) != hipSuccess) return e;
    // the auxiliary sources take the full dt after the chemistry, at the time after the step
    if (aux_on(c) && (e = enqueue_aux(c, 0.0, c->d_dt, 0.0, c->d_time)) != hipSuccess) return e;
    // without a source term the refresh would recompute identical primitives (QcToqf of an
    // unchanged Qc), so it is only enqueued when a source has changed the state
    if ((react || aux_on(c)) && (e = enqueue_refresh(c)) != hipSuccess) return e;
    return hipSuccess;
}

void drop_graphs(fcfd_ctx *c) {
    for (auto &g : c->graph)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

}  // namespace

// ---- scratch of the allocation-free batched chemistry entry point ---------------------------------
// Work-queue counter (offset 0), stiff-list length (offset 64) and stiff list (offset 256, one int
// per cell) of fcfd_chem_batch_noalloc, in device memory the library owns for the process lifetime
// (each GPU holds its own copy).  Host code never uses the array's name as an address: the device
// address of each GPU's copy is looked up once, with that GPU current, and kept per device.
constexpr int NOALLOC_MAX_CELLS = 1 << 20;
constexpr size_t NOALLOC_HEAD = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_HEAD + sizeof(int) * size_t(NOALLOC_MAX_CELLS);   // 4 MiB + 256 B
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 256;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// This device's copy of scratch (dev must be the current device).
int noalloc_scratch(int dev, unsigned char **out) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { set_error("device index out of range"); return FCFD_E_ARG; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return FCFD_E_HIP;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return FCFD_OK;
}

}  // namespace

extern "C" {

const char *fcfd_last_error(void) { return g_err.c_str(); }

int fcfd_mech_load(const char *chem_bin, const char *tran_bin, fcfd_mech **out) {
    if (!out) { set_error("out is NULL"); return FCFD_E_ARG; }
    *out = nullptr;
    auto m = std::make_unique<fcfd_mech>();
    std::string err;
    int rc = load_mech(chem_bin, tran_bin, m->host, err);
    if (rc != 0) { set_error(err); return rc; }
    *out = m.release();
    return FCFD_OK;
}

int fcfd_mech_load_text(const char *chem_inp, const char *therm_dat, const char *tran_bin, fcfd_mech **out) {
    if (!out) { set_error("out is NULL"); return FCFD_E_ARG; }
    *out = nullptr;
    auto m = std::make_unique<fcfd_mech>();
    std::string err;
    int rc = load_mech_text(chem_inp, therm_dat, m->host, err);
    if (rc == 0 && tran_bin) rc = load_tran_bin(tran_bin, m->host, err);
    if (rc != 0) { set_error(err); return rc; }
    *out = m.release();
    return FCFD_OK;
}

int fcfd_mech_thermo(const fcfd_mech *m, int k, double *temps, double *coef) {
    if (!m || k < 0 || k >= m->host.kk || !temps || !coef) { set_error("fcfd_mech_thermo: invalid argument"); return FCFD_E_ARG; }
    for (int c = 0; c < 3; ++c) temps[c] = m->host.tt[size_t(k) * 3 + c];
    for (int c = 0; c < 14; ++c) coef[c] = m->host.aa[size_t(k) * 14 + c];
    return FCFD_OK;
}

int fcfd_mech_reaction(const fcfd_mech *m, int i, double *par, int *nu, int *nk, int *ns, int *nr) {
    if (!m || i < 0 || i >= m->host.ii) { set_error("fcfd_mech_reaction: invalid argument"); return FCFD_E_ARG; }
    const Mech &h = m->host;
    if (par) for (int c = 0; c < 3; ++c) par[c] = h.par[size_t(i) * (h.npar + 1) + c];
    for (int q = 0; q < 6; ++q) {
        if (nu) nu[q] = h.nu[size_t(i) * 6 + q];
        if (nk) nk[q] = h.nk[size_t(i) * 6 + q];
    }
    if (ns) *ns = h.ns[i];
    if (nr) *nr = h.nr[i];
    return FCFD_OK;
}

int fcfd_mech_falloff(const fcfd_mech *m, int i, int *ifo, int *kfl, double *fpar) {
    if (!m || i < 0 || i >= m->host.ii || !ifo) { set_error("fcfd_mech_falloff: invalid argument"); return FCFD_E_ARG; }
    const Mech &h = m->host;
    *ifo = 0;
    for (size_t f = 0; f < h.ifl.size(); ++f)
        if (h.ifl[f] == i + 1) {
            *ifo = h.ifo[f];
            if (kfl) *kfl = h.kfl[f];
            if (fpar) for (int c = 0; c < 8 && c < h.nfar; ++c) fpar[c] = h.fpar[f * h.nfar + c];
        }
    return FCFD_OK;
}

int fcfd_mech_thirdbody(const fcfd_mech *m, int i, int *ntbs, int *nktb, double *aik) {
    if (!m || i < 0 || i >= m->host.ii || !ntbs) { set_error("fcfd_mech_thirdbody: invalid argument"); return FCFD_E_ARG; }
    const Mech &h = m->host;
    *ntbs = -1;
    for (size_t t = 0; t < h.itb.size(); ++t)
        if (h.itb[t] == i + 1) {
            *ntbs = h.ntbs[t];
            for (int l = 0; l < h.ntbs[t]; ++l) {
                if (nktb) nktb[l] = h.nktb[t * h.mxtb + l];
                if (aik) aik[l] = h.aik[t * h.mxtb + l];
            }
        }
    return FCFD_OK;
}

void fcfd_mech_free(fcfd_mech *m) {
    if (!m) return;
    (void)hipFree(m->d_thermo);
    (void)hipFree(m->d_kin);
    (void)hipFree(m->d_counter);
    (void)hipFree(m->d_list);
    (void)hipFree(m->d_list_n);
    delete m;
}

int fcfd_mech_species(const fcfd_mech *m, int k, char *name, int len, double *wt, int *ncf) {
    if (!m || k < 0 || k >= m->host.kk) { set_error("fcfd_mech_species: invalid argument"); return FCFD_E_ARG; }
    if (name && len > 0) {
        std::strncpy(name, m->host.names[k].c_str(), size_t(len) - 1);
        name[len - 1] = '\0';
    }
    if (wt) *wt = m->host.wt[k];
    if (ncf)
        for (int e = 0; e < m->host.mm; ++e) ncf[e] = m->host.ncf[size_t(k) * m->host.mm + e];
    return FCFD_OK;
}

int fcfd_mech_elements(const fcfd_mech *m, int *mm) {
    if (!m || !mm) { set_error("fcfd_mech_elements: invalid argument"); return FCFD_E_ARG; }
    *mm = m->host.mm;
    return FCFD_OK;
}

int fcfd_mech_dims(const fcfd_mech *m, int *kk, int *ii) {
    if (!m) { set_error("mech is NULL"); return FCFD_E_ARG; }
    if (kk) *kk = m->host.kk;
    if (ii) *ii = m->host.ii;
    return FCFD_OK;
}

namespace {

// Shared by fcfd_create (dxb empty: dx in every block) and fcfd_create_smr.
int create_impl(const fcfd_mech *m, const fcfd_opts *o, int nblk, int ncell, double dx,
                const std::vector<double> &dxb, const std::vector<double> &xm_in, fcfd_ctx **out) {
    *out = nullptr;
    int rc = check_opts(*o, m->host.kk);
    if (rc) return rc;
    // every error return below releases what was allocated so far through fcfd_destroy
    std::unique_ptr<fcfd_ctx, void (*)(fcfd_ctx *)> c(new fcfd_ctx(), fcfd_destroy);
    c->mech = m; c->o = *o; c->nblk = nblk; c->ncell = ncell; c->dx = dx;
    c->lv = lvir_of(*o);
    if (ncell < c->lv) { set_error("ncell must be at least lvir"); return FCFD_E_ARG; }
    c->S = ncell + 2 * c->lv; c->kk = m->host.kk; c->nequ = 3 + c->kk;
    c->P = round_up(long(nblk) * c->S, 64);
    c->PF = round_up(long(nblk) * (ncell + 1), 64);
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&c->thermo, sizeof(DevThermo)));
    HIP_TRY(hipMemcpy(c->thermo, &m->host.thermo, sizeof(DevThermo), hipMemcpyHostToDevice));
    const size_t sb = sizeof(double) * size_t(c->nequ) * c->P;
    for (auto &q : c->Q) { HIP_TRY(hipMalloc(&q, sb)); HIP_TRY(hipMemset(q, 0, sb)); }
    HIP_TRY(hipMalloc(&c->qf, sb));
    HIP_TRY(hipMemset(c->qf, 0, sb));
    HIP_TRY(hipMalloc(&c->F, sizeof(double) * size_t(c->nequ) * c->PF));
    if (o->select_time == 2) {
        HIP_TRY(hipMalloc(&c->acc, sb));
        HIP_TRY(hipMemset(c->acc, 0, sb));
    }
    HIP_TRY(hipMalloc(&c->god, sizeof(double) * 3 * c->PF));
    HIP_TRY(hipMemset(c->god, 0, sizeof(double) * 3 * c->PF));
    HIP_TRY(hipMalloc(&c->d_dt, sizeof(double)));
    HIP_TRY(hipMalloc(&c->d_time, sizeof(double)));
    HIP_TRY(hipMalloc(&c->d_cvel, sizeof(unsigned long long)));
    HIP_TRY(hipMemset(c->d_cvel, 0, sizeof(unsigned long long)));
    HIP_TRY(hipMemset(c->d_dt, 0, sizeof(double)));
    if (!dxb.empty()) {
        c->h_dxb = dxb;
        HIP_TRY(hipMalloc(&c->d_dxb, sizeof(double) * nblk));
        HIP_TRY(hipMemcpy(c->d_dxb, dxb.data(), sizeof(double) * nblk, hipMemcpyHostToDevice));
        std::vector<int> ifb;
        for (int b = 0; b + 1 < nblk; ++b)
            if (dxb[b] != dxb[b + 1]) ifb.push_back(b);
        c->n_smr_if = int(ifb.size());
        if (!ifb.empty()) {
            HIP_TRY(hipMalloc(&c->d_smr_ifb, sizeof(int) * ifb.size()));
            HIP_TRY(hipMemcpy(c->d_smr_ifb, ifb.data(), sizeof(int) * ifb.size(), hipMemcpyHostToDevice));
        }
        while (c->smr_left < nblk && dxb[c->smr_left] != dx) ++c->smr_left;
        while (c->smr_right < nblk - c->smr_left && dxb[nblk - 1 - c->smr_right] != dx) ++c->smr_right;
    }
    {
        // cell centres of the uniform block layout (SRC/block.f90:264-305 without refinement): a
        // running sum of dx from 0, shifted by the half cell length_smr = -dx/2
        std::vector<double> xm(size_t(nblk) * ncell);
        double xl = 0.0;
        const double length_smr = 0.0 - dx / 2.0;
        for (size_t g = 0; g < xm.size(); ++g) {
            if (g > 0) xl = xl + dx;
            xm[g] = xl - length_smr;
        }
        if (!xm_in.empty()) xm = xm_in;
        HIP_TRY(hipMalloc(&c->d_xm, sizeof(double) * xm.size()));
        HIP_TRY(hipMemcpy(c->d_xm, xm.data(), sizeof(double) * xm.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&c->d_aux, sizeof(AuxParams)));
        HIP_TRY(hipMemcpy(c->d_aux, &c->aux, sizeof(AuxParams), hipMemcpyHostToDevice));
        if (o->flag_lbound == 1 || o->flag_rbound == 1) HIP_TRY(hipMalloc(&c->d_bc, sizeof(FlowBC)));
    }
    if (o->do_tran) {
        if (!m->host.has_tran) { set_error("do_tran = 1 needs the transport link file (tran.bin)"); return FCFD_E_ARG; }
        if (m->host.no > 4) { set_error("transport fits of order above 4 are not supported"); return FCFD_E_UNSUPPORTED; }
        auto t = std::make_unique<DevTran>();
        std::memset(t.get(), 0, sizeof(DevTran));
        build_dev_tran(m->host, *t);
        HIP_TRY(hipMalloc(&c->d_tran, sizeof(DevTran)));
        HIP_TRY(hipMemcpy(c->d_tran, t.get(), sizeof(DevTran), hipMemcpyHostToDevice));
        c->PP = round_up(long(nblk) * (ncell + 2), 64);
        HIP_TRY(hipMalloc(&c->props, sizeof(double) * size_t(5 + 2 * c->kk) * c->PP));
    }
    if (o->do_reaction) {
        if (int rc2 = ensure_dev_tables(m)) return rc2;
        const size_t nint = size_t(nblk) * ncell;
        HIP_TRY(hipMalloc(&c->d_counter, sizeof(unsigned int)));
        HIP_TRY(hipMalloc(&c->d_hrr, sizeof(double) * nint));
        HIP_TRY(hipMalloc(&c->d_hrr_dot, sizeof(double) * nint));
        HIP_TRY(hipMalloc(&c->d_stats, sizeof(ChemCellStats) * nint));
        HIP_TRY(hipMalloc(&c->d_accepted, sizeof(int) * nint));
        HIP_TRY(hipMemset(c->d_accepted, 0, sizeof(int) * nint));
        HIP_TRY(hipMalloc(&c->d_list, sizeof(int) * nint));
        HIP_TRY(hipMalloc(&c->d_order, sizeof(int) * nint));
        HIP_TRY(hipMalloc(&c->d_hmem, sizeof(double) * nint));
        HIP_TRY(hipMemset(c->d_hmem, 0, sizeof(double) * nint));
        HIP_TRY(hipMalloc(&c->d_list_n, sizeof(unsigned)));
        HIP_TRY(hipMalloc(&c->d_list2, sizeof(int) * nint));
        HIP_TRY(hipMalloc(&c->d_list2_n, sizeof(unsigned)));
        HIP_TRY(hipMalloc(&c->d_retry, sizeof(int)));
        HIP_TRY(hipMalloc(&c->d_f0, sizeof(double) * size_t(c->kk + 1) * nint));
        HIP_TRY(hipMemset(c->d_retry, 0, sizeof(int)));
        HIP_TRY(hipMemset(c->d_hrr, 0, sizeof(double) * nint));
        HIP_TRY(hipMemset(c->d_hrr_dot, 0, sizeof(double) * nint));
        HIP_TRY(hipMemset(c->d_stats, 0, sizeof(ChemCellStats) * nint));
    }
    *out = c.release();
    return FCFD_OK;
}

}  // namespace

int fcfd_create(const fcfd_mech *m, const fcfd_opts *o, int nblk, int ncell, double dx, fcfd_ctx **out) {
    if (!m || !o || !out || nblk < 1 || ncell < 1 || !(dx > 0.0)) { set_error("fcfd_create: invalid argument"); return FCFD_E_ARG; }
    return create_impl(m, o, nblk, ncell, dx, {}, {}, out);
}

int fcfd_create_smr(const fcfd_mech *m, const fcfd_opts *o, int nblk_comp, int ncell, double dx_comp,
                    int lev_smr_0, int num_smr_0, int lev_smr_n, int num_smr_n, fcfd_ctx **out) {
    if (!m || !o || !out || nblk_comp < 1 || ncell < 1 || !(dx_comp > 0.0) || lev_smr_0 < 1 || lev_smr_n < 1 ||
        num_smr_0 < 0 || num_smr_n < 0 || lev_smr_0 > 40 || lev_smr_n > 40) {
        set_error("fcfd_create_smr: invalid argument");
        return FCFD_E_ARG;
    }
    if (ncell < 2 * lvir_of(*o)) { set_error("fcfd_create_smr: ncell must be at least twice lvir"); return FCFD_E_ARG; }
    if (o->select_time == 2) { set_error("fcfd_create_smr: a refined mesh is supported with select_time 0 or 1"); return FCFD_E_UNSUPPORTED; }
    const int n0 = (lev_smr_0 - 1) * num_smr_0, nn = (lev_smr_n - 1) * num_smr_n;
    const int nblk = nblk_comp + n0 + nn;                         // SRC/block.f90:136-146
    // block spacings (:160-218): the left refinement blocks coarsest first, the right ones coarsest
    // last, num_smr blocks per level, doubling per level away from the fine region
    std::vector<double> dxb(nblk, dx_comp);
    double length_smr_0 = 0.0;
    int b = 0;
    for (int lev = 2; lev <= lev_smr_0; ++lev)
        for (int k = 0; k < num_smr_0; ++k, ++b) {
            dxb[b] = dx_comp * double(1LL << (lev_smr_0 - lev + 1));
            length_smr_0 = length_smr_0 + dxb[b] * ncell;
        }
    b = 0;
    for (int lev = 2; lev <= lev_smr_n; ++lev)
        for (int k = 0; k < num_smr_n; ++k, ++b) dxb[nblk - 1 - b] = dx_comp * double(1LL << (lev_smr_n - lev + 1));
    // cell centres (:264-305): a running position stepping by each cell's dx, half a coarse cell
    // further on a coarse -> fine step and a quarter back on fine -> coarse, shifted so that the
    // left refinement region ends at x = 0
    std::vector<double> xm(size_t(nblk) * ncell);
    double xl = 0.0, dxl = 0.0;
    for (int ib = 0; ib < nblk; ++ib)
        for (int i = 0; i < ncell; ++i) {
            const double h = dxb[ib];
            double &x = xm[size_t(ib) * ncell + i];
            if (ib == 0 && i == 0) {
                x = 0.0; xl = 0.0; dxl = 0.0;
            } else {
                x = xl + h;
                xl = xl + h;
                if (dxl > h) { x = x + h / 2.0; xl = xl + h / 2.0; }
                if (dxl < h) { x = x - h / 4.0; xl = xl - h / 4.0; }
            }
            dxl = h;
        }
    const double length_smr = length_smr_0 - dxb[0] / 2.0;
    for (auto &x : xm) x = x - length_smr;
    int rc = create_impl(m, o, nblk, ncell, dx_comp, dxb, xm, out);
    if (rc == FCFD_OK) (*out)->idx0 = 1 - ncell * n0;            // nx_all_0, SRC/block.f90:128
    return rc;
}

int fcfd_nblocks(const fcfd_ctx *c) { return c ? c->nblk : FCFD_E_ARG; }

int fcfd_block_dx(const fcfd_ctx *c, double *dx_block) {
    if (!c || !dx_block) { set_error("fcfd_block_dx: invalid argument"); return FCFD_E_ARG; }
    for (int b = 0; b < c->nblk; ++b) dx_block[b] = c->h_dxb.empty() ? c->dx : c->h_dxb[b];
    return FCFD_OK;
}

int fcfd_get_cell_centres(fcfd_ctx *c, double *xm) {
    if (!c || !xm) { set_error("fcfd_get_cell_centres: invalid argument"); return FCFD_E_ARG; }
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(xm, c->d_xm, sizeof(double) * size_t(c->nblk) * c->ncell, hipMemcpyDeviceToHost));
    return FCFD_OK;
}

void fcfd_destroy(fcfd_ctx *c) {
    if (!c) return;
    drop_graphs(c);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto q : c->Q) (void)hipFree(q);
    (void)hipFree(c->qf); (void)hipFree(c->F); (void)hipFree(c->god); (void)hipFree(c->cells);
    (void)hipFree(c->d_dt); (void)hipFree(c->d_time); (void)hipFree(c->d_cvel); (void)hipFree(c->thermo);
    (void)hipFree(c->d_counter); (void)hipFree(c->d_hrr); (void)hipFree(c->d_hrr_dot); (void)hipFree(c->d_stats);
    (void)hipFree(c->d_accepted); (void)hipFree(c->d_list); (void)hipFree(c->d_order); (void)hipFree(c->d_hmem); (void)hipFree(c->d_list_n);
    (void)hipFree(c->d_list2); (void)hipFree(c->d_list2_n); (void)hipFree(c->d_retry); (void)hipFree(c->d_f0);
    (void)hipFree(c->d_tran); (void)hipFree(c->props); (void)hipFree(c->acc); (void)hipFree(c->d_dth);
    (void)hipFree(c->d_bc); (void)hipFree(c->d_aux); (void)hipFree(c->d_xm); (void)hipFree(c->d_qrad); (void)hipFree(c->d_qw);
    (void)hipFree(c->d_dxb); (void)hipFree(c->d_smr_ifb);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int fcfd_lvir(const fcfd_ctx *c) { return c ? c->lv : FCFD_E_ARG; }

int fcfd_upload(fcfd_ctx *c, int ib, const double *Qc, const double *qf) {
    if (!c || !Qc || ib < 0 || ib >= c->nblk) { set_error("fcfd_upload: invalid argument"); return FCFD_E_ARG; }
    const int S = c->S, ne = c->nequ;
    std::vector<double> h(size_t(ne) * S);
    auto put = [&](const double *src, double *dst) -> int {
        for (int v = 0; v < ne; ++v)
            for (int j = 0; j < S; ++j) h[size_t(v) * S + j] = src[size_t(j) * ne + v];
        for (int v = 0; v < ne; ++v)
            HIP_TRY(hipMemcpyAsync(dst + long(v) * c->P + long(ib) * S, h.data() + size_t(v) * S, sizeof(double) * S,
                                   hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return 0;
    };
    if (c->d_hmem) {   // a new state: the integrator's step-size memory starts over
        HIP_TRY(hipMemsetAsync(c->d_hmem, 0, sizeof(double) * size_t(c->nblk) * c->ncell, c->stream));
    }
    if (int rc = put(Qc, c->Q[c->cur])) return rc;
    if (qf) {
        if (int rc = put(qf, c->qf)) return rc;
    } else {
        HIP_TRY(launch_qc_to_qf(c->thermo, c->kk, S, c->Q[c->cur] + long(ib) * S, c->qf + long(ib) * S, c->P, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return FCFD_OK;
}

int fcfd_download(fcfd_ctx *c, int ib, double *Qc, double *qf) {
    if (!c || ib < 0 || ib >= c->nblk) { set_error("fcfd_download: invalid argument"); return FCFD_E_ARG; }
    const int S = c->S, ne = c->nequ;
    std::vector<double> h(size_t(ne) * S);
    auto get = [&](const double *src, double *dst) -> int {
        // one contiguous copy per variable (block ib's S columns of it)
        for (int v = 0; v < ne; ++v)
            HIP_TRY(hipMemcpyAsync(h.data() + size_t(v) * S, src + long(v) * c->P + long(ib) * S, sizeof(double) * S,
                                   hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int v = 0; v < ne; ++v)
            for (int j = 0; j < S; ++j) dst[size_t(j) * ne + v] = h[size_t(v) * S + j];
        return 0;
    };
    if (Qc)
        if (int rc = get(c->Q[c->cur], Qc)) return rc;
    if (qf)
        if (int rc = get(c->qf, qf)) return rc;
    return FCFD_OK;
}

int fcfd_init_cells(fcfd_ctx *c, const double *cells) {
    if (!c || !cells) { set_error("fcfd_init_cells: invalid argument"); return FCFD_E_ARG; }
    const size_t n = size_t(c->nblk) * c->ncell * (3 + c->kk);
    if (!c->cells) HIP_TRY(hipMalloc(&c->cells, sizeof(double) * n));
    HIP_TRY(hipMemcpyAsync(c->cells, cells, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    InitArgs a{};
    a.thermo = c->thermo; a.cells = c->cells; a.Qc = c->Q[c->cur]; a.qf = c->qf;
    a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P;
    HIP_TRY(launch_init(a, c->kk, c->stream));
    if (c->d_hmem) HIP_TRY(hipMemsetAsync(c->d_hmem, 0, sizeof(double) * size_t(c->nblk) * c->ncell, c->stream));
    // ghost fill + QcToqf on ghosts only (SRC/main.f90:96-114)
    StageArgs s = stage_args(c, 2, nullptr, c->Q[c->cur], c->Q[c->cur], 0.0, nullptr, 1.0, 1.0);
    s.ghosts_only = 1;
    if ((c->o.flag_lbound == 1 || c->o.flag_rbound == 1) && !c->bc_set) {
        set_error("fcfd_init_cells: inflow / outflow boundaries need fcfd_set_flow_boundary first");
        return FCFD_E_ARG;
    }
    HIP_TRY(launch_stage_end(s, c->kk, c->stream));
    HIP_TRY(smr_halo(c, c->Q[c->cur]));
    HIP_TRY(open_bc(c, c->Q[c->cur]));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_calc_dt(fcfd_ctx *c, double cfl, double *dt) {
    if (!c || !dt) { set_error("fcfd_calc_dt: invalid argument"); return FCFD_E_ARG; }
    HIP_TRY(enqueue_dt(c, cfl, nullptr));
    HIP_TRY(hipMemcpyAsync(dt, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(collect_timing(c));
    return FCFD_OK;
}

int fcfd_solve_fluid(fcfd_ctx *c, double dt) {
    if (!c) { set_error("fcfd_solve_fluid: invalid argument"); return FCFD_E_ARG; }
    if (c->o.do_hydro != 1 && c->o.do_tran != 1) return FCFD_OK;
    HIP_TRY(enqueue_fluid(c, dt, nullptr));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(collect_timing(c));
    return FCFD_OK;
}

int fcfd_refresh(fcfd_ctx *c) {
    if (!c) { set_error("fcfd_refresh: invalid argument"); return FCFD_E_ARG; }
    HIP_TRY(enqueue_refresh(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(collect_timing(c));
    return FCFD_OK;
}

int fcfd_advance(fcfd_ctx *c, double cfl, int nsteps, double *time) {
    if (!c || nsteps < 0 || !(cfl > 0.0)) { set_error("fcfd_advance: invalid argument"); return FCFD_E_ARG; }
    const double t0 = time ? *time : 0.0;
    HIP_TRY(hipMemcpyAsync(c->d_time, &t0, sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (c->timing) {
        for (int s = 0; s < nsteps; ++s) HIP_TRY(enqueue_step(c, cfl));
        HIP_TRY(collect_timing(c));
    } else {
        const int gen = tune_gen();
        if (cfl != c->graph_cfl || gen != c->graph_tune_gen) {
            drop_graphs(c);
            c->graph_cfl = cfl;
            c->graph_tune_gen = gen;
        }
        for (int s = 0; s < nsteps; ++s) {
            const int cur = c->cur;
            if (!c->graph[cur]) {
                hipGraph_t g;
                HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                hipError_t e = enqueue_step(c, cfl);
                hipError_t e2 = hipStreamEndCapture(c->stream, &g);
                HIP_TRY(e);
                HIP_TRY(e2);
                c->graph_next[cur] = c->cur;
                c->cur = cur;                       // capture only recorded the step
                HIP_TRY(hipGraphInstantiate(&c->graph[cur], g, nullptr, nullptr, 0));
                HIP_TRY(hipGraphDestroy(g));
            }
            HIP_TRY(hipGraphLaunch(c->graph[cur], c->stream));
            c->cur = c->graph_next[cur];
        }
    }
    if (time) HIP_TRY(hipMemcpyAsync(time, c->d_time, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_solve_source(fcfd_ctx *c, double dt, double time, double *hrr, double *hrr_dot) {
    if (!c) { set_error("fcfd_solve_source: invalid argument"); return FCFD_E_ARG; }
    if (c->o.do_reaction == 1) HIP_TRY(enqueue_source(c, dt, nullptr, hrr || hrr_dot));
    if (aux_on(c)) HIP_TRY(enqueue_aux(c, dt, nullptr, time, nullptr));
    if (c->o.do_reaction != 1) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(collect_timing(c));
        return FCFD_OK;
    }
    const size_t nint = size_t(c->nblk) * c->ncell;
    if (hrr) HIP_TRY(hipMemcpyAsync(hrr, c->d_hrr, sizeof(double) * nint, hipMemcpyDeviceToHost, c->stream));
    if (hrr_dot) HIP_TRY(hipMemcpyAsync(hrr_dot, c->d_hrr_dot, sizeof(double) * nint, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(collect_timing(c));
    return FCFD_OK;
}

int fcfd_set_step(fcfd_ctx *c, int strang, int hrr_each_step) {
    if (!c || (strang != 0 && strang != 1) || (hrr_each_step != 0 && hrr_each_step != 1)) {
        set_error("fcfd_set_step: strang and hrr_each_step must be 0 or 1");
        return FCFD_E_ARG;
    }
    if (strang && c->o.do_reaction != 1) { set_error("fcfd_set_step: Strang splitting needs do_reaction = 1"); return FCFD_E_ARG; }
    if (strang && !c->d_dth) HIP_TRY(hipMalloc(&c->d_dth, sizeof(double)));
    drop_graphs(c);
    c->strang = strang;
    c->step_hrr = hrr_each_step;
    return FCFD_OK;
}

// ---- open boundaries and auxiliary sources -----------------------------------------------------

namespace {

int ensure_tran(fcfd_ctx *c) {
    if (c->d_tran) return FCFD_OK;
    const fcfd_mech *m = c->mech;
    if (!m->host.has_tran) { set_error("the wall heat loss needs the transport link file (tran.bin)"); return FCFD_E_ARG; }
    if (m->host.no > 4) { set_error("transport fits of order above 4 are not supported"); return FCFD_E_UNSUPPORTED; }
    auto t = std::make_unique<DevTran>();
    std::memset(t.get(), 0, sizeof(DevTran));
    build_dev_tran(m->host, *t);
    DevBuf b;
    HIP_TRY(hipMalloc(&b.p, sizeof(DevTran)));
    HIP_TRY(hipMemcpy(b.p, t.get(), sizeof(DevTran), hipMemcpyHostToDevice));
    c->d_tran = static_cast<DevTran *>(b.p);
    b.p = nullptr;
    return FCFD_OK;
}

int push_aux(fcfd_ctx *c) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->d_aux, &c->aux, sizeof(AuxParams), hipMemcpyHostToDevice));
    drop_graphs(c);                      // the captured step enqueues the sources that are on
    return FCFD_OK;
}

int ensure_cells_buffer(fcfd_ctx *c, double **p) {
    if (*p) return FCFD_OK;
    const size_t n = size_t(c->nblk) * c->ncell;
    HIP_TRY(hipMalloc(p, sizeof(double) * n));
    HIP_TRY(hipMemset(*p, 0, sizeof(double) * n));
    return FCFD_OK;
}

}  // namespace

int fcfd_set_flow_boundary(fcfd_ctx *c, double T_in, double u_in_cm_s, const double *x_in, double p_out_atm) {
    if (!c) { set_error("fcfd_set_flow_boundary: invalid argument"); return FCFD_E_ARG; }
    if (c->o.flag_lbound != 1 && c->o.flag_rbound != 1) {
        set_error("fcfd_set_flow_boundary: the context has walls on both ends (flag_lbound = flag_rbound = 0)");
        return FCFD_E_ARG;
    }
    if (c->o.flag_lbound == 1 && (!x_in || !(T_in > 0.0))) {
        set_error("fcfd_set_flow_boundary: an inflow needs T_in > 0 and the inflow mole fractions");
        return FCFD_E_ARG;
    }
    if (c->o.flag_rbound == 1 && !(p_out_atm > 0.0)) { set_error("fcfd_set_flow_boundary: p_out must be positive"); return FCFD_E_ARG; }
    FlowBC h{};
    h.lbound = c->o.flag_lbound; h.rbound = c->o.flag_rbound;
    h.T_in = T_in;
    h.p_out = p_out_atm * 101325.0;                                // SRC/flow_bound.f90:51
    if (x_in)
        for (int k = 0; k < c->kk; ++k) h.x_in[k] = x_in[k];
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->d_bc, &h, sizeof(FlowBC), hipMemcpyHostToDevice));
    if (c->o.flag_lbound == 1) HIP_TRY(launch_flow_bc_setup(c->thermo, c->d_bc, u_in_cm_s * 1.e-2, c->kk, c->stream));
    c->bc_set = true;
    drop_graphs(c);
    // refill the ghost cells of the current state as the start-up does (SRC/main.f90:96-114)
    StageArgs s = stage_args(c, 2, nullptr, c->Q[c->cur], c->Q[c->cur], 0.0, nullptr, 1.0, 1.0);
    s.ghosts_only = 1;
    HIP_TRY(launch_stage_end(s, c->kk, c->stream));
    HIP_TRY(smr_halo(c, c->Q[c->cur]));
    HIP_TRY(open_bc(c, c->Q[c->cur]));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_set_radiation(fcfd_ctx *c, int n, const int *species, const double *pmac) {
    if (!c || n < 0 || n > MAX_RAD || (n > 0 && (!species || !pmac))) {
        set_error("fcfd_set_radiation: invalid argument (at most " + std::to_string(MAX_RAD) + " species)");
        return FCFD_E_ARG;
    }
    for (int r = 0; r < n; ++r)
        if (species[r] < 0 || species[r] >= c->kk) { set_error("fcfd_set_radiation: species index out of range"); return FCFD_E_ARG; }
    if (n > 0)
        if (int rc = ensure_cells_buffer(c, &c->d_qrad)) return rc;
    for (int r = 0; r < n; ++r) {
        c->aux.rad_idx[r] = species[r];
        for (int j = 0; j < 6; ++j) c->aux.pmac[r][j] = pmac[6 * r + j];
    }
    c->aux.nrad = n;
    return push_aux(c);
}

int fcfd_set_wall_heat(fcfd_ctx *c, int on, double Nu, double d_mm, double Tw_min, double Tw_max, double alpha,
                       double beta) {
    if (!c || (on != 0 && on != 1) || (on && !(d_mm > 0.0))) { set_error("fcfd_set_wall_heat: invalid argument"); return FCFD_E_ARG; }
    if (on) {
        if (int rc = ensure_tran(c)) return rc;
        if (int rc = ensure_cells_buffer(c, &c->d_qw)) return rc;
    }
    c->aux.wall_heat = on;
    c->aux.Nu = Nu; c->aux.d_m = d_mm * 1.0e-3;                     // SRC/mfr.f90:36
    c->aux.Tw_min = Tw_min; c->aux.Tw_max = Tw_max; c->aux.alpha = alpha; c->aux.beta = beta;
    return push_aux(c);
}

int fcfd_set_ignition(fcfd_ctx *c, int type, int n, const double *E_mJ, const double *r_um, const double *t_us,
                      const double *t0_us) {
    if (!c || type < 0 || type > 2 || n < 0 || n > MAX_IG || (type != 0 && (n == 0 || !E_mJ || !r_um || !t_us || !t0_us))) {
        set_error("fcfd_set_ignition: invalid argument (type 0, 1 or 2; 1 to " + std::to_string(MAX_IG) + " sources)");
        return FCFD_E_ARG;
    }
    const double pi = 3.14159265359;                                // SRC/ignitionsource.f90:27
    for (int s = 0; s < (type ? n : 0); ++s) {                      // :57-61, :121-125
        c->aux.E_j[s] = E_mJ[s] * 1.e-3;
        c->aux.r_m[s] = r_um[s] * 1.e-6;
        c->aux.t_s[s] = t_us[s] * 1.e-6;
        c->aux.t0_s[s] = t0_us[s] * 1.e-6;
        c->aux.vol[s] = 4.0 * pi * (c->aux.r_m[s] * c->aux.r_m[s] * c->aux.r_m[s]) / 3.0;
    }
    c->aux.ig_type = type;
    c->aux.n_ig = type ? n : 0;
    return push_aux(c);
}

int fcfd_set_cell_centres(fcfd_ctx *c, const double *xm) {
    if (!c || !xm) { set_error("fcfd_set_cell_centres: invalid argument"); return FCFD_E_ARG; }
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(c->d_xm, xm, sizeof(double) * size_t(c->nblk) * c->ncell, hipMemcpyHostToDevice));
    return FCFD_OK;
}

int fcfd_eos_fields(fcfd_ctx *c, double *out) {
    if (!c || !out) { set_error("fcfd_eos_fields: invalid argument"); return FCFD_E_ARG; }
    const size_t n = size_t(c->nblk) * c->ncell;
    DevBuf d;
    HIP_TRY(hipMalloc(&d.p, sizeof(double) * n * (15 + c->kk)));
    EosArgs a{};
    a.thermo = c->thermo; a.par = c->d_aux; a.wall_heat = c->aux.wall_heat; a.Q = c->Q[c->cur]; a.qf = c->qf;
    a.xm = c->d_xm; a.dxb = c->d_dxb; a.dx = c->dx;
    a.hrr = c->d_hrr; a.hrd = c->d_hrr_dot; a.qrad = c->d_qrad; a.qw = c->d_qw;
    a.out = static_cast<double *>(d.p); a.idx0 = c->idx0;
    a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P;
    HIP_TRY(launch_eos_fields(a, c->kk, c->stream));
    HIP_TRY(hipMemcpyAsync(out, d.p, sizeof(double) * n * (15 + c->kk), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_init_restart(fcfd_ctx *c, const double *fields) {
    if (!c || !fields) { set_error("fcfd_init_restart: invalid argument"); return FCFD_E_ARG; }
    if ((c->o.flag_lbound == 1 || c->o.flag_rbound == 1) && !c->bc_set) {
        set_error("fcfd_init_restart: inflow / outflow boundaries need fcfd_set_flow_boundary first");
        return FCFD_E_ARG;
    }
    const size_t n = size_t(c->nblk) * c->ncell;
    DevBuf d;
    HIP_TRY(hipMalloc(&d.p, sizeof(double) * n * (4 + c->kk)));
    HIP_TRY(hipMemcpyAsync(d.p, fields, sizeof(double) * n * (4 + c->kk), hipMemcpyHostToDevice, c->stream));
    EosArgs a{};
    a.thermo = c->thermo; a.Q = c->Q[c->cur]; a.qf = c->qf; a.out = static_cast<double *>(d.p);
    a.nblk = c->nblk; a.ncell = c->ncell; a.lv = c->lv; a.S = c->S; a.P = c->P;
    HIP_TRY(launch_restart(a, c->kk, c->stream));
    if (c->d_hmem) HIP_TRY(hipMemsetAsync(c->d_hmem, 0, sizeof(double) * n, c->stream));
    StageArgs s = stage_args(c, 2, nullptr, c->Q[c->cur], c->Q[c->cur], 0.0, nullptr, 1.0, 1.0);
    s.ghosts_only = 1;
    HIP_TRY(launch_stage_end(s, c->kk, c->stream));
    HIP_TRY(smr_halo(c, c->Q[c->cur]));
    HIP_TRY(open_bc(c, c->Q[c->cur]));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_get_aux(fcfd_ctx *c, double *qrad, double *qw) {
    if (!c) { set_error("fcfd_get_aux: invalid argument"); return FCFD_E_ARG; }
    const size_t n = size_t(c->nblk) * c->ncell;
    if (qrad) {
        if (c->d_qrad) HIP_TRY(hipMemcpyAsync(qrad, c->d_qrad, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        else std::memset(qrad, 0, sizeof(double) * n);
    }
    if (qw) {
        if (c->d_qw) HIP_TRY(hipMemcpyAsync(qw, c->d_qw, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        else std::memset(qw, 0, sizeof(double) * n);
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_get_hrr(fcfd_ctx *c, double *hrr, double *hrr_dot) {
    if (!c || !c->d_hrr) { set_error("fcfd_get_hrr: no chemistry on this context"); return FCFD_E_ARG; }
    const size_t nint = size_t(c->nblk) * c->ncell;
    if (hrr) HIP_TRY(hipMemcpyAsync(hrr, c->d_hrr, sizeof(double) * nint, hipMemcpyDeviceToHost, c->stream));
    if (hrr_dot) HIP_TRY(hipMemcpyAsync(hrr_dot, c->d_hrr_dot, sizeof(double) * nint, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_chem_stats(fcfd_ctx *c, int *per_cell) {
    if (!c || !per_cell || !c->d_stats) { set_error("fcfd_chem_stats: no chemistry on this context"); return FCFD_E_ARG; }
    HIP_TRY(hipMemcpyAsync(per_cell, c->d_stats, sizeof(ChemCellStats) * size_t(c->nblk) * c->ncell,
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FCFD_OK;
}

int fcfd_chem_batch(const fcfd_mech *m, const fcfd_opts *o, int n, const double *rho, const double *dt,
                    double *T, double *Y, int *stats, void *stream) {
    if (!m || !o || n < 0 || (n > 0 && (!rho || !dt || !T || !Y))) { set_error("fcfd_chem_batch: invalid argument"); return FCFD_E_ARG; }
    if (!(o->rtol > 0.0 && o->atol > 0.0)) { set_error("fcfd_chem_batch: atol and rtol must be positive"); return FCFD_E_ARG; }
    if (n == 0) return FCFD_OK;
    if (int rc = check_opts(fcfd_opts{1, 0, 1, 1, 2, 2, 0, 2, 0, 0, 0, o->atol, o->rtol}, m->host.kk)) return rc;
    if (int rc = ensure_dev_tables(m)) return rc;
    ChemArgs a{};
    a.thermo = m->d_thermo; a.kin = static_cast<const DevKin *>(m->d_kin); a.counter = m->d_counter;
    a.ncells = n; a.rtol = o->rtol; a.atol = o->atol;
    a.b_rho = rho; a.b_dt = dt; a.b_T = T; a.b_y = Y;
    a.stats = reinterpret_cast<ChemCellStats *>(stats);
    a.layout = 1;
    fcfd_mech *mm = const_cast<fcfd_mech *>(m);
    if (size_t(n) > mm->list_cap) {
        if (mm->d_list) HIP_TRY(hipFree(mm->d_list));
        mm->d_list = nullptr;
        mm->list_cap = 0;
        HIP_TRY(hipMalloc(&mm->d_list, sizeof(int) * size_t(n)));
        mm->list_cap = size_t(n);
    }
    if (!mm->d_list_n) HIP_TRY(hipMalloc(&mm->d_list_n, sizeof(unsigned)));
    HIP_TRY(launch_chem_full(a, m->host.kk, chem_workgroups(m), mm->d_list, mm->d_list_n,
                             static_cast<hipStream_t>(stream)));
    HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return FCFD_OK;
}

int fcfd_mech_prepare(const fcfd_mech *m) {
    if (!m) { set_error("fcfd_mech_prepare: invalid argument"); return FCFD_E_ARG; }
    if (int rc = ensure_dev_tables(m)) return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    unsigned char *base = nullptr;
    return noalloc_scratch(dev, &base);
}

int fcfd_chem_batch_noalloc(const fcfd_mech *m, const fcfd_opts *o, int n, const double *rho, const double *dt,
                            double *T, double *Y, int *stats, void *stream) {
    if (!m || !o || n < 0 || (n > 0 && (!rho || !dt || !T || !Y))) { set_error("fcfd_chem_batch_noalloc: invalid argument"); return FCFD_E_ARG; }
    if (!(o->rtol > 0.0 && o->atol > 0.0)) { set_error("fcfd_chem_batch_noalloc: atol and rtol must be positive"); return FCFD_E_ARG; }
    if (n > NOALLOC_MAX_CELLS) {
        set_error("fcfd_chem_batch_noalloc: n = " + std::to_string(n) + " exceeds the allocation-free limit of " +
                  std::to_string(NOALLOC_MAX_CELLS) + " cells (use fcfd_chem_batch, or split the batch)");
        return FCFD_E_ARG;
    }
    if (n == 0) return FCFD_OK;
    if (int rc = check_opts(fcfd_opts{1, 0, 1, 1, 2, 2, 0, 2, 0, 0, 0, o->atol, o->rtol}, m->host.kk)) return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (!m->d_thermo || m->device != dev) {
        set_error("fcfd_chem_batch_noalloc: the mechanism's tables are not on this device; call fcfd_mech_prepare "
                  "once (it allocates) before the allocation-free calls");
        return FCFD_E_ARG;
    }
    unsigned char *base = nullptr;
    if (int rc = noalloc_scratch(dev, &base)) return rc;
    ChemArgs a{};
    a.thermo = m->d_thermo; a.kin = static_cast<const DevKin *>(m->d_kin);
    a.counter = reinterpret_cast<unsigned int *>(base);                 // work-queue head
    a.ncells = n; a.rtol = o->rtol; a.atol = o->atol;
    a.b_rho = rho; a.b_dt = dt; a.b_T = T; a.b_y = Y;
    a.stats = reinterpret_cast<ChemCellStats *>(stats);
    a.layout = 1;
    unsigned *list_n = reinterpret_cast<unsigned *>(base + 64);          // stiff-list length
    int *list = reinterpret_cast<int *>(base + NOALLOC_HEAD);            // stiff list
    // both counters are cleared by kernels on the stream before use (launch_chem_full, launch_chem)
    HIP_TRY(launch_chem_full(a, m->host.kk, chem_workgroups(m), list, list_n, static_cast<hipStream_t>(stream)));
    HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return FCFD_OK;
}

int fcfd_set_surrogate(fcfd_ctx *c, fcfd_surrogate *s) {
    if (!c) { set_error("fcfd_set_surrogate: invalid argument"); return FCFD_E_ARG; }
    if (s && c->o.do_reaction != 1) { set_error("fcfd_set_surrogate: the context has no chemistry"); return FCFD_E_ARG; }
    if (s && s->mech != c->mech) { set_error("fcfd_set_surrogate: surrogate and context use different mechanisms"); return FCFD_E_ARG; }
    if (s) {
        const int n = c->nblk * c->ncell;
        HIP_TRY(s->reserve(int(round_up(n, 64)), false));
    }
    drop_graphs(c);
    c->sur = s;
    return FCFD_OK;
}

int fcfd_set_launch(const char *key, int nsp, int value) {
    static const char *const keys[] = {"flux_variant", "stage_variant", "chem_wg_per_cu", "chem_four_wave_factor",
                                       "chem_lu"};
    static const int lo[] = {0, 0, 1, 0, 0}, hi[] = {2, 3, 64, 1 << 20, 2};
    if (!key) { set_error("fcfd_set_launch: invalid key"); return FCFD_E_ARG; }
    for (int i = 0; i < 5; ++i)
        if (std::strcmp(key, keys[i]) == 0) {
            if (value < lo[i] || value > hi[i]) { set_error(std::string("fcfd_set_launch: ") + key + " out of range"); return FCFD_E_ARG; }
            tuned(nsp, key, 0);                       // loads the file first, so it does not overwrite this later
            TuneTable &t = tune_table();
            std::lock_guard<std::mutex> g(t.mu);
            t.v[std::to_string(nsp) + ":" + key] = value;
            ++t.gen;
            return FCFD_OK;
        }
    set_error(std::string("fcfd_set_launch: unknown key ") + key);
    return FCFD_E_ARG;
}

int fcfd_get_launch(const char *key, int nsp, int *value) {
    if (!key || !value) { set_error("fcfd_get_launch: invalid argument"); return FCFD_E_ARG; }
    *value = tuned(nsp, key, launch_default(key));
    return FCFD_OK;
}

int fcfd_load_tuning(const char *path) {
    TuneTable &t = tune_table();
    std::lock_guard<std::mutex> g(t.mu);
    if (!load_tuning_locked(t, path ? std::string(path) : default_tune_path())) {
        set_error(std::string("fcfd_load_tuning: cannot read ") + (path ? path : t.source.c_str()));
        return FCFD_E_IO;
    }
    return FCFD_OK;
}

int fcfd_set_chem_workgroups(fcfd_ctx *c, int n) {
    if (!c || n < 0 || n > (1 << 24)) { set_error("fcfd_set_chem_workgroups: n must be in [0, 2^24]"); return FCFD_E_ARG; }
    c->chem_wg = n;
    drop_graphs(c);   // captured steps hold the old grid
    return FCFD_OK;
}

int fcfd_source_accepted(fcfd_ctx *c, int *per_cell, int *n_fallback) {
    if (!c || !c->d_accepted) { set_error("fcfd_source_accepted: the context has no chemistry"); return FCFD_E_ARG; }
    const size_t n = size_t(c->nblk) * c->ncell;
    std::vector<int> h(n);
    HIP_TRY(hipMemcpyAsync(h.data(), c->d_accepted, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (per_cell) std::memcpy(per_cell, h.data(), sizeof(int) * n);
    if (n_fallback) {
        long nf = 0;
        for (int v : h) nf += v == 0;
        *n_fallback = int(nf);
    }
    return FCFD_OK;
}

int fcfd_set_timing(fcfd_ctx *c, int enable) {
    if (!c) { set_error("fcfd_set_timing: invalid argument"); return FCFD_E_ARG; }
    c->timing = enable != 0;
    for (int i = 0; i < NTIMERS; ++i) { c->ms[i] = 0.0; c->cnt[i] = 0; }
    return FCFD_OK;
}

int fcfd_get_timing(fcfd_ctx *c, int id, double *ms_total, long long *launches) {
    if (!c || id < 0 || id >= NTIMERS) { set_error("fcfd_get_timing: invalid argument"); return FCFD_E_ARG; }
    if (ms_total) *ms_total = c->ms[id];
    if (launches) *launches = c->cnt[id];
    return FCFD_OK;
}

}  // extern "C"

// ---- stateless host-array entry points ----------------------------------------------------
namespace {

// rows x cols (row-major, AoS) -> cols x rows (SoA)
std::vector<double> transpose(const double *in, long rows, int cols) {
    std::vector<double> out(size_t(rows) * cols);
    for (long r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) out[size_t(c) * rows + r] = in[size_t(r) * cols + c];
    return out;
}

void untranspose(const std::vector<double> &in, long rows, int cols, double *out) {
    for (long r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) out[size_t(r) * cols + c] = in[size_t(c) * rows + r];
}

}  // namespace

extern "C" {

int fcfd_flux_tran_host(const fcfd_mech *m, int n, int lv, double dx, const double *Qc, double *Fplus) {
    if (!m || !Qc || !Fplus || n < 1 || lv < 1 || !(dx > 0.0)) { set_error("fcfd_flux_tran_host: invalid argument"); return FCFD_E_ARG; }
    if (!m->host.has_tran) { set_error("fcfd_flux_tran_host: the mechanism was loaded without tran.bin"); return FCFD_E_ARG; }
    if (m->host.no > 4) { set_error("transport fits of order above 4 are not supported"); return FCFD_E_UNSUPPORTED; }
    if (int rc = check_opts(fcfd_opts{1, 1, 0, 1, 2, 2, 0, 2, 0, 0, 0, 1e-13, 1e-5}, m->host.kk)) return rc;
    const int ne = 3 + m->host.kk;
    const long cols = n + 2L * lv, nif = n + 1;
    DevBuf th, dt, dq, dF, dp;
    HIP_TRY(hipMalloc(&th.p, sizeof(DevThermo)));
    HIP_TRY(hipMemcpy(th.p, &m->host.thermo, sizeof(DevThermo), hipMemcpyHostToDevice));
    auto t = std::make_unique<DevTran>();
    std::memset(t.get(), 0, sizeof(DevTran));
    build_dev_tran(m->host, *t);
    HIP_TRY(hipMalloc(&dt.p, sizeof(DevTran)));
    HIP_TRY(hipMemcpy(dt.p, t.get(), sizeof(DevTran), hipMemcpyHostToDevice));
    auto hq = transpose(Qc, cols, ne);
    HIP_TRY(hipMalloc(&dq.p, sizeof(double) * hq.size()));
    HIP_TRY(hipMemcpy(dq.p, hq.data(), sizeof(double) * hq.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&dF.p, sizeof(double) * ne * nif));
    HIP_TRY(hipMalloc(&dp.p, sizeof(double) * (5 + 2 * m->host.kk) * (n + 2)));
    TranArgs a{};
    a.thermo = static_cast<DevThermo *>(th.p); a.tran = static_cast<DevTran *>(dt.p);
    a.Qc = static_cast<double *>(dq.p); a.props = static_cast<double *>(dp.p); a.F = static_cast<double *>(dF.p);
    a.nblk = 1; a.ncell = n; a.lv = lv; a.S = int(cols); a.P = cols; a.PF = nif; a.PP = n + 2; a.dx = dx;
    HIP_TRY(launch_flux_tran(a, m->host.kk, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    std::vector<double> hF(size_t(ne) * nif);
    HIP_TRY(hipMemcpy(hF.data(), dF.p, sizeof(double) * hF.size(), hipMemcpyDeviceToHost));
    untranspose(hF, nif, ne, Fplus);
    return FCFD_OK;
}

int fcfd_flux_advection_host(const fcfd_mech *m, const fcfd_opts *o, int n, int lv, double dx,
                             const double *qf, double *Fplus, double *godunov) {
    if (!m || !o || !qf || !Fplus || n < 1 || lv < 1) { set_error("fcfd_flux_advection_host: invalid argument"); return FCFD_E_ARG; }
    fcfd_opts oo = *o;
    oo.do_tran = 0;
    if (int rc = check_opts(oo, m->host.kk)) return rc;
    if (lv < lvir_of(oo)) { set_error("lv smaller than the scheme's stencil half-width"); return FCFD_E_ARG; }
    (void)dx;
    const int ne = 3 + m->host.kk;
    const long cols = n + 2L * lv, nif = n + 1;
    DevBuf th, dq, dF, dG;
    HIP_TRY(hipMalloc(&th.p, sizeof(DevThermo)));
    HIP_TRY(hipMemcpy(th.p, &m->host.thermo, sizeof(DevThermo), hipMemcpyHostToDevice));
    auto hq = transpose(qf, cols, ne);
    HIP_TRY(hipMalloc(&dq.p, sizeof(double) * hq.size()));
    HIP_TRY(hipMemcpy(dq.p, hq.data(), sizeof(double) * hq.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&dF.p, sizeof(double) * ne * nif));
    HIP_TRY(hipMalloc(&dG.p, sizeof(double) * 3 * nif));
    FluxArgs a{};
    a.thermo = static_cast<DevThermo *>(th.p); a.qf = static_cast<double *>(dq.p);
    a.F = static_cast<double *>(dF.p); a.godunov = static_cast<double *>(dG.p);
    a.nblk = 1; a.ncell = n; a.lv = lv; a.S = int(cols); a.P = cols; a.PF = nif;
    a.muscl = oo.flag_muscl; a.riemann = oo.select_riemann; a.spherical = oo.is_spherical == 1;
    HIP_TRY(launch_flux(a, m->host.kk, oo.flag_acu, oo.flag_weno, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    std::vector<double> hF(size_t(ne) * nif), hG(size_t(3) * nif);
    HIP_TRY(hipMemcpy(hF.data(), dF.p, sizeof(double) * hF.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hG.data(), dG.p, sizeof(double) * hG.size(), hipMemcpyDeviceToHost));
    untranspose(hF, nif, ne, Fplus);
    if (godunov) untranspose(hG, nif, 3, godunov);
    return FCFD_OK;
}

static int conv_host(const fcfd_mech *m, int n, const double *in, double *out, bool to_qc) {
    if (!m || !in || !out || n < 1) { set_error("conversion: invalid argument"); return FCFD_E_ARG; }
    const int ne = 3 + m->host.kk;
    DevBuf th, di, dout;
    HIP_TRY(hipMalloc(&th.p, sizeof(DevThermo)));
    HIP_TRY(hipMemcpy(th.p, &m->host.thermo, sizeof(DevThermo), hipMemcpyHostToDevice));
    auto h = transpose(in, n, ne);
    HIP_TRY(hipMalloc(&di.p, sizeof(double) * h.size()));
    HIP_TRY(hipMalloc(&dout.p, sizeof(double) * h.size()));
    HIP_TRY(hipMemcpy(di.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    const DevThermo *t = static_cast<DevThermo *>(th.p);
    if (to_qc) HIP_TRY(launch_qf_to_qc(t, m->host.kk, n, static_cast<double *>(di.p), static_cast<double *>(dout.p), n, nullptr));
    else HIP_TRY(launch_qc_to_qf(t, m->host.kk, n, static_cast<double *>(di.p), static_cast<double *>(dout.p), n, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(h.data(), dout.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    untranspose(h, n, ne, out);
    return FCFD_OK;
}

int fcfd_qf_to_qc_host(const fcfd_mech *m, int n, const double *qf, double *Qc) { return conv_host(m, n, qf, Qc, true); }
int fcfd_qc_to_qf_host(const fcfd_mech *m, int n, const double *Qc, double *qf) { return conv_host(m, n, Qc, qf, false); }

int fcfd_chem_step_host(const fcfd_mech *m, const fcfd_opts *o, int n, const double *dt,
                        const double *rho, double *yT, int *stats) {
    if (!m || !o || n < 1 || !dt || !rho || !yT) { set_error("fcfd_chem_step_host: invalid argument"); return FCFD_E_ARG; }
    const int kk = m->host.kk;
    DevBuf dr, dd, dT, dY, ds;
    HIP_TRY(hipMalloc(&dr.p, sizeof(double) * n));
    HIP_TRY(hipMalloc(&dd.p, sizeof(double) * n));
    HIP_TRY(hipMalloc(&dT.p, sizeof(double) * n));
    HIP_TRY(hipMalloc(&dY.p, sizeof(double) * n * kk));
    HIP_TRY(hipMalloc(&ds.p, sizeof(ChemCellStats) * n));
    std::vector<double> hT(n), hY(size_t(n) * kk);
    for (int c = 0; c < n; ++c) {
        hT[c] = yT[size_t(c) * (kk + 1) + kk];
        for (int k = 0; k < kk; ++k) hY[size_t(k) * n + c] = yT[size_t(c) * (kk + 1) + k];
    }
    HIP_TRY(hipMemcpy(dr.p, rho, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dd.p, dt, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dT.p, hT.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dY.p, hY.data(), sizeof(double) * hY.size(), hipMemcpyHostToDevice));
    int rc = fcfd_chem_batch(m, o, n, static_cast<double *>(dr.p), static_cast<double *>(dd.p),
                             static_cast<double *>(dT.p), static_cast<double *>(dY.p),
                             static_cast<int *>(ds.p), nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(hT.data(), dT.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hY.data(), dY.p, sizeof(double) * hY.size(), hipMemcpyDeviceToHost));
    if (stats) HIP_TRY(hipMemcpy(stats, ds.p, sizeof(ChemCellStats) * n, hipMemcpyDeviceToHost));
    for (int c = 0; c < n; ++c) {
        yT[size_t(c) * (kk + 1) + kk] = hT[c];
        for (int k = 0; k < kk; ++k) yT[size_t(c) * (kk + 1) + k] = hY[size_t(k) * n + c];
    }
    return FCFD_OK;
}

int fcfd_rates_host(const fcfd_mech *m, int n, const double *rho, const double *T, const double *Y,
                    double *wdot, double *cdot, double *ddot, double *qf, double *qr) {
    if (!m || n < 1 || !rho || !T || !Y || !wdot || !cdot || !ddot || !qf || !qr) {
        set_error("fcfd_rates_host: invalid argument");
        return FCFD_E_ARG;
    }
    if (int rc = ensure_dev_tables(m)) return rc;
    const int kk = m->host.kk, ii = m->host.ii;
    const size_t nk = size_t(n) * kk, ni = size_t(n) * ii;
    DevBuf dr, dT, dY, dw, dc, dd, dqf, dqr;
    HIP_TRY(hipMalloc(&dr.p, sizeof(double) * n));
    HIP_TRY(hipMalloc(&dT.p, sizeof(double) * n));
    HIP_TRY(hipMalloc(&dY.p, sizeof(double) * nk));
    HIP_TRY(hipMalloc(&dw.p, sizeof(double) * nk));
    HIP_TRY(hipMalloc(&dc.p, sizeof(double) * nk));
    HIP_TRY(hipMalloc(&dd.p, sizeof(double) * nk));
    HIP_TRY(hipMalloc(&dqf.p, sizeof(double) * ni));
    HIP_TRY(hipMalloc(&dqr.p, sizeof(double) * ni));
    HIP_TRY(hipMemcpy(dr.p, rho, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dT.p, T, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dY.p, Y, sizeof(double) * nk, hipMemcpyHostToDevice));
    auto d = [](DevBuf &b) { return static_cast<double *>(b.p); };
    HIP_TRY(launch_rates(m->d_thermo, static_cast<const DevKin *>(m->d_kin), kk, n, d(dr), d(dT), d(dY), d(dw),
                         d(dc), d(dd), d(dqf), d(dqr), nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(wdot, dw.p, sizeof(double) * nk, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cdot, dc.p, sizeof(double) * nk, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ddot, dd.p, sizeof(double) * nk, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(qf, dqf.p, sizeof(double) * ni, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(qr, dqr.p, sizeof(double) * ni, hipMemcpyDeviceToHost));
    return FCFD_OK;
}

int fcfd_recon_host(int n, const double *s, double *out) {
    if (!s || !out || n < 1) { set_error("fcfd_recon_host: invalid argument"); return FCFD_E_ARG; }
    DevBuf ds, dout;
    HIP_TRY(hipMalloc(&ds.p, sizeof(double) * 8 * n));
    HIP_TRY(hipMalloc(&dout.p, sizeof(double) * 18 * n));
    HIP_TRY(hipMemcpy(ds.p, s, sizeof(double) * 8 * n, hipMemcpyHostToDevice));
    HIP_TRY(launch_recon(n, static_cast<double *>(ds.p), static_cast<double *>(dout.p), nullptr));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, dout.p, sizeof(double) * 18 * n, hipMemcpyDeviceToHost));
    return FCFD_OK;
}

}  // extern "C"
