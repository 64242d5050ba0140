// hfx_api.hip -- C ABI (include/hfx.h) over the gfx950 kernels.
//
// Owns the device-resident tables of a basis (hfx_basis) and of one SCF job's
// Fock builds (hfx_fock).  Every HIP call is checked; errors come back as
// status codes with a thread-local message, never as exit().
#include <hip/hip_runtime.h>

#include <atomic>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../../include/hfx.h"
#include "devmem.hpp"
#include "hfx_launch.hpp"
#include "plan.hpp"

using namespace hfx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return fail(_e == hipErrorOutOfMemory ? HFX_ERR_NOMEM : HFX_ERR_DEVICE,           \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

// Process-wide device constants: Boys tables (FP64 and FP32) and the Hermite
// index table in constant memory.  Built on first use.
struct DeviceConsts {
    std::once_flag once;
    int status = 0;
    std::string err;
    double* boys_d = nullptr;
    float* boys_f = nullptr;
    HermTab* herm = nullptr;
};
DeviceConsts g_consts;

int ensure_consts() {
    std::call_once(g_consts.once, []() {
        std::vector<double> tab;
        boys_table(tab);
        std::vector<float> tabf(tab.begin(), tab.end());
        hipError_t e = hipMalloc(&g_consts.boys_d, tab.size() * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&g_consts.boys_f, tabf.size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(g_consts.boys_d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(g_consts.boys_f, tabf.data(), tabf.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&g_consts.herm, sizeof(HermTab));
        const HermTab ht = herm_tab();
        if (e == hipSuccess) e = hipMemcpy(g_consts.herm, &ht, sizeof(HermTab), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            g_consts.status = HFX_ERR_DEVICE;
            g_consts.err = std::string("device constant setup: ") + hipGetErrorString(e);
        }
    });
    if (g_consts.status) return fail(g_consts.status, g_consts.err);
    return 0;
}

// Per-job buffers come from the block cache (devmem.hpp): no device-wide
// synchronization when a job's objects are released.
template <class T>
int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) return 0;
    hipError_t e = hfx_detail::dev_alloc((void**)p, n * sizeof(T));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? HFX_ERR_NOMEM : HFX_ERR_DEVICE,
                    std::string("device allocation: ") + hipGetErrorString(e));
    return 0;
}

template <class T>
void dfree(T*& p) {
    if (p) hfx_detail::dev_free(p);
    p = nullptr;
}

}  // namespace

thread_local bool t_no_graphs = false;   // see fock_graphs_on()

// Streams besides the caller's that the class launches of one plan spread
// over: four launch queues in all, the number of hardware queues HIP gives a
// process by default.
constexpr int NSIDE = 3;

namespace hfx_detail {
void set_error(const std::string& msg) { g_err = msg; }
void disable_graphs_in_this_thread() { t_no_graphs = true; }
// FP64 Boys table on the device (built on first use); nullptr with the error set on failure.
const double* device_boys_table() { return ensure_consts() == 0 ? g_consts.boys_d : nullptr; }
}  // namespace hfx_detail

// Quartet lists of one plan, resident on the device, with per-class work counts.
struct DevLists {
    uint2* d = nullptr;
    size_t off[NQCLASS][2] = {};
    int cnt[NQCLASS][2] = {};            // work units per list (a quartet or a primitive slice of one)
    long long nquart[NQCLASS][2] = {};   // shell quartets per list
    int bucket[NQCLASS][2][NGROUP] = {};
    long long nprimq[NQCLASS][2] = {};   // sum over quartets of kept bra x ket primitive pairs
    bool valid = false;
    double threshold = -1.0;
    int mixed = 0;
    double split = 0.0;
};

struct hfx_basis {
    BasisPlan plan;
    PairHdr* d_hdr = nullptr;
    PrimPair<double>* d_ppd = nullptr;
    PrimPair<float>* d_ppf = nullptr;
    double* d_etd = nullptr;         // Hermite -> Cartesian matrices of the primitive pairs (family M),
    float* d_etf = nullptr;          // bra form and (etk) ket form
    double* d_etkd = nullptr;
    float* d_etkf = nullptr;
    double* d_fscale = nullptr;
    double* d_bounds = nullptr;      // function-pair Schwarz bounds (P)
    int2* d_fnpairs = nullptr;       // function pair id -> (i, j)
    int2* d_shells = nullptr;        // shell -> (first function, function count)
    double* d_qpair = nullptr;       // shell-pair Schwarz maxima by pair id (with the bounds)
    bool bounds_ready = false;
    std::vector<double> bounds;      // host copy
    std::vector<double> Qsh;         // shell-pair maxima
    DevLists unique_plan;            // cached FP64 plan of the last unique-ERI build
    hipStream_t stream = nullptr;
    // class launches of one plan fan out over the caller's stream and these
    // (fork / join through events), so the small launches of one build overlap
    hipStream_t side[NSIDE] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[NSIDE] = {};
};

namespace {

// Optional per-class timing of ERI launches (HIP events on the launch stream).
// A measurement hook for one job at a time: while it is on, the events and
// `used` flags belong to whichever build ran last; with it off (the default)
// launch_plan touches nothing but the atomic flag, so concurrent builds from
// several threads share no state here.
struct ClassTimer {
    std::atomic<bool> on{false};
    std::atomic<bool> serial{false};   // every list on the caller's stream: each class has the device to itself
    hipEvent_t ev[NQCLASS][2][2] = {};
    bool used[NQCLASS][2] = {};
    bool created = false;
};
ClassTimer g_timer;

}  // namespace

struct hfx_fock {
    hfx_basis* b = nullptr;
    hfx_fock_opts opts{};
    int mode = HFX_MODE_STORED;
    size_t bytes = 0;
    size_t nq = 0;
    double* d_eri = nullptr;                 // stored mode
    DevLists plan;                           // direct mode: all lists back to back
    double *d_Dj = nullptr, *d_Da = nullptr, *d_Db = nullptr;
    double *d_J = nullptr, *d_Ka = nullptr, *d_Kb = nullptr;
    // the build captured as a HIP graph for the last (densities, outputs, stream)
    hipGraphExec_t graph = nullptr;
    const void* graph_key[6] = {};
    // recorded after every build on the build's stream: destroy waits for it
    hipEvent_t ev_done = nullptr;
    // incremental direct builds: previous densities, density differences,
    // accumulated (symmetrised) J / K, shell-block maxima of |dP|
    double* d_inc = nullptr;
    // direct builds: the densities scaled by the function component scales
    // (Dj, Da, Db, each diag(fs) D diag(fs))
    double* d_scl = nullptr;
    // pinned staging of the host-matrix entry point (Pa, Pb in; J, Ka, Kb out)
    double* h_stage = nullptr;
    double* h_stage_dev = nullptr;
    bool have_prev = false;
    int prev_nspin = 0;
    int since_full = 0;
    // work-unit autotuner of direct builds: the first NTUNE builds each run a
    // full (non-incremental, eager) build with one candidate slice target,
    // timed with HIP events; the fastest plan is kept for the rest
    static constexpr int NTUNE = 3;
    bool tuning = false;
    bool tune_warm = true;           // the first build (kernel loading) is not timed
    int tune_i = 0;                  // candidate of the next / pending timed build
    bool tune_pending = false;
    long long tune_target[NTUNE] = {};
    float tune_ms[NTUNE] = {};
    long long plan_target = 0;       // slice target of the current plan
    hipEvent_t tune_ev[2] = {};
};

namespace {

size_t n_unique(int K) {
    const size_t P = (size_t)K * (K + 1) / 2;
    return P * (P + 1) / 2;
}

EpiArgs base_epi(const hfx_basis* b) {
    EpiArgs ep{};
    ep.K = b->plan.K;
    ep.bounds = b->d_bounds;
    ep.fscale = b->d_fscale;
    ep.nspin = 1;
    return ep;
}

// Screening threshold of the density-difference terms of incremental builds:
// HFX_INCREMENTAL_TAU in the environment when set, else the option, else
// 1e-3 x the Schwarz threshold (the skipped terms of one build then stay three
// orders below what the Schwarz screen itself drops).
double incremental_tau(const hfx_fock* f) {
    static const double env = [] {
        const char* e = std::getenv("HFX_INCREMENTAL_TAU");
        return e ? std::atof(e) : -1.0;
    }();
    return env >= 0.0 ? env : f->opts.incremental_tau > 0.0 ? f->opts.incremental_tau : 1e-3 * f->opts.threshold;
}

// Zeroes the J / K accumulators: one launch when they lie back to back (the
// context's own matrices do), else one per matrix.
hipError_t zero_jk(double* J, double* Ka, double* Kb, int nspin, size_t K2, hipStream_t st) {
    if (Ka == J + K2 && (nspin == 1 || Kb == Ka + K2)) return launch_zero(J, (size_t)(1 + nspin) * K2, st);
    hipError_t e = launch_zero(J, K2, st);
    if (e == hipSuccess) e = launch_zero(Ka, K2, st);
    if (e == hipSuccess && nspin == 2) e = launch_zero(Kb, K2, st);
    return e;
}

// Relative cost of one list: kept primitive quartets times the work of one
// (Hermite x Hermite plus Cartesian x Cartesian products of the class).
double list_cost(const DevLists& dl, int qc, int p) {
    const int pb = qc_bra(qc), pk = qc_ket(qc);
    const int la = pc_la(pb), lb = pc_lb(pb), lc = pc_la(pk), ld = pc_lb(pk);
    auto nc = [](int l) { return (l + 1) * (l + 2) / 2; };
    auto nh = [](int l) { return (l + 1) * (l + 2) * (l + 3) / 6; };
    return (double)dl.nprimq[qc][p] * (nh(la + lb) * nh(lc + ld) + nc(la) * nc(lb) * nc(lc) * nc(ld));
}

int launch_plan(const hfx_basis* b, const DevLists& dl, const EpiArgs& ep, hipStream_t st) {
    // Lists go, largest first, to the least loaded of the caller's stream and
    // the basis's side streams; the caller's stream then waits for the others.
    // Class timing brackets each list with events on the stream it runs on
    // (serial timing: all lists on the caller's stream).
    static const int nlaunch = [] {
        const char* e = std::getenv("HFX_LAUNCH_STREAMS");
        const int v = e ? std::atoi(e) : NSIDE + 1;
        return v < 1 ? 1 : v > NSIDE + 1 ? NSIDE + 1 : v;
    }();
    const bool timing = g_timer.on.load();
    const bool fan = b->ev_fork && nlaunch > 1 && !(timing && g_timer.serial.load());
    hipStream_t lane_st[NSIDE + 1] = {st};
    for (int i = 0; i < NSIDE; ++i) lane_st[i + 1] = b->side[i];
    int order[NQCLASS * 2], nl = 0;
    for (int l = 0; l < 2 * NQCLASS; ++l)
        if (dl.cnt[l / 2][l % 2]) order[nl++] = l;
    if (fan)
        std::stable_sort(order, order + nl, [&](int x, int y) {
            return list_cost(dl, x / 2, x % 2) > list_cost(dl, y / 2, y % 2);
        });
    double load[NSIDE + 1] = {};
    bool used_side[NSIDE] = {};
    if (fan) HIPCHK(hipEventRecord(b->ev_fork, st));
    if (timing)
        for (int qc = 0; qc < NQCLASS; ++qc) g_timer.used[qc][0] = g_timer.used[qc][1] = false;
    for (int k = 0; k < nl; ++k) {
        const int qc = order[k] / 2, p = order[k] % 2;
        int lane = 0;
        if (fan) {
            for (int i = 1; i < nlaunch; ++i)
                if (load[i] < load[lane]) lane = i;
            load[lane] += list_cost(dl, qc, p) + 1.0;
            if (lane > 0 && !used_side[lane - 1]) {
                HIPCHK(hipStreamWaitEvent(lane_st[lane], b->ev_fork, 0));
                used_side[lane - 1] = true;
            }
        }
        hipStream_t ls = lane_st[lane];
        {
            if (timing) HIPCHK(hipEventRecord(g_timer.ev[qc][p][0], ls));
            size_t off = dl.off[qc][p];
            for (int g = 0; g < NGROUP; ++g) {
                const int n = dl.bucket[qc][p][g];
                if (n == 0) continue;
                ClassLaunch L;
                L.list = dl.d + off;
                L.n = n;
                L.hdr = b->d_hdr;
                L.prim = p ? (const void*)b->d_ppf : (const void*)b->d_ppd;
                L.boys = p ? (const void*)g_consts.boys_f : (const void*)g_consts.boys_d;
                L.herm = g_consts.herm;
                L.etab = p ? (const void*)b->d_etf : (const void*)b->d_etd;
                L.etabk = p ? (const void*)b->d_etkf : (const void*)b->d_etkd;
                L.prim64 = b->d_ppd;
                L.boys64 = g_consts.boys_d;
                L.G = GROUP_SIZE[g];
                L.ep = ep;
                HIPCHK(launch_eri_class(qc, p == 1, L, ls));
                off += n;
            }
            if (timing) {
                HIPCHK(hipEventRecord(g_timer.ev[qc][p][1], ls));
                g_timer.used[qc][p] = true;
            }
        }
    }
    for (int i = 0; i < NSIDE; ++i)
        if (used_side[i]) {
            HIPCHK(hipEventRecord(b->ev_join[i], b->side[i]));
            HIPCHK(hipStreamWaitEvent(st, b->ev_join[i], 0));
        }
    return 0;
}

// Wavefront target of the primitive-slice policy (expand_units):
// HFX_SLICE_TARGET in the environment when set (0 disables slicing), else
// SLICE_TARGET.  Read once.  Measured on benzene/cc-pVDZ and on small direct
// builds (NaCl/cc-pVDZ): slicing family-B launches pays in every plan;
// slicing the register-resident family-A lane groups pays in J/K plans
// (launches of few quartets) but slowed the unique-array build, where the
// strided primitive loop rebuilds the bra tables more often and the slices
// add through atomics, so unique-array plans slice family B only.
long long slice_target() {
    static const long long t = [] {
        const char* e = std::getenv("HFX_SLICE_TARGET");
        return e ? std::atoll(e) : SLICE_TARGET;
    }();
    return t;
}

struct SliceTargets {
    long long fam_a, fam_b;
};

// Work-unit autotuning of direct builds (hfx_set_autotune; HFX_AUTOTUNE=1 in
// the environment turns it on).  Off by default: measured on the n-octane and
// (H2O)4 direct SCFs (~30 builds), three full, re-planned tuning builds cost
// more than the plan they find saves.
std::atomic<int> g_autotune{-1};
bool autotune_on() {
    int v = g_autotune.load();
    if (v < 0) {
        const char* e = std::getenv("HFX_AUTOTUNE");
        v = (e && e[0] == '1') ? 1 : 0;
        g_autotune.store(v);
    }
    return v == 1;
}
SliceTargets jk_slices() { return {slice_target(), slice_target()}; }
// Direct J/K plans: a quarter of the resident-unit target (fewer, longer
// slices; with group_bucket_jk measured n-octane / cc-pVDZ 10.1 -> 10.0 ms and
// (H2O)4 / cc-pVTZ 23.9 -> 23.8 ms against half).
SliceTargets jk_plan_slices() { return {slice_target() / 4, slice_target() / 4}; }
SliceTargets unique_slices() { return {0, slice_target()}; }

// Uploads the work units of the quartet lists (expand_units with the given
// wavefront targets per family; 0 for the Schwarz diagonal, whose epilogue
// needs whole quartets) into one device buffer; fills offsets / counts /
// buckets in units.
// Expands each (class, precision) list into work units (lists in parallel on
// large plans) and uploads them in list order through pinned staging.
int upload_lists(const hfx_basis* b, const QuartetLists& ql, SliceTargets target, DevLists& dl) {
    constexpr int NL = 2 * NQCLASS;
    const int T = plan_threads(ql.total());
    std::vector<std::vector<uint2>> units(NL);
    parallel_tasks(NL, T, [&](int l) {
        const int qc = l / 2, p = l % 2;
        const std::vector<uint2>& v = ql.list[qc][p];
        std::vector<uint2>& u = units[l];
        u.reserve(v.size());
        size_t pos = 0;
        for (int g = 0; g < NGROUP; ++g) {
            const size_t before = u.size();
            expand_units(b->plan, qc_family_a(qc), GROUP_SIZE[g], v.data() + pos, ql.bucket[qc][p][g],
                         qc_family_a(qc) ? target.fam_a : target.fam_b, u, qc_famb_kb(qc));
            pos += ql.bucket[qc][p][g];
            dl.bucket[qc][p][g] = (int)(u.size() - before);
        }
        long long s = 0;
        for (const uint2& e : v) s += (long long)b->plan.hdr[e.x].nprim * b->plan.hdr[e.y].nprim;
        dl.nprimq[qc][p] = s;
    });
    size_t total = 0;
    for (int l = 0; l < NL; ++l) {
        const int qc = l / 2, p = l % 2;
        dl.off[qc][p] = total;
        dl.cnt[qc][p] = (int)units[l].size();
        dl.nquart[qc][p] = (long long)ql.list[qc][p].size();
        total += units[l].size();
    }
    if (int rc = dalloc(&dl.d, total)) return rc;
    if (total == 0) return 0;
    void* h = nullptr;
    HIPCHK(hfx_detail::host_alloc(&h, total * sizeof(uint2)));
    uint2* hs = static_cast<uint2*>(h);
    parallel_tasks(NL, T, [&](int l) {
        if (!units[l].empty())
            std::memcpy(hs + dl.off[l / 2][l % 2], units[l].data(), units[l].size() * sizeof(uint2));
    });
    hipError_t e = hipMemcpyAsync(dl.d, hs, total * sizeof(uint2), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    hfx_detail::host_free(h);
    if (e != hipSuccess) return fail(HFX_ERR_DEVICE, std::string("quartet list upload: ") + hipGetErrorString(e));
    return 0;
}

// Unique-array plans whose total primitive work is at most this many primitive
// quartets also slice their family-A quartets and take the small-plan lane
// groups (group_bucket_small): such builds (the serving stream's molecules)
// are set by their longest quartet, while on large ones (benzene / cc-pVDZ:
// 1.1e8) family-A slices cost more than they balance.
constexpr long long SMALL_PLAN_WORK = 1LL << 25;

bool small_plan(const BasisPlan& bp, const QuartetLists& ql) {
    long long s = 0;
    for (int qc = 0; qc < NQCLASS; ++qc)
        for (int p = 0; p < 2; ++p)
            for (const uint2& e : ql.list[qc][p]) {
                s += (long long)bp.hdr[e.x].nprim * bp.hdr[e.y].nprim;
                if (s > SMALL_PLAN_WORK) return false;
            }
    return true;
}

// Plans the screened quartets of a basis and uploads them (cached per basis
// for the FP64 unique build).  jk_order: the J/K ordering of plan_quartets
// (direct builds).
int make_plan(hfx_basis* b, double thr, int mixed, double split, SliceTargets target, DevLists& dl,
              bool jk_order = false) {
    dfree(dl.d);
    dl.valid = false;
    // host work in the C ABI's call: failures come back as status codes
    try {
        QuartetLists ql;
        plan_quartets(b->plan, b->Qsh, thr, mixed != 0, split, ql, 0, jk_order);
        if (jk_order) rebucket(b->plan, ql, group_bucket_jk);
        if (target.fam_a == 0 && small_plan(b->plan, ql)) {
            target.fam_a = slice_target();
            rebucket(b->plan, ql, group_bucket_small);
        }
        if (int rc = upload_lists(b, ql, target, dl)) return rc;
    } catch (const std::bad_alloc&) {
        dfree(dl.d);
        return fail(HFX_ERR_NOMEM, "host memory for the quartet plan");
    } catch (const std::exception& e) {
        dfree(dl.d);
        return fail(HFX_ERR_DEVICE, std::string("quartet plan: ") + e.what());
    }
    dl.threshold = thr;
    dl.mixed = mixed;
    dl.split = split;
    dl.valid = true;
    return 0;
}

// Waits for the streams a plan's launches may still run on (the caller's and
// the basis' side streams) before their blocks go back to the cache after a
// failure part-way through a launch sequence.
void drain(const hfx_basis* b, hipStream_t st) {
    if (st) (void)hipStreamSynchronize(st);
    if (b->stream && b->stream != st) (void)hipStreamSynchronize(b->stream);
    for (int i = 0; i < NSIDE; ++i)
        if (b->side[i]) (void)hipStreamSynchronize(b->side[i]);
}

int compute_bounds_impl(hfx_basis* b);

// Host work reached from the C ABI turns allocation failures and other
// exceptions into status codes instead of terminating the process.
int compute_bounds(hfx_basis* b) {
    if (b->bounds_ready) return 0;
    try {
        return compute_bounds_impl(b);
    } catch (const std::bad_alloc&) {
        drain(b, b->stream);
        return fail(HFX_ERR_NOMEM, "host memory for the Schwarz bounds");
    } catch (const std::exception& e) {
        drain(b, b->stream);
        return fail(HFX_ERR_DEVICE, std::string("Schwarz bounds: ") + e.what());
    }
}

int compute_bounds_impl(hfx_basis* b) {
    const int P = b->plan.K * (b->plan.K + 1) / 2;
    QuartetLists ql;
    std::vector<uint2> diag[NQCLASS];
    plan_diag(b->plan, diag);
    for (int c = 0; c < NQCLASS; ++c) {
        ql.list[c][0] = diag[c];
        const bool fa = qc_family_a(c);
        for (const uint2& e : diag[c])
            ++ql.bucket[c][0][fa ? group_bucket_small(b->plan.hdr[e.x].nprim * b->plan.hdr[e.y].nprim) : NGROUP - 1];
        // diag lists are built in pair order: sort them like plan_quartets does
        std::stable_sort(ql.list[c][0].begin(), ql.list[c][0].end(), [&](const uint2& x, const uint2& y) {
            return b->plan.hdr[x.x].nprim * b->plan.hdr[x.y].nprim > b->plan.hdr[y.x].nprim * b->plan.hdr[y.y].nprim;
        });
    }
    // Quartets longer than a resident slot's share are cut into primitive
    // slices (a 12-primitive s shell alone gives 20,736 primitive quartets):
    // the epilogue leaves (mu nu|mu nu), slices adding into the zeroed
    // output, and the square roots follow in one pass.
    DevLists dl;
    if (int rc = upload_lists(b, ql, jk_slices(), dl)) return rc;
    HIPCHK(launch_zero(b->d_bounds, P, b->stream));
    EpiArgs ep = base_epi(b);
    ep.mode = EPI_DIAG;
    ep.out = b->d_bounds;
    int rc = launch_plan(b, dl, ep, b->stream);
    if (rc == 0) {
        const hipError_t e = launch_sqrt_clamp(b->d_bounds, P, b->stream);
        if (e != hipSuccess) rc = fail(HFX_ERR_DEVICE, std::string("schwarz bounds: ") + hipGetErrorString(e));
    }
    if (rc == 0) {
        b->bounds.resize(P);
        hipError_t e = hipMemcpyAsync(b->bounds.data(), b->d_bounds, P * sizeof(double), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) rc = fail(HFX_ERR_DEVICE, std::string("schwarz bounds: ") + hipGetErrorString(e));
    }
    if (rc) drain(b, b->stream);
    dfree(dl.d);
    if (rc) return rc;
    shell_pair_bounds(b->plan, b->bounds.data(), b->Qsh);
    if (int rc = dalloc(&b->d_qpair, b->Qsh.size())) return rc;
    if (!b->Qsh.empty()) {
        HIPCHK(hipMemcpyAsync(b->d_qpair, b->Qsh.data(), b->Qsh.size() * sizeof(double), hipMemcpyHostToDevice,
                              b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    b->bounds_ready = true;
    return 0;
}

// The unique array of a plan (FP64, or mixed: FP32 for its small-bound lists).
int eri_unique_plan(hfx_basis* b, const DevLists& dl, double threshold, double* d_out, hipStream_t st) {
    HIPCHK(launch_zero(d_out, n_unique(b->plan.K), st));
    EpiArgs ep = base_epi(b);
    ep.mode = EPI_UNIQUE;
    ep.threshold = threshold;
    ep.out = d_out;
    if (int rc = launch_plan(b, dl, ep, st)) return rc;
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int eri_unique_into(hfx_basis* b, double threshold, double* d_out, hipStream_t st) {
    if (int rc = compute_bounds(b)) return rc;
    DevLists& dl = b->unique_plan;
    if (!dl.valid || dl.threshold != threshold)
        if (int rc = make_plan(b, threshold, 0, 0.0, unique_slices(), dl)) return rc;
    return eri_unique_plan(b, dl, threshold, d_out, st);
}

// J/K from a device unique array into device accumulators, then symmetrised.
int jk_from_unique(hfx_basis* b, const double* d_eri, const double* d_Dj, const double* d_Da, const double* d_Db,
                   double* d_J, double* d_Ka, double* d_Kb, int nspin, hipStream_t st) {
    const int K = b->plan.K;
    const size_t K2 = (size_t)K * K;
    HIPCHK(zero_jk(d_J, d_Ka, d_Kb, nspin, K2, st));
    HIPCHK(launch_jk_rows(d_eri, b->d_fnpairs, K, nspin, d_Dj, d_Da, d_Db, d_J, d_Ka, d_Kb, st));
    HIPCHK(launch_symmetrize_n(d_J, d_Ka, d_Kb, 1 + nspin, K, st));
    return 0;
}

}  // namespace

extern "C" {

const char* hfx_last_error(void) { return g_err.c_str(); }

int hfx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int hfx_basis_create(int K, const double* xyz, const int* lmn, const int* poff, const double* alpha,
                     const double* coeff, const double* norm, hfx_basis** out) {
    if (!out || !xyz || !lmn || !poff || !alpha || !coeff || !norm) return fail(HFX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (int rc = ensure_consts()) return rc;
    hfx_basis* b = new hfx_basis();
    std::string err;
    int rc = 0;
    try {
        rc = build_plan(K, xyz, lmn, poff, alpha, coeff, norm, b->plan, err);
    } catch (const std::bad_alloc&) {
        delete b;
        return fail(HFX_ERR_NOMEM, "host memory for the basis tables");
    }
    if (rc) {
        delete b;
        return fail(rc == -4 ? HFX_ERR_UNSUPPORTED : HFX_ERR_INVALID, err);
    }
    const BasisPlan& p = b->plan;
    const int P = K * (K + 1) / 2;
    rc = dalloc(&b->d_hdr, p.hdr.size());
    if (!rc) rc = dalloc(&b->d_ppd, std::max<size_t>(1, p.ppd.size()));
    if (!rc) rc = dalloc(&b->d_ppf, std::max<size_t>(1, p.ppf.size()));
    if (!rc) rc = dalloc(&b->d_etd, std::max<size_t>(1, p.etd.size()));
    if (!rc) rc = dalloc(&b->d_etf, std::max<size_t>(1, p.etf.size()));
    if (!rc) rc = dalloc(&b->d_etkd, std::max<size_t>(1, p.etkd.size()));
    if (!rc) rc = dalloc(&b->d_etkf, std::max<size_t>(1, p.etkf.size()));
    if (!rc) rc = dalloc(&b->d_fscale, (size_t)K);
    if (!rc) rc = dalloc(&b->d_bounds, (size_t)P);
    if (!rc) rc = dalloc(&b->d_fnpairs, (size_t)P);
    std::vector<int2> shells(p.nshell);
    for (int i = 0; i < p.nshell; ++i) shells[i] = make_int2(p.shells[i].fn0, ncart(p.shells[i].L));
    if (!rc) rc = dalloc(&b->d_shells, shells.size());
    hipError_t e = hipSuccess;
    if (!rc) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int i = 0; i < NSIDE; ++i) {
        if (!rc && e == hipSuccess) e = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking);
        if (!rc && e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming);
    }
    if (!rc && e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    auto h2d = [&](void* dst, const void* src, size_t n) {
        if (!rc && e == hipSuccess && n) e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, b->stream);
    };
    h2d(b->d_hdr, p.hdr.data(), p.hdr.size() * sizeof(PairHdr));
    h2d(b->d_ppd, p.ppd.data(), p.ppd.size() * sizeof(PrimPair<double>));
    h2d(b->d_ppf, p.ppf.data(), p.ppf.size() * sizeof(PrimPair<float>));
    h2d(b->d_etd, p.etd.data(), p.etd.size() * sizeof(double));
    h2d(b->d_etf, p.etf.data(), p.etf.size() * sizeof(float));
    h2d(b->d_etkd, p.etkd.data(), p.etkd.size() * sizeof(double));
    h2d(b->d_etkf, p.etkf.data(), p.etkf.size() * sizeof(float));
    h2d(b->d_fscale, p.fscale.data(), K * sizeof(double));
    h2d(b->d_fnpairs, p.fnpairs.data(), P * sizeof(int2));
    h2d(b->d_shells, shells.data(), shells.size() * sizeof(int2));
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (!rc && e != hipSuccess) rc = fail(HFX_ERR_DEVICE, std::string("basis upload: ") + hipGetErrorString(e));
    if (rc) {
        hfx_basis_destroy(b);
        return rc;
    }
    *out = b;
    return 0;
}

void hfx_basis_destroy(hfx_basis* b) {
    if (!b) return;
    // the basis' own streams: its buffers are back in the cache only once idle
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (int i = 0; i < NSIDE; ++i)
        if (b->side[i]) (void)hipStreamSynchronize(b->side[i]);
    dfree(b->d_hdr);
    dfree(b->d_ppd);
    dfree(b->d_ppf);
    dfree(b->d_etd);
    dfree(b->d_etf);
    dfree(b->d_etkd);
    dfree(b->d_etkf);
    dfree(b->d_fscale);
    dfree(b->d_bounds);
    dfree(b->d_fnpairs);
    dfree(b->d_shells);
    dfree(b->d_qpair);
    dfree(b->unique_plan.d);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (int i = 0; i < NSIDE; ++i) {
        if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
        if (b->ev_join[i]) (void)hipEventDestroy(b->ev_join[i]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    delete b;
}

int hfx_basis_info(const hfx_basis* b, int* nshell, int* npair, int* nprimpair) {
    if (!b) return fail(HFX_ERR_INVALID, "null basis");
    if (nshell) *nshell = b->plan.nshell;
    if (npair) *npair = b->plan.npair;
    if (nprimpair) *nprimpair = (int)b->plan.ppd.size();
    return 0;
}

int hfx_schwarz_bounds(hfx_basis* b, double* bounds) {
    if (!b || !bounds) return fail(HFX_ERR_INVALID, "null argument");
    if (int rc = compute_bounds(b)) return rc;
    std::memcpy(bounds, b->bounds.data(), b->bounds.size() * sizeof(double));
    return 0;
}

int hfx_eri_unique_device(hfx_basis* b, double threshold, double* d_out, size_t Q, void* stream) {
    if (!b || !d_out) return fail(HFX_ERR_INVALID, "null argument");
    if (Q != n_unique(b->plan.K)) return fail(HFX_ERR_INVALID, "Q must equal P(P+1)/2 with P = K(K+1)/2");
    return eri_unique_into(b, threshold, d_out, stream ? (hipStream_t)stream : b->stream);
}

namespace {

// Device-to-host copy of a large array into the caller's (pageable) memory:
// the DMA fills two pinned chunks in turn while host threads move the chunk
// that landed before into place, so the copy engine and the host copies
// overlap.  A chunk is refilled only after its host copy has returned.
// Arrays up to one chunk take a single plain copy.
int copy_to_host(double* out, const double* d, size_t n, hipStream_t st) {
    constexpr size_t CH = size_t(32) << 20;
    const size_t bytes = n * sizeof(double);
    if (bytes <= CH) {
        HIPCHK(hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    void* h = nullptr;
    HIPCHK(hfx_detail::host_alloc(&h, 2 * CH));
    char* stage[2] = {static_cast<char*>(h), static_cast<char*>(h) + CH};
    hipEvent_t ev[2] = {};
    hipError_t e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    const size_t nch = (bytes + CH - 1) / CH;
    auto len = [&](size_t c) { return std::min(CH, bytes - c * CH); };
    const char* src = reinterpret_cast<const char*>(d);
    char* dst = reinterpret_cast<char*>(out);
    constexpr int NCOPY = 8;   // host threads per chunk copy
    for (size_t c = 0; c <= nch && e == hipSuccess; ++c) {
        if (c < nch) {
            e = hipMemcpyAsync(stage[c % 2], src + c * CH, len(c), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipEventRecord(ev[c % 2], st);
        }
        if (c >= 1 && e == hipSuccess) {
            const size_t p = c - 1, lp = len(p);
            e = hipEventSynchronize(ev[p % 2]);
            if (e != hipSuccess) break;
            const size_t slice = (lp + NCOPY - 1) / NCOPY;
            try {
                parallel_tasks(NCOPY, NCOPY, [&](int t) {
                    const size_t o = (size_t)t * slice;
                    if (o < lp) std::memcpy(dst + p * CH + o, stage[p % 2] + o, std::min(slice, lp - o));
                });
            } catch (...) {
                for (size_t o = 0; o < lp; o += slice)   // no threads: copy on this one
                    std::memcpy(dst + p * CH + o, stage[p % 2] + o, std::min(slice, lp - o));
            }
        }
    }
    if (e != hipSuccess) (void)hipStreamSynchronize(st);   // no copy may still target the staging
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    hfx_detail::host_free(h);
    if (e != hipSuccess) return fail(HFX_ERR_DEVICE, std::string("D2H: ") + hipGetErrorString(e));
    return 0;
}

}  // namespace

int hfx_eri_unique(hfx_basis* b, double threshold, double* out, size_t Q) {
    if (!b || !out) return fail(HFX_ERR_INVALID, "null argument");
    if (Q != n_unique(b->plan.K)) return fail(HFX_ERR_INVALID, "Q must equal P(P+1)/2 with P = K(K+1)/2");
    double* d = nullptr;
    if (int rc = dalloc(&d, Q)) return rc;
    int rc = eri_unique_into(b, threshold, d, b->stream);
    if (rc == 0) rc = copy_to_host(out, d, Q, b->stream);
    if (rc) drain(b, b->stream);
    dfree(d);
    return rc;
}

int hfx_coulomb_exchange(hfx_basis* b, const double* d_eri, const double* Pa, const double* Pb, double* J,
                         double* Ka, double* Kb) {
    if (!b || !d_eri || !Pa) return fail(HFX_ERR_INVALID, "null argument");
    const int K = b->plan.K;
    const size_t K2 = (size_t)K * K;
    const int nspin = Pb ? 2 : 1;
    std::vector<double> Dj(Pa, Pa + K2);
    if (Pb)
        for (size_t i = 0; i < K2; ++i) Dj[i] += Pb[i];
    double* buf = nullptr;
    if (int rc = dalloc(&buf, 6 * K2)) return rc;
    double *dDj = buf, *dDa = buf + K2, *dDb = buf + 2 * K2, *dJ = buf + 3 * K2, *dKa = buf + 4 * K2, *dKb = buf + 5 * K2;
    int rc = 0;
    hipError_t e = hipMemcpyAsync(dDj, Dj.data(), K2 * sizeof(double), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dDa, Pa, K2 * sizeof(double), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess && Pb) e = hipMemcpyAsync(dDb, Pb, K2 * sizeof(double), hipMemcpyHostToDevice, b->stream);
    if (e != hipSuccess) rc = fail(HFX_ERR_DEVICE, std::string("H2D: ") + hipGetErrorString(e));
    if (!rc) rc = jk_from_unique(b, d_eri, dDj, dDa, dDb, dJ, dKa, dKb, nspin, b->stream);
    if (!rc) {
        if (J) e = hipMemcpyAsync(J, dJ, K2 * sizeof(double), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess && Ka) e = hipMemcpyAsync(Ka, dKa, K2 * sizeof(double), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess && Kb && Pb) e = hipMemcpyAsync(Kb, dKb, K2 * sizeof(double), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) rc = fail(HFX_ERR_DEVICE, std::string("J/K: ") + hipGetErrorString(e));
    }
    if (rc) drain(b, b->stream);
    dfree(buf);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free J/K
// Scratch of hfx_coulomb_exchange_noalloc: the J density Pa + Pb of a UHF
// build (K x K doubles, K <= HFX_NOALLOC_MAX_K).  Each device holds its own
// copy for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns on that device.
namespace {

constexpr int NOALLOC_MAX_K = HFX_NOALLOC_MAX_K;
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[(size_t)NOALLOC_MAX_K * NOALLOC_MAX_K];

std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up once per device with
// that device current.
int noalloc_scratch(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(HFX_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HFX_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

int hfx_coulomb_exchange_noalloc(hfx_basis* b, const double* d_eri, const double* d_Pa, const double* d_Pb,
                                 double* d_J, double* d_Ka, double* d_Kb, void* stream) {
    if (!b || !d_eri || !d_Pa || !d_J || !d_Ka) return fail(HFX_ERR_INVALID, "null argument");
    if (d_Pb && !d_Kb) return fail(HFX_ERR_INVALID, "UHF build needs Kb");
    const int K = b->plan.K;
    const int nspin = d_Pb ? 2 : 1;
    if (nspin == 2 && K > NOALLOC_MAX_K)
        return fail(HFX_ERR_UNSUPPORTED, "hfx_coulomb_exchange_noalloc: K = " + std::to_string(K) +
                                             " exceeds the scratch maximum of " + std::to_string(NOALLOC_MAX_K) +
                                             " functions for UHF");
    hipStream_t st = stream ? (hipStream_t)stream : b->stream;
    double* dj = nullptr;
    if (int rc = noalloc_scratch(&dj)) return rc;
    const size_t K2 = (size_t)K * K;
    const double* Dj = d_Pa;
    if (nspin == 2) {   // every call rewrites the K x K part of the scratch it reads
        HIPCHK(launch_add(dj, d_Pa, d_Pb, K2, st));
        Dj = dj;
    }
    HIPCHK(zero_jk(d_J, d_Ka, d_Kb, nspin, K2, st));
    HIPCHK(launch_jk_rows(d_eri, b->d_fnpairs, K, nspin, Dj, d_Pa, d_Pb, d_J, d_Ka, d_Kb, st));
    HIPCHK(launch_symmetrize_n(d_J, d_Ka, d_Kb, 1 + nspin, K, st));
    return 0;
}

int hfx_repulsion_matrix(hfx_basis* b, const double* d_eri, const double* P, double* G) {
    if (!b || !G) return fail(HFX_ERR_INVALID, "null argument");
    const int K = b->plan.K;
    std::vector<double> J((size_t)K * K), Kx((size_t)K * K);
    if (int rc = hfx_coulomb_exchange(b, d_eri, P, nullptr, J.data(), Kx.data(), nullptr)) return rc;
    for (size_t i = 0; i < J.size(); ++i) G[i] = J[i] - 0.5 * Kx[i];
    return 0;
}

int hfx_fock_create(hfx_basis* b, const hfx_fock_opts* opts, hfx_fock** out) {
    if (!b || !opts || !out) return fail(HFX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (opts->mode < 0 || opts->mode > 2 || opts->precision < 0 || opts->precision > 1)
        return fail(HFX_ERR_INVALID, "invalid mode or precision");
    if (int rc = compute_bounds(b)) return rc;
    const int K = b->plan.K;
    const size_t K2 = (size_t)K * K;
    const size_t Q = n_unique(K);
    hfx_fock* f = new hfx_fock();
    f->b = b;
    f->opts = *opts;
    if (f->opts.mixed_split <= 0.0) f->opts.mixed_split = 1e-5;
    const size_t mats = 6 * K2 * sizeof(double);
    const size_t stored_bytes = Q * sizeof(double) + mats;
    const int mixed = opts->precision == HFX_PREC_MIXED;
    int mode = opts->mode;
    if (mode == HFX_MODE_AUTO)
        mode = (opts->hbm_budget == 0 || stored_bytes <= opts->hbm_budget) ? HFX_MODE_STORED : HFX_MODE_DIRECT;
    // direct: the J/K plan; stored mixed: the plan of its unique array (FP32
    // for the small-bound quartets); stored FP64 builds from the basis' cached
    // unique plan (made here if missing) and keeps only its counts
    int prc = 0;
    if (mode == HFX_MODE_STORED && !mixed) {
        DevLists& u = b->unique_plan;
        if (!u.valid || u.threshold != opts->threshold)
            prc = make_plan(b, opts->threshold, 0, 0.0, unique_slices(), u);
        if (!prc) {
            f->plan = u;
            f->plan.d = nullptr;   // the lists stay the basis'
        }
    } else {
        prc = make_plan(b, opts->threshold, mixed, f->opts.mixed_split,
                        mode == HFX_MODE_STORED ? unique_slices() : jk_plan_slices(), f->plan,
                        mode != HFX_MODE_STORED);
    }
    if (prc) {
        hfx_fock_destroy(f);
        return prc;
    }
    for (int qc = 0; qc < NQCLASS; ++qc) f->nq += f->plan.nquart[qc][0] + f->plan.nquart[qc][1];
    const size_t direct_bytes = f->nq * sizeof(uint2) + mats + 3 * K2 * sizeof(double);
    f->mode = mode;
    f->bytes = mode == HFX_MODE_STORED ? stored_bytes : direct_bytes;
    if (opts->hbm_budget && f->bytes > opts->hbm_budget) {
        hfx_fock_destroy(f);
        return fail(HFX_ERR_NOMEM, "job does not fit its HBM budget");
    }
    double* buf = nullptr;
    int rc = dalloc(&buf, 6 * K2);
    if (!rc && hipEventCreateWithFlags(&f->ev_done, hipEventDisableTiming) != hipSuccess)
        rc = fail(HFX_ERR_DEVICE, "hipEventCreate");
    if (!rc && opts->incremental && mode == HFX_MODE_DIRECT)
        rc = dalloc(&f->d_inc, 8 * K2 + (size_t)b->plan.nshell * b->plan.nshell);
    if (!rc && mode == HFX_MODE_DIRECT) rc = dalloc(&f->d_scl, 3 * K2);
    f->plan_target = jk_plan_slices().fam_a;   // the direct plan's target (stored plans do not re-plan)
    if (!rc && mode == HFX_MODE_DIRECT && autotune_on() && f->plan_target > 0) {
        f->tune_target[0] = f->plan_target;   // the plan just made
        f->tune_target[1] = std::max<long long>(1, f->plan_target / 4);
        f->tune_target[2] = f->plan_target * 4;
        if (hipEventCreate(&f->tune_ev[0]) == hipSuccess && hipEventCreate(&f->tune_ev[1]) == hipSuccess)
            f->tuning = true;
    }
    if (!rc) {
        f->d_Dj = buf;
        f->d_Da = buf + K2;
        f->d_Db = buf + 2 * K2;
        f->d_J = buf + 3 * K2;
        f->d_Ka = buf + 4 * K2;
        f->d_Kb = buf + 5 * K2;
    }
    if (!rc && mode == HFX_MODE_STORED) {
        rc = dalloc(&f->d_eri, Q);
        if (!rc) rc = mixed ? eri_unique_plan(b, f->plan, opts->threshold, f->d_eri, b->stream)
                            : eri_unique_into(b, opts->threshold, f->d_eri, b->stream);
        dfree(f->plan.d);
    }
    if (rc) {
        drain(b, b->stream);
        hfx_fock_destroy(f);
        return rc;
    }
    *out = f;
    return 0;
}

void hfx_fock_destroy(hfx_fock* f) {
    if (!f) return;
    if (f->ev_done) {
        (void)hipEventSynchronize(f->ev_done);
        (void)hipEventDestroy(f->ev_done);
    }
    if (f->b && f->b->stream) (void)hipStreamSynchronize(f->b->stream);
    if (f->graph) (void)hipGraphExecDestroy(f->graph);
    for (hipEvent_t e : f->tune_ev)
        if (e) (void)hipEventDestroy(e);
    dfree(f->d_eri);
    dfree(f->d_inc);
    dfree(f->d_scl);
    if (f->h_stage) hfx_detail::host_free(f->h_stage);
    dfree(f->plan.d);
    dfree(f->d_Dj);   // one allocation for all six matrices
    delete f;
}

int hfx_fock_info(const hfx_fock* f, int* mode, size_t* device_bytes, size_t* n_quartets) {
    if (!f) return fail(HFX_ERR_INVALID, "null fock");
    if (mode) *mode = f->mode;
    if (device_bytes) *device_bytes = f->bytes;
    if (n_quartets) *n_quartets = f->nq;
    return 0;
}

int hfx_set_autotune(int on) {
    g_autotune.store(on ? 1 : 0);
    return 0;
}

int hfx_fock_autotune_info(const hfx_fock* f, int* done, long long* target, long long* targets, float* ms) {
    if (!f) return fail(HFX_ERR_INVALID, "null fock");
    if (done) *done = f->tune_target[0] != 0 && !f->tuning ? 1 : 0;
    if (target) *target = f->plan_target;
    for (int i = 0; i < hfx_fock::NTUNE; ++i) {
        if (targets) targets[i] = f->tune_target[i];
        if (ms) ms[i] = f->tune_ms[i];
    }
    return 0;
}

int hfx_fock_class_counts(const hfx_fock* f, long long* n64, long long* n32) {
    if (!f) return fail(HFX_ERR_INVALID, "null fock");
    for (int qc = 0; qc < NQCLASS; ++qc) {
        if (n64) n64[qc] = f->plan.nquart[qc][0];
        if (n32) n32[qc] = f->plan.nquart[qc][1];
    }
    return 0;
}

namespace {

// Graph replay of the Fock build: on by default, HFX_FOCK_GRAPH=0 in the
// environment or hfx_set_fock_graphs(0) turns it off.  Threads that run
// concurrently with other threads' synchronous HIP calls (the server's
// workers) turn it off for themselves: a stream capture in flight makes the
// runtime reject legacy-stream operations issued by other threads.
std::atomic<int> g_fock_graphs{-1};
bool fock_graphs_on() {
    if (t_no_graphs) return false;
    int v = g_fock_graphs.load();
    if (v < 0) {
        const char* e = std::getenv("HFX_FOCK_GRAPH");
        v = (e && std::atoi(e) == 0) ? 0 : 1;
        g_fock_graphs.store(v);
    }
    return v == 1;
}

// Enqueues one build on `st` (eager, or recorded under stream capture): only
// kernel launches, no memset / memcpy nodes.
// Direct J/K from densities (Dj for J; Da, Db for K) into zeroed, then
// symmetrised accumulators; dpmax != nullptr screens units by the density
// difference bound.  The J/K epilogues work in the basis of the unscaled
// Cartesian components: with F = diag(fscale) the block is v = F4 raw, so
// J = F J'(F Dj F) F and K = F K'(F D F) F where J', K' contract the raw
// block.  The densities are scaled once here and the outputs while being
// symmetrised, so no epilogue loads a function scale.
int enqueue_direct(hfx_fock* f, const double* Dj, const double* d_Pa, const double* d_Pb, double* d_J, double* d_Ka,
                   double* d_Kb, const double* dpmax, hipStream_t st) {
    hfx_basis* b = f->b;
    const int K = b->plan.K;
    const size_t K2 = (size_t)K * K;
    const int nspin = d_Pb ? 2 : 1;
    HIPCHK(zero_jk(d_J, d_Ka, d_Kb, nspin, K2, st));
    double* sDj = f->d_scl;
    double* sDa = f->d_scl + K2;
    double* sDb = f->d_scl + 2 * K2;
    HIPCHK(launch_scale_n(sDj, sDa, sDb, Dj, d_Pa, d_Pb, 1 + nspin, b->d_fscale, K, st));
    EpiArgs ep = base_epi(b);
    ep.mode = EPI_JK;
    ep.threshold = f->opts.threshold;
    ep.nspin = nspin;
    ep.Dj = sDj;
    ep.Da = sDa;
    ep.Db = nspin == 2 ? sDb : nullptr;
    ep.J = d_J;
    ep.Ka = d_Ka;
    ep.Kb = d_Kb;
    if (dpmax) {
        ep.dpmax = dpmax;
        ep.qpair = b->d_qpair;
        ep.nshell = b->plan.nshell;
        ep.dthr = incremental_tau(f);
    }
    if (int rc = launch_plan(b, f->plan, ep, st)) return rc;
    HIPCHK(launch_symmetrize_n(d_J, d_Ka, d_Kb, 1 + nspin, K, st, b->d_fscale));
    return 0;
}

int enqueue_build(hfx_fock* f, const double* d_Pa, const double* d_Pb, double* d_J, double* d_Ka, double* d_Kb,
                  hipStream_t st) {
    hfx_basis* b = f->b;
    const size_t K2 = (size_t)b->plan.K * b->plan.K;
    const int nspin = d_Pb ? 2 : 1;
    // Dj = Pa (+ Pb) for the Coulomb part
    const double* Dj = d_Pa;
    if (nspin == 2) {
        HIPCHK(launch_add(f->d_Dj, d_Pa, d_Pb, K2, st));
        Dj = f->d_Dj;
    }
    if (f->mode == HFX_MODE_STORED) return jk_from_unique(b, f->d_eri, Dj, d_Pa, d_Pb, d_J, d_Ka, d_Kb, nspin, st);
    return enqueue_direct(f, Dj, d_Pa, d_Pb, d_J, d_Ka, d_Kb, nullptr, st);
}

// Incremental direct build: G(P) = G(P_prev) + G(P - P_prev), the second term
// screened by the shell-block maxima of |P - P_prev|; J / K are linear in the
// densities and symmetrisation is linear, so the accumulated matrices are the
// full build's up to the screened terms.  A full build every
// INCREMENTAL_FULL_EVERY calls (and on the first) bounds how long skipped
// terms accumulate.
constexpr int INCREMENTAL_FULL_EVERY = 8;

int enqueue_incremental(hfx_fock* f, const double* d_Pa, const double* d_Pb, double* d_J, double* d_Ka, double* d_Kb,
                        hipStream_t st) {
    hfx_basis* b = f->b;
    const int K = b->plan.K;
    const size_t K2 = (size_t)K * K;
    const int nspin = d_Pb ? 2 : 1;
    double* Pa0 = f->d_inc;
    double* Pb0 = Pa0 + K2;
    double* dPa = Pa0 + 2 * K2;
    double* dPb = Pa0 + 3 * K2;
    double* dPj = Pa0 + 4 * K2;
    double* Jacc = Pa0 + 5 * K2;
    double* Kaacc = Pa0 + 6 * K2;
    double* Kbacc = Pa0 + 7 * K2;
    double* dpm = Pa0 + 8 * K2;
    if (!f->have_prev || f->prev_nspin != nspin || f->since_full + 1 >= INCREMENTAL_FULL_EVERY) {
        if (int rc = enqueue_build(f, d_Pa, d_Pb, d_J, d_Ka, d_Kb, st)) return rc;
        HIPCHK(launch_copy(Pa0, d_Pa, K2, st));
        HIPCHK(launch_copy(Jacc, d_J, K2, st));
        HIPCHK(launch_copy(Kaacc, d_Ka, K2, st));
        if (nspin == 2) {
            HIPCHK(launch_copy(Pb0, d_Pb, K2, st));
            HIPCHK(launch_copy(Kbacc, d_Kb, K2, st));
        }
        f->have_prev = true;
        f->prev_nspin = nspin;
        f->since_full = 0;
        return 0;
    }
    HIPCHK(launch_delta(dPa, d_Pa, Pa0, K2, st));
    if (nspin == 2) {
        HIPCHK(launch_delta(dPb, d_Pb, Pb0, K2, st));
        HIPCHK(launch_add(dPj, dPa, dPb, K2, st));
    }
    HIPCHK(launch_shell_dpmax(b->d_shells, b->plan.nshell, K, dPa, nspin == 2 ? dPb : nullptr,
                              nspin == 2 ? dPj : nullptr, dpm, st));
    if (int rc = enqueue_direct(f, nspin == 2 ? dPj : dPa, dPa, nspin == 2 ? dPb : nullptr, d_J, d_Ka, d_Kb, dpm, st))
        return rc;
    HIPCHK(launch_accum(d_J, Jacc, K2, st));
    HIPCHK(launch_accum(d_Ka, Kaacc, K2, st));
    if (nspin == 2) HIPCHK(launch_accum(d_Kb, Kbacc, K2, st));
    ++f->since_full;
    return 0;
}

}  // namespace

namespace {

// Re-plans a direct context's work units for slice target t; the previous
// plan's builds must have completed (its lists go back to the cache).
int replan(hfx_fock* f, long long t) {
    // the lists about to be replaced may still be read by the last build
    // (ev_done is recorded after every build, on its stream)
    if (f->ev_done) HIPCHK(hipEventSynchronize(f->ev_done));
    if (f->graph) {
        HIPCHK(hipGraphExecDestroy(f->graph));
        f->graph = nullptr;
    }
    if (int rc = make_plan(f->b, f->opts.threshold, f->opts.precision == HFX_PREC_MIXED, f->opts.mixed_split,
                           SliceTargets{t, t}, f->plan, f->mode != HFX_MODE_STORED))
        return rc;
    f->plan_target = t;
    return 0;
}

// One autotuner step at the start of a direct build: collects the timing of
// the previous tuning build; returns 1 when this build is a tuning build
// (then it is enqueued full and eager between the tuning events by the
// caller), 0 otherwise (the tuner finished: the fastest plan is in place).
int tune_step(hfx_fock* f, int& is_tuning) {
    is_tuning = 0;
    if (!f->tuning) return 0;
    if (f->tune_warm) {
        f->tune_warm = false;
        return 0;
    }
    if (f->tune_pending) {
        HIPCHK(hipEventSynchronize(f->tune_ev[1]));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, f->tune_ev[0], f->tune_ev[1]));
        f->tune_ms[f->tune_i] = ms;
        f->tune_pending = false;
        ++f->tune_i;
    }
    if (f->tune_i < hfx_fock::NTUNE) {
        if (f->plan_target != f->tune_target[f->tune_i])
            if (int rc = replan(f, f->tune_target[f->tune_i])) return rc;
        f->have_prev = false;   // a full build, comparable across candidates
        is_tuning = 1;
        return 0;
    }
    int best = 0;
    for (int i = 1; i < hfx_fock::NTUNE; ++i)
        if (f->tune_ms[i] < f->tune_ms[best]) best = i;
    if (f->plan_target != f->tune_target[best])
        if (int rc = replan(f, f->tune_target[best])) return rc;
    f->tuning = false;
    return 0;
}

}  // namespace

int hfx_fock_build_device(hfx_fock* f, const double* d_Pa, const double* d_Pb, double* d_J, double* d_Ka,
                          double* d_Kb, void* stream) {
    if (!f || !d_Pa || !d_J || !d_Ka) return fail(HFX_ERR_INVALID, "null argument");
    if (d_Pb && !d_Kb) return fail(HFX_ERR_INVALID, "UHF build needs Kb");
    hipStream_t st = stream ? (hipStream_t)stream : f->b->stream;
    if (f->mode == HFX_MODE_DIRECT && !g_timer.on) {
        int is_tuning = 0;
        if (int rc = tune_step(f, is_tuning)) return rc;
        if (is_tuning) {
            HIPCHK(hipEventRecord(f->tune_ev[0], st));
            const int rc = f->d_inc ? enqueue_incremental(f, d_Pa, d_Pb, d_J, d_Ka, d_Kb, st)
                                    : enqueue_build(f, d_Pa, d_Pb, d_J, d_Ka, d_Kb, st);
            if (rc) return rc;
            HIPCHK(hipEventRecord(f->tune_ev[1], st));
            f->tune_pending = true;
            HIPCHK(hipEventRecord(f->ev_done, st));
            return 0;
        }
    }
    if (f->d_inc && !g_timer.on) {   // incremental direct builds: eager
        if (int rc = enqueue_incremental(f, d_Pa, d_Pb, d_J, d_Ka, d_Kb, st)) return rc;
        HIPCHK(hipEventRecord(f->ev_done, st));
        return 0;
    }
    // class timing records events between launches: always eager; a stored
    // build is a handful of launches: no graph
    if (g_timer.on || f->mode == HFX_MODE_STORED || !fock_graphs_on() || !st) {
        if (int rc = enqueue_build(f, d_Pa, d_Pb, d_J, d_Ka, d_Kb, st)) return rc;
        HIPCHK(hipEventRecord(f->ev_done, st));
        return 0;
    }
    const void* key[6] = {d_Pa, d_Pb, d_J, d_Ka, d_Kb, st};
    if (!f->graph || !std::equal(key, key + 6, f->graph_key)) {
        if (f->graph) {
            HIPCHK(hipGraphExecDestroy(f->graph));
            f->graph = nullptr;
        }
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_build(f, d_Pa, d_Pb, d_J, d_Ka, d_Kb, st);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        HIPCHK(e);
        const hipError_t ei = hipGraphInstantiate(&f->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ei);
        std::copy(key, key + 6, f->graph_key);
    }
    HIPCHK(hipGraphLaunch(f->graph, st));
    HIPCHK(hipEventRecord(f->ev_done, st));
    return 0;
}

int hfx_set_fock_graphs(int on) {
    g_fock_graphs.store(on ? 1 : 0);
    return 0;
}

// Host matrices travel through a pinned staging block of the context: the
// densities are copied into it on the CPU and read by a copy kernel over the
// host link, the results written back the same way.  Copies of the caller's
// pageable memory would each take an SDMA round trip (measured ~0.1 ms apiece
// from about 10 KB on, the bulk of a small molecule's build).
int hfx_fock_build(hfx_fock* f, const double* Pa, const double* Pb, double* J, double* Ka, double* Kb) {
    if (!f || !Pa) return fail(HFX_ERR_INVALID, "null argument");
    hfx_basis* b = f->b;
    const int K = b->plan.K;
    const size_t K2 = (size_t)K * K;
    hipStream_t st = b->stream;
    if (!f->h_stage) {
        void* h = nullptr;
        HIPCHK(hfx_detail::host_alloc(&h, 5 * K2 * sizeof(double)));
        f->h_stage = static_cast<double*>(h);
        void* dv = nullptr;
        const hipError_t e = hipHostGetDevicePointer(&dv, h, 0);
        if (e != hipSuccess) {
            hfx_detail::host_free(h);
            f->h_stage = nullptr;
            HIPCHK(e);
        }
        f->h_stage_dev = static_cast<double*>(dv);
    }
    double* h = f->h_stage;
    double* hd = f->h_stage_dev;   // the same block, as the device addresses it
    // staging [Pa | Pb | J | Ka | Kb] and the context's [Da | Db] and
    // [J | Ka | Kb] are back to back: one copy launch each way
    std::memcpy(h, Pa, K2 * sizeof(double));
    if (Pb) std::memcpy(h + K2, Pb, K2 * sizeof(double));
    HIPCHK(launch_copy(f->d_Da, hd, (Pb ? 2 : 1) * K2, st));
    if (int rc = hfx_fock_build_device(f, f->d_Da, Pb ? f->d_Db : nullptr, f->d_J, f->d_Ka, f->d_Kb, st)) return rc;
    HIPCHK(launch_copy(hd + 2 * K2, f->d_J, (Pb ? 3 : 2) * K2, st));
    HIPCHK(hipStreamSynchronize(st));
    if (J) std::memcpy(J, h + 2 * K2, K2 * sizeof(double));
    if (Ka) std::memcpy(Ka, h + 3 * K2, K2 * sizeof(double));
    if (Kb && Pb) std::memcpy(Kb, h + 4 * K2, K2 * sizeof(double));
    return 0;
}

int hfx_set_class_timing(int on) {
    if (on && !g_timer.created) {
        for (int qc = 0; qc < NQCLASS; ++qc)
            for (int p = 0; p < 2; ++p)
                for (int k = 0; k < 2; ++k) HIPCHK(hipEventCreate(&g_timer.ev[qc][p][k]));
        g_timer.created = true;
    }
    g_timer.serial.store(on == 2);
    g_timer.on.store(on != 0);
    return 0;
}

int hfx_class_times(float* ms) {
    for (int qc = 0; qc < NQCLASS; ++qc)
        for (int p = 0; p < 2; ++p) {
            float t = 0.0f;
            if (g_timer.used[qc][p]) HIPCHK(hipEventElapsedTime(&t, g_timer.ev[qc][p][0], g_timer.ev[qc][p][1]));
            ms[2 * qc + p] = t;
        }
    return 0;
}

int hfx_plan_stats(hfx_basis* b, double threshold, int precision, double mixed_split, long long* nq,
                   long long* nprimq) {
    if (!b) return fail(HFX_ERR_INVALID, "null basis");
    if (int rc = compute_bounds(b)) return rc;
    try {
        QuartetLists ql;
        plan_quartets(b->plan, b->Qsh, threshold, precision == HFX_PREC_MIXED, mixed_split > 0 ? mixed_split : 1e-5,
                      ql);
        for (int qc = 0; qc < NQCLASS; ++qc)
            for (int p = 0; p < 2; ++p) {
                long long s = 0;
                for (const uint2& e : ql.list[qc][p]) s += (long long)b->plan.hdr[e.x].nprim * b->plan.hdr[e.y].nprim;
                if (nq) nq[2 * qc + p] = (long long)ql.list[qc][p].size();
                if (nprimq) nprimq[2 * qc + p] = s;
            }
    } catch (const std::bad_alloc&) {
        return fail(HFX_ERR_NOMEM, "host memory for the quartet plan");
    } catch (const std::exception& e) {
        return fail(HFX_ERR_DEVICE, std::string("quartet plan: ") + e.what());
    }
    return 0;
}

}  // extern "C"
