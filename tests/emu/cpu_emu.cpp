// cpu_emu.cpp -- test-only CPU build of the product's ERI kernel bodies.
//
// Compiles hartree_fock_cpp_amd/csrc/device/{plan,eri_body}.hpp for the host
// and runs every shell quartet with one "lane" (nlanes = 1), so the planner,
// class dispatch, Hermite/R recursions, contractions and all three epilogues
// can be checked against the oracle without a GPU.  It is never loaded by the
// product library, smoke() or bench.py.
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <algorithm>
#include <vector>

#include "../../hartree_fock_cpp_amd/csrc/device/plan.hpp"

using namespace hfx;

namespace {

std::string g_err;

struct EmuCtx {
    BasisPlan plan;
    std::vector<double> boysd;
    std::vector<float> boysf;
    HermTab ht;
    std::vector<double> bounds, Qsh;
};

// Slicing of the emulated launches: the device's policy (expand_units with
// SLICE_TARGET), or a forced slice count per quartet (tests exercise the
// partial-epilogue path on small inputs), or none (0 target).  A negative
// count -c cuts family-B quartets into c ket chunks per bra primitive pair
// (famb_range), family A as for c.
int g_force_slices = 0;

std::vector<uint2> emu_units(const EmuCtx& c, bool fam_a, int G, const uint2* q, int n, const EpiArgs& ep, int kb) {
    std::vector<uint2> out;
    // the device slices J/K and Schwarz-diagonal plans, and family B only in
    // (large) unique-array plans
    if (g_force_slices == 1 || (g_force_slices == 0 && ep.mode == EPI_UNIQUE && fam_a)) {
        out.assign(q, q + n);
    } else if (g_force_slices < 0) {
        for (int k = 0; k < n; ++k) {
            const int nb = c.plan.hdr[q[k].x].nprim, nk = c.plan.hdr[q[k].y].nprim;
            int S;
            if (fam_a) {
                S = std::min(-g_force_slices, std::max(1, std::min(nb * nk / G, nb)));
            } else {
                int nkc = std::min(-g_force_slices, (nk + kb - 1) / kb);
                while (nkc > 1 && nb * nkc > SLICE_MAX) --nkc;
                S = nkc > 1 ? nb * nkc : 1;
            }
            for (int sl = 0; sl < S; ++sl) out.push_back(make_uint2(q[k].x | ((unsigned)sl << 24), q[k].y | ((unsigned)S << 24)));
        }
    } else if (g_force_slices > 1) {
        for (int k = 0; k < n; ++k) {
            const int nb = c.plan.hdr[q[k].x].nprim;
            const int cap = fam_a ? std::max(1, std::min(nb * c.plan.hdr[q[k].y].nprim / G, nb)) : nb;
            const int S = std::min(g_force_slices, cap);
            for (int sl = 0; sl < S; ++sl) out.push_back(make_uint2(q[k].x | ((unsigned)sl << 24), q[k].y | ((unsigned)S << 24)));
        }
    } else {
        expand_units(c.plan, fam_a, G, q, n, SLICE_TARGET, out, kb);
    }
    return out;
}

template <int QC, class Real>
void run_class(const EmuCtx& c, const std::vector<uint2>& list, const int* bucket, const EpiArgs& ep) {
    using C = QClass<QC>;
    constexpr bool fam_a = runs_family_a(C::LA, C::LB, C::LC, C::LD);
    const PrimPair<Real>* pp;
    const Real* tab;
    if constexpr (sizeof(Real) == 8) {
        pp = c.plan.ppd.data();
        tab = c.boysd.data();
    } else {
        pp = c.plan.ppf.data();
        tab = c.boysf.data();
    }
    size_t pos = 0;
    auto smj = fam_a ? nullptr : std::make_unique<FamBSmem<C::LA, C::LB, C::LC, C::LD, Real, true>>();
    auto smw = fam_a ? nullptr : std::make_unique<FamBSmem<C::LA, C::LB, C::LC, C::LD, Real, false>>();
    for (int g = 0; g < NGROUP; ++g) {
        const std::vector<uint2> units = emu_units(c, fam_a, GROUP_SIZE[g], list.data() + pos, bucket[g], ep,
                                                   famb_kb(C::LA + C::LB + C::LC + C::LD));
        pos += bucket[g];
        for (const uint2& e : units) {
            if constexpr (fam_a) {
                // replay the device's lane-group partition unit by unit
                family_a_group_emu<C::LA, C::LB, C::LC, C::LD, Real>(e, GROUP_SIZE[g], c.plan.hdr.data(), pp, tab, ep);
            } else {
                auto sync = []() {};
                constexpr int NCH = FamBShape<C::LA, C::LB, C::LC, C::LD>::NCHUNK;
                for (int ch = 0; ch < NCH; ++ch) {   // one device wavefront per (unit, chunk)
                    if (ep.mode == EPI_JK)
                        family_b_one<C::LA, C::LB, C::LC, C::LD, Real, true>(e, ch, c.plan.hdr.data(), pp, tab, ep,
                                                                             *smj, c.ht, 0, 1, sync);
                    else
                        family_b_one<C::LA, C::LB, C::LC, C::LD, Real, false>(e, ch, c.plan.hdr.data(), pp, tab, ep,
                                                                              *smw, c.ht, 0, 1, sync);
                }
            }
        }
    }
}

using RunFn = void (*)(const EmuCtx&, const std::vector<uint2>&, const int*, const EpiArgs&);

template <class Real, int... QC>
RunFn pick(std::integer_sequence<int, QC...>, int qc) {
    static const RunFn t[] = {&run_class<QC, Real>...};
    return t[qc];
}

void run_lists(const EmuCtx& c, const QuartetLists& ql, const EpiArgs& ep) {
    for (int qc = 0; qc < NQCLASS; ++qc) {
        if (!ql.list[qc][0].empty())
            pick<double>(std::make_integer_sequence<int, NQCLASS>{}, qc)(c, ql.list[qc][0], ql.bucket[qc][0], ep);
        if (!ql.list[qc][1].empty())
            pick<float>(std::make_integer_sequence<int, NQCLASS>{}, qc)(c, ql.list[qc][1], ql.bucket[qc][1], ep);
    }
}

EpiArgs base_epi(const EmuCtx& c) {
    EpiArgs ep{};
    ep.K = c.plan.K;
    ep.bounds = c.bounds.data();
    ep.fscale = c.plan.fscale.data();
    ep.nspin = 1;
    return ep;
}

int make_ctx(int K, const double* xyz, const int* lmn, const int* poff, const double* alpha, const double* coeff,
             const double* norm, EmuCtx& c) {
    if (build_plan(K, xyz, lmn, poff, alpha, coeff, norm, c.plan, g_err)) return -1;
    boys_table(c.boysd);
    c.boysf.assign(c.boysd.begin(), c.boysd.end());
    c.ht = herm_tab();
    const int P = K * (K + 1) / 2;
    c.bounds.assign(P, 0.0);
    std::vector<uint2> diag[NQCLASS];
    plan_diag(c.plan, diag);
    QuartetLists ql;
    for (int q = 0; q < NQCLASS; ++q) {
        ql.list[q][0] = diag[q];
        ql.bucket[q][0][NGROUP - 1] = (int)diag[q].size();   // one lane per diagonal quartet
    }
    std::vector<double> out(P, 0.0);
    EpiArgs ep = base_epi(c);
    ep.mode = EPI_DIAG;
    ep.out = out.data();
    run_lists(c, ql, ep);
    for (double& v : out) v = v > 0.0 ? std::sqrt(v) : 0.0;   // the epilogue leaves (mu nu|mu nu)
    c.bounds = out;
    shell_pair_bounds(c.plan, c.bounds.data(), c.Qsh);
    return 0;
}

}  // namespace

#define EMU_BASIS_ARGS int K, const double *xyz, const int *lmn, const int *poff, const double *alpha, \
                       const double *coeff, const double *norm

extern "C" {

const char* emu_last_error() { return g_err.c_str(); }

// Forces S primitive slices per quartet in the unique and J/K builds (0: the
// device's own choice).
void emu_set_slices(int S) { g_force_slices = S; }

int emu_schwarz(EMU_BASIS_ARGS, double* bounds) {
    EmuCtx c;
    if (make_ctx(K, xyz, lmn, poff, alpha, coeff, norm, c)) return -1;
    std::memcpy(bounds, c.bounds.data(), c.bounds.size() * sizeof(double));
    return 0;
}

// Full unique array (Q doubles, zero-filled by the caller).
int emu_eri_unique(EMU_BASIS_ARGS, double threshold, double* out) {
    EmuCtx c;
    if (make_ctx(K, xyz, lmn, poff, alpha, coeff, norm, c)) return -1;
    QuartetLists ql;
    plan_quartets(c.plan, c.Qsh, threshold, false, 0.0, ql);
    EpiArgs ep = base_epi(c);
    ep.mode = EPI_UNIQUE;
    ep.threshold = threshold;
    ep.out = out;
    run_lists(c, ql, ep);
    return 0;
}

// Direct J/K (symmetrised).  Pb == NULL: RHF (J and Ka from Pa).
int emu_fock_direct(EMU_BASIS_ARGS, double threshold, int mixed, double split, const double* Pa, const double* Pb,
                    double* J, double* Ka, double* Kb) {
    EmuCtx c;
    if (make_ctx(K, xyz, lmn, poff, alpha, coeff, norm, c)) return -1;
    QuartetLists ql;
    plan_quartets(c.plan, c.Qsh, threshold, mixed != 0, split, ql, 0, true);   // the device's J/K order
    const size_t K2 = (size_t)K * K;
    // the device's convention: densities scaled by the function scales
    // (diag(fs) D diag(fs)), outputs scaled back while symmetrised
    const double* fs = c.plan.fscale.data();
    auto scaled = [&](const double* M) {
        std::vector<double> o(K2);
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j) o[(size_t)i * K + j] = fs[i] * fs[j] * M[(size_t)i * K + j];
        return o;
    };
    std::vector<double> Dj(Pa, Pa + K2);
    if (Pb)
        for (size_t i = 0; i < K2; ++i) Dj[i] += Pb[i];
    Dj = scaled(Dj.data());
    const std::vector<double> sPa = scaled(Pa);
    const std::vector<double> sPb = Pb ? scaled(Pb) : std::vector<double>();
    std::memset(J, 0, K2 * sizeof(double));
    std::memset(Ka, 0, K2 * sizeof(double));
    if (Pb) std::memset(Kb, 0, K2 * sizeof(double));
    EpiArgs ep = base_epi(c);
    ep.mode = EPI_JK;
    ep.threshold = threshold;
    ep.nspin = Pb ? 2 : 1;
    ep.Dj = Dj.data();
    ep.Da = sPa.data();
    ep.Db = Pb ? sPb.data() : nullptr;
    ep.J = J;
    ep.Ka = Ka;
    ep.Kb = Kb;
    run_lists(c, ql, ep);
    auto sym = [&](double* M) {
        for (int i = 0; i < K; ++i)
            for (int j = 0; j <= i; ++j) {
                const double s = (M[i * K + j] + M[j * K + i]) * fs[i] * fs[j];
                M[i * K + j] = s;
                M[j * K + i] = s;
            }
    };
    sym(J);
    sym(Ka);
    if (Pb) sym(Kb);
    return 0;
}

// Number of shell quartets per class (fp64, fp32) for a threshold/split.
int emu_class_counts(EMU_BASIS_ARGS, double threshold, int mixed, double split, long long* n64, long long* n32) {
    EmuCtx c;
    if (make_ctx(K, xyz, lmn, poff, alpha, coeff, norm, c)) return -1;
    QuartetLists ql;
    plan_quartets(c.plan, c.Qsh, threshold, mixed != 0, split, ql);
    for (int q = 0; q < NQCLASS; ++q) {
        n64[q] = (long long)ql.list[q][0].size();
        n32[q] = (long long)ql.list[q][1].size();
    }
    return 0;
}

// Plans the screened quartets on one thread and on `threads` threads and
// compares the lists and lane-group buckets entry by entry; returns the number
// of (class, precision) lists that differ (0: identical) and the total quartet
// count in *nq.
int emu_plan_threads_agree(EMU_BASIS_ARGS, double threshold, int mixed, double split, int threads, long long* nq) {
    EmuCtx c;
    if (make_ctx(K, xyz, lmn, poff, alpha, coeff, norm, c)) return -1;
    QuartetLists one, many;
    plan_quartets(c.plan, c.Qsh, threshold, mixed != 0, split, one, 1);
    plan_quartets(c.plan, c.Qsh, threshold, mixed != 0, split, many, threads);
    int bad = 0;
    *nq = (long long)one.total();
    for (int q = 0; q < NQCLASS; ++q)
        for (int p = 0; p < 2; ++p) {
            const auto& a = one.list[q][p];
            const auto& b = many.list[q][p];
            bool same = a.size() == b.size();
            for (size_t i = 0; same && i < a.size(); ++i) same = a[i].x == b[i].x && a[i].y == b[i].y;
            for (int g = 0; same && g < NGROUP; ++g) same = one.bucket[q][p][g] == many.bucket[q][p][g];
            bad += same ? 0 : 1;
        }
    return bad;
}

// Plans the screened quartets in the default order and in the J/K order (on
// `threads` threads) and checks every J/K list: the same quartets and
// lane-group buckets as the default list, primitive work w descending, and
// within equal w ascending (bra pair, ket first shell, ket pair) with no
// repeats.  Returns the number of lists that fail (0: all pass) and the total
// quartet count in *nq.
int emu_plan_jk_order_check(EMU_BASIS_ARGS, double threshold, int mixed, double split, int threads, long long* nq) {
    EmuCtx c;
    if (make_ctx(K, xyz, lmn, poff, alpha, coeff, norm, c)) return -1;
    QuartetLists def, jk;
    plan_quartets(c.plan, c.Qsh, threshold, mixed != 0, split, def, threads);
    plan_quartets(c.plan, c.Qsh, threshold, mixed != 0, split, jk, threads, true);
    const auto& H = c.plan.hdr;
    auto w = [&](const uint2& e) { return (long long)H[e.x].nprim * H[e.y].nprim; };
    auto key = [&](const uint2& e) { return std::make_tuple(-w(e), e.x, H[e.y].s1, e.y); };
    int bad = 0;
    *nq = (long long)jk.total();
    for (int q = 0; q < NQCLASS; ++q)
        for (int p = 0; p < 2; ++p) {
            std::vector<uint2> a = def.list[q][p];
            const std::vector<uint2>& b = jk.list[q][p];
            bool ok = a.size() == b.size();
            for (int g = 0; ok && g < NGROUP; ++g) ok = def.bucket[q][p][g] == jk.bucket[q][p][g];
            for (size_t i = 1; ok && i < b.size(); ++i) ok = key(b[i - 1]) < key(b[i]);
            if (ok) {
                std::sort(a.begin(), a.end(), [&](const uint2& x, const uint2& y) { return key(x) < key(y); });
                for (size_t i = 0; ok && i < a.size(); ++i) ok = a[i].x == b[i].x && a[i].y == b[i].y;
            }
            bad += ok ? 0 : 1;
        }
    return bad;
}

// Runs parallel_tasks over ntask tasks on `threads` threads; task throw_at
// (if in range) throws.  Returns 1 when the exception reached the caller, 0
// when none was thrown, -1 when one was expected and lost; *ran counts the
// tasks that completed.
int emu_parallel_tasks_check(int ntask, int threads, int throw_at, int* ran) {
    std::atomic<int> n{0};
    try {
        parallel_tasks(ntask, threads, [&](int t) {
            if (t == throw_at) throw std::runtime_error("task failed");
            n.fetch_add(1);
        });
    } catch (const std::runtime_error& e) {
        *ran = n.load();
        return std::string(e.what()) == "task failed" ? 1 : -1;
    }
    *ran = n.load();
    return (throw_at >= 0 && throw_at < ntask) ? -1 : 0;
}

// Boys values from the product's table path (FP64 and FP32).
void emu_boys(int mmax, double T, double* F64, float* F32) {
    static std::vector<double> tab;
    static std::vector<float> tabf;
    if (tab.empty()) {
        boys_table(tab);
        tabf.assign(tab.begin(), tab.end());
    }
    double Fd[BOYS_MMAX + 1];
    float Ff[BOYS_MMAX + 1];
    // dispatch on mmax at run time through a small switch of template instances
    switch (mmax) {
#define EMU_B(M)                                          \
    case M:                                               \
        boys_fm<M, double>(T, tab.data(), Fd);            \
        boys_fm<M, float>((float)T, tabf.data(), Ff);     \
        break;
        EMU_B(0) EMU_B(1) EMU_B(2) EMU_B(3) EMU_B(4) EMU_B(5) EMU_B(6) EMU_B(7) EMU_B(8) EMU_B(9) EMU_B(10) EMU_B(11)
        EMU_B(12)
#undef EMU_B
        default:
            return;
    }
    for (int m = 0; m <= mmax; ++m) {
        F64[m] = Fd[m];
        F32[m] = Ff[m];
    }
}

}  // extern "C"
