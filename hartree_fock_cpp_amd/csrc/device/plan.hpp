// plan.hpp -- host-side preparation of the device tables (shells, shell pairs,
// primitive pairs, quartet lists, Boys table).  Pure host C++, shared by the
// HIP library and the CPU emulation build used by the tests.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "eri_body.hpp"
#include "eri_classes.hpp"

namespace hfx {

struct ShellH {
    double c[3];
    int L;
    int nprim;
    int poff;   // into BasisPlan::alpha / w
    int fn0;    // first basis function
};

struct BasisPlan {
    int K = 0;
    int nshell = 0;
    int npair = 0;
    std::vector<ShellH> shells;
    std::vector<double> alpha, w;       // shell primitives: exponent, coeff * radial norm
    std::vector<double> fscale;         // per function: component norm / radial norm
    std::vector<PairHdr> hdr;           // indexed by canonical pair id
    std::vector<int> pclass;            // pair class per canonical pair id
    std::vector<PrimPair<double>> ppd;
    std::vector<PrimPair<float>> ppf;
    std::vector<int2> fnpairs;          // function pair id -> (i, j), i >= j
    // Hermite -> Cartesian matrices of every kept primitive pair (PairHdr::eoff):
    // per primitive pair of a pair of class (la, lb), nherm(la + lb) rows
    // (packed tuv) of ncart(la) ncart(lb) columns (ab = a ncart(lb) + b, so
    // the 16 lanes of an MFMA operand row read 16 consecutive values), entry
    // E^x_{a_x b_x, t} E^y_{a_y b_y, u} E^z_{a_z b_z, v} (zero where an
    // order exceeds a_k + b_k); FP64 and an FP32 copy.  etkd / etkf: the same
    // with column tuv scaled by (-1)^{t+u+v}, the sign the pair carries on the
    // ket side of (ab|cd)
    std::vector<double> etd, etkd;
    std::vector<float> etf, etkf;
};

// 1-D Hermite expansion E[a][b][t] of one primitive pair along one axis, for
// runtime a <= la, b <= lb (the recursion of Herm1D with runtime bounds).
inline void herm1d_rt(int la, int lb, double i2p, double PA, double PB, double e[4][4][7]) {
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int t = 0; t < 7; ++t) e[a][b][t] = 0.0;
    e[0][0][0] = 1.0;
    for (int a = 1; a <= la; ++a)
        for (int t = 0; t <= a; ++t) {
            double v = PA * e[a - 1][0][t];
            if (t > 0) v += i2p * e[a - 1][0][t - 1];
            if (t + 1 <= a - 1) v += (t + 1) * e[a - 1][0][t + 1];
            e[a][0][t] = v;
        }
    for (int b = 1; b <= lb; ++b)
        for (int a = 0; a <= la; ++a)
            for (int t = 0; t <= a + b; ++t) {
                double v = (t <= a + b - 1) ? PB * e[a][b - 1][t] : 0.0;
                if (t > 0) v += i2p * e[a][b - 1][t - 1];
                if (t + 1 <= a + b - 1) v += (t + 1) * e[a][b - 1][t + 1];
                e[a][b][t] = v;
            }
}

// Appends the Hermite -> Cartesian matrix of one primitive pair (layout in BasisPlan::etd).
inline void append_emat(int la, int lb, const PrimPair<double>& d, std::vector<double>& etd,
                        std::vector<double>& etkd) {
    double ex[3][4][4][7];
    for (int x = 0; x < 3; ++x) herm1d_rt(la, lb, d.i2p, d.PA[x], d.PB[x], ex[x]);
    const int L = la + lb, nab = ncart(la) * ncart(lb);
    const size_t base = etd.size();
    etd.resize(base + (size_t)nab * nherm(L));
    etkd.resize(etd.size());
    for (int a = 0; a < ncart(la); ++a)
        for (int b = 0; b < ncart(lb); ++b) {
            const int ax = cart_x(la, a), ay = cart_y(la, a), az = cart_z(la, a);
            const int bx = cart_x(lb, b), by = cart_y(lb, b), bz = cart_z(lb, b);
            const int ab = a * ncart(lb) + b;
            int tuv = 0;
            for (int s = 0; s <= L; ++s)
                for (int t = 0; t <= s; ++t)
                    for (int u = 0; u <= s - t; ++u, ++tuv) {
                        const int v = s - t - u;
                        const bool nz = t <= ax + bx && u <= ay + by && v <= az + bz;
                        const double x = nz ? ex[0][ax][bx][t] * ex[1][ay][by][u] * ex[2][az][bz][v] : 0.0;
                        etd[base + (size_t)tuv * nab + ab] = x;
                        etkd[base + (size_t)tuv * nab + ab] = (s & 1) ? -x : x;
                    }
        }
}

// Drop primitive pairs whose Gaussian-product prefactor exp(-mu |AB|^2) is
// below e^-60 (~9e-27): their contribution to any integral is < 1e-22.
constexpr double PRIM_PAIR_CUT = 60.0;

inline int build_plan(int K, const double* xyz, const int* lmn, const int* poff, const double* alpha,
                      const double* coeff, const double* norm, BasisPlan& bp, std::string& err) {
    bp = BasisPlan();
    bp.K = K;
    bp.fscale.assign(K, 1.0);
    if (K <= 0) {
        err = "basis has no functions";
        return -1;
    }
    int mu = 0;
    while (mu < K) {
        const int L = lmn[3 * mu] + lmn[3 * mu + 1] + lmn[3 * mu + 2];
        if (L > MAX_L) {
            err = "angular momentum above f (L > 3) is not supported by the GPU path";
            return -4;
        }
        const int n = ncart(L);
        const int np = poff[mu + 1] - poff[mu];
        if (mu + n > K || np <= 0) {
            err = "basis functions must come in complete Cartesian shells (basis_construction order)";
            return -1;
        }
        ShellH sh;
        sh.c[0] = xyz[3 * mu];
        sh.c[1] = xyz[3 * mu + 1];
        sh.c[2] = xyz[3 * mu + 2];
        sh.L = L;
        sh.nprim = np;
        sh.poff = (int)bp.alpha.size();
        sh.fn0 = mu;
        for (int k = 0; k < n; ++k) {
            const int f = mu + k;
            const bool okc = lmn[3 * f] == cart_x(L, k) && lmn[3 * f + 1] == cart_y(L, k) && lmn[3 * f + 2] == cart_z(L, k);
            const bool okp = poff[f + 1] - poff[f] == np && xyz[3 * f] == sh.c[0] && xyz[3 * f + 1] == sh.c[1] &&
                             xyz[3 * f + 2] == sh.c[2];
            if (!okc || !okp) {
                err = "basis functions must come in complete Cartesian shells in angmom_partition order";
                return -1;
            }
            for (int i = 0; i < np; ++i) {
                if (alpha[poff[f] + i] != alpha[poff[mu] + i] || coeff[poff[f] + i] != coeff[poff[mu] + i]) {
                    err = "components of one shell must share exponents and coefficients";
                    return -1;
                }
            }
            // normalisation must factor as radial(alpha_i) * component scale
            const double s = norm[poff[f]] / norm[poff[mu]];
            for (int i = 0; i < np; ++i) {
                const double r = norm[poff[mu] + i] * s;
                if (std::fabs(r - norm[poff[f] + i]) > 1e-13 * std::fabs(norm[poff[f] + i])) {
                    err = "primitive normalisation does not factor into radial and component parts";
                    return -1;
                }
            }
            bp.fscale[f] = s;
        }
        for (int i = 0; i < np; ++i) {
            if (!(alpha[poff[mu] + i] > 0.0)) {
                err = "primitive exponents must be positive";
                return -1;
            }
            bp.alpha.push_back(alpha[poff[mu] + i]);
            bp.w.push_back(coeff[poff[mu] + i] * norm[poff[mu] + i]);
        }
        bp.shells.push_back(sh);
        mu += n;
    }
    bp.nshell = (int)bp.shells.size();
    bp.npair = bp.nshell * (bp.nshell + 1) / 2;
    bp.hdr.resize(bp.npair);
    bp.pclass.resize(bp.npair);
    for (int A = 0; A < bp.nshell; ++A)
        for (int B = 0; B <= A; ++B) {
            const int x = A * (A + 1) / 2 + B;
            int s1 = A, s2 = B;
            if (bp.shells[A].L < bp.shells[B].L) std::swap(s1, s2);
            const ShellH& S1 = bp.shells[s1];
            const ShellH& S2 = bp.shells[s2];
            PairHdr h;
            h.s1 = s1;
            h.s2 = s2;
            h.fn1 = S1.fn0;
            h.fn2 = S2.fn0;
            h.diag = (A == B);
            h.canon = x;
            h.prim_off = (int)bp.ppd.size();
            h.eoff = (long long)bp.etd.size();
            const double AB2 = (S1.c[0] - S2.c[0]) * (S1.c[0] - S2.c[0]) + (S1.c[1] - S2.c[1]) * (S1.c[1] - S2.c[1]) +
                               (S1.c[2] - S2.c[2]) * (S1.c[2] - S2.c[2]);
            int kept = 0;
            for (int i = 0; i < S1.nprim; ++i)
                for (int j = 0; j < S2.nprim; ++j) {
                    const double a = bp.alpha[S1.poff + i], b = bp.alpha[S2.poff + j];
                    const double p = a + b, m = a * b / p;
                    if (m * AB2 > PRIM_PAIR_CUT) continue;
                    PrimPair<double> d;
                    d.p = p;
                    d.i2p = 0.5 / p;
                    d.w = bp.w[S1.poff + i] * bp.w[S2.poff + j] * std::exp(-m * AB2) / p;
                    for (int x3 = 0; x3 < 3; ++x3) {
                        d.P[x3] = (a * S1.c[x3] + b * S2.c[x3]) / p;
                        d.PA[x3] = d.P[x3] - S1.c[x3];
                        d.PB[x3] = d.P[x3] - S2.c[x3];
                    }
                    bp.ppd.push_back(d);
                    append_emat(S1.L, S2.L, d, bp.etd, bp.etkd);
                    PrimPair<float> f;
                    f.p = (float)d.p;
                    f.i2p = (float)d.i2p;
                    f.w = (float)d.w;
                    for (int x3 = 0; x3 < 3; ++x3) {
                        f.P[x3] = (float)d.P[x3];
                        f.PA[x3] = (float)d.PA[x3];
                        f.PB[x3] = (float)d.PB[x3];
                    }
                    bp.ppf.push_back(f);
                    ++kept;
                }
            h.nprim = kept;
            bp.hdr[x] = h;
            bp.pclass[x] = pair_class(S1.L, S2.L);
        }
    bp.etf.assign(bp.etd.begin(), bp.etd.end());
    bp.etkf.assign(bp.etkd.begin(), bp.etkd.end());
    const int P = K * (K + 1) / 2;
    bp.fnpairs.resize(P);
    for (int i = 0; i < K; ++i)
        for (int j = 0; j <= i; ++j) bp.fnpairs[i * (i + 1) / 2 + j] = make_int2(i, j);
    return 0;
}

// Shell-pair maximum of the function-pair Schwarz bounds.
inline void shell_pair_bounds(const BasisPlan& bp, const double* fnb, std::vector<double>& Q) {
    Q.assign(bp.npair, 0.0);
    for (int x = 0; x < bp.npair; ++x) {
        const PairHdr& h = bp.hdr[x];
        const int n1 = ncart(bp.shells[h.s1].L), n2 = ncart(bp.shells[h.s2].L);
        double m = 0.0;
        for (int a = 0; a < n1; ++a)
            for (int b = 0; b < n2; ++b) m = std::max(m, fnb[pair_of(h.fn1 + a, h.fn2 + b)]);
        Q[x] = m;
    }
}

// Lane-group sizes of family A: a quartet with w = (bra x ket kept primitive
// pairs) runs on G lanes, G = 64 for w >= 1024, 16 for w >= 128, 4 for w >= 16,
// else 1 (measured against 512 / 64 / 8 and 2048 / 256 / 32 on the benzene
// unique build and the n-octane and (H2O)4 direct builds: 2.09 / 11.2 / 24.3 ms
// -> 2.00 / 10.6 / 24.2 ms, and 2.07 / 10.4 / 24.3 ms).  Lists are sorted by w
// descending, so the buckets are contiguous ranges [G=64 | G=16 | G=4 | G=1].
constexpr int NGROUP = 4;
constexpr int GROUP_SIZE[NGROUP] = {64, 16, 4, 1};
inline int group_bucket(int w) { return w >= 1024 ? 0 : w >= 128 ? 1 : w >= 16 ? 2 : 3; }
// Small plans (the Schwarz diagonal, the serving stream's molecules) are set
// by their longest units rather than by throughput: there wider groups pay
// (measured: the README molecules' stored SCFs 5-10 % slower with the
// thresholds above).
inline int group_bucket_small(int w) { return w >= 512 ? 0 : w >= 64 ? 1 : w >= 8 ? 2 : 3; }
// Direct J/K plans: narrower groups still (more units per wavefront share the
// J/K run merging; measured n-octane / cc-pVDZ direct build 10.6 -> 10.2 ms,
// (H2O)4 / cc-pVTZ 24.2 -> 24.0 ms with 2048 / 256 / 32 and the halved J/K
// slice target, then 10.2 -> 9.9 ms and 24.0 -> 23.9 ms with 4096 / 512 / 64).
inline int group_bucket_jk(int w) { return w >= 4096 ? 0 : w >= 512 ? 1 : w >= 64 ? 2 : 3; }

inline bool qc_family_a(int qc) {
    const int pb = qc_bra(qc), pk = qc_ket(qc);
    return runs_family_a(pc_la(pb), pc_lb(pb), pc_la(pk), pc_lb(pk));
}

struct QuartetLists {
    std::vector<uint2> list[NQCLASS][2];   // [class][0 = FP64, 1 = FP32]
    int bucket[NQCLASS][2][NGROUP] = {};   // quartets per lane-group bucket (family A; family B uses bucket 3)
    size_t total() const {
        size_t n = 0;
        for (int c = 0; c < NQCLASS; ++c) n += list[c][0].size() + list[c][1].size();
        return n;
    }
};

// Recounts the lane-group buckets of every family-A list with bucket(w) (the
// lists are sorted by w, so any monotone thresholds give contiguous buckets).
template <class Bucket>
inline void rebucket(const BasisPlan& bp, QuartetLists& ql, Bucket bucket) {
    for (int c = 0; c < NQCLASS; ++c) {
        if (!qc_family_a(c)) continue;
        for (int p = 0; p < 2; ++p) {
            for (int g = 0; g < NGROUP; ++g) ql.bucket[c][p][g] = 0;
            for (const uint2& e : ql.list[c][p]) ++ql.bucket[c][p][bucket(bp.hdr[e.x].nprim * bp.hdr[e.y].nprim)];
        }
    }
}

// Host threads for planning: one below 2^18 work items (thread start-up
// would cost more than it saves), else up to 16.
inline int plan_threads(size_t work) {
    if (work < ((size_t)1 << 18)) return 1;
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::clamp<unsigned>(hc ? hc : 1u, 1u, 16u);
}

// fn(task) for every task in [0, ntask), spread over up to nthreads threads
// (the caller's thread included) by a shared counter.  If a thread cannot be
// started, the threads already running take its tasks.  The first exception
// a task throws stops the hand-out of further tasks and is rethrown on the
// caller's thread once all threads have joined.
template <class F>
inline void parallel_tasks(int ntask, int nthreads, F&& fn) {
    if (nthreads <= 1 || ntask <= 1) {
        for (int t = 0; t < ntask; ++t) fn(t);
        return;
    }
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::exception_ptr err;
    auto worker = [&] {
        for (int t; (t = next.fetch_add(1)) < ntask;) {
            try {
                fn(t);
            } catch (...) {
                if (!failed.exchange(true)) err = std::current_exception();
                next.store(ntask);
                return;
            }
        }
    };
    std::vector<std::thread> th;
    const int n = std::min(nthreads, ntask);
    for (int i = 1; i < n; ++i) {
        try {
            th.emplace_back(worker);
        } catch (...) {
            break;
        }
    }
    worker();
    for (auto& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

// Screened canonical shell quartets (x >= y over canonical pair ids), split by
// class and precision.  A quartet is kept when Q_x Q_y >= threshold (no
// function quartet inside it can pass the function-level mask otherwise).
// In mixed precision, quartets with Q_x Q_y < split go to the FP32 list.
// Within a class, quartets are ordered by primitive work w (descending) so
// that lanes of a wavefront run loops of similar length.  With jk_order (J/K
// plans), quartets of equal w are further ordered by bra pair, then by the
// ket pair's first shell (kernel order), then by ket pair: neighbouring units
// of a family-A wavefront then share the J_ab and K_ac / K_bc targets, which
// the J/K epilogue adds up along the wavefront before one atomic per run.
//
// Two passes over the pair triangle (count, then place) and a stable counting
// sort on w per list: linear in the number of quartets (13.6 M for
// n-octane/cc-pVDZ at 1e-12), with no per-element reallocation.  Large
// triangles are cut into row ranges of equal area, one per thread; a thread's
// entries of a list follow those of the threads before it, so the lists come
// out in the same order as from one thread.  threads > 0 overrides the
// automatic thread count (tests compare thread counts).
inline void plan_quartets(const BasisPlan& bp, const std::vector<double>& Q, double thr, bool mixed, double split,
                          QuartetLists& ql, int threads = 0, bool jk_order = false) {
    constexpr int NL = 2 * NQCLASS;
    const int np = bp.npair;
    std::vector<int> nprim(np), pcls(np);
    for (int x = 0; x < np; ++x) {
        nprim[x] = bp.hdr[x].nprim;
        pcls[x] = bp.pclass[x];
    }
    // list id of a kept quartet: 2 * class + precision; -1 when screened
    auto list_of = [&](int x, int y, unsigned& bra, unsigned& ket) -> int {
        if (nprim[x] == 0 || nprim[y] == 0) return -1;
        const double qq = Q[x] * Q[y];
        if (qq < thr) return -1;
        int b = x, k = y;
        if (pcls[x] < pcls[y]) std::swap(b, k);
        bra = (unsigned)b;
        ket = (unsigned)k;
        return 2 * quartet_class(pcls[b], pcls[k]) + ((mixed && qq < split) ? 1 : 0);
    };
    const int T = threads > 0 ? threads : plan_threads((size_t)np * (np + 1) / 2);
    std::vector<int> row0(T + 1);   // thread t takes rows [row0[t], row0[t+1])
    for (int t = 0; t <= T; ++t) row0[t] = std::min(np, (int)std::lround(np * std::sqrt((double)t / T)));
    row0[T] = np;
    std::vector<size_t> cnt((size_t)T * NL, 0);
    parallel_tasks(T, T, [&](int t) {
        size_t* c = &cnt[(size_t)t * NL];
        for (int x = row0[t]; x < row0[t + 1]; ++x)
            for (int y = 0; y <= x; ++y) {
                unsigned b, k;
                const int l = list_of(x, y, b, k);
                if (l >= 0) ++c[l];
            }
    });
    // one buffer for all lists; the start of thread t's run in list l
    std::vector<size_t> lstart(NL + 1, 0), pos((size_t)T * NL);
    for (int l = 0; l < NL; ++l) {
        size_t run = lstart[l];
        for (int t = 0; t < T; ++t) {
            pos[(size_t)t * NL + l] = run;
            run += cnt[(size_t)t * NL + l];
        }
        lstart[l + 1] = run;
    }
    std::unique_ptr<uint2[]> raw(new uint2[std::max<size_t>(lstart[NL], 1)]);
    parallel_tasks(T, T, [&](int t) {
        size_t* p = &pos[(size_t)t * NL];
        for (int x = row0[t]; x < row0[t + 1]; ++x)
            for (int y = 0; y <= x; ++y) {
                unsigned b, k;
                const int l = list_of(x, y, b, k);
                if (l >= 0) raw[p[l]++] = make_uint2(b, k);
            }
    });
    int maxw = 0;
    for (int x = 0; x < np; ++x) maxw = std::max(maxw, nprim[x]);
    maxw = maxw * maxw;
    // jk_order: rank of each pair in (first shell, pair id) order
    std::vector<int> kr;
    if (jk_order) {
        std::vector<int> start((size_t)bp.nshell + 1, 0);
        for (int x = 0; x < np; ++x) ++start[(size_t)bp.hdr[x].s1 + 1];
        for (int sh = 0; sh < bp.nshell; ++sh) start[(size_t)sh + 1] += start[sh];
        kr.resize(np);
        for (int x = 0; x < np; ++x) kr[x] = start[bp.hdr[x].s1]++;
    }
    // stable counting sort of n entries from src to dst by key(e) in [0, nkey)
    auto counting_sort = [](const uint2* src, uint2* dst, size_t n, size_t nkey, auto key) {
        std::vector<size_t> hist(nkey + 1, 0);
        for (size_t i = 0; i < n; ++i) ++hist[key(src[i]) + 1];
        for (size_t k = 0; k < nkey; ++k) hist[k + 1] += hist[k];
        for (size_t i = 0; i < n; ++i) dst[hist[key(src[i])]++] = src[i];
    };
    parallel_tasks(NL, T, [&](int l) {
        const int c = l / 2, p = l % 2;
        const uint2* v = raw.get() + lstart[l];
        const size_t n = lstart[l + 1] - lstart[l];
        std::vector<uint2>& out = ql.list[c][p];
        out.resize(n);
        std::vector<uint2> byket, bybra;
        if (jk_order && n > 1) {   // least significant key first: ket rank, then bra
            byket.resize(n);
            bybra.resize(n);
            counting_sort(v, byket.data(), n, (size_t)np, [&](const uint2& e) { return (size_t)kr[e.y]; });
            counting_sort(byket.data(), bybra.data(), n, (size_t)np, [](const uint2& e) { return (size_t)e.x; });
            v = bybra.data();
        }
        // stable counting sort, w descending: slot of w starts after all larger w
        std::vector<size_t> hist((size_t)maxw + 2, 0);
        for (size_t i = 0; i < n; ++i) ++hist[(size_t)(maxw - nprim[v[i].x] * nprim[v[i].y])];
        size_t run = 0;
        for (size_t k = 0; k < hist.size(); ++k) {
            const size_t h = hist[k];
            hist[k] = run;
            run += h;
        }
        for (size_t i = 0; i < n; ++i) out[hist[(size_t)(maxw - nprim[v[i].x] * nprim[v[i].y])]++] = v[i];
        for (int g = 0; g < NGROUP; ++g) ql.bucket[c][p][g] = 0;
        const bool fa = qc_family_a(c);
        for (const uint2& e : out) ++ql.bucket[c][p][fa ? group_bucket(nprim[e.x] * nprim[e.y]) : NGROUP - 1];
    });
}

// Work units of one launch (one lane-group bucket of n quartets, sorted by
// primitive work w = bra x ket kept primitive pairs).  A launch lasts at
// least as long as its longest unit: a few very contracted quartets, or a
// launch of few quartets, leave most of the 256 CUs idle.  With R units
// resident on the device at once (`target` wavefronts: one unit per wavefront
// in family B, 64 / G units per wavefront in family A) the launch needs about
// total work / R per slot, so a quartet is cut into nslice = ceil(w / W)
// slices of its bra primitive loop (at most one per bra primitive pair) with
//   W = max(total work / R, w_min),
// w_min = one bra primitive (family B) or four primitive quartets per lane
// (family A).  Launches that fill the device keep their quartets whole
// except for outliers; slices add partial blocks in the epilogue (atomics).
// target <= 0 disables slicing.
//
// A unit is the quartet's uint2 list entry with the slice index in the top
// 8 bits of .x and the slice count in the top 8 bits of .y (pair ids < 2^24;
// a count of 0 reads as 1, so plain entries are whole quartets).
constexpr long long SLICE_TARGET = 2048;
constexpr int SLICE_MAX = 255;

// Ket primitive pairs per step of class qc's family-B kernel.
inline int qc_famb_kb(int qc) {
    const int pb = qc_bra(qc), pk = qc_ket(qc);
    return famb_kb(pc_la(pb) + pc_lb(pb) + pc_la(pk) + pc_lb(pk));
}

inline void expand_units(const BasisPlan& bp, bool fam_a, int G, const uint2* q, int n, long long target,
                         std::vector<uint2>& out, int kb = 1) {
    if (n <= 0) return;
    long long total = 0, wmax = 0;
    for (int k = 0; k < n; ++k) {
        const long long w = (long long)bp.hdr[q[k].x].nprim * bp.hdr[q[k].y].nprim;
        total += w;
        wmax = std::max(wmax, w);
    }
    const long long U = target <= 0 ? 1 : fam_a ? target * 64 / G : target;
    const long long W = std::max<long long>((total + U - 1) / U, 1);
    if (target <= 0 || W >= wmax) {   // no quartet is longer than a slot's share: keep all whole
        out.insert(out.end(), q, q + n);
        return;
    }
    auto emit = [&](const uint2& e, long long S) {
        if (S <= 1) {
            out.push_back(e);
            return;
        }
        for (long long sl = 0; sl < S; ++sl)
            out.push_back(make_uint2(e.x | ((unsigned)sl << 24), e.y | ((unsigned)S << 24)));
    };
    for (int k = 0; k < n; ++k) {
        const int nb = bp.hdr[q[k].x].nprim, nk = bp.hdr[q[k].y].nprim;
        const long long w = (long long)nb * nk;
        if (fam_a) {
            const long long Wq = std::max<long long>(W, 4LL * G);
            const long long cap = std::max<long long>(1, std::min<long long>(w / (4LL * G), nb));
            emit(q[k], std::max<long long>(1, std::min<long long>(std::min<long long>((w + Wq - 1) / Wq, cap), SLICE_MAX)));
            continue;
        }
        // family B: whole ket loops per bra pair while a slot's share covers
        // one; below that, ket chunks of at least two steps (famb_range)
        if (W >= nk || nk <= 2 * kb) {
            const long long Wq = std::max<long long>(W, nk);
            emit(q[k], std::max<long long>(1, std::min<long long>(std::min<long long>((w + Wq - 1) / Wq, nb), SLICE_MAX)));
            continue;
        }
        long long chunk = std::max<long long>(2 * kb, (W + kb - 1) / kb * kb);
        long long nkc = (nk + chunk - 1) / chunk;
        while (nkc > 1 && (long long)nb * nkc > SLICE_MAX) nkc = (nkc + 1) / 2;
        emit(q[k], nkc > 1 ? nb * nkc : std::min<long long>(nb, SLICE_MAX));
    }
}

// Diagonal quartets (x|x) for the Schwarz bounds.
inline void plan_diag(const BasisPlan& bp, std::vector<uint2> (&lists)[NQCLASS]) {
    for (int c = 0; c < NQCLASS; ++c) lists[c].clear();
    for (int x = 0; x < bp.npair; ++x) {
        if (bp.hdr[x].nprim == 0) continue;
        lists[quartet_class(bp.pclass[x], bp.pclass[x])].push_back(make_uint2((unsigned)x, (unsigned)x));
    }
}

// F_n(T_i) on the grid T_i = i h, n = 0..BOYS_NCOL-1, in long double: the top
// column from the positive series
//   F_n(T) = e^{-T} sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)),
// the others by the (stable) downward recursion F_n = (2T F_{n+1} + e^{-T}) / (2n+1).
inline void boys_table(std::vector<double>& tab) {
    tab.assign((size_t)BOYS_NGRID * BOYS_NCOL, 0.0);
    for (int i = 0; i < BOYS_NGRID; ++i) {
        const long double T = (long double)i * (long double)BOYS_H;
        const long double e = expl(-T);
        const int n = BOYS_NCOL - 1;
        long double term = 1.0L / (2.0L * n + 1.0L), sum = term;
        for (int k = 1; k < 4000; ++k) {
            term *= 2.0L * T / (2.0L * n + 2.0L * k + 1.0L);
            sum += term;
            if (term < 1e-24L * sum) break;
        }
        long double f = e * sum;
        tab[(size_t)i * BOYS_NCOL + n] = (double)f;
        for (int m = n - 1; m >= 0; --m) {
            f = (2.0L * T * f + e) / (2.0L * m + 1.0L);
            tab[(size_t)i * BOYS_NCOL + m] = (double)f;
        }
    }
}

inline HermTab herm_tab() {
    HermTab h{};
    int w = 0;
    for (int s = 0; s <= 12; ++s)
        for (int t = 0; t <= s; ++t)
            for (int u = 0; u <= s - t; ++u) {
                h.t[w] = (unsigned char)t;
                h.u[w] = (unsigned char)u;
                h.v[w] = (unsigned char)(s - t - u);
                ++w;
            }
    return h;
}

}  // namespace hfx
