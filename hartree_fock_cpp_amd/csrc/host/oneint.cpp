// oneint.cpp -- see oneint.hpp.
#include "oneint.hpp"

#include <cmath>
#include <mutex>
#include <vector>

namespace hfxh {

namespace {

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr int NCOL = 17 + 8;        // F_0 .. F_24 tabulated
constexpr double H = 0.05;
constexpr int NGRID = 801;          // T = 0 .. 40

const std::vector<double>& boys_tab() {
    static std::vector<double> tab;
    static std::once_flag once;
    std::call_once(once, []() {
        tab.assign((size_t)NGRID * NCOL, 0.0);
        for (int i = 0; i < NGRID; ++i) {
            const long double T = (long double)i * (long double)H;
            const long double e = expl(-T);
            for (int n = 0; n < NCOL; ++n) {
                long double term = 1.0L / (2.0L * n + 1.0L), sum = term;
                for (int k = 1; k < 4000; ++k) {
                    term *= 2.0L * T / (2.0L * n + 2.0L * k + 1.0L);
                    sum += term;
                    if (term < 1e-24L * sum) break;
                }
                tab[(size_t)i * NCOL + n] = (double)(e * sum);
            }
        }
    });
    return tab;
}

// 1-D Hermite expansion E[i][j][t] (i <= la, j <= lb, t <= i + j) including the
// Gaussian prefactor exp(-mu X_AB^2) in E[0][0][0].
struct E1D {
    int la, lb;
    std::vector<double> e;   // (la+1)(lb+1)(la+lb+1)
    double& at(int i, int j, int t) { return e[((size_t)i * (lb + 1) + j) * (la + lb + 1) + t]; }
    double get(int i, int j, int t) const {
        if (i < 0 || j < 0 || t < 0 || t > i + j) return 0.0;
        return e[((size_t)i * (lb + 1) + j) * (la + lb + 1) + t];
    }
    E1D(int la_, int lb_, double a, double b, double A, double B) : la(la_), lb(lb_) {
        e.assign((size_t)(la + 1) * (lb + 1) * (la + lb + 1), 0.0);
        const double p = a + b, mu = a * b / p, X = A - B;
        const double PA = -b / p * X, PB = a / p * X, i2p = 0.5 / p;
        at(0, 0, 0) = std::exp(-mu * X * X);
        for (int i = 1; i <= la; ++i)
            for (int t = 0; t <= i; ++t)
                at(i, 0, t) = i2p * get(i - 1, 0, t - 1) + PA * get(i - 1, 0, t) + (t + 1) * get(i - 1, 0, t + 1);
        for (int j = 1; j <= lb; ++j)
            for (int i = 0; i <= la; ++i)
                for (int t = 0; t <= i + j; ++t)
                    at(i, j, t) = i2p * get(i, j - 1, t - 1) + PB * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1);
    }
};

inline int hidx(int t, int u, int v) {
    const int s = t + u + v;
    int off = s * (s + 1) * (s + 2) / 6;
    for (int a = 0; a < t; ++a) off += (s - a + 1);
    return off + u;
}

// R_tuv for t+u+v <= L with R^{(n)}_000 = (-2 rho)^n F_n(rho |X|^2)
void hermite_R(int L, double rho, double X, double Y, double Z, std::vector<double>& R) {
    const int NR = (L + 1) * (L + 2) * (L + 3) / 6;
    R.assign(NR, 0.0);
    double F[32];
    boys_host(L, rho * (X * X + Y * Y + Z * Z), F);
    std::vector<double> a(L + 1);
    double f = 1.0;
    for (int n = 0; n <= L; ++n) {
        a[n] = f * F[n];
        f *= -2.0 * rho;
    }
    R[0] = a[L];
    for (int n = L - 1; n >= 0; --n) {
        for (int s = L - n; s >= 1; --s)
            for (int t = s; t >= 0; --t)
                for (int u = s - t; u >= 0; --u) {
                    const int v = s - t - u;
                    double val;
                    if (t > 0)
                        val = X * R[hidx(t - 1, u, v)] + (t > 1 ? (t - 1) * R[hidx(t - 2, u, v)] : 0.0);
                    else if (u > 0)
                        val = Y * R[hidx(t, u - 1, v)] + (u > 1 ? (u - 1) * R[hidx(t, u - 2, v)] : 0.0);
                    else
                        val = Z * R[hidx(t, u, v - 1)] + (v > 1 ? (v - 1) * R[hidx(t, u, v - 2)] : 0.0);
                    R[hidx(t, u, v)] = val;
                }
        R[0] = a[n];
    }
}

template <class F>
void for_pairs(const FlatBasis& b, double* M, F&& prim) {
    const int K = b.K;
    for (int m = 0; m < K; ++m)
        for (int n = 0; n <= m; ++n) {
            double s = 0.0;
            for (int i = b.poff[m]; i < b.poff[m + 1]; ++i)
                for (int j = b.poff[n]; j < b.poff[n + 1]; ++j)
                    s += b.coeff[i] * b.coeff[j] * b.norm[i] * b.norm[j] * prim(m, n, i, j);
            M[(size_t)m * K + n] = s;
            M[(size_t)n * K + m] = s;
        }
}

}  // namespace

void boys_host(int mmax, double T, double* F) {
    const std::vector<double>& tab = boys_tab();
    if (T < 40.0) {
        const int i = (int)(T * 20.0 + 0.5);
        const double d = i * H - T;
        const double* row = &tab[(size_t)i * NCOL + mmax];
        double f = row[7];
        for (int k = 6; k >= 0; --k) f = row[k] + f * d / (k + 1);
        F[mmax] = f;
        const double e = std::exp(-T);
        for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        F[0] = 0.5 * std::sqrt(PI / T);
        const double e = std::exp(-T);
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
    }
}

void overlap_matrix(const FlatBasis& b, double* S) {
    for_pairs(b, S, [&](int m, int n, int i, int j) {
        const double a = b.alpha[i], c = b.alpha[j], p = a + c;
        double v = std::pow(PI / p, 1.5);
        for (int x = 0; x < 3; ++x) {
            const int la = b.lmn[3 * m + x], lb = b.lmn[3 * n + x];
            E1D E(la, lb, a, c, b.xyz[3 * m + x], b.xyz[3 * n + x]);
            v *= E.get(la, lb, 0);
        }
        return v;
    });
}

void kinetic_matrix(const FlatBasis& b, double* T) {
    for_pairs(b, T, [&](int m, int n, int i, int j) {
        const double a = b.alpha[i], c = b.alpha[j], p = a + c;
        const double sq = std::sqrt(PI / p);
        double S1[3], T1[3];
        for (int x = 0; x < 3; ++x) {
            const int l = b.lmn[3 * m + x], k = b.lmn[3 * n + x];
            E1D E(l, k + 2, a, c, b.xyz[3 * m + x], b.xyz[3 * n + x]);
            auto s1 = [&](int kk) { return kk < 0 ? 0.0 : E.get(l, kk, 0) * sq; };
            S1[x] = s1(k);
            T1[x] = c * (2 * k + 1) * s1(k) - 2.0 * c * c * s1(k + 2) - 0.5 * k * (k - 1) * s1(k - 2);
        }
        return T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2];
    });
}

void attraction_matrix(const FlatBasis& b, int natom, const int* Z, const double* Rn, double* V) {
    std::vector<double> R;
    for_pairs(b, V, [&](int m, int n, int i, int j) {
        const double a = b.alpha[i], c = b.alpha[j], p = a + c;
        const int* lm = &b.lmn[3 * m];
        const int* ln = &b.lmn[3 * n];
        E1D Ex(lm[0], ln[0], a, c, b.xyz[3 * m], b.xyz[3 * n]);
        E1D Ey(lm[1], ln[1], a, c, b.xyz[3 * m + 1], b.xyz[3 * n + 1]);
        E1D Ez(lm[2], ln[2], a, c, b.xyz[3 * m + 2], b.xyz[3 * n + 2]);
        double P[3];
        for (int x = 0; x < 3; ++x) P[x] = (a * b.xyz[3 * m + x] + c * b.xyz[3 * n + x]) / p;
        const int L = lm[0] + ln[0] + lm[1] + ln[1] + lm[2] + ln[2];
        double tot = 0.0;
        for (int at = 0; at < natom; ++at) {
            hermite_R(L, p, P[0] - Rn[3 * at], P[1] - Rn[3 * at + 1], P[2] - Rn[3 * at + 2], R);
            double s = 0.0;
            for (int t = 0; t <= lm[0] + ln[0]; ++t)
                for (int u = 0; u <= lm[1] + ln[1]; ++u)
                    for (int v = 0; v <= lm[2] + ln[2]; ++v)
                        s += Ex.get(lm[0], ln[0], t) * Ey.get(lm[1], ln[1], u) * Ez.get(lm[2], ln[2], v) * R[hidx(t, u, v)];
            tot += -Z[at] * s;
        }
        return 2.0 * PI / p * tot;
    });
}

}  // namespace hfxh
