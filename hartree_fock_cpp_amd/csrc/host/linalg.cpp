// linalg.cpp -- see linalg.hpp.
#include "linalg.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace hfxh {

void gemm(int n, const double* A, bool ta, const double* B, bool tb, double* C) {
    // i-k-j loop with explicit transposes for cache-friendly inner loops
    std::vector<double> At, Bt;
    const double* a = A;
    const double* b = B;
    if (ta) {
        At.resize((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) At[(size_t)i * n + j] = A[(size_t)j * n + i];
        a = At.data();
    }
    if (tb) {
        Bt.resize((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) Bt[(size_t)i * n + j] = B[(size_t)j * n + i];
        b = Bt.data();
    }
    // C must not overlap A or B (callers pass distinct matrices): the inner
    // loop then vectorises
    std::fill(C, C + (size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double* __restrict__ c = C + (size_t)i * n;
        for (int k = 0; k < n; ++k) {
            const double aik = a[(size_t)i * n + k];
            if (aik == 0.0) continue;
            const double* __restrict__ bk = b + (size_t)k * n;
            for (int j = 0; j < n; ++j) c[j] += aik * bk[j];
        }
    }
}

namespace {

// Householder reduction of symmetric A (row-major, n x n, overwritten) to
// tridiagonal form T = Q^T A Q.  d: diagonal, e[i] = T(i, i+1).  Q is
// accumulated column-major into Qc (Qc[j*n + i] = Q(i, j)) so that the later
// Givens updates touch contiguous memory.
void tridiagonalize(int n, std::vector<double>& A, std::vector<double>& d, std::vector<double>& e,
                    std::vector<double>& Qc) {
    d.assign(n, 0.0);
    e.assign(n, 0.0);
    Qc.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Qc[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), w(n), s(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;   // length of x = A[k+1.., k]
        double xn = 0.0;
        for (int i = 0; i < m; ++i) {
            v[i] = A[(size_t)(k + 1 + i) * n + k];
            xn += v[i] * v[i];
        }
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const double alpha = v[0] > 0 ? -xn : xn;
        v[0] -= alpha;
        double vv = 0.0;
        for (int i = 0; i < m; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        // p = beta * A' v  (A' = trailing block; four partial sums so the
        // dot products vectorise)
        const double* __restrict__ vp = v.data();
        for (int i = 0; i < m; ++i) {
            const double* __restrict__ row = &A[(size_t)(k + 1 + i) * n + k + 1];
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int j = 0;
            for (; j + 4 <= m; j += 4) {
                s0 += row[j] * vp[j];
                s1 += row[j + 1] * vp[j + 1];
                s2 += row[j + 2] * vp[j + 2];
                s3 += row[j + 3] * vp[j + 3];
            }
            for (; j < m; ++j) s0 += row[j] * vp[j];
            p[i] = beta * ((s0 + s1) + (s2 + s3));
        }
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * v[i];
        const double c = 0.5 * beta * pv;
        for (int i = 0; i < m; ++i) w[i] = p[i] - c * v[i];
        // A' -= v w^T + w v^T
        const double* __restrict__ wp = w.data();
        for (int i = 0; i < m; ++i) {
            double* __restrict__ row = &A[(size_t)(k + 1 + i) * n + k + 1];
            const double vi = vp[i], wi = wp[i];
            for (int j = 0; j < m; ++j) row[j] -= vi * wp[j] + wi * vp[j];
        }
        A[(size_t)(k + 1) * n + k] = alpha;
        A[(size_t)k * n + k + 1] = alpha;
        for (int i = 1; i < m; ++i) {
            A[(size_t)(k + 1 + i) * n + k] = 0.0;
            A[(size_t)k * n + k + 1 + i] = 0.0;
        }
        // Q <- Q H, H = I - beta v v^T acting on columns k+1..n-1: s = Q v, then
        // Q -= beta s v^T, both passes streaming whole columns of Qc
        std::fill(s.begin(), s.end(), 0.0);
        double* __restrict__ sp = s.data();
        for (int j = 0; j < m; ++j) {
            const double vj = vp[j];
            const double* __restrict__ col = &Qc[(size_t)(k + 1 + j) * n];
            for (int r = 0; r < n; ++r) sp[r] += col[r] * vj;
        }
        for (int j = 0; j < m; ++j) {
            const double bv = beta * vp[j];
            double* __restrict__ col = &Qc[(size_t)(k + 1 + j) * n];
            for (int r = 0; r < n; ++r) col[r] -= sp[r] * bv;
        }
    }
    for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
    for (int i = 0; i + 1 < n; ++i) e[i] = A[(size_t)i * n + i + 1];
}

// Implicit symmetric QR with Wilkinson shift on (d, e); rotations applied to
// the columns of Qc (column-major).
void tridiag_qr(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& Qc) {
    const double eps = 2.220446049250313e-16;
    int iter = 0;
    const int max_iter = 60 * std::max(n, 1);
    int q = n - 1;
    while (q > 0 && iter < max_iter) {
        // deflate from the bottom
        for (int i = 0; i < q; ++i)
            if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
        while (q > 0 && e[q - 1] == 0.0) --q;
        if (q == 0) break;
        int p = q - 1;
        while (p > 0 && e[p - 1] != 0.0) --p;
        // Wilkinson shift from the trailing 2x2 of block [p..q]
        const double a = d[q - 1], b = e[q - 1], c = d[q];
        const double delta = 0.5 * (a - c);
        const double sgn = delta >= 0 ? 1.0 : -1.0;
        const double mu = c - b * b / (delta + sgn * std::hypot(delta, b));
        double x = d[p] - mu, z = e[p];
        for (int k = p; k < q; ++k) {
            // Givens rotation G(k, k+1) with [c s; -s c]^T [x; z] = [r; 0]
            double cs, sn;
            if (z == 0.0) {
                cs = 1.0;
                sn = 0.0;
            } else if (std::fabs(z) > std::fabs(x)) {
                const double t = -x / z;
                sn = 1.0 / std::sqrt(1.0 + t * t);
                cs = sn * t;
            } else {
                const double t = -z / x;
                cs = 1.0 / std::sqrt(1.0 + t * t);
                sn = cs * t;
            }
            // apply G^T T G to rows/cols k, k+1
            if (k > p) e[k - 1] = cs * e[k - 1] - sn * z;
            const double dk = d[k], dk1 = d[k + 1], ek = e[k];
            const double t1 = cs * cs * dk - 2.0 * cs * sn * ek + sn * sn * dk1;
            const double t2 = sn * sn * dk + 2.0 * cs * sn * ek + cs * cs * dk1;
            const double t3 = cs * sn * (dk - dk1) + (cs * cs - sn * sn) * ek;
            d[k] = t1;
            d[k + 1] = t2;
            e[k] = t3;
            if (k + 1 < q) {
                x = e[k];
                z = -sn * e[k + 1];
                e[k + 1] = cs * e[k + 1];
            }
            // Q <- Q G (distinct columns: no aliasing, the loop vectorises)
            double* __restrict__ qk = &Qc[(size_t)k * n];
            double* __restrict__ qk1 = &Qc[(size_t)(k + 1) * n];
            for (int r = 0; r < n; ++r) {
                const double u = qk[r], w = qk1[r];
                qk[r] = cs * u - sn * w;
                qk1[r] = sn * u + cs * w;
            }
        }
        ++iter;
    }
}

}  // namespace

void sym_eig(int n, const double* A, double* w, double* V) {
    std::vector<double> a(A, A + (size_t)n * n), d, e, Qc;
    // symmetrise against round-off in the caller's matrix
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double s = 0.5 * (a[(size_t)i * n + j] + a[(size_t)j * n + i]);
            a[(size_t)i * n + j] = s;
            a[(size_t)j * n + i] = s;
        }
    tridiagonalize(n, a, d, e, Qc);
    tridiag_qr(n, d, e, Qc);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int k = 0; k < n; ++k) {
        const int src = order[k];
        w[k] = d[src];
        for (int r = 0; r < n; ++r) V[(size_t)r * n + k] = Qc[(size_t)src * n + r];
    }
}

void solve_fullpiv(int n, const double* A, const double* b, double* x) {
    std::vector<double> a(A, A + (size_t)n * n), rhs(b, b + n);
    std::vector<int> colp(n);
    std::iota(colp.begin(), colp.end(), 0);
    double maxpiv = 0.0;
    int rank = n;
    for (int k = 0; k < n; ++k) {
        int pr = k, pc = k;
        double best = -1.0;
        for (int i = k; i < n; ++i)
            for (int j = k; j < n; ++j) {
                const double v = std::fabs(a[(size_t)i * n + j]);
                if (v > best) {
                    best = v;
                    pr = i;
                    pc = j;
                }
            }
        if (k == 0) maxpiv = best;
        if (best <= 1e-15 * maxpiv || best == 0.0) {
            rank = k;
            break;
        }
        if (pr != k) {
            for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)pr * n + j]);
            std::swap(rhs[k], rhs[pr]);
        }
        if (pc != k) {
            for (int i = 0; i < n; ++i) std::swap(a[(size_t)i * n + k], a[(size_t)i * n + pc]);
            std::swap(colp[k], colp[pc]);
        }
        const double piv = a[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const double f = a[(size_t)i * n + k] / piv;
            if (f == 0.0) continue;
            a[(size_t)i * n + k] = 0.0;
            for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
            rhs[i] -= f * rhs[k];
        }
    }
    std::vector<double> y(n, 0.0);
    for (int k = rank - 1; k >= 0; --k) {
        double s = rhs[k];
        for (int j = k + 1; j < rank; ++j) s -= a[(size_t)k * n + j] * y[j];
        y[k] = s / a[(size_t)k * n + k];
    }
    for (int k = 0; k < n; ++k) x[colp[k]] = y[k];
}

}  // namespace hfxh
