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
// With refl != nullptr, Q is not formed: the reflector of step k (v over rows
// k+1..n-1, scale beta[k]; beta 0 = identity) is kept in row k of refl.
// Only the lower triangle of A is read and updated.
void tridiagonalize(int n, std::vector<double>& A, std::vector<double>& d, std::vector<double>& e,
                    std::vector<double>& Qc, std::vector<double>* refl = nullptr, std::vector<double>* betas = nullptr) {
    d.assign(n, 0.0);
    e.assign(n, 0.0);
    if (refl) {
        refl->assign((size_t)n * n, 0.0);
        betas->assign(n, 0.0);
    } else {
        Qc.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) Qc[(size_t)i * n + i] = 1.0;
    }
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
        // p = beta * A' v from the lower triangle of the trailing block A':
        // row i gives its dot with v_0..v_i and, transposed, adds A'_ij v_i
        // to p_j for j < i (one pass over half the block)
        const double* __restrict__ vp = v.data();
        double* __restrict__ pp = p.data();
        std::fill(pp, pp + m, 0.0);
        for (int i = 0; i < m; ++i) {
            const double* __restrict__ row = &A[(size_t)(k + 1 + i) * n + k + 1];
            const double vi = vp[i];
            double s0 = 0.0, s1 = 0.0;
            int j = 0;
            for (; j + 2 <= i; j += 2) {
                s0 += row[j] * vp[j];
                s1 += row[j + 1] * vp[j + 1];
                pp[j] += row[j] * vi;
                pp[j + 1] += row[j + 1] * vi;
            }
            for (; j < i; ++j) {
                s0 += row[j] * vp[j];
                pp[j] += row[j] * vi;
            }
            pp[i] += (s0 + s1) + row[i] * vi;
        }
        for (int i = 0; i < m; ++i) pp[i] *= beta;
        double pv = 0.0;
        for (int i = 0; i < m; ++i) pv += p[i] * v[i];
        const double c = 0.5 * beta * pv;
        for (int i = 0; i < m; ++i) w[i] = p[i] - c * v[i];
        // A' -= v w^T + w v^T on the lower triangle (the upper one is stale
        // from here on and never read)
        const double* __restrict__ wp = w.data();
        for (int i = 0; i < m; ++i) {
            double* __restrict__ row = &A[(size_t)(k + 1 + i) * n + k + 1];
            const double vi = vp[i], wi = wp[i];
            for (int j = 0; j <= i; ++j) row[j] -= vi * wp[j] + wi * vp[j];
        }
        A[(size_t)(k + 1) * n + k] = alpha;
        for (int i = 1; i < m; ++i) A[(size_t)(k + 1 + i) * n + k] = 0.0;
        if (refl) {
            std::copy(v.begin(), v.begin() + m, refl->begin() + (size_t)k * n + k + 1);
            (*betas)[k] = beta;
            continue;
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
    for (int i = 0; i + 1 < n; ++i) e[i] = A[(size_t)(i + 1) * n + i];
}

// Implicit symmetric QR with Wilkinson shift on (d, e); rotations applied to
// the columns of Qc (column-major) unless vectors is false.
void tridiag_qr(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& Qc, bool vectors = true) {
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
            if (!vectors) continue;
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

// Eigenvalues of the symmetric tridiagonal (d, e) by the root-free implicit
// QR step of Pal, Walker and Kahan (B. N. Parlett, The Symmetric Eigenvalue
// Problem, sec. 8-15; the scheme of LAPACK dsterf): the sweep carries the
// squared off-diagonals and the squared rotation cosine / sine, so it takes
// no square root per rotation (one per sweep, for the Wilkinson shift of the
// trailing 2 x 2).  Eigenvalues overwrite d, unsorted; e is destroyed.
void tridiag_eigenvalues(int n, std::vector<double>& d, std::vector<double>& e) {
    const double eps = 2.220446049250313e-16;
    const double eps2 = eps * eps;
    std::vector<double>& e2 = e;
    for (int i = 0; i + 1 < n; ++i) e2[i] *= e2[i];
    int iter = 0;
    const int max_iter = 60 * std::max(n, 1);
    int q = n - 1;
    while (q > 0 && iter < max_iter) {
        // deflate: |e_i| <= eps (|d_i| + |d_i+1|), squared
        for (int i = 0; i < q; ++i) {
            const double t = std::fabs(d[i]) + std::fabs(d[i + 1]);
            if (e2[i] <= eps2 * t * t) e2[i] = 0.0;
        }
        while (q > 0 && e2[q - 1] == 0.0) --q;
        if (q == 0) break;
        int p = q - 1;
        while (p > 0 && e2[p - 1] != 0.0) --p;
        // Wilkinson shift from the trailing 2x2 of block [p..q]
        const double a = d[q - 1], c = d[q], b2 = e2[q - 1];
        const double delta = 0.5 * (a - c);
        const double sgn = delta >= 0 ? 1.0 : -1.0;
        const double sigma = c - b2 / (delta + sgn * std::hypot(delta, std::sqrt(b2)));
        // e2[k] != 0 inside the block, so R > 0 below
        double cs = 1.0, sn = 0.0;
        double gamma = d[p] - sigma;
        double P = gamma * gamma;
        for (int k = p; k < q; ++k) {
            const double bb = e2[k];
            const double R = P + bb;
            if (k > p) e2[k - 1] = sn * R;
            const double oldc = cs;
            cs = P / R;
            sn = bb / R;
            const double oldgam = gamma;
            const double alpha = d[k + 1];
            gamma = cs * (alpha - sigma) - sn * oldgam;
            d[k] = oldgam + (alpha - gamma);
            P = cs != 0.0 ? gamma * gamma / cs : oldc * bb;
        }
        e2[q - 1] = sn * P;
        d[q] = gamma + sigma;
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

namespace {

// LU factorisation with partial pivoting of the tridiagonal T - shift I
// (diagonal d, off-diagonal e) and solves with it (the dgttrf / dgttrs
// scheme: row interchanges fill a second superdiagonal).  Zero pivots are
// replaced by `tiny` so that inverse iteration can solve at an eigenvalue.
struct TriLU {
    std::vector<double> dl, dd, du, du2;
    std::vector<int> swap;   // 1 when rows i, i+1 were interchanged
    int n = 0;
    void factor(const std::vector<double>& d, const std::vector<double>& e, double shift, double tiny) {
        n = (int)d.size();
        dd.resize(n);
        dl.assign(n, 0.0);
        du.assign(n, 0.0);
        du2.assign(n, 0.0);
        swap.assign(n, 0);
        for (int i = 0; i < n; ++i) dd[i] = d[i] - shift;
        for (int i = 0; i + 1 < n; ++i) dl[i] = du[i] = e[i];
        for (int i = 0; i + 1 < n; ++i) {
            if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
                if (dd[i] == 0.0) dd[i] = tiny;
                const double f = dl[i] / dd[i];
                dl[i] = f;
                dd[i + 1] -= f * du[i];
            } else {
                const double f = dd[i] / dl[i];
                dd[i] = dl[i];
                dl[i] = f;
                const double t = du[i];
                du[i] = dd[i + 1];
                dd[i + 1] = t - f * dd[i + 1];
                if (i + 2 < n) {
                    du2[i] = du[i + 1];
                    du[i + 1] = -f * du[i + 1];
                }
                swap[i] = 1;
            }
        }
        if (n > 0 && dd[n - 1] == 0.0) dd[n - 1] = tiny;
    }
    void solve(double* b) const {
        for (int i = 0; i + 1 < n; ++i) {
            if (swap[i]) {
                const double t = b[i];
                b[i] = b[i + 1];
                b[i + 1] = t - dl[i] * b[i];
            } else {
                b[i + 1] -= dl[i] * b[i];
            }
        }
        for (int i = n - 1; i >= 0; --i) {
            double x = b[i];
            if (i + 1 < n) x -= du[i] * b[i + 1];
            if (i + 2 < n) x -= du2[i] * b[i + 2];
            b[i] = x / dd[i];
        }
    }
};

}  // namespace

void sym_eig_lowest(int n, const double* A, int m, double* w, double* V) {
    m = std::max(0, std::min(m, n));
    if (n <= 2 || m == n) {
        sym_eig(n, A, w, V);
        return;
    }
    std::vector<double> a(A, A + (size_t)n * n), d, e, Qc, refl, betas;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double s = 0.5 * (a[(size_t)i * n + j] + a[(size_t)j * n + i]);
            a[(size_t)i * n + j] = s;
            a[(size_t)j * n + i] = s;
        }
    tridiagonalize(n, a, d, e, Qc, &refl, &betas);
    // eigenvalues: root-free QR on a copy of T
    std::vector<double> dq(d), eq(e);
    tridiag_eigenvalues(n, dq, eq);
    std::sort(dq.begin(), dq.end());
    for (int k = 0; k < n; ++k) w[k] = dq[k];
    // eigenvectors of T for the m lowest eigenvalues by inverse iteration,
    // orthogonalised within clusters (the dstein scheme)
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i)
        tnorm = std::max(tnorm, std::fabs(d[i]) + (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0));
    if (tnorm == 0.0) tnorm = 1.0;
    const double eps = 2.220446049250313e-16;
    const double ortol = 1e-3 * tnorm, pertol = 10.0 * eps * tnorm;
    std::vector<double> Z((size_t)m * n), x(n);
    TriLU lu;
    int c0 = 0;                 // first index of the current cluster
    double shift_prev = 0.0;
    for (int k = 0; k < m; ++k) {
        double shift = w[k];
        if (k > 0) {
            if (w[k] - w[k - 1] > ortol) c0 = k;
            if (shift - shift_prev < pertol) shift = shift_prev + pertol;
        }
        shift_prev = shift;
        lu.factor(d, e, shift, eps * tnorm);
        // deterministic start vector
        unsigned long long r = 0x9E3779B97F4A7C15ull * (unsigned long long)(k + 1);
        for (int i = 0; i < n; ++i) {
            r ^= r << 13;
            r ^= r >> 7;
            r ^= r << 17;
            x[i] = (double)(r >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
        for (int it = 0; it < 4; ++it) {
            double xm = 0.0;
            for (int i = 0; i < n; ++i) xm = std::max(xm, std::fabs(x[i]));
            if (xm > 0.0)
                for (int i = 0; i < n; ++i) x[i] *= 1.0 / xm;
            lu.solve(x.data());
            for (int j = c0; j < k; ++j) {   // modified Gram-Schmidt against the cluster
                const double* z = &Z[(size_t)j * n];
                double dot = 0.0;
                for (int i = 0; i < n; ++i) dot += z[i] * x[i];
                for (int i = 0; i < n; ++i) x[i] -= dot * z[i];
            }
            double nrm = 0.0;
            for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
            nrm = std::sqrt(nrm);
            if (nrm == 0.0) {
                x.assign(n, 0.0);
                x[k] = 1.0;
                continue;
            }
            for (int i = 0; i < n; ++i) x[i] /= nrm;
        }
        std::copy(x.begin(), x.end(), Z.begin() + (size_t)k * n);
    }
    // back to A's basis: q = H_0 H_1 ... H_{n-3} z, four vectors per pass
    // over each reflector (four independent dot products in flight)
    int k = 0;
    for (; k + 4 <= m; k += 4) {
        double* __restrict__ z0 = &Z[(size_t)k * n];
        double* __restrict__ z1 = z0 + n;
        double* __restrict__ z2 = z1 + n;
        double* __restrict__ z3 = z2 + n;
        for (int j = n - 3; j >= 0; --j) {
            const double beta = betas[j];
            if (beta == 0.0) continue;
            const double* __restrict__ v = &refl[(size_t)j * n];
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            for (int i = j + 1; i < n; ++i) {
                s0 += v[i] * z0[i];
                s1 += v[i] * z1[i];
                s2 += v[i] * z2[i];
                s3 += v[i] * z3[i];
            }
            s0 *= beta;
            s1 *= beta;
            s2 *= beta;
            s3 *= beta;
            for (int i = j + 1; i < n; ++i) {
                z0[i] -= s0 * v[i];
                z1[i] -= s1 * v[i];
                z2[i] -= s2 * v[i];
                z3[i] -= s3 * v[i];
            }
        }
    }
    for (; k < m; ++k) {
        double* __restrict__ z = &Z[(size_t)k * n];
        for (int j = n - 3; j >= 0; --j) {
            const double beta = betas[j];
            if (beta == 0.0) continue;
            const double* __restrict__ v = &refl[(size_t)j * n];
            double s = 0.0;
            for (int i = j + 1; i < n; ++i) s += v[i] * z[i];
            s *= beta;
            for (int i = j + 1; i < n; ++i) z[i] -= s * v[i];
        }
    }
    for (size_t i = 0; i < (size_t)n * n; ++i) V[i] = 0.0;
    for (int k = 0; k < m; ++k)
        for (int r = 0; r < n; ++r) V[(size_t)r * n + k] = Z[(size_t)k * n + r];
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
