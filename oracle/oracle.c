/*
 * oracle.c -- CPU restatement of the reference's ERI / one-electron / Fock-build
 * arithmetic, used ONLY as test infrastructure (the checker).  Nothing in the
 * product (hartree_fock_cpp_amd/) links or calls this file; tests/, the smoke
 * check and bench.py's cpu_baseline leg are its only users.
 *
 * Every function restates one reference function, in the reference's own
 * operation order, so that its results agree with the compiled reference to
 * the last few ulps.  Citations are to KLYYCYNN/hartree_fock_cpp.
 *
 * Parity pins (what checks this file, see tests/test_oracle.py):
 *   orc_boys_ref          -> compiled reference boys() (oracle/_ref), mpmath table
 *   orc_boys_exact        -> mpmath table (tests/golden/boys_mpmath.json)
 *   orc_hermite_E         -> closed form E^{00}_0, 1-D overlap identity
 *   orc_coulomb_R         -> brute-force derivatives of F_0 (tests), reference
 *   orc_basis_repulsion   -> golden ERIs from the compiled reference, Szabo-Ostlund
 *                            H2/STO-3G values, (ss|ss) closed form, 8-fold symmetry
 *   orc_schwarz_bounds    -> golden bounds from the compiled reference
 *   orc_overlap/kinetic/attraction_matrix -> golden S, T, V
 *   orc_repulsion_matrix, orc_coulomb_matrix, orc_exchange_matrix
 *                         -> brute-force einsum on the expanded tensor
 *
 * Basis description (shared with the product's C ABI): K functions, each with
 * centre xyz[3*mu..], Cartesian exponents lmn[3*mu..], primitives
 * [poff[mu], poff[mu+1]) in alpha/coeff/norm.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_R_ORDER 12                     /* integrals.hpp:325-326 */
#define ORC_N_R ((ORC_MAX_R_ORDER + 1) * (ORC_MAX_R_ORDER + 2) * (ORC_MAX_R_ORDER + 3) / 6)
#define ORC_PI 3.141592653589793238462643383279502884

/* 0 = reference Boys (integrals.hpp:223-266), 1 = exact Boys (long double series) */
static int g_boys_mode = 0;
void orc_set_boys_mode(int mode) { g_boys_mode = mode; }

/* ---------------------------------------------------------------- indexing */
/* indexing.hpp:33-39 */
static inline size_t pair_index(int a, int b) {
    if (a < b) { int t = a; a = b; b = t; }
    return (size_t)a * (a + 1) / 2 + b;
}
/* indexing.hpp:42-52 */
size_t orc_eri_index(int i, int j, int k, int l) {
    size_t IJ = pair_index(i, j), KL = pair_index(k, l);
    if (IJ < KL) { size_t t = IJ; IJ = KL; KL = t; }
    return IJ * (IJ + 1) / 2 + KL;
}
/* indexing.hpp:55-77 */
static void quartet_to_pairs(size_t qid, int *p1, int *p2) {
    int pid1 = (int)((sqrt(8.0 * (double)qid + 1.0) - 1.0) * 0.5);
    size_t tri = (size_t)pid1 * (pid1 + 1) / 2;
    if (tri > qid) { --pid1; tri = (size_t)pid1 * (pid1 + 1) / 2; }
    while ((size_t)(pid1 + 1) * (pid1 + 2) / 2 <= qid) { ++pid1; tri = (size_t)pid1 * (pid1 + 1) / 2; }
    *p1 = pid1;
    *p2 = (int)(qid - tri);
}
/* indexing.hpp:80-86 */
static void pair_to_indices(int pid, int *a, int *b) {
    *a = (int)((sqrt(8.0 * pid + 1.0) - 1.0) / 2.0);
    *b = pid - *a * (*a + 1) / 2;
}

/* ---------------------------------------------------------------- Boys */
/* integrals.hpp:223-266 (series for T < 1e-4, erf + upward recursion else) */
double orc_boys_ref(int m, double T) {
    if (T < 1e-4) {
        double term = 1.0, sum = 0.0;
        for (int k = 0; k < 50; ++k) {
            double add = term / (2.0 * m + 2.0 * k + 1.0);
            sum += add;
            double s = fabs(sum) > 1.0 ? fabs(sum) : 1.0;
            if (fabs(add) < 1e-16 * s) break;
            term *= -T / (double)(k + 1);
        }
        return sum;
    }
    double sqrtT = sqrt(T);
    double F = 0.5 * sqrt(ORC_PI / T) * erf(sqrtT);
    double expT = exp(-T);
    for (int k = 0; k < m; ++k) F = ((2.0 * k + 1.0) * F - expT) / (2.0 * T);
    return F;
}

/* Exact Boys value in long double: F_m(T) = e^{-T} sum_k (2T)^k / prod_{i=0..k}(2m+2i+1)
 * (all terms positive, no cancellation) for T <= 60; beyond that F_0 from erf
 * and the upward recursion, which is stable for T > m. */
double orc_boys_exact(int m, double Td) {
    long double T = Td;
    if (T <= 60.0L) {
        long double term = 1.0L / (2.0L * m + 1.0L), sum = term;
        for (int k = 1; k < 2000; ++k) {
            term *= 2.0L * T / (2.0L * m + 2.0L * k + 1.0L);
            sum += term;
            if (term < 1e-22L * sum) break;
        }
        return (double)(expl(-T) * sum);
    }
    long double F = 0.5L * sqrtl((long double)ORC_PI / T) * erfl(sqrtl(T));
    long double e = expl(-T);
    for (int k = 0; k < m; ++k) F = ((2.0L * k + 1.0L) * F - e) / (2.0L * T);
    return (double)F;
}

static inline double boys(int m, double T) {
    return g_boys_mode ? orc_boys_exact(m, T) : orc_boys_ref(m, T);
}

/* ---------------------------------------------------------------- Hermite E */
/* integrals.hpp:269-322; out[0..i+j] */
int orc_hermite_E(int i, int j, double alpha, double beta, double A, double B, double *out) {
    enum { ML = 4, ME = 2 * ML + 1 };
    if (i > ML || j > ML || i + j >= ME) return -1;
    double E[ML + 1][ML + 1][ME];
    memset(E, 0, sizeof(E));
    double p = alpha + beta, q = alpha * beta / p, AB = A - B;
    E[0][0][0] = exp(-q * AB * AB);
    for (int ia = 1; ia <= i; ++ia)
        for (int tt = 0; tt <= ia; ++tt) {
            double t1 = (tt > 0) ? E[ia - 1][0][tt - 1] / (2.0 * p) : 0.0;
            double t2 = -(beta / p) * AB * E[ia - 1][0][tt];
            double t3 = (tt + 1 < ME) ? (tt + 1) * E[ia - 1][0][tt + 1] : 0.0;
            E[ia][0][tt] = t1 + t2 + t3;
        }
    for (int ia = 0; ia <= i; ++ia)
        for (int jb = 1; jb <= j; ++jb)
            for (int tt = 0; tt <= ia + jb; ++tt) {
                double t1 = (tt > 0) ? E[ia][jb - 1][tt - 1] / (2.0 * p) : 0.0;
                double t2 = +(alpha / p) * AB * E[ia][jb - 1][tt];
                double t3 = (tt + 1 < ME) ? (tt + 1) * E[ia][jb - 1][tt + 1] : 0.0;
                E[ia][jb][tt] = t1 + t2 + t3;
            }
    for (int tt = 0; tt <= i + j; ++tt) out[tt] = E[i][j][tt];
    return 0;
}

/* ---------------------------------------------------------------- R_tuv */
/* integrals.hpp:343-358 */
static inline int packed_r(int a, int b, int c) {
    int s = a + b + c, off = s * (s + 1) * (s + 2) / 6, oa = 0;
    for (int aa = 0; aa < a; ++aa) oa += (s - aa + 1);
    return off + oa + b;
}

/* Shared body of coulomb_R (integrals.hpp:541-618) and coulomb_R_attraction
 * (integrals.hpp:362-439): the two differ only in rho and T. */
static double hermite_R(int tx, int ty, int tz, int n, double rho, double X, double Y, double Z) {
    int max_order = tx + ty + tz + n;
    if (max_order > ORC_MAX_R_ORDER || tx < 0 || ty < 0 || tz < 0) return 0.0;
    double T = rho * (X * X + Y * Y + Z * Z);
    double bufA[ORC_N_R], bufB[ORC_N_R];
    double *Rnext = bufA, *Rcurr = bufB;
    memset(bufA, 0, sizeof(bufA));
    Rnext[0] = pow(-2.0 * rho, (double)max_order) * boys(max_order, T);
    for (int m = max_order - 1; m >= n; --m) {
        memset(Rcurr, 0, sizeof(double) * ORC_N_R);
        Rcurr[0] = pow(-2.0 * rho, (double)m) * boys(m, T);
        for (int s = 1; s <= tx + ty + tz; ++s)
            for (int a = 0; a <= s; ++a)
                for (int b = 0; b <= s - a; ++b) {
                    int c = s - a - b;
                    if (a > tx || b > ty || c > tz) continue;
                    double val = 0.0;
                    if (a > 0) {
                        val = X * Rnext[packed_r(a - 1, b, c)];
                        if (a >= 2) val += (a - 1) * Rnext[packed_r(a - 2, b, c)];
                    } else if (b > 0) {
                        val = Y * Rnext[packed_r(a, b - 1, c)];
                        if (b >= 2) val += (b - 1) * Rnext[packed_r(a, b - 2, c)];
                    } else if (c > 0) {
                        val = Z * Rnext[packed_r(a, b, c - 1)];
                        if (c >= 2) val += (c - 1) * Rnext[packed_r(a, b, c - 2)];
                    }
                    Rcurr[packed_r(a, b, c)] = val;
                }
        double *t = Rnext; Rnext = Rcurr; Rcurr = t;
    }
    return Rnext[packed_r(tx, ty, tz)];
}

/* integrals.hpp:541-618 */
double orc_coulomb_R(int tx, int ty, int tz, int n, double p, double q,
                     const double *P, const double *Q) {
    double rho = p * q / (p + q);
    return hermite_R(tx, ty, tz, n, rho, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]);
}

/* ---------------------------------------------------------------- basis view */
typedef struct {
    int K;
    const double *xyz;   /* 3K */
    const int *lmn;      /* 3K */
    const int *poff;     /* K+1 */
    const double *alpha, *coeff, *norm;
} orc_basis;

/* ---------------------------------------------------------------- ERI */
/* integrals.hpp:624-699 */
static double primitive_repulsion(const orc_basis *bs, int f1, int f2, int f3, int f4,
                                  int i1, int i2, int i3, int i4) {
    const double *A = bs->xyz + 3 * f1, *B = bs->xyz + 3 * f2, *C = bs->xyz + 3 * f3, *D = bs->xyz + 3 * f4;
    const int *l1 = bs->lmn + 3 * f1, *l2 = bs->lmn + 3 * f2, *l3 = bs->lmn + 3 * f3, *l4 = bs->lmn + 3 * f4;
    double alpha = bs->alpha[i1], beta = bs->alpha[i2], gamma = bs->alpha[i3], delta = bs->alpha[i4];
    double p = alpha + beta, q = gamma + delta;
    double P[3], Q[3];
    for (int x = 0; x < 3; ++x) {
        P[x] = (alpha * A[x] + beta * B[x]) / p;
        Q[x] = (gamma * C[x] + delta * D[x]) / q;
    }
    int t = l1[0] + l2[0], u = l1[1] + l2[1], v = l1[2] + l2[2];
    int tau = l3[0] + l4[0], nu = l3[1] + l4[1], phi = l3[2] + l4[2];
    double E1[9], E2[9], E3[9], E4[9], E5[9], E6[9];
    orc_hermite_E(l1[0], l2[0], alpha, beta, A[0], B[0], E1);
    orc_hermite_E(l1[1], l2[1], alpha, beta, A[1], B[1], E2);
    orc_hermite_E(l1[2], l2[2], alpha, beta, A[2], B[2], E3);
    orc_hermite_E(l3[0], l4[0], gamma, delta, C[0], D[0], E4);
    orc_hermite_E(l3[1], l4[1], gamma, delta, C[1], D[1], E5);
    orc_hermite_E(l3[2], l4[2], gamma, delta, C[2], D[2], E6);
    double eri = 0.0;
    for (int ti = 0; ti <= t; ++ti)
        for (int ui = 0; ui <= u; ++ui)
            for (int vi = 0; vi <= v; ++vi)
                for (int taui = 0; taui <= tau; ++taui)
                    for (int nui = 0; nui <= nu; ++nui)
                        for (int phii = 0; phii <= phi; ++phii) {
                            double sign = ((taui + nui + phii) % 2 == 0) ? 1.0 : -1.0;
                            eri += sign * E1[ti] * E2[ui] * E3[vi] * E4[taui] * E5[nui] * E6[phii] *
                                   orc_coulomb_R(ti + taui, ui + nui, vi + phii, 0, p, q, P, Q);
                        }
    return 2 * pow(ORC_PI, 2.5) * pow(p + q, -0.5) * eri / (p * q);
}

/* integrals.hpp:701-729 */
static double basis_repulsion(const orc_basis *bs, int f1, int f2, int f3, int f4) {
    double eri = 0.0;
    for (int i = bs->poff[f1]; i < bs->poff[f1 + 1]; ++i)
        for (int j = bs->poff[f2]; j < bs->poff[f2 + 1]; ++j)
            for (int k = bs->poff[f3]; k < bs->poff[f3 + 1]; ++k)
                for (int l = bs->poff[f4]; l < bs->poff[f4 + 1]; ++l)
                    eri += bs->coeff[i] * bs->coeff[j] * bs->coeff[k] * bs->coeff[l] *
                           bs->norm[i] * bs->norm[j] * bs->norm[k] * bs->norm[l] *
                           primitive_repulsion(bs, f1, f2, f3, f4, i, j, k, l);
    return eri;
}

#define ORC_BASIS_ARGS int K, const double *xyz, const int *lmn, const int *poff, \
                       const double *alpha, const double *coeff, const double *norm
#define ORC_BASIS_INIT orc_basis bs = {K, xyz, lmn, poff, alpha, coeff, norm}

double orc_basis_repulsion(ORC_BASIS_ARGS, int f1, int f2, int f3, int f4) {
    ORC_BASIS_INIT;
    return basis_repulsion(&bs, f1, f2, f3, f4);
}

/* integrals.hpp:732-752 */
void orc_schwarz_bounds(ORC_BASIS_ARGS, double *bounds) {
    ORC_BASIS_INIT;
    int n_pairs = K * (K + 1) / 2;
    for (int i = 0; i < n_pairs; ++i) {
        int a, b;
        pair_to_indices(i, &a, &b);
        double v = basis_repulsion(&bs, a, b, a, b);
        bounds[i] = sqrt(v > 0.0 ? v : 0.0);
    }
}

/* integrals.hpp:754-787 restricted to a list of unique quartet ids (qids):
 * out[n] = 0 when bounds[IJ]*bounds[KL] < threshold, else the contracted ERI.
 * With qids == NULL the full unique array of n = Q entries is produced. */
void orc_eri_quartets(ORC_BASIS_ARGS, const double *bounds, double threshold,
                      const uint64_t *qids, size_t n, double *out) {
    ORC_BASIS_INIT;
    for (size_t s = 0; s < n; ++s) {
        size_t i = qids ? (size_t)qids[s] : s;
        int p1, p2, mu, nu, lam, sig;
        quartet_to_pairs(i, &p1, &p2);
        pair_to_indices(p1, &mu, &nu);
        pair_to_indices(p2, &lam, &sig);
        if (bounds[p1] * bounds[p2] < threshold) { out[s] = 0.0; continue; }
        out[s] = basis_repulsion(&bs, mu, nu, lam, sig);
    }
}

/* ---------------------------------------------------------------- one-electron */
/* integrals.hpp:13-54 */
static double overlap_1d(double alpha, double beta, int l, int m, double A, double B) {
    if (l < 0 || m < 0) return 0.0;
    double p = alpha + beta, mu = alpha * beta / p, P = (alpha * A + beta * B) / p;
    if (l == 0 && m == 0) return sqrt(ORC_PI / p) * exp(-mu * (A - B) * (A - B));
    if (l > 0)
        return (P - A) * overlap_1d(alpha, beta, l - 1, m, A, B) +
               ((l - 1) / (2.0 * p)) * overlap_1d(alpha, beta, l - 2, m, A, B) +
               (m / (2.0 * p)) * overlap_1d(alpha, beta, l - 1, m - 1, A, B);
    return (P - B) * overlap_1d(alpha, beta, l, m - 1, A, B) +
           (l / (2.0 * p)) * overlap_1d(alpha, beta, l - 1, m - 1, A, B) +
           ((m - 1) / (2.0 * p)) * overlap_1d(alpha, beta, l, m - 2, A, B);
}
/* integrals.hpp:122-142 */
static double kinetic_1d(double alpha, double beta, int l, int m, double A, double B) {
    double t1 = beta * (2 * m + 1) * overlap_1d(alpha, beta, l, m, A, B);
    double t2 = -2 * beta * beta * overlap_1d(alpha, beta, l, m + 2, A, B);
    double t3 = -0.5 * m * (m - 1) * overlap_1d(alpha, beta, l, m - 2, A, B);
    return t1 + t2 + t3;
}

/* integrals.hpp:58-118 */
void orc_overlap_matrix(ORC_BASIS_ARGS, double *S) {
    ORC_BASIS_INIT;
    for (int a = 0; a < K; ++a)
        for (int b = 0; b <= a; ++b) {
            const double *A = xyz + 3 * a, *B = xyz + 3 * b;
            const int *la = lmn + 3 * a, *lb = lmn + 3 * b;
            double s = 0.0;
            for (int i = bs.poff[a]; i < bs.poff[a + 1]; ++i)
                for (int j = bs.poff[b]; j < bs.poff[b + 1]; ++j) {
                    double nrm = norm[i] * norm[j];
                    double Sx = overlap_1d(alpha[i], alpha[j], la[0], lb[0], A[0], B[0]);
                    double Sy = overlap_1d(alpha[i], alpha[j], la[1], lb[1], A[1], B[1]);
                    double Sz = overlap_1d(alpha[i], alpha[j], la[2], lb[2], A[2], B[2]);
                    s += coeff[i] * coeff[j] * (nrm * Sx * Sy * Sz);
                }
            S[a * K + b] = s;
            S[b * K + a] = s;
        }
}

/* integrals.hpp:144-218 */
void orc_kinetic_matrix(ORC_BASIS_ARGS, double *Tm) {
    ORC_BASIS_INIT;
    for (int a = 0; a < K; ++a)
        for (int b = 0; b <= a; ++b) {
            const double *A = xyz + 3 * a, *B = xyz + 3 * b;
            const int *l = lmn + 3 * a, *m = lmn + 3 * b;
            double t = 0.0;
            for (int i = bs.poff[a]; i < bs.poff[a + 1]; ++i)
                for (int j = bs.poff[b]; j < bs.poff[b + 1]; ++j) {
                    double al = alpha[i], be = alpha[j], nrm = norm[i] * norm[j];
                    double t1 = kinetic_1d(al, be, l[0], m[0], A[0], B[0]) *
                                overlap_1d(al, be, l[1], m[1], A[1], B[1]) *
                                overlap_1d(al, be, l[2], m[2], A[2], B[2]);
                    double t2 = kinetic_1d(al, be, l[1], m[1], A[1], B[1]) *
                                overlap_1d(al, be, l[0], m[0], A[0], B[0]) *
                                overlap_1d(al, be, l[2], m[2], A[2], B[2]);
                    double t3 = kinetic_1d(al, be, l[2], m[2], A[2], B[2]) *
                                overlap_1d(al, be, l[1], m[1], A[1], B[1]) *
                                overlap_1d(al, be, l[0], m[0], A[0], B[0]);
                    t += coeff[i] * coeff[j] * (nrm * (t1 + t2 + t3));
                }
            Tm[a * K + b] = t;
            Tm[b * K + a] = t;
        }
}

/* integrals.hpp:443-536; atoms: natom charges Z[] and positions R[3*natom] */
void orc_attraction_matrix(ORC_BASIS_ARGS, int natom, const int *Z, const double *R, double *V) {
    ORC_BASIS_INIT;
    for (int a = 0; a < K; ++a)
        for (int b = 0; b <= a; ++b) {
            const double *A = xyz + 3 * a, *B = xyz + 3 * b;
            const int *l = lmn + 3 * a, *m = lmn + 3 * b;
            double v = 0.0;
            for (int i = bs.poff[a]; i < bs.poff[a + 1]; ++i)
                for (int j = bs.poff[b]; j < bs.poff[b + 1]; ++j)
                    for (int c = 0; c < natom; ++c) {
                        double al = alpha[i], be = alpha[j], p = al + be;
                        double P[3];
                        for (int x = 0; x < 3; ++x) P[x] = (al * A[x] + be * B[x]) / p;
                        double E1[9], E2[9], E3[9];
                        orc_hermite_E(l[0], m[0], al, be, A[0], B[0], E1);
                        orc_hermite_E(l[1], m[1], al, be, A[1], B[1], E2);
                        orc_hermite_E(l[2], m[2], al, be, A[2], B[2], E3);
                        double val = 0.0;
                        for (int t = 0; t <= l[0] + m[0]; ++t)
                            for (int u = 0; u <= l[1] + m[1]; ++u)
                                for (int w = 0; w <= l[2] + m[2]; ++w)
                                    val += E1[t] * E2[u] * E3[w] *
                                           hermite_R(t, u, w, 0, p, P[0] - R[3 * c], P[1] - R[3 * c + 1],
                                                     P[2] - R[3 * c + 2]);
                        double prim = -(2 * ORC_PI / p) * norm[i] * norm[j] * Z[c] * val;
                        v += coeff[i] * coeff[j] * prim;
                    }
            V[a * K + b] = v;
            V[b * K + a] = v;
        }
}

/* ---------------------------------------------------------------- Fock build */
/* linalg.hpp:100-129: G = J - 1/2 K from the unique array */
void orc_repulsion_matrix(int K, const double *P, const double *eri, double *G) {
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k)
                for (int l = 0; l < K; ++l) {
                    double J = P[k * K + l] * eri[orc_eri_index(i, j, k, l)];
                    double Kx = P[k * K + l] * eri[orc_eri_index(i, k, j, l)];
                    acc += J - 0.5 * Kx;
                }
            G[i * K + j] = acc;
        }
}

/* linalg.hpp:215-257 */
void orc_coulomb_matrix(int K, const double *P1, const double *P2, const double *eri, double *J) {
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k)
                for (int l = 0; l < K; ++l)
                    acc += (P1[k * K + l] + P2[k * K + l]) * eri[orc_eri_index(i, j, k, l)];
            J[i * K + j] = acc;
        }
}

/* linalg.hpp:260-296 */
void orc_exchange_matrix(int K, const double *P, const double *eri, double *Kx) {
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k)
                for (int l = 0; l < K; ++l)
                    acc += P[k * K + l] * eri[orc_eri_index(i, k, j, l)];
            Kx[i * K + j] = acc;
        }
}
