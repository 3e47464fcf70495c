/*
 * neorx_oracle.c -- CPU restatement of the NEO-RXCHF geminal-integral and
 * Fock-build path, used ONLY as the checker of the HIP library.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  The product path
 * (libneorx_hip.so) never links, loads or calls it.
 *
 * What it restates (reference file:line in each function comment):
 *   - Boys function with the reference's two-branch semantics
 *     (AC_Integrals/common.f:317-365, XGAMMAF/iboys2).
 *   - McMurchie-Davidson evaluation of every elementary geminal integral
 *     the path uses: per-particle Hermite expansion (ghec,
 *     CS_Integrals/XC1_int_tools.f:136-227), integration of the product of
 *     Hermite Gaussians and Gaussian geminals (the reference does this with
 *     Boys' lemma, SFBoys2 XC1_int_tools.f:230-270, and DGK123/DKGx
 *     XC1_G3_MD.f:1876-2179; here the same integral is done with the
 *     n-particle exponent matrix), Coulomb factors through R_tuv
 *     (RTUV, XC1_int_tools.f:341-464) and the kinetic operator as the second
 *     derivative of the ket (G3_MD_T1d XC1_G3_MD.f:2680-2868, pgi_Lapx1
 *     AC_Integrals/g2.f).
 *   - Term assembly of one primitive tuple: xcalc_GAM1_MD (gam_1_OMP.f:398),
 *     xcalc_GAM2_MD (gam_2_OMP.f:546-1077), xcalc_GAM3_MD (gam_3_OMP.f:617-983).
 *   - Contracted builders and symmetrisation: GAM1_OMP_MD (gam_1_OMP.f:2),
 *     GAM2_IC1 + contract_omega2_conv (gam2_IC1_OMP.f:2-337,
 *     gam2_CONV_OMP.f:388-579), GAM3_IC1 + contract_omega3_conv
 *     (gam3_IC1_OMP.f:2-431, gam3_CONV_OMP.f:479-705), calc_GAM_ee
 *     (gam_ee.f:2-168), GAM4_ICR/symm_gam4 (gam4_ICR_OMP.f:2-270,
 *     gam_4_OMP.f:712-1192).
 *   - Fock/energy contractions EG2IC1, EG3IC1, EG4ICR
 *     (gam2_IC1_OMP.f:339-570, gam3_IC1_OMP.f:433-725, gam4_ICR_OMP.f:273-490).
 *
 * Array conventions follow the reference's Fortran arrays exactly: all index
 * arrays are 64-bit, matrices are column-major (ELCAM(npebf,3) is stored as
 * ELCAM[k*npebf+i]), cat(3,nat) as cat[3*iat+k], contracted tensors in the
 * pack_GAM.f layouts (CS_Integrals/Spot_Check/pack_GAM.f:72-240).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* ------------------------------------------------------------------ */
/* Boys function, reference semantics (common.f:317-341 XGAMMAF,       */
/* common.f:351-365 iboys2: each order from its own series)           */
/* ------------------------------------------------------------------ */
static void xgammaf(double *F, double T, int M)
{
    const double TMAX = 20.0, ACCY = 1.0e-12;
    const double SQRPI = 0.88622692545275801365;
    double pt5pm = 0.5 + M;
    if (T > TMAX) {
        F[0] = SQRPI / sqrt(T);
        for (int l = 1; l <= M; ++l) F[l] = F[l - 1] * (1.0 * l - 0.5) / T;
    } else {
        double ti2 = T * 2.0, expti = exp(-T);
        double term = 0.5 / pt5pm, sumk = term;
        for (int k = 1; k <= 1000; ++k) {
            term = term * T / (pt5pm + k);
            sumk = sumk + term;
            if (term < ACCY) break;
        }
        F[M] = expti * sumk;
        for (int l = M - 1; l >= 0; --l) F[l] = (F[l + 1] * ti2 + expti) / (2.0 * l + 1.0);
    }
}

double orc_boys(i64 m, double T)
{
    double F[64];
    xgammaf(F, T, (int)m);
    return F[m];
}

static void underflow(double *x)
{ /* AC_Integrals/underflow.f:2-12 */
    if (fabs(*x) < 1.0e-100) *x = 0.0;
}

/* ------------------------------------------------------------------ */
/* Generic n-particle Gaussian-geminal integral (MD formulation)       */
/* ------------------------------------------------------------------ */
#define MAXN 4
#define MAXT 7 /* per-dimension Hermite order per particle + 1 (l<=3) */

typedef struct {
    int l[3];
    double a;
    double A[3];
} prim_t;

enum { OP_OVL = 0, OP_NUC = 1, OP_COUL = 2, OP_KIN = 3 };

typedef struct {
    int n;
    prim_t bra[MAXN], ket[MAXN];
    double gam[MAXN][MAXN];  /* geminal exponent between particles i,j, bra+ket */
    double gamk[MAXN][MAXN]; /* ket-side geminal exponents (kinetic operator) */
    int op, a, b;
    double C[3];
    double mass;
    double pi; /* value of pi in the normalisation constants */
} gint_t;

/* Hermite expansion coefficients, ghec (XC1_int_tools.f:136-227):
 * G_i(x;a,A) G_j(x;b,B) = sum_t E_t Lambda_t(x;p,P), E_0^{00} = K_AB. */
static void hermite_E(int il, int jl, double a, double A, double b, double B, double *Eout)
{
    double E[MAXT][MAXT][2 * MAXT + 2];
    memset(E, 0, sizeof(E));
    double p = a + b, q = a * b / p, Qx = A - B;
    double x1 = 1.0 / (2.0 * p);
    E[0][0][0] = exp(-q * Qx * Qx);
    double xpa = -q * Qx / a;
    for (int i = 1; i <= il; ++i) {
        E[i][0][0] = xpa * E[i - 1][0][0] + E[i - 1][0][1];
        for (int t = 1; t <= i; ++t)
            E[i][0][t] = x1 * E[i - 1][0][t - 1] + xpa * E[i - 1][0][t] + (t + 1) * E[i - 1][0][t + 1];
    }
    double xpb = q * Qx / b;
    for (int j = 1; j <= jl; ++j)
        for (int i = 0; i <= il; ++i) {
            E[i][j][0] = xpb * E[i][j - 1][0] + E[i][j - 1][1];
            for (int t = 1; t <= i + j; ++t)
                E[i][j][t] = x1 * E[i][j - 1][t - 1] + xpb * E[i][j - 1][t] + (t + 1) * E[i][j - 1][t + 1];
        }
    for (int t = 0; t <= il + jl; ++t) Eout[t] = E[il][jl][t];
}

/* small dense inverse (Gauss-Jordan with partial pivoting); returns det */
static double inv_small(int n, const double *M, double *Mi)
{
    double a[MAXN][2 * MAXN];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            a[i][j] = M[i * n + j];
            a[i][n + j] = (i == j) ? 1.0 : 0.0;
        }
    double det = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (piv != c) {
            for (int k = 0; k < 2 * n; ++k) {
                double t = a[c][k];
                a[c][k] = a[piv][k];
                a[piv][k] = t;
            }
            det = -det;
        }
        double d = a[c][c];
        det *= d;
        for (int k = 0; k < 2 * n; ++k) a[c][k] /= d;
        for (int r = 0; r < n; ++r)
            if (r != c) {
                double f = a[r][c];
                if (f != 0.0)
                    for (int k = 0; k < 2 * n; ++k) a[r][k] -= f * a[c][k];
            }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Mi[i * n + j] = a[i][n + j];
    return det;
}

static double binom(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

/* R_tuv^(0) for exponent alpha and vector D, Boys values via iboys2 for each
 * order (RTUV, XC1_int_tools.f:341-464). */
/* R_tuv^(0) for exponent alpha and vector D, Boys values via iboys2 for each
 * order (RTUV, XC1_int_tools.f:341-464).  R is (L+1)^3, index [t][u][v]. */
static void rtuv_table(int L, double alpha, const double *D, double *R)
{
    int S = L + 2, S2 = S * S, S3 = S2 * S;
    double *Rn = (double *)calloc((size_t)(L + 2) * S3, sizeof(double));
#define RN(n, t, u, v) Rn[(size_t)(n) * S3 + (t) * S2 + (u) * S + (v)]
    double T = alpha * (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    for (int n = 0; n <= L; ++n) {
        double F[128];
        xgammaf(F, T, n);
        RN(n, 0, 0, 0) = pow(-2.0 * alpha, (double)n) * F[n];
    }
    for (int n = L; n >= 0; --n) {
        int m = L - n;
        for (int t = 0; t <= m; ++t)
            for (int u = 0; u <= m - t; ++u)
                for (int v = 0; v <= m - t - u; ++v) {
                    if (t == 0 && u == 0 && v == 0) continue;
                    double val;
                    if (t > 0) {
                        val = D[0] * RN(n + 1, t - 1, u, v);
                        if (t > 1) val += (t - 1) * RN(n + 1, t - 2, u, v);
                    } else if (u > 0) {
                        val = D[1] * RN(n + 1, t, u - 1, v);
                        if (u > 1) val += (u - 1) * RN(n + 1, t, u - 2, v);
                    } else {
                        val = D[2] * RN(n + 1, t, u, v - 1);
                        if (v > 1) val += (v - 1) * RN(n + 1, t, u, v - 2);
                    }
                    RN(n, t, u, v) = val;
                }
    }
    int L1 = L + 1;
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u <= L - t; ++u)
            for (int v = 0; v <= L - t - u; ++v) R[(t * L1 + u) * L1 + v] = RN(0, t, u, v);
#undef RN
    free(Rn);
}

typedef struct {
    int n;
    double p[MAXN], P[3][MAXN];
    double Mi[MAXN * MAXN], Q[MAXN * MAXN];
    double detM;
} gsetup_t;

static void gsetup(const gint_t *g, gsetup_t *s)
{
    int n = g->n;
    s->n = n;
    double M[MAXN * MAXN];
    for (int i = 0; i < n; ++i) {
        double a = g->bra[i].a, b = g->ket[i].a;
        s->p[i] = a + b;
        for (int d = 0; d < 3; ++d) s->P[d][i] = (a * g->bra[i].A[d] + b * g->ket[i].A[d]) / (a + b);
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (i == j) {
                double sgm = s->p[i];
                for (int k = 0; k < n; ++k)
                    if (k != i) sgm += g->gam[i][k];
                M[i * n + j] = sgm;
            } else
                M[i * n + j] = -g->gam[i][j];
        }
    s->detM = inv_small(n, M, s->Mi);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) s->Q[i * n + j] = (i == j ? s->p[i] : 0.0) - s->p[i] * s->Mi[i * n + j] * s->p[j];
}

/* multi-index helpers over n particles with radix r[i] */
static int mi_size(int n, const int *r)
{
    int s = 1;
    for (int i = 0; i < n; ++i) s *= r[i];
    return s;
}
static void mi_decode(int n, const int *r, int idx, int *t)
{
    for (int i = n - 1; i >= 0; --i) {
        t[i] = idx % r[i];
        idx /= r[i];
    }
}
static int mi_encode(int n, const int *r, const int *t)
{
    int idx = 0;
    for (int i = 0; i < n; ++i) idx = idx * r[i] + t[i];
    return idx;
}

/* Hermite-Gaussian derivative polynomials: H[s] = d^s exp(-P'QP) / exp(-P'QP) */
static void gauss_derivs(int n, const int *r, const double *Q, const double *P, double *H)
{
    int N = mi_size(n, r);
    double QP[MAXN];
    for (int i = 0; i < n; ++i) {
        QP[i] = 0.0;
        for (int j = 0; j < n; ++j) QP[i] += Q[i * n + j] * P[j];
    }
    int t[MAXN];
    H[0] = 1.0;
    for (int idx = 1; idx < N; ++idx) {
        mi_decode(n, r, idx, t);
        int i = 0;
        while (t[i] == 0) ++i;
        t[i] -= 1;
        double val = -2.0 * QP[i] * H[mi_encode(n, r, t)];
        for (int j = 0; j < n; ++j)
            if (t[j] > 0) {
                t[j] -= 1;
                val -= 2.0 * Q[i * n + j] * (t[j] + 1) * H[mi_encode(n, r, t)];
                t[j] += 1;
            }
        H[idx] = val;
    }
}

/* per-dimension Hermite sums: V[k] = g * sum_t prod E_t sum_{s<=t,|s|=k} C(t,s) kap^s H[t-s] */
static void dim_sums(const gint_t *g, const gsetup_t *s, int d, const int *lb, const int *lk, const double *kap,
                     int kmax, double *V)
{
    int n = g->n, r[MAXN];
    double E[MAXN][2 * MAXT];
    for (int i = 0; i < n; ++i) {
        r[i] = lb[i] + lk[i] + 1;
        hermite_E(lb[i], lk[i], g->bra[i].a, g->bra[i].A[d], g->ket[i].a, g->ket[i].A[d], E[i]);
    }
    /* shift centres relative to particle 0 (translation invariance Q.1 = 0) */
    double P[MAXN];
    for (int i = 0; i < n; ++i) P[i] = s->P[d][i] - s->P[d][0];
    int N = mi_size(n, r);
    double *H = (double *)malloc(sizeof(double) * N);
    gauss_derivs(n, r, s->Q, P, H);
    double qf = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) qf += P[i] * s->Q[i * n + j] * P[j];
    double gval = exp(-qf);
    for (int k = 0; k <= kmax; ++k) V[k] = 0.0;
    int t[MAXN], sv[MAXN], rs[MAXN], ts[MAXN];
    for (int idx = 0; idx < N; ++idx) {
        mi_decode(n, r, idx, t);
        double ep = 1.0;
        for (int i = 0; i < n; ++i) ep *= E[i][t[i]];
        if (ep == 0.0) continue;
        if (kmax == 0) {
            V[0] += ep * H[idx];
            continue;
        }
        for (int i = 0; i < n; ++i) rs[i] = t[i] + 1;
        int Ns = mi_size(n, rs);
        for (int js = 0; js < Ns; ++js) {
            mi_decode(n, rs, js, sv);
            int ks = 0;
            double c = 1.0;
            for (int i = 0; i < n; ++i) {
                ks += sv[i];
                c *= binom(t[i], sv[i]) * pow(kap[i], sv[i]);
                ts[i] = t[i] - sv[i];
            }
            if (ks > kmax) continue;
            V[ks] += ep * c * H[mi_encode(n, r, ts)];
        }
    }
    for (int k = 0; k <= kmax; ++k) V[k] *= gval;
    free(H);
}

static const double PI_CS = 3.14159265358979; /* XC1_G3_MD.f:125 */

/* 1-D overlap with explicit bra/ket powers (per-dimension factor of the
 * n-particle overlap; the constant pi^{n/2} det(M)^{-1/2} is included). */
static double ovl1d(const gint_t *g, const gsetup_t *s, int d, const int *lb, const int *lk)
{
    for (int i = 0; i < g->n; ++i)
        if (lk[i] < 0 || lb[i] < 0) return 0.0;
    double V[1];
    dim_sums(g, s, d, lb, lk, NULL, 0, V);
    return V[0] * pow(g->pi, 0.5 * g->n) / sqrt(s->detM);
}

/* Second derivative of the ket along dimension d for particle a, written as
 * overlaps with raised/lowered ket powers.  General form of G3_MD_T1d
 * (XC1_G3_MD.f:2680-2868) / pgi_Lapx1 (AC_Integrals/g2.f): with
 * X = x_a - B_a, h = beta X + sum_c g_c (x_a - x_c):
 *   ket'' / ket = L(L-1) X^-2 - 4 L h X^-1 + 4 h^2 - 2 (beta + sum_c g_c). */
static double kin1d(const gint_t *g, const gsetup_t *s, int d)
{
    int n = g->n, a = g->a;
    int lb[MAXN], lk[MAXN];
    for (int i = 0; i < n; ++i) {
        lb[i] = g->bra[i].l[d];
        lk[i] = g->ket[i].l[d];
    }
    int L = lk[a];
    double beta = g->ket[a].a;
    double gc[MAXN], dc[MAXN];
    double alpha = beta, delta = 0.0;
    for (int c = 0; c < n; ++c) {
        gc[c] = (c == a) ? 0.0 : g->gamk[a][c];
        dc[c] = g->ket[a].A[d] - g->ket[c].A[d];
        alpha += gc[c];
        delta += gc[c] * dc[c];
    }
    /* S(ea, c1, e1, c2, e2): overlap with ket power of a shifted by ea and of
       particles c1,c2 raised by e1,e2 */
#define SOV(ea, c1, e1, c2, e2)                                                                                    \
    ({                                                                                                             \
        int _k[MAXN];                                                                                              \
        for (int _i = 0; _i < n; ++_i) _k[_i] = lk[_i];                                                            \
        _k[a] += (ea);                                                                                             \
        if ((c1) >= 0) _k[(c1)] += (e1);                                                                           \
        if ((c2) >= 0) _k[(c2)] += (e2);                                                                           \
        ovl1d(g, s, d, lb, _k);                                                                                    \
    })
    double ans = 0.0;
    double s0 = SOV(0, -1, 0, -1, 0);
    if (L >= 2) ans += (double)(L * (L - 1)) * SOV(-2, -1, 0, -1, 0);
    if (L >= 1) {
        double t = alpha * s0 + delta * SOV(-1, -1, 0, -1, 0);
        for (int c = 0; c < n; ++c)
            if (gc[c] != 0.0) t -= gc[c] * SOV(-1, c, 1, -1, 0);
        ans -= 4.0 * L * t;
    }
    double h2 = alpha * alpha * SOV(2, -1, 0, -1, 0) + delta * delta * s0 + 2.0 * alpha * delta * SOV(1, -1, 0, -1, 0);
    for (int c = 0; c < n; ++c) {
        if (gc[c] == 0.0) continue;
        h2 -= 2.0 * alpha * gc[c] * SOV(1, c, 1, -1, 0);
        h2 -= 2.0 * delta * gc[c] * SOV(0, c, 1, -1, 0);
        for (int c2 = 0; c2 < n; ++c2) {
            if (gc[c2] == 0.0) continue;
            if (c2 == c)
                h2 += gc[c] * gc[c] * SOV(0, c, 2, -1, 0);
            else
                h2 += gc[c] * gc[c2] * SOV(0, c, 1, c2, 1);
        }
    }
    ans += 4.0 * h2 - 2.0 * alpha * s0;
#undef SOV
    return ans;
}

static double gint_eval(const gint_t *g)
{
    gsetup_t s;
    gsetup(g, &s);
    int n = g->n;
    if (g->op == OP_OVL || g->op == OP_KIN) {
        double S[3], T[3];
        for (int d = 0; d < 3; ++d) {
            int lb[MAXN], lk[MAXN];
            for (int i = 0; i < n; ++i) {
                lb[i] = g->bra[i].l[d];
                lk[i] = g->ket[i].l[d];
            }
            S[d] = ovl1d(g, &s, d, lb, lk);
        }
        if (g->op == OP_OVL) return S[0] * S[1] * S[2];
        for (int d = 0; d < 3; ++d) T[d] = kin1d(g, &s, d);
        double t = T[0] * S[1] * S[2] + S[0] * T[1] * S[2] + S[0] * S[1] * T[2];
        return -0.5 * t / g->mass;
    }
    /* Coulomb-type: e = e_a (NUC, c0 = C) or e_a - e_b (COUL) */
    double e[MAXN] = {0}, w[MAXN], kap[MAXN];
    e[g->a] = 1.0;
    if (g->op == OP_COUL) e[g->b] = -1.0;
    double sden = 0.0;
    for (int i = 0; i < n; ++i) {
        w[i] = 0.0;
        for (int j = 0; j < n; ++j) w[i] += s.Mi[i * n + j] * e[j];
        sden += e[i] * w[i];
    }
    double rho = 1.0 / sden;
    for (int i = 0; i < n; ++i) kap[i] = s.p[i] * w[i];
    double D[3];
    int kmax[3];
    double V[3][4 * MAXN * MAXT];
    for (int d = 0; d < 3; ++d) {
        int lb[MAXN], lk[MAXN];
        kmax[d] = 0;
        for (int i = 0; i < n; ++i) {
            lb[i] = g->bra[i].l[d];
            lk[i] = g->ket[i].l[d];
            kmax[d] += lb[i] + lk[i];
        }
        dim_sums(g, &s, d, lb, lk, kap, kmax[d], V[d]);
        /* D = kap.P - c0 evaluated with the same origin shift as dim_sums */
        double dd = 0.0;
        for (int i = 0; i < n; ++i) dd += kap[i] * (s.P[d][i] - s.P[d][0]);
        double sk = 0.0;
        for (int i = 0; i < n; ++i) sk += kap[i];
        dd += sk * s.P[d][0];
        if (g->op == OP_NUC) dd -= g->C[d];
        D[d] = dd;
    }
    int L = kmax[0] + kmax[1] + kmax[2];
    int L1 = L + 1;
    double *R = (double *)malloc(sizeof(double) * L1 * L1 * L1);
    rtuv_table(L, rho, D, R);
    double sum = 0.0;
    for (int kx = 0; kx <= kmax[0]; ++kx)
        for (int ky = 0; ky <= kmax[1]; ++ky)
            for (int kz = 0; kz <= kmax[2]; ++kz) sum += V[0][kx] * V[1][ky] * V[2][kz] * R[(kx * L1 + ky) * L1 + kz];
    free(R);
    double pref = pow(g->pi, 1.5 * n) / (s.detM * sqrt(s.detM)) * 2.0 * sqrt(rho / g->pi);
    return pref * sum;
}

/* ------------------------------------------------------------------ */
/* Reference-shaped wrappers of the elementary integrals               */
/* ------------------------------------------------------------------ */
static void gi_init(gint_t *g, int n)
{
    memset(g, 0, sizeof(*g));
    g->n = n;
    g->mass = 1.0;
    g->pi = PI_CS;
}
static void gi_gem(gint_t *g, int i, int j, double gA, double gB)
{
    g->gam[i][j] += gA + gB;
    g->gam[j][i] += gA + gB;
    g->gamk[i][j] += gB;
    g->gamk[j][i] += gB;
}

/* G3_MD_xggs / G3_MD_KE / G3_MD_xgVxCg / G3_MD_xgVeeg (XC1_G3_MD.f:2, 2546,
 * 711, 1376): particles in argument order 1,2,3. */
static double g3(int op, const prim_t *b1, const prim_t *b2, const prim_t *b3, const prim_t *k1, const prim_t *k2,
                 const prim_t *k3, double gA12, double gA13, double gA23, double gB12, double gB13, double gB23,
                 const double *C, double xmass)
{
    gint_t g;
    gi_init(&g, 3);
    g.bra[0] = *b1; g.bra[1] = *b2; g.bra[2] = *b3;
    g.ket[0] = *k1; g.ket[1] = *k2; g.ket[2] = *k3;
    gi_gem(&g, 0, 1, gA12, gB12);
    gi_gem(&g, 0, 2, gA13, gB13);
    gi_gem(&g, 1, 2, gA23, gB23);
    g.op = op;
    g.a = 0;
    g.b = 1;
    g.mass = xmass;
    if (C) memcpy(g.C, C, sizeof(g.C));
    return gint_eval(&g);
}

/* G2_MD_xggs / G2_MD_xggvec / G2_MD_xggvee / pgike1 (XC1_G2_MD.f:3,190,630;
 * AC_Integrals/g2.f:545). */
static double g2(int op, const prim_t *b1, const prim_t *b2, const prim_t *k1, const prim_t *k2, double gA, double gB,
                 const double *C, double xmass)
{
    gint_t g;
    gi_init(&g, 2);
    g.bra[0] = *b1; g.bra[1] = *b2;
    g.ket[0] = *k1; g.ket[1] = *k2;
    gi_gem(&g, 0, 1, gA, gB);
    g.op = op;
    g.a = 0;
    g.b = 1;
    g.mass = xmass;
    if (C) memcpy(g.C, C, sizeof(g.C));
    return gint_eval(&g);
}

/* pgike1 (AC_Integrals/g2.f:545-653): kinetic energy of particle 1 with the
 * ket geminal; its overlaps (pgi_hint0_1D, g2.f:200-236) use the constant
 * PI = 3.141592650d0 (g2.f:216), which is kept here for parity. */
static const double PI_AC_G2 = 3.141592650;
static double pgike1(const prim_t *b1, const prim_t *b2, const prim_t *k1, const prim_t *k2, double gA, double gB,
                     double xmass)
{
    gint_t g;
    gi_init(&g, 2);
    g.bra[0] = *b1; g.bra[1] = *b2;
    g.ket[0] = *k1; g.ket[1] = *k2;
    gi_gem(&g, 0, 1, gA, gB);
    g.op = OP_KIN;
    g.a = 0;
    g.mass = xmass;
    g.pi = PI_AC_G2;
    return gint_eval(&g);
}

/* gfvee (vee.f:3), gfke (ke.f:2), gfvec (vec.f:2) */
static double gfvee(const prim_t *b1, const prim_t *b2, const prim_t *k1, const prim_t *k2)
{
    return g2(OP_COUL, b1, b2, k1, k2, 0.0, 0.0, NULL, 1.0);
}
static double g1(int op, const prim_t *b, const prim_t *k, const double *C, double xmass)
{
    gint_t g;
    gi_init(&g, 1);
    g.bra[0] = *b;
    g.ket[0] = *k;
    g.op = op;
    g.mass = xmass;
    if (C) memcpy(g.C, C, sizeof(g.C));
    return gint_eval(&g);
}

/* G4_MD_xgVeeg (XC1_G4_MD.f:796): 1/r12 with geminals g14, g34.
 * G4_MD_xgVepg (XC1_G4_MD.f:1559): 1/r34 with geminals g14, g24. */
static double g4(int which, const prim_t *b, const prim_t *k, double gx, double gy)
{
    gint_t g;
    gi_init(&g, 4);
    for (int i = 0; i < 4; ++i) {
        g.bra[i] = b[i];
        g.ket[i] = k[i];
    }
    g.op = OP_COUL;
    if (which == 0) { /* Veeg */
        gi_gem(&g, 0, 3, gx, 0.0);
        gi_gem(&g, 2, 3, gy, 0.0);
        g.a = 0;
        g.b = 1;
    } else { /* Vepg */
        gi_gem(&g, 0, 3, gx, 0.0);
        gi_gem(&g, 1, 3, gy, 0.0);
        g.a = 2;
        g.b = 3;
    }
    return gint_eval(&g);
}

/* ------------------------------------------------------------------ */
/* Primitive tuple term assembly                                      */
/* ------------------------------------------------------------------ */
typedef struct {
    i64 nat, ngtg;
    double pmass;
    const double *cat, *zan, *bcoef, *gamma;
} sys_t;

/* xcalc_GAM1_MD (gam_1_OMP.f:398-839): electron e1 = particle 1, proton = 2.
 * With mult set: RXCHFmult_xcalc_GAM1_MD (rxchfmult/rxchfmult_gam1.f:399-840),
 * which has no one-geminal terms (tegx, tpgx, vcegx, vcpgx, vmgx, texg, tpxg,
 * ogx stay zero). */
static void gam1_prim_impl(const sys_t *S, int mult, const prim_t *e1b, const prim_t *pb, const prim_t *e1k,
                           const prim_t *pk, double *ansE, double *ansS)
{
    double tegx = 0, tpgx = 0, texg = 0, tpxg = 0, tegg = 0, tpgg = 0;
    double vcegx = 0, vcegg = 0, vcpgx = 0, vcpgg = 0, vmgx = 0, vmgg = 0;
    for (i64 k = 0; k < S->ngtg; ++k) {
        double gv = S->gamma[k], bv = S->bcoef[k];
        if (!mult) {
            tegx += bv * pgike1(e1b, pb, e1k, pk, gv, 0.0, 1.0);
            tpgx += bv * pgike1(pb, e1b, pk, e1k, gv, 0.0, S->pmass);
            double vce = 0, vcp = 0;
            for (i64 c = 0; c < S->nat; ++c) {
                const double *C = S->cat + 3 * c;
                vce += g2(OP_NUC, e1b, pb, e1k, pk, gv, 0.0, C, 1.0) * S->zan[c];
                vcp += g2(OP_NUC, pb, e1b, pk, e1k, gv, 0.0, C, 1.0) * S->zan[c];
            }
            double vm = g2(OP_COUL, e1b, pb, e1k, pk, gv, 0.0, NULL, 1.0);
            vcegx += bv * vce;
            vcpgx += bv * vcp;
            vmgx += bv * vm;
            texg += bv * pgike1(e1b, pb, e1k, pk, 0.0, gv, 1.0);
            tpxg += bv * pgike1(pb, e1b, pk, e1k, 0.0, gv, S->pmass);
        }
        for (i64 l = 0; l < S->ngtg; ++l) {
            double gw = S->gamma[l], bw = S->bcoef[l];
            tegg += bv * bw * pgike1(e1b, pb, e1k, pk, gv, gw, 1.0);
            tpgg += bv * bw * pgike1(pb, e1b, pk, e1k, gv, gw, S->pmass);
            double vce2 = 0, vcp2 = 0;
            for (i64 c = 0; c < S->nat; ++c) {
                const double *C = S->cat + 3 * c;
                vce2 += g2(OP_NUC, e1b, pb, e1k, pk, gv, gw, C, 1.0) * S->zan[c];
                vcp2 += g2(OP_NUC, pb, e1b, pk, e1k, gv, gw, C, 1.0) * S->zan[c];
            }
            double vm2 = g2(OP_COUL, e1b, pb, e1k, pk, gv, gw, NULL, 1.0);
            vcegg += bv * bw * vce2;
            vcpgg += bv * bw * vcp2;
            vmgg += bv * bw * vm2;
        }
    }
    double te = tegx + texg + tegg, tp = tpgx + tpxg + tpgg;
    double vce = 2.0 * vcegx + vcegg, vcp = 2.0 * vcpgx + vcpgg, vmx = 2.0 * vmgx + vmgg;
    *ansE = te + tp - vce + vcp - vmx;
    double ogx = 0, ogg = 0;
    for (i64 k = 0; k < S->ngtg; ++k) {
        double gv = S->gamma[k], bv = S->bcoef[k];
        if (!mult) ogx += bv * g2(OP_OVL, e1b, pb, e1k, pk, gv, 0.0, NULL, 1.0);
        for (i64 l = 0; l < S->ngtg; ++l)
            ogg += bv * S->bcoef[l] * g2(OP_OVL, e1b, pb, e1k, pk, gv, S->gamma[l], NULL, 1.0);
    }
    *ansS = 0.0 + 2.0 * ogx + ogg;
}

static void gam1_prim(const sys_t *S, const prim_t *e1b, const prim_t *pb, const prim_t *e1k, const prim_t *pk,
                      double *ansE, double *ansS)
{
    gam1_prim_impl(S, 0, e1b, pb, e1k, pk, ansE, ansS);
}

/* xcalc_GAM2_MD (gam_2_OMP.f:546-1077): particles e1, e2, p.
 * With mult set: RXCHFmult_GAM2_MD (rxchfmult/rxchfmult_gam2.f:627-1162), the
 * same two-geminal loop without the one-geminal terms, xgVEEg1 with the
 * geminals on e2-p (gamA23, gamB23, :1067-1068), and the outputs
 * ansE1 (:1149-1155) in ansE, ansE2 = gVEPg2 + gVEEg1 (:1156), ansS = gsg (:1158). */
static void gam2_prim_impl(const sys_t *S, const prim_t *b, const prim_t *k, int mult, double *ansE, double *ansS,
                           double *ansE2, double *ansE3)
{
    const prim_t *b1 = &b[0], *b2 = &b[1], *b3 = &b[2], *k1 = &k[0], *k2 = &k[1], *k3 = &k[2];
    double gTE = 0, gVEC = 0, gVEE = 0, gVEP = 0;
    for (i64 ik = 0; ik < (mult ? 0 : S->ngtg); ++ik) {
        double gk = S->gamma[ik], bk = S->bcoef[ik];
        double xgTE = g3(OP_KIN, b1, b2, b3, k1, k2, k3, 0, 0, gk, 0, 0, 0, NULL, 1.0);
        double xgVEC = 0.0;
        for (i64 c = 0; c < S->nat; ++c) {
            double v = g3(OP_NUC, b1, b2, b3, k1, k2, k3, 0, 0, gk, 0, 0, 0, S->cat + 3 * c, 1.0);
            underflow(&v);
            xgVEC -= S->zan[c] * v;
        }
        double xgVEP = g3(OP_COUL, b1, b3, b2, k1, k3, k2, 0, 0, gk, 0, 0, 0, NULL, 1.0);
        underflow(&xgVEP);
        xgVEP = -xgVEP;
        double xgVEE = g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, 0, gk, 0, 0, 0, NULL, 1.0);
        underflow(&xgVEE);
        gTE += bk * xgTE;
        gVEC += bk * xgVEC;
        gVEP += bk * xgVEP;
        gVEE += bk * xgVEE;
    }
    double gTPg = 0, gTEg1 = 0, gTEg2 = 0, gTEg3 = 0, gVPCg = 0, gVECg1 = 0, gVECg2 = 0;
    double gVEPg1 = 0, gVEPg2 = 0, gVEEg1 = 0, gVEEg2 = 0, gsg = 0;
    for (i64 ik = 0; ik < S->ngtg; ++ik)
        for (i64 il = 0; il < S->ngtg; ++il) {
            double gk = S->gamma[ik], gl = S->gamma[il];
            double bb = S->bcoef[ik] * S->bcoef[il];
            double xgTPg = g3(OP_KIN, b3, b1, b2, k3, k1, k2, gk, 0, 0, 0, gl, 0, NULL, S->pmass);
            double xgTEg1 = g3(OP_KIN, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, NULL, 1.0);
            double xgTEg2 = g3(OP_KIN, b1, b2, b3, k1, k2, k3, 0, 0, gk, 0, gl, 0, NULL, 1.0);
            double xgTEg3 = g3(OP_KIN, b1, b2, b3, k1, k2, k3, 0, 0, gk, 0, 0, gl, NULL, 1.0);
            double xgVPCg = 0, xgVECg1 = 0, xgVECg2 = 0;
            for (i64 c = 0; c < S->nat; ++c) {
                const double *C = S->cat + 3 * c;
                double z = S->zan[c], v;
                v = g3(OP_NUC, b3, b1, b2, k3, k1, k2, gk, 0, 0, 0, gl, 0, C, 1.0);
                underflow(&v);
                xgVPCg += z * v;
                v = g3(OP_NUC, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, C, 1.0);
                underflow(&v);
                xgVECg1 -= z * v;
                v = g3(OP_NUC, b1, b2, b3, k1, k2, k3, 0, 0, gk, 0, 0, gl, C, 1.0);
                underflow(&v);
                xgVECg2 -= z * v;
            }
            double xgVEPg1 = g3(OP_COUL, b1, b3, b2, k1, k3, k2, gk, 0, 0, 0, 0, gl, NULL, 1.0);
            underflow(&xgVEPg1);
            xgVEPg1 = -xgVEPg1;
            double xgVEPg2 = g3(OP_COUL, b1, b3, b2, k1, k3, k2, 0, 0, gk, 0, 0, gl, NULL, 1.0);
            underflow(&xgVEPg2);
            xgVEPg2 = -xgVEPg2;
            double xgVEEg1 = mult ? g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, 0, gk, 0, 0, gl, NULL, 1.0)
                                  : g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, gl, 0, NULL, 1.0);
            underflow(&xgVEEg1);
            double xgVEEg2 = g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, NULL, 1.0);
            underflow(&xgVEEg2);
            double xgsg = g3(OP_OVL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, NULL, 1.0);
            gTPg += bb * xgTPg;
            gTEg1 += bb * xgTEg1;
            gTEg2 += bb * xgTEg2;
            gTEg3 += bb * xgTEg3;
            gVPCg += bb * xgVPCg;
            gVECg1 += bb * xgVECg1;
            gVECg2 += bb * xgVECg2;
            gVEPg1 += bb * xgVEPg1;
            gVEPg2 += bb * xgVEPg2;
            gVEEg1 += bb * xgVEEg1;
            gVEEg2 += bb * xgVEEg2;
            gsg += bb * xgsg;
        }
    if (mult)
        *ansE = gTPg + gVPCg + gTEg1 + gTEg3 + 2.0 * (gVEPg1 + gVECg1) + gTEg2 + gVEPg2 + gVECg2 + 2.0 * gVEEg1 * 0.5 +
                2.0 * gVEEg2 * 0.5;
    else
        *ansE = 2.0 * (gTE + gVEC + gVEP) + 4.0 * gVEE * 0.5 + gTPg + gVPCg + gTEg1 + gTEg3 +
                2.0 * (gVEPg1 + gVECg1) + gTEg2 + gVEPg2 + gVECg2 + 2.0 * gVEEg1 * 0.5 + 2.0 * gVEEg2 * 0.5;
    *ansS = gsg;
    if (ansE2) *ansE2 = gVEPg2 + gVEEg1;
    if (ansE3) *ansE3 = gVEEg2; /* RXCHFmult_xcalc_GAM2_MDex (rxchfmult_gam2ex.f:559-1097) */
}

static void gam2_prim(const sys_t *S, const prim_t *b, const prim_t *k, double *ansE, double *ansS)
{
    gam2_prim_impl(S, b, k, 0, ansE, ansS, NULL, NULL);
}

/* RXCHFmult_INT_GAM3_MD (rxchfmult/rxchfmult_gam3-int.f:661-1054): particles
 * e1 (regular basis), e2, e3 (special basis), p.  ans3, ans4 (may be NULL):
 * the extra outputs of RXCHFmult_xcalc_GAM3_MDex (rxchfmult_gam3ex.f:712-1111). */
static void gam3int_prim_ex(const sys_t *S, const prim_t *b, const prim_t *k, double *ans, double *ans3,
                            double *ans4)
{
    double gVEEg1 = 0, gVEEg2 = 0, gVEPg1 = 0;
    for (i64 ik = 0; ik < S->ngtg; ++ik)
        for (i64 il = 0; il < S->ngtg; ++il) {
            double gk = S->gamma[ik], gl = S->gamma[il];
            /* G4_MD_xgVepg(e2, e3, e1, p; gamA14, gamB24): 1/r(e1,p), geminals e2-p, e3-p (:901) */
            prim_t bv[4] = {b[1], b[2], b[0], b[3]}, kv[4] = {k[1], k[2], k[0], k[3]};
            double zz = g4(1, bv, kv, gk, gl);
            underflow(&zz);
            double xgVEPg1 = zz * (-1.0);
            /* G2_MD_xggs(e2, p; gamA, gamB) * gfvee(e1, e3) (:941, :954) */
            double xx = g2(OP_OVL, &b[1], &b[3], &k[1], &k[3], gk, gl, NULL, 1.0);
            double yy = gfvee(&b[0], &b[2], &k[0], &k[2]);
            underflow(&xx);
            underflow(&yy);
            double xgVEEg1 = xx * yy;
            /* G4_MD_xgVeeg(e2, e1, e3, p; gamA14, gamB34): 1/r(e2,e1) (:998), no underflow */
            prim_t bw[4] = {b[1], b[0], b[2], b[3]}, kw[4] = {k[1], k[0], k[2], k[3]};
            double xgVEEg2 = g4(0, bw, kw, gk, gl);
            double bb = S->bcoef[ik] * S->bcoef[il];
            gVEEg1 = gVEEg1 + (bb * xgVEEg1);
            gVEEg2 = gVEEg2 + (bb * xgVEEg2);
            gVEPg1 = gVEPg1 + (bb * xgVEPg1);
        }
    *ans = gVEPg1 + gVEEg1 + 2.0 * gVEEg2;
    if (ans3) *ans3 = 2.0 * gVEEg2;
    if (ans4) *ans4 = gVEEg1;
}

static void gam3int_prim(const sys_t *S, const prim_t *b, const prim_t *k, double *ans)
{
    gam3int_prim_ex(S, b, k, ans, NULL, NULL);
}

/* RXCHFne_xcalc_GAM2_MD (rxchfne/rxchfne_gam_2_OMP.f:2-530): two outputs,
 * ansE1 = gVEPg2 + gVEEg1 and ansE2 = gVEEg2, summed over all (k,l). */
static void rxne_prim(const sys_t *S, const prim_t *b, const prim_t *k, double *ansE1, double *ansE2)
{
    const prim_t *b1 = &b[0], *b2 = &b[1], *b3 = &b[2], *k1 = &k[0], *k2 = &k[1], *k3 = &k[2];
    double gVEPg2 = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (i64 ik = 0; ik < S->ngtg; ++ik)
        for (i64 il = 0; il < S->ngtg; ++il) {
            double gk = S->gamma[ik], gl = S->gamma[il], bb = S->bcoef[ik] * S->bcoef[il];
            /* G3_MD_xgVeeg(p, e2, e1): 1/r(p,e2), geminal p-e1 on bra (gk) and ket (gl) */
            double vep = g3(OP_COUL, b3, b2, b1, k3, k2, k1, 0, gk, 0, 0, gl, 0, NULL, 1.0);
            underflow(&vep);
            vep = -vep;
            double vee1 = g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, gl, 0, NULL, 1.0);
            underflow(&vee1);
            double vee2 = g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, NULL, 1.0);
            underflow(&vee2);
            gVEPg2 += bb * vep;
            gVEEg1 += bb * vee1;
            gVEEg2 += bb * vee2;
        }
    *ansE1 = gVEPg2 + gVEEg1;
    *ansE2 = gVEEg2;
}

/* RXCHF_xcalc_GAM2_MD (rxchf/rxchf_gam_2_OMP.f:546-1120): particle 1 the
 * special electron, 2 a regular electron, 3 the proton; the two-geminal
 * terms only (:700-1106), in three outputs (:1108-1118). */
static void rxchf2_prim(const sys_t *S, const prim_t *b, const prim_t *k, double *ansE1, double *ansE2,
                        double *ansS)
{
    const prim_t *b1 = &b[0], *b2 = &b[1], *b3 = &b[2], *k1 = &k[0], *k2 = &k[1], *k3 = &k[2];
    double gTPg = 0, gTEg1 = 0, gTEg2 = 0, gTEg3 = 0, gVPCg = 0, gVECg1 = 0, gVECg2 = 0, gVECg3 = 0;
    double gVEPg1 = 0, gVEPg2 = 0, gVEPg3 = 0, gVEEg1 = 0, gVEEg2 = 0, gsg = 0;
    for (i64 ik = 0; ik < S->ngtg; ++ik)
        for (i64 il = 0; il < S->ngtg; ++il) {
            double gk = S->gamma[ik], gl = S->gamma[il];
            /* G3_MD_KE(3,1,2; gamA12, gamB13), pmass (:722) */
            double xgTPg = g3(OP_KIN, b3, b1, b2, k3, k1, k2, gk, 0, 0, 0, gl, 0, NULL, S->pmass);
            /* G3_MD_KE(1,2,3; gamA13, gamB23) (:737) */
            double xgTEg1 = g3(OP_KIN, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, NULL, 1.0);
            /* G3_MD_KE(2,1,3; gamA23, gamB13) (:754) */
            double xgTEg2 = g3(OP_KIN, b2, b1, b3, k2, k1, k3, 0, 0, gk, 0, gl, 0, NULL, 1.0);
            /* G3_MD_KE(2,1,3; gamA23, gamB23) (:770) */
            double xgTEg3 = g3(OP_KIN, b2, b1, b3, k2, k1, k3, 0, 0, gk, 0, 0, gl, NULL, 1.0);
            double xgVPCg = 0, xgVECg1 = 0, xgVECg2 = 0, xgVECg3 = 0;
            for (i64 c = 0; c < S->nat; ++c) {
                const double *C = S->cat + 3 * c;
                double z = S->zan[c], v;
                v = g3(OP_NUC, b3, b1, b2, k3, k1, k2, gk, 0, 0, 0, gl, 0, C, 1.0); /* :802 */
                underflow(&v);
                xgVPCg += z * v;
                v = g3(OP_NUC, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, C, 1.0); /* :821 */
                underflow(&v);
                xgVECg1 -= z * v;
                v = g3(OP_NUC, b2, b1, b3, k2, k1, k3, 0, 0, gk, 0, gl, 0, C, 1.0); /* :838 */
                underflow(&v);
                xgVECg2 -= z * v;
                v = g3(OP_NUC, b2, b1, b3, k2, k1, k3, 0, 0, gk, 0, 0, gl, C, 1.0); /* :857 */
                underflow(&v);
                xgVECg3 -= z * v;
            }
            /* G3_MD_xgVeeg(1,3,2; gamA12, gamB23): 1/r(e1,p) (:876) */
            double xgVEPg1 = g3(OP_COUL, b1, b3, b2, k1, k3, k2, gk, 0, 0, 0, 0, gl, NULL, 1.0);
            underflow(&xgVEPg1);
            xgVEPg1 = -xgVEPg1;
            /* G3_MD_xgVeeg(2,3,1; gamA23, gamB12): 1/r(e2,p) (:893) */
            double xgVEPg2 = g3(OP_COUL, b2, b3, b1, k2, k3, k1, 0, 0, gk, gl, 0, 0, NULL, 1.0);
            underflow(&xgVEPg2);
            xgVEPg2 = -xgVEPg2;
            /* G3_MD_xgVeeg(2,3,1; gamA23, gamB23) (:910) */
            double xgVEPg3 = g3(OP_COUL, b2, b3, b1, k2, k3, k1, 0, 0, gk, 0, 0, gl, NULL, 1.0);
            underflow(&xgVEPg3);
            xgVEPg3 = -xgVEPg3;
            /* G3_MD_xgVeeg(1,2,3; gamA13, gamB13): 1/r(e1,e2) (:929) */
            double xgVEEg1 = g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, gl, 0, NULL, 1.0);
            underflow(&xgVEEg1);
            /* G3_MD_xgVeeg(1,2,3; gamA13, gamB23) (:945) */
            double xgVEEg2 = g3(OP_COUL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, NULL, 1.0);
            underflow(&xgVEEg2);
            /* G3_MD_xggs(1,2,3; gamA13, gamB23) (:962) */
            double xgsg = g3(OP_OVL, b1, b2, b3, k1, k2, k3, 0, gk, 0, 0, 0, gl, NULL, 1.0);
            double bb = S->bcoef[ik] * S->bcoef[il];
            gTPg += bb * xgTPg;
            gTEg1 += bb * xgTEg1;
            gTEg2 += bb * xgTEg2;
            gTEg3 += bb * xgTEg3;
            gVPCg += bb * xgVPCg;
            gVECg1 += bb * xgVECg1;
            gVECg2 += bb * xgVECg2;
            gVECg3 += bb * xgVECg3;
            gVEPg1 += bb * xgVEPg1;
            gVEPg2 += bb * xgVEPg2;
            gVEPg3 += bb * xgVEPg3;
            gVEEg1 += bb * xgVEEg1;
            gVEEg2 += bb * xgVEEg2;
            gsg += bb * xgsg;
        }
    *ansE1 = gTEg3 + gVECg3 + gVEPg3 + gVEEg1;
    *ansE2 = gTPg + gVPCg + gTEg1 + gVECg1 + gTEg2 + gVECg2 + gVEPg1 + gVEPg2 + gVEEg2;
    *ansS = gsg;
}

/* xcalc_GAM3_MD (gam_3_OMP.f:617-983): particles e1, e2, e3, p */
static void gam3_prim(const sys_t *S, const prim_t *b, const prim_t *k, double *ans)
{
    double gVEE = 0.0;
    for (i64 ik = 0; ik < S->ngtg; ++ik) {
        double xx = g2(OP_OVL, &b[0], &b[3], &k[0], &k[3], S->gamma[ik], 0.0, NULL, 1.0);
        double yy = gfvee(&b[1], &b[2], &k[1], &k[2]);
        underflow(&xx);
        underflow(&yy);
        gVEE += S->bcoef[ik] * (xx * yy);
    }
    double gHEg = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (i64 ik = 0; ik < S->ngtg; ++ik)
        for (i64 il = 0; il < S->ngtg; ++il) {
            double gk = S->gamma[ik], gl = S->gamma[il];
            double xke = g1(OP_KIN, &b[2], &k[2], NULL, 1.0);
            double vc = 0.0;
            for (i64 c = 0; c < S->nat; ++c) vc += S->zan[c] * g1(OP_NUC, &b[2], &k[2], S->cat + 3 * c, 1.0);
            double xx = xke + vc * (-1.0);
            double yy = g3(OP_OVL, &b[0], &b[1], &b[3], &k[0], &k[1], &k[3], 0, gk, 0, 0, 0, gl, NULL, 1.0);
            double zz = g4(1, b, k, gk, gl);
            underflow(&xx);
            underflow(&yy);
            underflow(&zz);
            double xgHEg = (xx * yy) + (zz * (-1.0));
            xx = g2(OP_OVL, &b[0], &b[3], &k[0], &k[3], gk, gl, NULL, 1.0);
            yy = gfvee(&b[1], &b[2], &k[1], &k[2]);
            underflow(&xx);
            underflow(&yy);
            double xgVEEg1 = xx * yy;
            double xgVEEg2 = g4(0, b, k, gk, gl);
            double bb = S->bcoef[ik] * S->bcoef[il];
            gHEg += bb * xgHEg;
            gVEEg1 += bb * xgVEEg1;
            gVEEg2 += bb * xgVEEg2;
        }
    *ans = 2.0 * gVEE * 0.5 + gHEg + gVEEg1 * 0.5 + 4.0 * gVEEg2 * 0.5;
}

/* RXCHFmult_XCHF_GAM3_MD (rxchfmult/rxchfmult_gam3-xchf.f:630-1023): no
 * one-geminal gVEE term; the Coulomb terms of the (k,l) loop with the
 * electrons relabelled: G4_MD_xgVepg(e2, e3, e1, p) (:870), G2_MD_xggs(e2, p)
 * times gfvee(e1, e3) (:910, :923), G4_MD_xgVeeg(e2, e1, e3, p) (:967). */
static void gam3xchf_prim(const sys_t *S, const prim_t *b, const prim_t *k, double *ans)
{
    double gHEg = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (i64 ik = 0; ik < S->ngtg; ++ik)
        for (i64 il = 0; il < S->ngtg; ++il) {
            double gk = S->gamma[ik], gl = S->gamma[il];
            double xke = g1(OP_KIN, &b[2], &k[2], NULL, 1.0);
            double vc = 0.0;
            for (i64 c = 0; c < S->nat; ++c) vc += S->zan[c] * g1(OP_NUC, &b[2], &k[2], S->cat + 3 * c, 1.0);
            double xx = xke + vc * (-1.0);
            double yy = g3(OP_OVL, &b[0], &b[1], &b[3], &k[0], &k[1], &k[3], 0, gk, 0, 0, 0, gl, NULL, 1.0);
            prim_t bv[4] = {b[1], b[2], b[0], b[3]}, kv[4] = {k[1], k[2], k[0], k[3]};
            double zz = g4(1, bv, kv, gk, gl);
            underflow(&xx);
            underflow(&yy);
            underflow(&zz);
            double xgHEg = (xx * yy) + (zz * (-1.0));
            xx = g2(OP_OVL, &b[1], &b[3], &k[1], &k[3], gk, gl, NULL, 1.0);
            yy = gfvee(&b[0], &b[2], &k[0], &k[2]);
            underflow(&xx);
            underflow(&yy);
            double xgVEEg1 = xx * yy;
            prim_t bw[4] = {b[1], b[0], b[2], b[3]}, kw[4] = {k[1], k[0], k[2], k[3]};
            double xgVEEg2 = g4(0, bw, kw, gk, gl);
            double bb = S->bcoef[ik] * S->bcoef[il];
            gHEg += bb * xgHEg;
            gVEEg1 += bb * xgVEEg1;
            gVEEg2 += bb * xgVEEg2;
        }
    *ans = gHEg + gVEEg1 * 0.5 + 4.0 * gVEEg2 * 0.5;
}

/* RXCHF_xcalc_GAM3_MD (rxchf/rxchf_gam_3_OMP.f:617-1032): e1 the special
 * electron, e2 and e3 regular, p the proton; the (k,l) terms only, two
 * outputs (:1010-1011):
 *   ans1 = gVEEg1: G2_MD_xggs(e1, p; gk, gl) x gfvee(e2, e3)       (:892-920)
 *   ans2 = gHEg + gVEEg2 + gVEEg3 with
 *     gHEg   = hcore(e3) G3_MD_xggs(e1, e2, p; g(e1,p) gk, g(e2,p) gl)
 *              - G4_MD_xgVepg(e1, e2, e3, p; gk, gl)                 (:766-889)
 *     gVEEg2 = G4_MD_xgVeeg(e1, e3, e2, p; gk, gl)                   (:922-945)
 *     gVEEg3 = G4_MD_xgVeeg(e2, e3, e1, p; gk, gl)                   (:956-984)
 * accumulated with bcoef(k) bcoef(l) (:999-1002) */
static void rxchf3_prim(const sys_t *S, const prim_t *b, const prim_t *k, double *ans1, double *ans2)
{
    double gHEg = 0, gVEEg1 = 0, gVEEg2 = 0, gVEEg3 = 0;
    for (i64 ik = 0; ik < S->ngtg; ++ik)
        for (i64 il = 0; il < S->ngtg; ++il) {
            double gk = S->gamma[ik], gl = S->gamma[il];
            double xke = g1(OP_KIN, &b[2], &k[2], NULL, 1.0);
            double vc = 0.0;
            for (i64 c = 0; c < S->nat; ++c) vc += S->zan[c] * g1(OP_NUC, &b[2], &k[2], S->cat + 3 * c, 1.0);
            double xx = xke + vc * (-1.0);
            double yy = g3(OP_OVL, &b[0], &b[1], &b[3], &k[0], &k[1], &k[3], 0, gk, 0, 0, 0, gl, NULL, 1.0);
            double zz = g4(1, b, k, gk, gl);
            underflow(&xx);
            underflow(&yy);
            underflow(&zz);
            double xgHEg = (xx * yy) + (zz * (-1.0));
            xx = g2(OP_OVL, &b[0], &b[3], &k[0], &k[3], gk, gl, NULL, 1.0);
            yy = gfvee(&b[1], &b[2], &k[1], &k[2]);
            underflow(&xx);
            underflow(&yy);
            double xgVEEg1 = xx * yy;
            prim_t b2[4] = {b[0], b[2], b[1], b[3]}, k2[4] = {k[0], k[2], k[1], k[3]};
            double xgVEEg2 = g4(0, b2, k2, gk, gl);
            prim_t b3[4] = {b[1], b[2], b[0], b[3]}, k3[4] = {k[1], k[2], k[0], k[3]};
            double xgVEEg3 = g4(0, b3, k3, gk, gl);
            double bb = S->bcoef[ik] * S->bcoef[il];
            gHEg += bb * xgHEg;
            gVEEg1 += bb * xgVEEg1;
            gVEEg2 += bb * xgVEEg2;
            gVEEg3 += bb * xgVEEg3;
        }
    *ans1 = gVEEg1;
    *ans2 = gHEg + gVEEg2 + gVEEg3;
}

/* ------------------------------------------------------------------ */
/* Basis access (Fortran arrays of the reference)                      */
/* ------------------------------------------------------------------ */
typedef struct {
    i64 nebf, npebf, npbf;
    const i64 *kpestr, *kpeend, *ampeb2c; /* 1-based */
    const double *agebfcc, *agnbfcc;
    const double *elcex, *nucex;
    const i64 *elcam, *nucam;
    const double *elcbfc, *nucbfc;
} basis_t;

static prim_t eprim(const basis_t *B, i64 i)
{ /* i: 0-based primitive index */
    prim_t p;
    for (int d = 0; d < 3; ++d) {
        p.l[d] = (int)B->elcam[d * B->npebf + i];
        p.A[d] = B->elcbfc[d * B->npebf + i];
    }
    p.a = B->elcex[i];
    return p;
}
static prim_t pprim(const basis_t *B, i64 i)
{
    prim_t p;
    for (int d = 0; d < 3; ++d) {
        p.l[d] = (int)B->nucam[d * B->npbf + i];
        p.A[d] = B->nucbfc[d * B->npbf + i];
    }
    p.a = B->nucex[i];
    return p;
}

/* ------------------------------------------------------------------ */
/* Exported entry points (ctypes)                                      */
/* ------------------------------------------------------------------ */
#define SYS_ARGS i64 nat, i64 ngtg, double pmass, const double *cat, const double *zan, const double *bcoef, const double *gamma
#define SYS_INIT sys_t S = {nat, ngtg, pmass, cat, zan, bcoef, gamma}
#define BAS_ARGS                                                                                                   \
    i64 nebf, i64 npebf, i64 npbf, const i64 *kpestr, const i64 *kpeend, const i64 *ampeb2c, const double *agebfcc, \
        const double *agnbfcc, const double *elcex, const double *nucex, const i64 *elcam, const i64 *nucam,       \
        const double *elcbfc, const double *nucbfc
#define BAS_INIT                                                                                                   \
    basis_t B = {nebf, npebf, npbf, kpestr, kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc}

/* Evaluate a list of primitive tuples.  idx holds, per tuple, 0-based
 * primitive indices: cls 1 -> (ie1, ip, je1, jp); cls 2 -> (ie1,ie2,ip,je1,je2,jp);
 * cls 3 -> (ie1,ie2,ie3,ip,je1,je2,je3,jp).  Outputs ansE (and ansS for
 * cls 1,2). */
static void prim_list_impl(const basis_t *Bp, const sys_t *Sp, i64 cls, i64 ntup, const i64 *idx, double *outE,
                           double *outS, double *outX)
{
    /* cls 4: the RXCHF-ne e-e-p tuple; cls 5, 6: the RXCHFmult e-e-p and
     * e-e-e-p tuples (primitive indices over one combined list) */
    int nel = (cls == 4 || cls == 5 || cls == 9) ? 2 : (cls == 6 || cls == 7 || cls == 10) ? 3 : cls == 8 ? 1 : (int)cls;
    int w = 2 * (nel + 1);
#pragma omp parallel for schedule(dynamic, 4)
    for (i64 t = 0; t < ntup; ++t) {
        const i64 *x = idx + t * w;
        prim_t b[4], k[4];
        int ne = nel;
        for (int i = 0; i < ne; ++i) {
            b[i] = eprim(Bp, x[i]);
            k[i] = eprim(Bp, x[ne + 1 + i]);
        }
        b[ne] = pprim(Bp, x[ne]);
        k[ne] = pprim(Bp, x[2 * ne + 1]);
        double e = 0, s = 0, e2 = 0;
        if (cls == 1)
            gam1_prim(Sp, &b[0], &b[1], &k[0], &k[1], &e, &s);
        else if (cls == 2)
            gam2_prim(Sp, b, k, &e, &s);
        else if (cls == 4)
            rxne_prim(Sp, b, k, &e, &s);
        else if (cls == 5)
            gam2_prim_impl(Sp, b, k, 1, &e, &s, &e2, NULL);
        else if (cls == 6)
            gam3int_prim_ex(Sp, b, k, &e, &s, &e2);
        else if (cls == 7)
            gam3xchf_prim(Sp, b, k, &e);
        else if (cls == 8)
            gam1_prim_impl(Sp, 1, &b[0], &b[1], &k[0], &k[1], &e, &s);
        else if (cls == 9) /* (ansE1, ansS, ansE2) */
            rxchf2_prim(Sp, b, k, &e, &e2, &s);
        else if (cls == 10) /* (ans1, ans2) */
            rxchf3_prim(Sp, b, k, &e, &s);
        else
            gam3_prim(Sp, b, k, &e);
        outE[t] = e;
        if (outS) outS[t] = s;
        if (outX) outX[t] = e2;
    }
}

void orc_prim_list(BAS_ARGS, SYS_ARGS, i64 cls, i64 ntup, const i64 *idx, double *outE, double *outS)
{
    BAS_INIT;
    SYS_INIT;
    prim_list_impl(&B, &S, cls, ntup, idx, outE, outS, NULL);
}

/* Same with the third output (ansE2 of cls 5, zero otherwise). */
void orc_prim_list3(BAS_ARGS, SYS_ARGS, i64 cls, i64 ntup, const i64 *idx, double *outE, double *outS, double *outX)
{
    BAS_INIT;
    SYS_INIT;
    prim_list_impl(&B, &S, cls, ntup, idx, outE, outS, outX);
}

/* Elementary integral probe: op (0 ovl,1 nuc,2 coul,3 kin), n particles,
 * powers[2n][3], exps[2n], centres[2n][3] (bra then ket), gem[n][n] bra and
 * gemk[n][n] ket geminal exponents, a,b,C,mass. */
double orc_gint(i64 op, i64 n, const i64 *pw, const double *ex, const double *cen, const double *gemA,
                const double *gemB, i64 a, i64 b, const double *C, double mass, double pi)
{
    gint_t g;
    gi_init(&g, (int)n);
    if (pi > 0.0) g.pi = pi;
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            g.bra[i].l[d] = (int)pw[3 * i + d];
            g.ket[i].l[d] = (int)pw[3 * (n + i) + d];
            g.bra[i].A[d] = cen[3 * i + d];
            g.ket[i].A[d] = cen[3 * (n + i) + d];
        }
        g.bra[i].a = ex[i];
        g.ket[i].a = ex[n + i];
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) gi_gem(&g, i, j, gemA[i * n + j], gemB[i * n + j]);
    g.op = (int)op;
    g.a = (int)a;
    g.b = (int)b;
    g.mass = mass;
    if (C) memcpy(g.C, C, sizeof(g.C));
    return gint_eval(&g);
}

/* index_GAM_2PK / index_GAM_3PK (pack_GAM.f:72-147), 0-based */
static inline i64 ix2(i64 ne, i64 np, i64 ip, i64 jp, i64 i1, i64 j1, i64 i2, i64 j2)
{
    return ((((ip * np + jp) * ne + i1) * ne + j1) * ne + i2) * ne + j2;
}
static inline i64 ix3(i64 ne, i64 np, i64 ip, i64 jp, i64 i1, i64 j1, i64 i2, i64 j2, i64 i3, i64 j3)
{
    return ix2(ne, np, ip, jp, i1, j1, i2, j2) * ne * ne + i3 * ne + j3;
}

/* GAM1_OMP_MD (gam_1_OMP.f:2-397): contracted e-p integrals, layout
 * pack_4D(nebf,nebf,npbf, jec1,iec1,jp,ip) */
static void gam1_impl(const basis_t *Bp, const sys_t *Sp, int mult, double *GAM1, double *GAM1s)
{
    basis_t B = *Bp;
    sys_t S = *Sp;
    i64 nebf = B.nebf, npebf = B.npebf, npbf = B.npbf;
    const i64 *ampeb2c = B.ampeb2c;
    const double *agebfcc = B.agebfcc, *agnbfcc = B.agnbfcc;
    i64 ng1 = nebf * nebf * npbf * npbf;
    memset(GAM1, 0, sizeof(double) * ng1);
    memset(GAM1s, 0, sizeof(double) * ng1);
    for (i64 ip = 0; ip < npbf; ++ip)
        for (i64 jp = 0; jp < npbf; ++jp)
            for (i64 ie = 0; ie < npebf; ++ie)
                for (i64 je = 0; je < npebf; ++je) {
                    prim_t eb = eprim(&B, ie), ek = eprim(&B, je), pb = pprim(&B, ip), pk = pprim(&B, jp);
                    double e, s;
                    gam1_prim_impl(&S, mult, &eb, &pb, &ek, &pk, &e, &s);
                    i64 ic = ampeb2c[ie] - 1, jc = ampeb2c[je] - 1;
                    double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[ie] * agebfcc[je];
                    i64 ia = ((ip * npbf + jp) * nebf + ic) * nebf + jc;
                    GAM1[ia] += e * cf;
                    GAM1s[ia] += s * cf;
                }
}

void orc_gam1(BAS_ARGS, SYS_ARGS, double *GAM1, double *GAM1s)
{
    BAS_INIT;
    SYS_INIT;
    gam1_impl(&B, &S, 0, GAM1, GAM1s);
}

/* RXCHFmult_GAM1_MD (rxchfmult/rxchfmult_gam1.f:2-398): the loops of
 * GAM1_OMP_MD over RXCHFmult_xcalc_GAM1_MD */
void orc_rxchfmult_gam1(BAS_ARGS, SYS_ARGS, double *GAM1, double *GAM1s)
{
    BAS_INIT;
    SYS_INIT;
    gam1_impl(&B, &S, 1, GAM1, GAM1s);
}

/* GAM2_IC1 (gam2_IC1_OMP.f:2-212) with contract_omega2_conv
 * (gam2_CONV_OMP.f:388-579) and the (x12+x21)/2 symmetrisation.  With flush
 * set every primitive value goes through underflow() first, as
 * contract_omega2 of the direct-SCF path does (gam2_DSCF_OMP.f:514-515). */
static void gam2_ic1_impl(const basis_t *Bp, const sys_t *Sp, int flush, double *GM2, double *GM2s)
{
    basis_t B = *Bp;
    sys_t S = *Sp;
    i64 nebf = B.nebf, npbf = B.npbf;
    const i64 *kpestr = B.kpestr, *kpeend = B.kpeend;
    const double *agebfcc = B.agebfcc, *agnbfcc = B.agnbfcc;
    i64 ne = nebf, np = npbf, ng2 = ne * ne * ne * ne * np * np;
    double *X = (double *)malloc(sizeof(double) * ng2), *Xs = (double *)malloc(sizeof(double) * ng2);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            double om = 0.0, oms = 0.0;
                            for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                    for (i64 a2 = kpestr[i2] - 1; a2 < kpeend[i2]; ++a2)
                                        for (i64 c2 = kpestr[j2] - 1; c2 < kpeend[j2]; ++c2) {
                                            prim_t b[3] = {eprim(&B, a1), eprim(&B, a2), pprim(&B, ip)};
                                            prim_t k[3] = {eprim(&B, c1), eprim(&B, c2), pprim(&B, jp)};
                                            double e, s;
                                            gam2_prim(&S, b, k, &e, &s);
                                            if (flush) {
                                                underflow(&e);
                                                underflow(&s);
                                            }
                                            double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] * agebfcc[c1] *
                                                        agebfcc[a2] * agebfcc[c2];
                                            om += e * cf;
                                            oms += s * cf;
                                        }
                            i64 ia = ix2(ne, np, ip, jp, i1, j1, i2, j2);
                            X[ia] = om;
                            Xs[ia] = oms;
                        }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            i64 a12 = ix2(ne, np, ip, jp, i1, j1, i2, j2), a21 = ix2(ne, np, ip, jp, i2, j2, i1, j1);
                            GM2[a12] = (X[a12] + X[a21]) / 2.0;
                            GM2s[a12] = (Xs[a12] + Xs[a21]) / 2.0;
                        }
    free(X);
    free(Xs);
}

void orc_gam2_ic1(BAS_ARGS, SYS_ARGS, double *GM2, double *GM2s)
{
    BAS_INIT;
    SYS_INIT;
    gam2_ic1_impl(&B, &S, 0, GM2, GM2s);
}

/* GAM3_IC1 (gam3_IC1_OMP.f:2-263) with contract_omega3_conv
 * (gam3_CONV_OMP.f:479-705) and the six-permutation average. */
typedef void (*prim3_fn)(const sys_t *, const prim_t *, const prim_t *, double *);

static void gam3_ic1_impl(const basis_t *Bp, const sys_t *Sp, prim3_fn prim, double *GM3)
{
    basis_t B = *Bp;
    sys_t S = *Sp;
    i64 nebf = B.nebf, npbf = B.npbf;
    const i64 *kpestr = B.kpestr, *kpeend = B.kpeend;
    const double *agebfcc = B.agebfcc, *agnbfcc = B.agnbfcc;
    i64 ne = nebf, np = npbf, ng3 = ne * ne * ne * ne * ne * ne * np * np;
    double *X = (double *)malloc(sizeof(double) * ng3);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3) {
                                    double om = 0.0;
                                    for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                        for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                            for (i64 a2 = kpestr[i2] - 1; a2 < kpeend[i2]; ++a2)
                                                for (i64 c2 = kpestr[j2] - 1; c2 < kpeend[j2]; ++c2)
                                                    for (i64 a3 = kpestr[i3] - 1; a3 < kpeend[i3]; ++a3)
                                                        for (i64 c3 = kpestr[j3] - 1; c3 < kpeend[j3]; ++c3) {
                                                            prim_t b[4] = {eprim(&B, a1), eprim(&B, a2), eprim(&B, a3),
                                                                           pprim(&B, ip)};
                                                            prim_t k[4] = {eprim(&B, c1), eprim(&B, c2), eprim(&B, c3),
                                                                           pprim(&B, jp)};
                                                            double e;
                                                            prim(&S, b, k, &e);
                                                            double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] *
                                                                        agebfcc[c1] * agebfcc[a2] * agebfcc[c2] *
                                                                        agebfcc[a3] * agebfcc[c3];
                                                            om += e * cf;
                                                        }
                                    X[ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3)] = om;
                                }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3) {
                                    double x = X[ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3)] +
                                               X[ix3(ne, np, ip, jp, i1, j1, i3, j3, i2, j2)] +
                                               X[ix3(ne, np, ip, jp, i2, j2, i1, j1, i3, j3)] +
                                               X[ix3(ne, np, ip, jp, i2, j2, i3, j3, i1, j1)] +
                                               X[ix3(ne, np, ip, jp, i3, j3, i1, j1, i2, j2)] +
                                               X[ix3(ne, np, ip, jp, i3, j3, i2, j2, i1, j1)];
                                    GM3[ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3)] = x / 6.0;
                                }
    free(X);
}

void orc_gam3_ic1(BAS_ARGS, SYS_ARGS, double *GM3)
{
    BAS_INIT;
    SYS_INIT;
    gam3_ic1_impl(&B, &S, gam3_prim, GM3);
}

/* RXCHFmult_GAM3_XCHF (rxchfmult/rxchfmult_gam3-xchf.f:2-268) with
 * RXCHFmult_contract_omega3_XCHF (:402-629) and the six-permutation average */
void orc_rxchfmult_gam3_xchf(BAS_ARGS, SYS_ARGS, double *GM3)
{
    BAS_INIT;
    SYS_INIT;
    gam3_ic1_impl(&B, &S, gam3xchf_prim, GM3);
}

/* calc_GAM_ee (gam_ee.f:2-168): contracted ERIs, layout
 * pack_4D(nebf,nebf,nebf, jec2,iec2,jec1,iec1) */
void orc_gam_ee(i64 nebf, i64 npebf, const i64 *ampeb2c, const double *agebfcc, const double *elcex,
                const i64 *elcam, const double *elcbfc, double *GEE)
{
    basis_t B;
    memset(&B, 0, sizeof(B));
    B.npebf = npebf;
    B.elcex = elcex;
    B.elcam = elcam;
    B.elcbfc = elcbfc;
    i64 ne = nebf;
    memset(GEE, 0, sizeof(double) * ne * ne * ne * ne);
    for (i64 ie1 = 0; ie1 < npebf; ++ie1)
        for (i64 je1 = 0; je1 < npebf; ++je1)
            for (i64 ie2 = 0; ie2 < npebf; ++ie2)
                for (i64 je2 = 0; je2 < npebf; ++je2) {
                    prim_t a1 = eprim(&B, ie1), a2 = eprim(&B, ie2), c1 = eprim(&B, je1), c2 = eprim(&B, je2);
                    double v = gfvee(&a1, &a2, &c1, &c2);
                    i64 ia = (((ampeb2c[ie1] - 1) * ne + ampeb2c[je1] - 1) * ne + ampeb2c[ie2] - 1) * ne +
                             ampeb2c[je2] - 1;
                    GEE[ia] += v * agebfcc[ie1] * agebfcc[je1] * agebfcc[ie2] * agebfcc[je2];
                }
}

/* calc_GAM_4_integral (gam_4_OMP.f:964-1192) */
static double gam4_int(i64 ne, i64 np, const double *G2s, const double *Gee, i64 ip, i64 jp, i64 ie1, i64 je1, i64 ie2,
                       i64 je2, i64 ie3, i64 je3, i64 ie4, i64 je4)
{
#define EE(a, b, c, d) Gee[(((a) * ne + (b)) * ne + (c)) * ne + (d)]
    double ee34 = EE(ie3, je3, ie4, je4), ee24 = EE(ie2, je2, ie4, je4), ee23 = EE(ie2, je2, ie3, je3);
    double ee14 = EE(ie1, je1, ie4, je4), ee13 = EE(ie1, je1, ie3, je3), ee12 = EE(ie1, je1, ie2, je2);
#undef EE
    underflow(&ee34); underflow(&ee24); underflow(&ee23);
    underflow(&ee14); underflow(&ee13); underflow(&ee12);
    double gs12 = G2s[ix2(ne, np, ip, jp, ie1, je1, ie2, je2)], gs13 = G2s[ix2(ne, np, ip, jp, ie1, je1, ie3, je3)];
    double gs14 = G2s[ix2(ne, np, ip, jp, ie1, je1, ie4, je4)], gs23 = G2s[ix2(ne, np, ip, jp, ie2, je2, ie3, je3)];
    double gs24 = G2s[ix2(ne, np, ip, jp, ie2, je2, ie4, je4)], gs34 = G2s[ix2(ne, np, ip, jp, ie3, je3, ie4, je4)];
    underflow(&gs12); underflow(&gs13); underflow(&gs14);
    underflow(&gs23); underflow(&gs24); underflow(&gs34);
    double x1 = 0.5 * ee34 * gs12, x2 = 0.5 * ee24 * gs13, x3 = 0.5 * ee23 * gs14;
    double x4 = 0.5 * ee14 * gs23, x5 = 0.5 * ee13 * gs24, x6 = 0.5 * ee12 * gs34;
    return (x1 + x2 + x3 + x4 + x5 + x6) / 6.0;
}

/* symm_gam4 (gam_4_OMP.f:712-918): 24 ket permutations with fixed weights.
 * Arguments are (ii,i),(jj,j),(kk,k),(ll,l) = (bra,ket) of electrons 1..4. */
static double symm_gam4(i64 ne, i64 np, const double *G2s, const double *Gee, i64 ip, i64 jp, i64 ii, i64 i, i64 jj,
                        i64 j, i64 kk, i64 k, i64 ll, i64 l)
{
    /* ket order for x1..x24 and their weights (gam_4_OMP.f:788-856) */
    static const int perm[24][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {2, 0, 1, 3}, {0, 2, 1, 3}, {1, 2, 0, 3}, {2, 1, 0, 3},
                                    {3, 1, 0, 2}, {1, 3, 0, 2}, {0, 3, 1, 2}, {3, 0, 1, 2}, {1, 0, 3, 2}, {0, 1, 3, 2},
                                    {0, 2, 3, 1}, {2, 0, 3, 1}, {3, 0, 2, 1}, {0, 3, 2, 1}, {2, 3, 0, 1}, {3, 2, 0, 1},
                                    {3, 2, 1, 0}, {2, 3, 1, 0}, {1, 3, 2, 0}, {3, 1, 2, 0}, {2, 1, 3, 0}, {1, 2, 3, 0}};
    static const double wt[24] = {16, -8, 4, -8, 4, -8, 4, -2, 4, -2, 4, -8, 4, -2, 4, -8, 4, -2, 4, -2, 4, -8, 4, -2};
    i64 kets[4] = {i, j, k, l};
    double ans = 0.0;
    for (int p = 0; p < 24; ++p) {
        double x = gam4_int(ne, np, G2s, Gee, ip, jp, ii, kets[perm[p][0]], jj, kets[perm[p][1]], kk, kets[perm[p][2]],
                            ll, kets[perm[p][3]]);
        ans += wt[p] * x;
    }
    return ans;
}

/* GAM4_ICR (gam4_ICR_OMP.f:2-270): in-core 5-particle tensor, int64 index
 * index_GAM_4PK2 (pack_GAM.f:196-240), value symm_gam4/16. */
void orc_gam4_icr(i64 nebf, i64 npbf, const double *GAM2s, const double *GAMee, double *GM4)
{
    i64 ne = nebf, np = npbf;
#pragma omp parallel for collapse(2)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3)
                                    for (i64 i4 = 0; i4 < ne; ++i4)
                                        for (i64 j4 = 0; j4 < ne; ++j4) {
                                            i64 ia = ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3) * ne * ne + i4 * ne + j4;
                                            GM4[ia] = symm_gam4(ne, np, GAM2s, GAMee, ip, jp, i1, j1, i2, j2, i3, j3,
                                                                i4, j4) /
                                                      16.0;
                                        }
}

/* EG2IC1 / thread_EG2ICR (gam2_IC1_OMP.f:339-570).  FE, FP accumulate
 * (add2fock); SE, SP are zeroed then filled; E, S overwritten. */
static void eg2_impl(i64 nebf, i64 npbf, const double *DE, const double *DP, const double *GM2, const double *GM2s,
                     double *FE, double *SE, double *FP, double *SP, double *E, double *S, int zero_s)
{
    i64 ne = nebf, np = npbf;
    double *xfe = (double *)calloc(ne * ne, sizeof(double)), *xfp = (double *)calloc(np * np, sizeof(double));
    if (zero_s) {
        memset(SE, 0, sizeof(double) * ne * ne);
        memset(SP, 0, sizeof(double) * np * np);
    }
    double e = 0.0, s = 0.0;
    /* Fortran DE(i,j) = DE[j*ne + i] (column-major) */
#define DEm(i, j) DE[(j) * ne + (i)]
#define DPm(i, j) DP[(j) * np + (i)]
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            i64 a12 = ix2(ne, np, ip, jp, i1, j1, i2, j2), a21 = ix2(ne, np, ip, jp, i1, j2, i2, j1);
                            double v = GM2[a12] - 0.5 * GM2[a21];
                            double vs = GM2s[a12] - 0.5 * GM2s[a21];
                            xfe[j1 * ne + i1] += DPm(ip, jp) * DEm(i2, j2) * 2.0 * v;
                            SE[j1 * ne + i1] += DPm(ip, jp) * DEm(i2, j2) * vs;
                            xfp[jp * np + ip] += DEm(i1, j1) * DEm(i2, j2) * v;
                            SP[jp * np + ip] += DEm(i1, j1) * DEm(i2, j2) * vs;
                            e += DEm(i1, j1) * DEm(i2, j2) * DPm(ip, jp) * v;
                            s += DEm(i1, j1) * DEm(i2, j2) * DPm(ip, jp) * vs;
                        }
    for (i64 i = 0; i < ne * ne; ++i) FE[i] += xfe[i];
    for (i64 i = 0; i < np * np; ++i) FP[i] += xfp[i];
    *E = e;
    *S = s;
    free(xfe);
    free(xfp);
}

void orc_eg2ic1(i64 nebf, i64 npbf, const double *DE, const double *DP, const double *GM2, const double *GM2s,
                double *FE, double *SE, double *FP, double *SP, double *E, double *S)
{
    eg2_impl(nebf, npbf, DE, DP, GM2, GM2s, FE, SE, FP, SP, E, S, 1);
}

/* EG3IC1 / thread_EG3ICR (gam3_IC1_OMP.f:433-725) */
void orc_eg3ic1(i64 nebf, i64 npbf, const double *DE, const double *DP, const double *GM3, double *FE, double *FP,
                double *E)
{
    i64 ne = nebf, np = npbf;
    double *xfe = (double *)calloc(ne * ne, sizeof(double)), *xfp = (double *)calloc(np * np, sizeof(double));
    double e = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3) {
                                    double o123 = GM3[ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3)];
                                    double o213 = GM3[ix3(ne, np, ip, jp, i1, j2, i2, j1, i3, j3)];
                                    double o312 = GM3[ix3(ne, np, ip, jp, i1, j3, i2, j1, i3, j2)];
                                    double o132 = GM3[ix3(ne, np, ip, jp, i1, j1, i2, j3, i3, j2)];
                                    double o231 = GM3[ix3(ne, np, ip, jp, i1, j2, i2, j3, i3, j1)];
                                    double o321 = GM3[ix3(ne, np, ip, jp, i1, j3, i2, j2, i3, j1)];
                                    double g = 8.0 * o123 - 4.0 * o213 + 2.0 * o312 - 4.0 * o132 + 2.0 * o231 -
                                               4.0 * o321;
                                    g = g / 8.0;
                                    xfe[j1 * ne + i1] += DPm(ip, jp) * DEm(i2, j2) * DEm(i3, j3) * 3.0 * g;
                                    xfp[jp * np + ip] += DEm(i1, j1) * DEm(i2, j2) * DEm(i3, j3) * g;
                                    e += DEm(i1, j1) * DEm(i2, j2) * DEm(i3, j3) * DPm(ip, jp) * g;
                                }
    for (i64 i = 0; i < ne * ne; ++i) FE[i] += xfe[i];
    for (i64 i = 0; i < np * np; ++i) FP[i] += xfp[i];
    *E = e;
    free(xfe);
    free(xfp);
}

/* EG4ICR / thread_EG4ICR (gam4_ICR_OMP.f:273-490) */
void orc_eg4icr(i64 nebf, i64 npbf, const double *DE, const double *DP, const double *GM4, double *FE, double *FP,
                double *E)
{
    i64 ne = nebf, np = npbf;
    double *xfe = (double *)calloc(ne * ne, sizeof(double)), *xfp = (double *)calloc(np * np, sizeof(double));
    double e = 0.0;
    i64 ia = 0;
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3)
                                    for (i64 i4 = 0; i4 < ne; ++i4)
                                        for (i64 j4 = 0; j4 < ne; ++j4, ++ia) {
                                            double g = GM4[ia];
                                            xfe[j1 * ne + i1] +=
                                                DPm(ip, jp) * DEm(i2, j2) * DEm(i3, j3) * DEm(i4, j4) * 4.0 * g;
                                            xfp[jp * np + ip] += DEm(i1, j1) * DEm(i2, j2) * DEm(i3, j3) * DEm(i4, j4) * g;
                                            e += DEm(i1, j1) * DEm(i2, j2) * DEm(i3, j3) * DEm(i4, j4) * DPm(ip, jp) * g;
                                        }
    for (i64 i = 0; i < ne * ne; ++i) FE[i] += xfe[i];
    for (i64 i = 0; i < np * np; ++i) FP[i] += xfp[i];
    *E = e;
    free(xfe);
    free(xfp);
#undef DEm
#undef DPm
}

/* ------------------------------------------------------------------------
 * Direct-SCF routines: the same contracted values as the in-core path,
 * recomputed and contracted in one call.
 * ------------------------------------------------------------------------ */

/* GAM2_DSCF (gam2_DSCF_OMP.f:2-139) with thread_gam2 (:141-343) and
 * contract_omega2 (:345-559): GAM2 = OMG2(1,2) - OMG2(exchange)/2 from
 * symmetrised contracted values whose primitives went through underflow();
 * focke, fockp, se2, sp2 accumulate (add2fock, :130-133), E, S overwritten. */
void orc_gam2_dscf(BAS_ARGS, SYS_ARGS, const double *DE, const double *DP, double *FE, double *FP, double *SE2,
                   double *SP2, double *E, double *S2)
{
    BAS_INIT;
    SYS_INIT;
    i64 ng2 = nebf * nebf * nebf * nebf * npbf * npbf;
    double *G = (double *)malloc(sizeof(double) * ng2), *Gs = (double *)malloc(sizeof(double) * ng2);
    gam2_ic1_impl(&B, &S, 1, G, Gs);
    eg2_impl(nebf, npbf, DE, DP, G, Gs, FE, SE2, FP, SP2, E, S2, 0);
    free(G);
    free(Gs);
}

/* GAM3_DSCF (gam3_DSCF_OMP.f:2-365): contract_omega3 (:367-717) is the
 * six-permutation average of contract_omega3_conv, combined and contracted
 * exactly as EG3IC1 does. */
void orc_gam3_dscf(BAS_ARGS, SYS_ARGS, const double *DE, const double *DP, double *FE, double *FP, double *E)
{
    i64 ng3 = nebf * nebf * nebf * nebf * nebf * nebf * npbf * npbf;
    double *G = (double *)malloc(sizeof(double) * ng3);
    orc_gam3_ic1(nebf, npebf, npbf, kpestr, kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc,
                 nucbfc, nat, ngtg, pmass, cat, zan, bcoef, gamma, G);
    orc_eg3ic1(nebf, npbf, DE, DP, G, FE, FP, E);
    free(G);
}

/* GAM4_DSCF (gam4_DSCF_OMP.f:2-320): symm_gam4/16 evaluated per entry from
 * GAM_2s and GAM_ee and contracted as EG4ICR does, without the ng4 tensor. */
void orc_gam4_dscf(i64 nebf, i64 npbf, const double *GAM2s, const double *GAMee, const double *DE, const double *DP,
                   double *FE, double *FP, double *E)
{
    i64 ne = nebf, np = npbf;
    double *xfe = (double *)calloc(ne * ne, sizeof(double)), *xfp = (double *)calloc(np * np, sizeof(double));
    double e = 0.0;
#define DEm(i, j) DE[(j) * ne + (i)]
#define DPm(i, j) DP[(j) * np + (i)]
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3)
                                    for (i64 i4 = 0; i4 < ne; ++i4)
                                        for (i64 j4 = 0; j4 < ne; ++j4) {
                                            double g = symm_gam4(ne, np, GAM2s, GAMee, ip, jp, i1, j1, i2, j2, i3, j3,
                                                                 i4, j4) /
                                                       16.0;
                                            xfe[j1 * ne + i1] +=
                                                DPm(ip, jp) * DEm(i2, j2) * DEm(i3, j3) * DEm(i4, j4) * 4.0 * g;
                                            xfp[jp * np + ip] += DEm(i1, j1) * DEm(i2, j2) * DEm(i3, j3) * DEm(i4, j4) * g;
                                            e += DEm(i1, j1) * DEm(i2, j2) * DEm(i3, j3) * DEm(i4, j4) * DPm(ip, jp) * g;
                                        }
#undef DEm
#undef DPm
    for (i64 i = 0; i < ne * ne; ++i) FE[i] += xfe[i];
    for (i64 i = 0; i < np * np; ++i) FP[i] += xfp[i];
    *E = e;
    free(xfe);
    free(xfp);
}

/* ------------------------------------------------------------------------
 * XCUHF contractions (source/xcuhf/xcuhf_omg{2,3,4}.f): spin-resolved
 * contractions of the same in-core tensors.  FAE_* add into FAE (the alpha
 * Fock matrix; the beta one is the same call with the spins swapped,
 * xcuhf_fock.f:221, :241); UFP_* add into FP.
 * ------------------------------------------------------------------------ */
#define DAm(i, j) DAE[(j) * ne + (i)]
#define DBm(i, j) DBE[(j) * ne + (i)]
#define DPm(i, j) DP[(j) * np + (i)]

/* FAE_OMG2 / thread_FAEOMG2 (xcuhf_omg2.f:2-300); SAE2 overwritten */
void orc_fae_omg2(i64 nae, i64 nbe, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                  const double *GM2, const double *GM2s, double *FAE, double *SAE2, double *E, double *S)
{
    i64 ne = nebf, np = npbf;
    int LAA = nae >= 2, LAB = nae >= 1 && nbe >= 1;
    double *xf = (double *)calloc(ne * ne, sizeof(double));
    memset(SAE2, 0, sizeof(double) * ne * ne);
    double e = 0.0, s = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            i64 a12 = ix2(ne, np, ip, jp, i1, j1, i2, j2), a21 = ix2(ne, np, ip, jp, i1, j2, i2, j1);
                            double o12 = GM2[a12], o21 = GM2[a21], so12 = GM2s[a12], so21 = GM2s[a21];
                            double c1 = LAA ? DAm(i2, j2) * DPm(ip, jp) : 0.0;
                            double c2 = LAB ? DBm(i2, j2) * DPm(ip, jp) : 0.0;
                            xf[j1 * ne + i1] += 2.0 * (c1 * (o12 - o21) + c2 * o12);
                            SAE2[j1 * ne + i1] += 2.0 * (c1 * (so12 - so21) + c2 * so12);
                            e += DAm(i1, j1) * (c1 * (o12 - o21) + c2 * o12);
                            s += DAm(i1, j1) * (c1 * (so12 - so21) + c2 * so12);
                        }
    for (i64 k = 0; k < ne * ne; ++k) FAE[k] += xf[k];
    *E = e;
    *S = s;
    free(xf);
}

/* UFP_OMG2 / thread_UFPOMG2 (xcuhf_omg2.f:302-560); SP2 overwritten */
void orc_ufp_omg2(i64 nae, i64 nbe, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                  const double *GM2, const double *GM2s, double *FP, double *SP2, double *E, double *S)
{
    i64 ne = nebf, np = npbf;
    int LAB = nae >= 1 && nbe >= 1, LAA = nae >= 2, LBB = nbe >= 2;
    double *xf = (double *)calloc(np * np, sizeof(double));
    memset(SP2, 0, sizeof(double) * np * np);
    double e = 0.0, s = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            i64 a12 = ix2(ne, np, ip, jp, i1, j1, i2, j2), a21 = ix2(ne, np, ip, jp, i1, j2, i2, j1);
                            double o12 = GM2[a12], o21 = GM2[a21], so12 = GM2s[a12], so21 = GM2s[a21];
                            double c1 = LAB ? DAm(i1, j1) * DBm(i2, j2) + DBm(i1, j1) * DAm(i2, j2) : 0.0;
                            double c2 = LAA ? DAm(i1, j1) * DAm(i2, j2) : 0.0;
                            double c3 = LBB ? DBm(i1, j1) * DBm(i2, j2) : 0.0;
                            double v = c1 * o12 + (c2 + c3) * (o12 - o21), vs = c1 * so12 + (c2 + c3) * (so12 - so21);
                            xf[jp * np + ip] += v;
                            SP2[jp * np + ip] += vs;
                            e += DPm(ip, jp) * v;
                            s += DPm(ip, jp) * vs;
                        }
    for (i64 k = 0; k < np * np; ++k) FP[k] += xf[k];
    *E = e;
    *S = s;
    free(xf);
}

/* the six ket-permuted reads of thread_FAEOMG3 / thread_UFPOMG3 */
static void omg3_reads(i64 ne, i64 np, const double *G, i64 ip, i64 jp, i64 i1, i64 j1, i64 i2, i64 j2, i64 i3, i64 j3,
                       double *o6, double *o23, double *o12, double *o13)
{
    double g123 = G[ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3)], g213 = G[ix3(ne, np, ip, jp, i1, j2, i2, j1, i3, j3)];
    double g312 = G[ix3(ne, np, ip, jp, i1, j3, i2, j1, i3, j2)], g132 = G[ix3(ne, np, ip, jp, i1, j1, i2, j3, i3, j2)];
    double g231 = G[ix3(ne, np, ip, jp, i1, j2, i2, j3, i3, j1)], g321 = G[ix3(ne, np, ip, jp, i1, j3, i2, j2, i3, j1)];
    *o6 = g123 - g213 + g312 - g132 + g231 - g321;
    *o23 = g123 - g132;
    *o12 = g123 - g213;
    *o13 = g123 - g321;
}

#define LOOP3                                                                                                      \
    for (i64 ip = 0; ip < np; ++ip)                                                                                \
        for (i64 jp = 0; jp < np; ++jp)                                                                            \
            for (i64 i1 = 0; i1 < ne; ++i1)                                                                        \
                for (i64 j1 = 0; j1 < ne; ++j1)                                                                    \
                    for (i64 i2 = 0; i2 < ne; ++i2)                                                                \
                        for (i64 j2 = 0; j2 < ne; ++j2)                                                            \
                            for (i64 i3 = 0; i3 < ne; ++i3)                                                        \
                                for (i64 j3 = 0; j3 < ne; ++j3)

/* FAE_OMG3 / thread_FAEOMG3 (xcuhf_omg3.f:2-282) */
void orc_fae_omg3(i64 nae, i64 nbe, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                  const double *GM3, double *FAE, double *E)
{
    i64 ne = nebf, np = npbf;
    int LAB = nbe >= 1 && nae >= 2, LBB = nbe >= 2, LAA = nae >= 3;
    double *xf = (double *)calloc(ne * ne, sizeof(double));
    double e = 0.0;
    LOOP3
    {
        double o6, o23, o12, o13;
        omg3_reads(ne, np, GM3, ip, jp, i1, j1, i2, j2, i3, j3, &o6, &o23, &o12, &o13);
        double xaa = LAA ? DAm(i2, j2) * DAm(i3, j3) * o6 : 0.0;
        double xbb = LBB ? DBm(i2, j2) * DBm(i3, j3) * o23 : 0.0;
        double xab = LAB ? DAm(i2, j2) * DBm(i3, j3) * o12 : 0.0;
        double xba = LAB ? DBm(i2, j2) * DAm(i3, j3) * o13 : 0.0;
        xf[j1 * ne + i1] += 3.0 * DPm(ip, jp) * (xaa + xbb + xab + xba);
        e += DPm(ip, jp) * DAm(i1, j1) * (xaa + xbb + xab + xba);
    }
    for (i64 k = 0; k < ne * ne; ++k) FAE[k] += xf[k];
    *E = e;
    free(xf);
}

/* UFP_OMG3 / thread_UFPOMG3 (xcuhf_omg3.f:284-560) */
void orc_ufp_omg3(i64 nae, i64 nbe, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                  const double *GM3, double *FP, double *E)
{
    i64 ne = nebf, np = npbf;
    int LAAA = nae >= 3, LBBB = nbe >= 3, LAAB = nae >= 2 && nbe >= 1, LBBA = nbe >= 2 && nae >= 1;
    double *xf = (double *)calloc(np * np, sizeof(double));
    double e = 0.0;
    LOOP3
    {
        double o6, o23, o12, o13;
        omg3_reads(ne, np, GM3, ip, jp, i1, j1, i2, j2, i3, j3, &o6, &o23, &o12, &o13);
        double caaa = LAAA ? DAm(i1, j1) * DAm(i2, j2) * DAm(i3, j3) : 0.0;
        double cbbb = LBBB ? DBm(i1, j1) * DBm(i2, j2) * DBm(i3, j3) : 0.0;
        double caab = 0, caba = 0, cbaa = 0, cbba = 0, cbab = 0, cabb = 0;
        if (LAAB) {
            caab = DAm(i1, j1) * DAm(i2, j2) * DBm(i3, j3);
            caba = DAm(i1, j1) * DBm(i2, j2) * DAm(i3, j3);
            cbaa = DBm(i1, j1) * DAm(i2, j2) * DAm(i3, j3);
        }
        if (LBBA) {
            cbba = DBm(i1, j1) * DBm(i2, j2) * DAm(i3, j3);
            cbab = DBm(i1, j1) * DAm(i2, j2) * DBm(i3, j3);
            cabb = DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3);
        }
        double v = (caaa + cbbb) * o6 + (caab + cbba) * o12 + (caba + cbab) * o13 + (cabb + cbaa) * o23;
        xf[jp * np + ip] += v;
        e += DPm(ip, jp) * v;
    }
    for (i64 k = 0; k < np * np; ++k) FP[k] += xf[k];
    *E = e;
    free(xf);
}

/* the 24 ket-permuted reads of calc_FAEOMG4 / calc_UFPOMG4
 * (xcuhf_omg4.f:234-510, :745-1078), in the reference's order:
 * OMG, OMG_23, OMG_12, OMG_12_23, OMG_13_23, OMG_13, OMG_24, OMG_12_24,
 * OMG_12_14, OMG_14, OMG_34, OMG_12_34, OMG_13_34, OMG_14_34, OMG_13_24,
 * OMG_14_23, OMG_23_34, OMG_24_34, OMG_12_23_34, OMG_12_23_24, OMG_12_13_34,
 * OMG_13_23_24, OMG_14_24_34, OMG_14_23_34 (ket of electron k = j[perm[k]]) */
static const int omg4_perm[24][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {1, 0, 2, 3}, {1, 2, 0, 3}, {2, 0, 1, 3}, {2, 1, 0, 3},
                                     {0, 3, 2, 1}, {1, 3, 2, 0}, {3, 0, 2, 1}, {3, 1, 2, 0}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                     {2, 1, 3, 0}, {3, 1, 0, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}, {0, 2, 3, 1}, {0, 3, 1, 2},
                                     {1, 2, 3, 0}, {1, 3, 0, 2}, {2, 0, 3, 1}, {2, 3, 1, 0}, {3, 0, 1, 2}, {3, 2, 0, 1}};
enum { O_, O23, O12, O12_23, O13_23, O13, O24, O12_24, O12_14, O14, O34, O12_34, O13_34, O14_34, O13_24, O14_23,
       O23_34, O24_34, O12_23_34, O12_23_24, O12_13_34, O13_23_24, O14_24_34, O14_23_34 };

static void omg4_reads(i64 ne, i64 np, const double *G, i64 ip, i64 jp, const i64 *ii, const i64 *jj, double *o)
{
    for (int p = 0; p < 24; ++p) {
        const int *q = omg4_perm[p];
        i64 ia = ix3(ne, np, ip, jp, ii[0], jj[q[0]], ii[1], jj[q[1]], ii[2], jj[q[2]]) * ne * ne + ii[3] * ne + jj[q[3]];
        o[p] = G[ia];
    }
}

static double omg4_anti(const double *o)
{
    return o[O_] - o[O23] - o[O12] - o[O13] - o[O24] - o[O14] - o[O34] + o[O12_23] + o[O13_23] + o[O12_24] +
           o[O12_14] + o[O12_34] + o[O13_34] + o[O14_34] + o[O13_24] + o[O14_23] + o[O23_34] + o[O24_34] -
           o[O12_23_34] - o[O12_23_24] - o[O12_13_34] - o[O13_23_24] - o[O14_24_34] - o[O14_23_34];
}
#define OM_AAB(o) (o[O_] - o[O23] - o[O12] + o[O12_23] + o[O13_23] - o[O13])
#define OM_ABA(o) (o[O_] - o[O24] - o[O12] + o[O12_24] + o[O12_14] - o[O14])
#define OM_BAA(o) (o[O_] - o[O34] - o[O13] + o[O13_34] + o[O14_34] - o[O14])
#define OM_ABB(o) (o[O_] - o[O34] - o[O12] + o[O12_34])
#define OM_BAB(o) (o[O_] - o[O24] - o[O13] + o[O13_24])
#define OM_BBA(o) (o[O_] - o[O23] - o[O14] + o[O14_23])
#define OM_BBB(o) (o[O_] - o[O34] - o[O23] + o[O23_34] + o[O24_34] - o[O24])

#define LOOP4                                                                                                      \
    LOOP3 for (i64 i4 = 0; i4 < ne; ++i4) for (i64 j4 = 0; j4 < ne; ++j4)

/* FAE_OMG4 / thread_FAEOMG4 / calc_FAEOMG4 (xcuhf_omg4.f:2-510) */
void orc_fae_omg4(i64 nae, i64 nbe, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                  const double *GM4, double *FAE, double *E)
{
    i64 ne = nebf, np = npbf;
    int LAAA = nae >= 4, LBBB = nbe >= 3, LAAB = nae >= 3 && nbe >= 1, LBBA = nbe >= 2 && nae >= 2;
    double *xf = (double *)calloc(ne * ne, sizeof(double));
    double e = 0.0;
    LOOP4
    {
        i64 ii[4] = {i1, i2, i3, i4}, jj[4] = {j1, j2, j3, j4};
        double o[24];
        omg4_reads(ne, np, GM4, ip, jp, ii, jj, o);
        double xaaa = LAAA ? DAm(i2, j2) * DAm(i3, j3) * DAm(i4, j4) * omg4_anti(o) : 0.0;
        double xaab = LAAB ? DAm(i2, j2) * DAm(i3, j3) * DBm(i4, j4) * OM_AAB(o) : 0.0;
        double xaba = LAAB ? DAm(i2, j2) * DBm(i3, j3) * DAm(i4, j4) * OM_ABA(o) : 0.0;
        double xbaa = LAAB ? DBm(i2, j2) * DAm(i3, j3) * DAm(i4, j4) * OM_BAA(o) : 0.0;
        double xabb = LBBA ? DAm(i2, j2) * DBm(i3, j3) * DBm(i4, j4) * OM_ABB(o) : 0.0;
        double xbab = LBBA ? DBm(i2, j2) * DAm(i3, j3) * DBm(i4, j4) * OM_BAB(o) : 0.0;
        double xbba = LBBA ? DBm(i2, j2) * DBm(i3, j3) * DAm(i4, j4) * OM_BBA(o) : 0.0;
        double xbbb = LBBB ? DBm(i2, j2) * DBm(i3, j3) * DBm(i4, j4) * OM_BBB(o) : 0.0;
        double ans = DPm(ip, jp) * (xaaa + xaab + xaba + xbaa + xabb + xbab + xbba + xbbb);
        xf[j1 * ne + i1] += 4.0 * ans;
        e += DAm(i1, j1) * ans;
    }
    for (i64 k = 0; k < ne * ne; ++k) FAE[k] += xf[k];
    *E = e;
    free(xf);
}

/* UFP_OMG4 / thread_UFPOMG4 / calc_UFPOMG4 (xcuhf_omg4.f:512-1078) */
void orc_ufp_omg4(i64 nae, i64 nbe, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                  const double *GM4, double *FP, double *E)
{
    i64 ne = nebf, np = npbf;
    int L4A = nae >= 4, L4B = nbe >= 4, L3A1B = nae >= 3 && nbe >= 1, L1A3B = nae >= 1 && nbe >= 3,
        L2A2B = nae >= 2 && nbe >= 2;
    double *xf = (double *)calloc(np * np, sizeof(double));
    double e = 0.0;
    LOOP4
    {
        i64 ii[4] = {i1, i2, i3, i4}, jj[4] = {j1, j2, j3, j4};
        double o[24];
        omg4_reads(ne, np, GM4, ip, jp, ii, jj, o);
        double A[4] = {DAm(i1, j1), DAm(i2, j2), DAm(i3, j3), DAm(i4, j4)};
        double B[4] = {DBm(i1, j1), DBm(i2, j2), DBm(i3, j3), DBm(i4, j4)};
        double anti = omg4_anti(o);
        double x = 0.0;
        x += L4A ? A[0] * A[1] * A[2] * A[3] * anti : 0.0;
        x += L3A1B ? A[0] * A[1] * A[2] * B[3] * OM_AAB(o) : 0.0;
        x += L3A1B ? A[0] * A[1] * B[2] * A[3] * OM_ABA(o) : 0.0;
        x += L3A1B ? A[0] * B[1] * A[2] * A[3] * OM_BAA(o) : 0.0;
        x += L2A2B ? A[0] * A[1] * B[2] * B[3] * OM_ABB(o) : 0.0;
        x += L2A2B ? A[0] * B[1] * A[2] * B[3] * OM_BAB(o) : 0.0;
        x += L2A2B ? A[0] * B[1] * B[2] * A[3] * OM_BBA(o) : 0.0;
        x += L1A3B ? A[0] * B[1] * B[2] * B[3] * OM_BBB(o) : 0.0;
        x += L4B ? B[0] * B[1] * B[2] * B[3] * anti : 0.0;
        x += L1A3B ? B[0] * B[1] * B[2] * A[3] * OM_AAB(o) : 0.0;
        x += L1A3B ? B[0] * B[1] * A[2] * B[3] * OM_ABA(o) : 0.0;
        x += L1A3B ? B[0] * A[1] * B[2] * B[3] * OM_BAA(o) : 0.0;
        x += L2A2B ? B[0] * B[1] * A[2] * A[3] * OM_ABB(o) : 0.0;
        x += L2A2B ? B[0] * A[1] * B[2] * A[3] * OM_BAB(o) : 0.0;
        x += L2A2B ? B[0] * A[1] * A[2] * B[3] * OM_BBA(o) : 0.0;
        x += L3A1B ? B[0] * A[1] * A[2] * A[3] * OM_BBB(o) : 0.0;
        xf[jp * np + ip] += x;
        e += DPm(ip, jp) * x;
    }
    for (i64 k = 0; k < np * np; ++k) FP[k] += xf[k];
    *E = e;
    free(xf);
}

/* OMG4_ICR / thread_omg4_IC (xcuhf_omg4.f:1080-1320): the unsymmetrised
 * calc_GAM_4_integral value per entry */
void orc_omg4_icr(i64 nebf, i64 npbf, const double *GAM2s, const double *GAMee, double *GM4)
{
    i64 ne = nebf, np = npbf;
#pragma omp parallel for collapse(2)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3)
                                    for (i64 i4 = 0; i4 < ne; ++i4)
                                        for (i64 j4 = 0; j4 < ne; ++j4) {
                                            i64 ia = ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3) * ne * ne + i4 * ne + j4;
                                            GM4[ia] = gam4_int(ne, np, GAM2s, GAMee, ip, jp, i1, j1, i2, j2, i3, j3, i4,
                                                               j4);
                                        }
}
#undef DAm
#undef DBm
#undef DPm

/* ------------------------------------------------------------------------
 * RXCHF-ne (source/rxchfne/): two-output e-e-p tuples, their contracted
 * tensors and the contractions of the RXCHF-ne / RXCUHF-ne Fock builds.
 * ------------------------------------------------------------------------ */
/* RXCHFne_GAM2_IC1 (rxchfne_gam2_IC1_OMP.f:2-243) with
 * RXCHFne_contract_omega2_conv (rxchfne_gam2_CONV_OMP.f:2-200):
 * GM2 = contracted ansE1 (not symmetrised), GM2ex = (x12 + x21)/2 of ansE2 */
void orc_rxchfne_gam2_ic1(BAS_ARGS, SYS_ARGS, double *GM2, double *GM2ex)
{
    BAS_INIT;
    SYS_INIT;
    i64 ne = nebf, np = npbf, ng2 = ne * ne * ne * ne * np * np;
    double *X = (double *)malloc(sizeof(double) * ng2);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            double om = 0.0, omx = 0.0;
                            for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                    for (i64 a2 = kpestr[i2] - 1; a2 < kpeend[i2]; ++a2)
                                        for (i64 c2 = kpestr[j2] - 1; c2 < kpeend[j2]; ++c2) {
                                            prim_t b[3] = {eprim(&B, a1), eprim(&B, a2), pprim(&B, ip)};
                                            prim_t k[3] = {eprim(&B, c1), eprim(&B, c2), pprim(&B, jp)};
                                            double e1, e2;
                                            rxne_prim(&S, b, k, &e1, &e2);
                                            double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] * agebfcc[c1] *
                                                        agebfcc[a2] * agebfcc[c2];
                                            om += e1 * cf;
                                            omx += e2 * cf;
                                        }
                            i64 ia = ix2(ne, np, ip, jp, i1, j1, i2, j2);
                            GM2[ia] = om;
                            X[ia] = omx;
                        }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            i64 a12 = ix2(ne, np, ip, jp, i1, j1, i2, j2), a21 = ix2(ne, np, ip, jp, i2, j2, i1, j1);
                            GM2ex[a12] = (X[a12] + X[a21]) / 2.0;
                        }
    free(X);
}

#define DAm(i, j) DAE[(j) * ne + (i)]
#define DBm(i, j) DBE[(j) * ne + (i)]
#define DPm(i, j) DP[(j) * np + (i)]
#define LOOP2                                                                                                      \
    for (i64 ip = 0; ip < np; ++ip)                                                                                \
        for (i64 jp = 0; jp < np; ++jp)                                                                            \
            for (i64 i1 = 0; i1 < ne; ++i1)                                                                        \
                for (i64 j1 = 0; j1 < ne; ++j1)                                                                    \
                    for (i64 i2 = 0; i2 < ne; ++i2)                                                                \
                        for (i64 j2 = 0; j2 < ne; ++j2)

/* RXCHFne_FAE_OMG2 / _FP_OMG2 / _FBE_OMG2 (rxchfne_omg2.f:2-600); which = 0,
 * 1, 2 selects the routine; F accumulates, E overwritten.  DAE: regular,
 * DBE: special electrons. */
void orc_rxchfne_omg2(i64 which, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                      const double *GM2, double *F, double *E)
{
    i64 ne = nebf, np = npbf;
    i64 nf = which == 1 ? np * np : ne * ne;
    double *xf = (double *)calloc(nf, sizeof(double));
    double e = 0.0;
    LOOP2
    {
        double val = GM2[ix2(ne, np, ip, jp, i1, j1, i2, j2)];
        if (which == 0) {
            xf[j2 * ne + i2] += DPm(ip, jp) * DBm(i1, j1) * val;
            e += DPm(ip, jp) * DAm(i2, j2) * DBm(i1, j1) * val;
        } else if (which == 1) {
            xf[jp * np + ip] += DAm(i2, j2) * DBm(i1, j1) * val;
            e += DPm(ip, jp) * DAm(i2, j2) * DBm(i1, j1) * val;
        } else {
            xf[j1 * ne + i1] += DAm(i2, j2) * DPm(ip, jp) * val;
            e += DPm(ip, jp) * DAm(i2, j2) * DBm(i1, j1) * val;
        }
    }
    for (i64 k = 0; k < nf; ++k) F[k] += xf[k];
    *E = e;
    free(xf);
}
#undef DAm

/* RXCUHFne_FE_OMG2 / _FP_OMG2 / _FBE_OMG2 (rxcuhfne_omg2.f:2-650):
 * val12 = GM2(i1 j1 i2 j2), val21 = GM2ex(i1 j2 i2 j1); DAa: regular alpha
 * density, DAt: regular total density.  which = 0 fills FAa and FAb (alpha
 * and beta regular Fock), 1 fills F = FP, 2 fills F = FBE. */
void orc_rxcuhfne_omg2(i64 which, i64 nebf, i64 npbf, const double *DAa, const double *DAt, const double *DBE,
                       const double *DP, const double *GM2, const double *GM2ex, double *F, double *F2, double *E)
{
    i64 ne = nebf, np = npbf;
    i64 nf = which == 1 ? np * np : ne * ne;
    double *xf = (double *)calloc(nf, sizeof(double)), *xf2 = (double *)calloc(ne * ne, sizeof(double));
    double e = 0.0;
#define DAam(i, j) DAa[(j) * ne + (i)]
#define DAtm(i, j) DAt[(j) * ne + (i)]
    LOOP2
    {
        double v12 = GM2[ix2(ne, np, ip, jp, i1, j1, i2, j2)], v21 = GM2ex[ix2(ne, np, ip, jp, i1, j2, i2, j1)];
        double t = DAtm(i2, j2) * v12 - DAam(i2, j2) * v21;
        if (which == 0) {
            xf[j2 * ne + i2] += DPm(ip, jp) * DBm(i1, j1) * (v12 - v21);
            xf2[j2 * ne + i2] += DPm(ip, jp) * DBm(i1, j1) * v12;
        } else if (which == 1) {
            xf[jp * np + ip] += DBm(i1, j1) * t;
        } else {
            xf[j1 * ne + i1] += DPm(ip, jp) * t;
        }
        e += DPm(ip, jp) * DBm(i1, j1) * t;
    }
#undef DAam
#undef DAtm
    for (i64 k = 0; k < nf; ++k) F[k] += xf[k];
    if (which == 0)
        for (i64 k = 0; k < ne * ne; ++k) F2[k] += xf2[k];
    *E = e;
    free(xf);
    free(xf2);
}
#undef DBm
#undef DPm

/* RXCHFne_FAE_GAMee (rxchfne_hfterms.f:2-60): HF Coulomb/exchange of the
 * regular electrons from GAM_ee (pack_4D order) */
void orc_rxchfne_fae_gamee(i64 nebf, const double *GAMee, const double *DAE, double *FAE, double *E)
{
    i64 ne = nebf;
    double *xf = (double *)calloc(ne * ne, sizeof(double));
    double e = 0.0;
#define EE(a, b, c, d) GAMee[(((a) * ne + (b)) * ne + (c)) * ne + (d)]
    for (i64 i2 = 0; i2 < ne; ++i2)
        for (i64 j2 = 0; j2 < ne; ++j2)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1) {
                    /* pack_4D(ne,ne,ne, je2,ie2,je1,ie1) and (je1,ie2,je2,ie1) */
                    double v1 = EE(i1, j1, i2, j2), v2 = EE(i1, j2, i2, j1);
                    double val = v1 - 0.5 * v2;
                    xf[j1 * ne + i1] += DAE[j2 * ne + i2] * val;
                    e += 0.5 * DAE[j1 * ne + i1] * DAE[j2 * ne + i2] * val;
                }
#undef EE
    for (i64 k = 0; k < ne * ne; ++k) FAE[k] += xf[k];
    *E = e;
    free(xf);
}

/* ------------------------------------------------------------------ */
/* RXCHFmult (rxchfmult/): regular basis B, special-electron basis BB   */
/* ------------------------------------------------------------------ */
#define MBAS_ARGS                                                                                                  \
    i64 nebf, i64 npebf, i64 nebfBE, i64 npebfBE, i64 npbf, const i64 *kpestr, const i64 *kpeend,                 \
        const double *agebfcc, const double *elcex, const i64 *elcam, const double *elcbfc, const i64 *kpestr_be,  \
        const i64 *kpeend_be, const double *agebfcc_be, const double *elcex_be, const i64 *elcam_be,              \
        const double *elcbfc_be, const double *agnbfcc, const double *nucex, const i64 *nucam, const double *nucbfc
#define MBAS_INIT                                                                                                  \
    basis_t B = {nebf, npebf, npbf, kpestr, kpeend, NULL, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc,     \
                 nucbfc};                                                                                          \
    basis_t BB = {nebfBE, npebfBE, npbf, kpestr_be, kpeend_be, NULL, agebfcc_be, agnbfcc, elcex_be, nucex,         \
                  elcam_be, nucam, elcbfc_be, nucbfc}

/* RXCHFmult_GAM_2PK / _3PK / _4PK (rxchfmult_basis.f:2-134), 0-based:
 * [ip][jp][ie1][je1] over the regular basis (na), the rest over nb */
static inline i64 mix2(i64 na, i64 nb, i64 np, i64 ip, i64 jp, i64 i1, i64 j1, i64 i2, i64 j2)
{
    return ((((ip * np + jp) * na + i1) * na + j1) * nb + i2) * nb + j2;
}
static inline i64 mix3(i64 na, i64 nb, i64 np, i64 ip, i64 jp, i64 i1, i64 j1, i64 i2, i64 j2, i64 i3, i64 j3)
{
    return (mix2(na, nb, np, ip, jp, i1, j1, i2, j2) * nb + i3) * nb + j3;
}

/* RXCHFmult_GAM2 (rxchfmult_gam2.f:2-265) with RXCHFmult_contract_omega2
 * (:413-626): XG2_1ICR, XG2_2ICR, XG2sICR over (ip,jp, iec1,jec1 regular,
 * iec2,jec2 special), then XCHF_GAM2, XCHF_GAM2s = (x12 + x21)/2 over special
 * indices (x12, x21 read through the two-basis packing, :173-228) and
 * INT_GAM2 = XG2_2ICR. */
void orc_rxchfmult_gam2(MBAS_ARGS, SYS_ARGS, double *INT_GAM2, double *XCHF_GAM2, double *XCHF_GAM2s)
{
    MBAS_INIT;
    SYS_INIT;
    i64 na = nebf, nb = nebfBE, np = npbf, ng = np * np * na * na * nb * nb;
    double *X1 = (double *)malloc(sizeof(double) * ng), *X2 = (double *)malloc(sizeof(double) * ng),
           *Xs = (double *)malloc(sizeof(double) * ng);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < na; ++i1)
                for (i64 j1 = 0; j1 < na; ++j1)
                    for (i64 i2 = 0; i2 < nb; ++i2)
                        for (i64 j2 = 0; j2 < nb; ++j2) {
                            double o1 = 0.0, o2 = 0.0, os = 0.0;
                            for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                    for (i64 a2 = kpestr_be[i2] - 1; a2 < kpeend_be[i2]; ++a2)
                                        for (i64 c2 = kpestr_be[j2] - 1; c2 < kpeend_be[j2]; ++c2) {
                                            prim_t b[3] = {eprim(&B, a1), eprim(&BB, a2), pprim(&B, ip)};
                                            prim_t k[3] = {eprim(&B, c1), eprim(&BB, c2), pprim(&B, jp)};
                                            double e1, e2, sv;
                                            gam2_prim_impl(&S, b, k, 1, &e1, &sv, &e2, NULL);
                                            double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] * agebfcc[c1] *
                                                        agebfcc_be[a2] * agebfcc_be[c2];
                                            o1 += e1 * cf;
                                            o2 += e2 * cf;
                                            os += sv * cf;
                                        }
                            i64 ia = mix2(na, nb, np, ip, jp, i1, j1, i2, j2);
                            X1[ia] = o1;
                            X2[ia] = o2;
                            Xs[ia] = os;
                        }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i2 = 0; i2 < nb; ++i2)
                for (i64 j2 = 0; j2 < nb; ++j2) {
                    for (i64 i1 = 0; i1 < nb; ++i1)
                        for (i64 j1 = 0; j1 < nb; ++j1) {
                            i64 a12 = mix2(na, nb, np, ip, jp, i1, j1, i2, j2);
                            i64 a21 = mix2(na, nb, np, ip, jp, i2, j2, i1, j1);
                            i64 ia = ix2(nb, np, ip, jp, i1, j1, i2, j2);
                            XCHF_GAM2[ia] = (X1[a12] + X1[a21]) / 2.0;
                            XCHF_GAM2s[ia] = (Xs[a12] + Xs[a21]) / 2.0;
                        }
                    for (i64 i1 = 0; i1 < na; ++i1)
                        for (i64 j1 = 0; j1 < na; ++j1) {
                            i64 ia = mix2(na, nb, np, ip, jp, i1, j1, i2, j2);
                            INT_GAM2[ia] = X2[ia];
                        }
                }
    free(X1);
    free(X2);
    free(Xs);
}

/* RXCHFmult_GAM3_INT (rxchfmult_gam3-int.f:2-254) with
 * RXCHFmult_contract_omega3_INT (:417-660) and the average over the exchange
 * of electrons 2 and 3 (:173-218). */
void orc_rxchfmult_gam3_int(MBAS_ARGS, SYS_ARGS, double *INT_GAM3)
{
    MBAS_INIT;
    SYS_INIT;
    i64 na = nebf, nb = nebfBE, np = npbf, ng = np * np * na * na * nb * nb * nb * nb;
    double *X = (double *)malloc(sizeof(double) * ng);
#pragma omp parallel for collapse(3) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < na; ++i1)
                for (i64 j1 = 0; j1 < na; ++j1)
                    for (i64 i2 = 0; i2 < nb; ++i2)
                        for (i64 j2 = 0; j2 < nb; ++j2)
                            for (i64 i3 = 0; i3 < nb; ++i3)
                                for (i64 j3 = 0; j3 < nb; ++j3) {
                                    double om = 0.0;
                                    for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                        for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                            for (i64 a2 = kpestr_be[i2] - 1; a2 < kpeend_be[i2]; ++a2)
                                                for (i64 c2 = kpestr_be[j2] - 1; c2 < kpeend_be[j2]; ++c2)
                                                    for (i64 a3 = kpestr_be[i3] - 1; a3 < kpeend_be[i3]; ++a3)
                                                        for (i64 c3 = kpestr_be[j3] - 1; c3 < kpeend_be[j3]; ++c3) {
                                                            prim_t b[4] = {eprim(&B, a1), eprim(&BB, a2),
                                                                           eprim(&BB, a3), pprim(&B, ip)};
                                                            prim_t k[4] = {eprim(&B, c1), eprim(&BB, c2),
                                                                           eprim(&BB, c3), pprim(&B, jp)};
                                                            double v;
                                                            gam3int_prim(&S, b, k, &v);
                                                            om += v * agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] *
                                                                  agebfcc[c1] * agebfcc_be[a2] * agebfcc_be[c2] *
                                                                  agebfcc_be[a3] * agebfcc_be[c3];
                                                        }
                                    X[mix3(na, nb, np, ip, jp, i1, j1, i2, j2, i3, j3)] = om;
                                }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < na; ++i1)
                for (i64 j1 = 0; j1 < na; ++j1)
                    for (i64 i2 = 0; i2 < nb; ++i2)
                        for (i64 j2 = 0; j2 < nb; ++j2)
                            for (i64 i3 = 0; i3 < nb; ++i3)
                                for (i64 j3 = 0; j3 < nb; ++j3) {
                                    i64 a123 = mix3(na, nb, np, ip, jp, i1, j1, i2, j2, i3, j3);
                                    i64 a132 = mix3(na, nb, np, ip, jp, i1, j1, i3, j3, i2, j2);
                                    double xx = (X[a123] + X[a132]) * 0.5;
                                    INT_GAM3[a123] = xx;
                                    INT_GAM3[a132] = xx;
                                }
    free(X);
}

/* RXCHFmult_calc_GAM_4_integral (rxchfmult_gam4-int.f:392-486), 0-based
 * indices: GAM_ee by pack_4D(nebf,nebf,nebf, je_x,ie_x,je1,ie1) over the
 * regular basis, GAM_2s by index_GAM_2PK(nebfBE, ...). */
static double mult_gam4_int(i64 na, i64 nb, i64 np, const double *G2s, const double *Gee, i64 ip, i64 jp, i64 ie1,
                            i64 je1, i64 ie2, i64 je2, i64 ie3, i64 je3, i64 ie4, i64 je4)
{
#define EE(a, b, c, d) Gee[(((a) * na + (b)) * na + (c)) * na + (d)]
    double ee14 = EE(ie1, je1, ie4, je4), ee13 = EE(ie1, je1, ie3, je3), ee12 = EE(ie1, je1, ie2, je2);
#undef EE
    underflow(&ee14);
    underflow(&ee13);
    underflow(&ee12);
    double gs23 = G2s[ix2(nb, np, ip, jp, ie2, je2, ie3, je3)];
    double gs24 = G2s[ix2(nb, np, ip, jp, ie2, je2, ie4, je4)];
    double gs34 = G2s[ix2(nb, np, ip, jp, ie3, je3, ie4, je4)];
    underflow(&gs23);
    underflow(&gs24);
    underflow(&gs34);
    double x1 = ee14 * gs23, x2 = ee13 * gs24, x3 = ee12 * gs34;
    return (x1 + x2 + x3) / 3.0;
}

/* RXCHFmult_GAM4_INT (rxchfmult_gam4-int.f:2-153) with RXCHFmult_symm_gam4
 * (:290-352) and RXCHFmult_i10 (:353-391). */
void orc_rxchfmult_gam4_int(i64 nebf, i64 nebfBE, i64 npbf, const double *GAM_2s, const double *GAM_ee, double *GAM_4)
{
    i64 na = nebf, nb = nebfBE, np = npbf;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 ii = 0; ii < na; ++ii)
                for (i64 i = 0; i < na; ++i)
                    for (i64 jj = 0; jj < nb; ++jj)
                        for (i64 j = 0; j < nb; ++j)
                            for (i64 kk = 0; kk < nb; ++kk)
                                for (i64 k = 0; k < nb; ++k)
                                    for (i64 ll = 0; ll < nb; ++ll)
                                        for (i64 l = 0; l < nb; ++l) {
#define I10(a, b, c, d) mult_gam4_int(na, nb, np, GAM_2s, GAM_ee, ip, jp, ii, a, jj, b, kk, c, ll, d)
                                            double x1 = I10(i, j, k, l), x2 = I10(i, j, l, k), x3 = I10(i, k, j, l);
                                            double x4 = I10(i, k, l, j), x5 = I10(i, l, j, k), x6 = I10(i, l, k, j);
#undef I10
                                            double ans = x1 - x2 / 2.0 - x3 / 2.0 + x4 / 4.0 + x5 / 4.0 - x6 / 2.0;
                                            i64 ia = ((mix2(na, nb, np, ip, jp, ii, i, jj, j) * nb + kk) * nb + k) *
                                                         nb * nb + ll * nb + l;
                                            GAM_4[ia] = ans;
                                        }
}

/* RXCHFmult_thread_{FAE,FP,FBE}_OMG{2,3,4} (rxchfmult_omg2.f:91-174,
 * :264-347, :437-520; rxchfmult_omg3.f:97-200, :296-399, :495-599;
 * rxchfmult_omg4.f:109-206, :314-411, :519-620) and add2fock.  order 2, 3, 4;
 * which 0 FAE (nebf^2), 1 FP (npbf^2), 2 FBE (nebfBE^2). */
void orc_rxchfmult_omg(i64 order, i64 which, i64 nebf, i64 nebfBE, i64 npbf, const double *DAE, const double *DBE,
                       const double *DP, const double *G, double *F, double *E)
{
    i64 na = nebf, nb = nebfBE, np = npbf;
    i64 nf = which == 0 ? na * na : which == 1 ? np * np : nb * nb;
    double *XF = (double *)calloc(nf, sizeof(double));
    double e = 0.0;
#define DAm(i, j) DAE[(j) * na + (i)]
#define DBm(i, j) DBE[(j) * nb + (i)]
#define DPm(i, j) DP[(j) * np + (i)]
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < na; ++i1)
                for (i64 j1 = 0; j1 < na; ++j1)
                    for (i64 i2 = 0; i2 < nb; ++i2)
                        for (i64 j2 = 0; j2 < nb; ++j2) {
                            if (order == 2) {
                                double val = G[mix2(na, nb, np, ip, jp, i1, j1, i2, j2)];
                                if (which == 0) XF[j1 * na + i1] += DPm(ip, jp) * DBm(i2, j2) * val;
                                if (which == 1) XF[jp * np + ip] += DAm(i1, j1) * DBm(i2, j2) * val;
                                if (which == 2) XF[j2 * nb + i2] += DAm(i1, j1) * DPm(ip, jp) * val;
                                e += DPm(ip, jp) * DAm(i1, j1) * DBm(i2, j2) * val;
                                continue;
                            }
                            for (i64 i3 = 0; i3 < nb; ++i3)
                                for (i64 j3 = 0; j3 < nb; ++j3) {
                                    if (order == 3) {
                                        double val = G[mix3(na, nb, np, ip, jp, i1, j1, i2, j2, i3, j3)] -
                                                     0.5 * G[mix3(na, nb, np, ip, jp, i1, j1, i2, j3, i3, j2)];
                                        if (which == 0)
                                            XF[j1 * na + i1] += DBm(i2, j2) * DBm(i3, j3) * DPm(ip, jp) * val;
                                        if (which == 1)
                                            XF[jp * np + ip] += DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3) * val;
                                        if (which == 2)
                                            XF[j2 * nb + i2] += 2.0 * DAm(i1, j1) * DBm(i3, j3) * DPm(ip, jp) * val;
                                        e += DPm(ip, jp) * DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3) * val;
                                        continue;
                                    }
                                    for (i64 i4 = 0; i4 < nb; ++i4)
                                        for (i64 j4 = 0; j4 < nb; ++j4) {
                                            i64 ia = (mix3(na, nb, np, ip, jp, i1, j1, i2, j2, i3, j3) * nb + i4) * nb +
                                                     j4;
                                            double val = G[ia];
                                            if (which == 0)
                                                XF[j1 * na + i1] += DPm(ip, jp) * DBm(i2, j2) * DBm(i3, j3) *
                                                                    DBm(i4, j4) * val;
                                            if (which == 1)
                                                XF[jp * np + ip] += DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3) *
                                                                    DBm(i4, j4) * val;
                                            if (which == 2)
                                                XF[j2 * nb + i2] += 3.0 * DAm(i1, j1) * DBm(i3, j3) * DBm(i4, j4) *
                                                                    DPm(ip, jp) * val;
                                            e += DPm(ip, jp) * DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3) *
                                                 DBm(i4, j4) * val;
                                        }
                                }
                        }
#undef DAm
#undef DBm
#undef DPm
    for (i64 t = 0; t < nf; ++t) F[t] += XF[t];
    *E = e;
    free(XF);
}

/* ---- RXCHFmult with exchange (LADDEXCH): one basis ---------------- */
/* RXCHFmult_GAM2_IC1ex (rxchfmult_gam2ex.f:2-229) with
 * RXCHFmult_contract_omega2_convex (:354-558): GM2_1ICR, GM2_3ICR, GM2sICR =
 * (x12 + x21)/2 of ansE1, ansE3, ansS; GM2_2ICR = ansE2 (:160-186). */
void orc_rxchfmult_gam2_ic1ex(BAS_ARGS, SYS_ARGS, double *GM2_1, double *GM2_2, double *GM2_3, double *GM2s)
{
    BAS_INIT;
    SYS_INIT;
    i64 ne = nebf, np = npbf, ng = np * np * ne * ne * ne * ne;
    double *X1 = (double *)malloc(sizeof(double) * ng), *X2 = (double *)malloc(sizeof(double) * ng),
           *X3 = (double *)malloc(sizeof(double) * ng), *Xs = (double *)malloc(sizeof(double) * ng);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            double o1 = 0, o2 = 0, o3 = 0, os = 0;
                            for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                    for (i64 a2 = kpestr[i2] - 1; a2 < kpeend[i2]; ++a2)
                                        for (i64 c2 = kpestr[j2] - 1; c2 < kpeend[j2]; ++c2) {
                                            prim_t b[3] = {eprim(&B, a1), eprim(&B, a2), pprim(&B, ip)};
                                            prim_t k[3] = {eprim(&B, c1), eprim(&B, c2), pprim(&B, jp)};
                                            double e1, e2, e3, sv;
                                            gam2_prim_impl(&S, b, k, 1, &e1, &sv, &e2, &e3);
                                            double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] * agebfcc[c1] *
                                                        agebfcc[a2] * agebfcc[c2];
                                            o1 += e1 * cf;
                                            o2 += e2 * cf;
                                            o3 += e3 * cf;
                                            os += sv * cf;
                                        }
                            i64 ia = ix2(ne, np, ip, jp, i1, j1, i2, j2);
                            X1[ia] = o1;
                            X2[ia] = o2;
                            X3[ia] = o3;
                            Xs[ia] = os;
                        }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            i64 a12 = ix2(ne, np, ip, jp, i1, j1, i2, j2), a21 = ix2(ne, np, ip, jp, i2, j2, i1, j1);
                            GM2_1[a12] = (X1[a12] + X1[a21]) / 2.0;
                            GM2_2[a12] = X2[a12];
                            GM2_3[a12] = (X3[a12] + X3[a21]) / 2.0;
                            GM2s[a12] = (Xs[a12] + Xs[a21]) / 2.0;
                        }
    free(X1);
    free(X2);
    free(X3);
    free(Xs);
}

/* RXCHFmult_GAM3_IC1ex (rxchfmult_gam3ex.f:2-316) with
 * RXCHFmult_contract_omega3_convex (:459-711) and the pair averages (:248-269) */
void orc_rxchfmult_gam3_ic1ex(BAS_ARGS, SYS_ARGS, double *GM3_2, double *GM3_3, double *GM3_4)
{
    BAS_INIT;
    SYS_INIT;
    i64 ne = nebf, np = npbf, ng = np * np * ne * ne * ne * ne * ne * ne;
    double *X2 = (double *)malloc(sizeof(double) * ng), *X3 = (double *)malloc(sizeof(double) * ng),
           *X4 = (double *)malloc(sizeof(double) * ng);
#pragma omp parallel for collapse(3) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3) {
                                    double o2 = 0, o3 = 0, o4 = 0;
                                    for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                        for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                            for (i64 a2 = kpestr[i2] - 1; a2 < kpeend[i2]; ++a2)
                                                for (i64 c2 = kpestr[j2] - 1; c2 < kpeend[j2]; ++c2)
                                                    for (i64 a3 = kpestr[i3] - 1; a3 < kpeend[i3]; ++a3)
                                                        for (i64 c3 = kpestr[j3] - 1; c3 < kpeend[j3]; ++c3) {
                                                            prim_t b[4] = {eprim(&B, a1), eprim(&B, a2), eprim(&B, a3),
                                                                           pprim(&B, ip)};
                                                            prim_t k[4] = {eprim(&B, c1), eprim(&B, c2), eprim(&B, c3),
                                                                           pprim(&B, jp)};
                                                            double v2, v3, v4;
                                                            gam3int_prim_ex(&S, b, k, &v2, &v3, &v4);
                                                            double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] *
                                                                        agebfcc[c1] * agebfcc[a2] * agebfcc[c2] *
                                                                        agebfcc[a3] * agebfcc[c3];
                                                            o2 += v2 * cf;
                                                            o3 += v3 * cf;
                                                            o4 += v4 * cf;
                                                        }
                                    i64 ia = ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3);
                                    X2[ia] = o2;
                                    X3[ia] = o3;
                                    X4[ia] = o4;
                                }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3) {
                                    i64 a123 = ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3);
                                    i64 a132 = ix3(ne, np, ip, jp, i1, j1, i3, j3, i2, j2);
                                    i64 a213 = ix3(ne, np, ip, jp, i2, j2, i1, j1, i3, j3);
                                    i64 a321 = ix3(ne, np, ip, jp, i3, j3, i2, j2, i1, j1);
                                    double x = (X2[a123] + X2[a132]) * 0.5;
                                    GM3_2[a123] = x;
                                    GM3_2[a132] = x;
                                    x = (X3[a123] + X3[a213]) * 0.5;
                                    GM3_3[a123] = x;
                                    GM3_3[a213] = x;
                                    x = (X4[a123] + X4[a321]) * 0.5;
                                    GM3_4[a123] = x;
                                    GM3_4[a321] = x;
                                }
    free(X2);
    free(X3);
    free(X4);
}

/* RXCHFmult_thread_{FAE,FP,FBE}_OMG2ex (rxchfmult_omg2ex.f:94-180, :273-359,
 * :452-538) and add2fock; which 0 FAE, 1 FP, 2 FBE. */
void orc_rxchfmult_omg2ex(i64 which, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                          const double *G, const double *Gx, double *F, double *E)
{
    i64 ne = nebf, np = npbf, nf = which == 1 ? np * np : ne * ne;
    double *XF = (double *)calloc(nf, sizeof(double));
    double e = 0.0;
#define DAm(i, j) DAE[(j) * ne + (i)]
#define DBm(i, j) DBE[(j) * ne + (i)]
#define DPm(i, j) DP[(j) * np + (i)]
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            double val = G[ix2(ne, np, ip, jp, i1, j1, i2, j2)] -
                                         0.5 * Gx[ix2(ne, np, ip, jp, i1, j2, i2, j1)];
                            if (which == 0) XF[j1 * ne + i1] += DPm(ip, jp) * DBm(i2, j2) * val;
                            if (which == 1) XF[jp * np + ip] += DAm(i1, j1) * DBm(i2, j2) * val;
                            if (which == 2) XF[j2 * ne + i2] += DAm(i1, j1) * DPm(ip, jp) * val;
                            e += DPm(ip, jp) * DAm(i1, j1) * DBm(i2, j2) * val;
                        }
    for (i64 t = 0; t < nf; ++t) F[t] += XF[t];
    *E = e;
    free(XF);
}

/* RXCHFmult_thread_{FAE,FP,FBE}_OMG3ex (rxchfmult_omg3ex.f:101-216, :316-431,
 * :531-663) and add2fock. */
void orc_rxchfmult_omg3ex(i64 which, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                          const double *G, const double *G1, const double *G2, double *F, double *E)
{
    i64 ne = nebf, np = npbf, nf = which == 1 ? np * np : ne * ne;
    double *XF = (double *)calloc(nf, sizeof(double));
    double e = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3) {
#define X(A, a, b, c, d, f, g) A[ix3(ne, np, ip, jp, a, b, c, d, f, g)]
                                    double val = X(G, i1, j1, i2, j2, i3, j3) - 0.5 * X(G, i1, j1, i2, j3, i3, j2);
                                    double d = DPm(ip, jp);
                                    if (which != 2) {
                                        double val2 = X(G1, i1, j2, i2, j1, i3, j3) - 0.5 * X(G1, i1, j3, i2, j1, i3, j2);
                                        double val3 = X(G2, i1, j3, i2, j2, i3, j1) - 0.5 * X(G2, i1, j2, i2, j3, i3, j1);
                                        double v = val - 0.5 * (val2 + val3);
                                        if (which == 0) XF[j1 * ne + i1] += v * DBm(i2, j2) * DBm(i3, j3) * d;
                                        else XF[jp * np + ip] += v * DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3);
                                        e += d * DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3) * v;
                                    } else {
                                        double val2 = X(G1, i1, j2, i2, j1, i3, j3) - 0.5 * X(G1, i1, j3, i2, j1, i3, j2) +
                                                      X(G2, i1, j3, i2, j2, i3, j1) - 0.5 * X(G2, i1, j2, i2, j3, i3, j1);
                                        double val3 = X(G1, i1, j3, i3, j1, i2, j2) - 0.5 * X(G1, i1, j2, i3, j1, i2, j3) +
                                                      X(G2, i1, j2, i3, j3, i2, j1) - 0.5 * X(G2, i1, j3, i3, j2, i2, j1);
                                        XF[j2 * ne + i2] += (2.0 * val - 0.5 * (val2 + val3)) * DAm(i1, j1) *
                                                            DBm(i3, j3) * d;
                                        e += d * DAm(i1, j1) * DBm(i2, j2) * DBm(i3, j3) * (val - 0.5 * val2);
                                    }
#undef X
                                }
#undef DAm
#undef DBm
#undef DPm
    for (i64 t = 0; t < nf; ++t) F[t] += XF[t];
    *E = e;
    free(XF);
}

/* ------------------------------------------------------------------------
 * NEO-XCHF RHF SCF side (source/xchf1_independent/): the one-particle and
 * e-p Coulomb integrals, the Fock build xchf1_fock with XCfock_correction,
 * ROOTHAN and the SCF loop of xcscf.
 * ------------------------------------------------------------------------ */
static const double PI_AC = 3.141592653589793238462643383279502884; /* dacos(-1) of ke.f, vec.f, ovlap.f, vee.f */

static double g1ac(int op, const prim_t *b, const prim_t *k, const double *C, double xmass)
{
    gint_t g;
    gi_init(&g, 1);
    g.bra[0] = *b;
    g.ket[0] = *k;
    g.op = op;
    g.mass = xmass;
    g.pi = PI_AC;
    if (C) memcpy(g.C, C, sizeof(g.C));
    return gint_eval(&g);
}

/* class_nuc_rep (class_nuc_rep.f:1-45) */
double orc_enuc(i64 nat, const double *cat, const double *zan)
{
    double e = 0.0;
    for (i64 i = 1; i < nat; ++i)
        for (i64 j = 0; j < i; ++j) {
            double r2 = 0.0;
            for (int k = 0; k < 3; ++k) r2 += (cat[3 * i + k] - cat[3 * j + k]) * (cat[3 * i + k] - cat[3 * j + k]);
            if (r2 != 0.0) e += zan[i] * zan[j] / sqrt(r2);
        }
    return e;
}

/* calc_GAM_epcore (gam_epcore.f:1-178) with xcalc_GAM_epcore (gfke + sum of
 * coul_sign Z gfvec), and elec_ovlap / nuc_ovlap (pure_ovlap.f:2-204):
 * He, Se in pack_2D(nebf) order [iec][jec], Hp, Sp in [ip][jp] */
void orc_one_body(BAS_ARGS, SYS_ARGS, double *He, double *Se, double *Hp, double *Sp)
{
    BAS_INIT;
    i64 ne = nebf, np = npbf;
    memset(He, 0, sizeof(double) * ne * ne);
    memset(Se, 0, sizeof(double) * ne * ne);
    for (i64 ie = 0; ie < npebf; ++ie)
        for (i64 je = 0; je < npebf; ++je) {
            prim_t b = eprim(&B, ie), k = eprim(&B, je);
            double ke = g1ac(OP_KIN, &b, &k, NULL, 1.0), v = 0.0;
            for (i64 c = 0; c < nat; ++c) v += -1.0 * zan[c] * g1ac(OP_NUC, &b, &k, cat + 3 * c, 1.0);
            i64 ia = (ampeb2c[ie] - 1) * ne + ampeb2c[je] - 1;
            He[ia] += (ke + v) * agebfcc[ie] * agebfcc[je];
            Se[ia] += g1ac(OP_OVL, &b, &k, NULL, 1.0) * agebfcc[ie] * agebfcc[je];
        }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp) {
            prim_t b = pprim(&B, ip), k = pprim(&B, jp);
            double ke = g1ac(OP_KIN, &b, &k, NULL, pmass), v = 0.0;
            for (i64 c = 0; c < nat; ++c) v += 1.0 * zan[c] * g1ac(OP_NUC, &b, &k, cat + 3 * c, 1.0);
            Hp[ip * np + jp] = (ke + v) * agnbfcc[ip] * agnbfcc[jp];
            Sp[ip * np + jp] = g1ac(OP_OVL, &b, &k, NULL, 1.0) * agnbfcc[ip] * agnbfcc[jp];
        }
    (void)ngtg; (void)bcoef; (void)gamma; (void)kpestr; (void)kpeend;
}

/* calc_GAM_ep (gam_ep.f:1-141) with xcalc_GAM_ep = -gfvee:
 * GAM_ep in pack_4D(nebf,nebf,npbf) order [ip][jp][iec][jec] */
void orc_gam_ep(BAS_ARGS, double *Gep)
{
    BAS_INIT;
    i64 ne = nebf, np = npbf;
    memset(Gep, 0, sizeof(double) * ne * ne * np * np);
#pragma omp parallel for collapse(2)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 ie = 0; ie < npebf; ++ie)
                for (i64 je = 0; je < npebf; ++je) {
                    prim_t b1 = eprim(&B, ie), k1 = eprim(&B, je), b2 = pprim(&B, ip), k2 = pprim(&B, jp);
                    gint_t g;
                    gi_init(&g, 2);
                    g.bra[0] = b1;
                    g.ket[0] = k1;
                    g.bra[1] = b2;
                    g.ket[1] = k2;
                    g.op = OP_COUL;
                    g.a = 0;
                    g.b = 1;
                    g.pi = PI_AC;
                    double ans = -gint_eval(&g);
                    i64 ia = ((ip * np + jp) * ne + ampeb2c[ie] - 1) * ne + ampeb2c[je] - 1;
                    Gep[ia] += ans * agebfcc[ie] * agebfcc[je] * agnbfcc[ip] * agnbfcc[jp];
                }
    (void)kpestr; (void)kpeend;
}

/* xchf1_fock (xchf1_fock.f:2-360) for the in-core path of xcneo_HYBRID
 * (LG2IC1, LG3IC1, LG4IC): E_from_GAM_ecore/pcore/ep/ee (:600-911),
 * S_from_GAM_1s and E_from_GAM_1 (:990-1130, reading the arrays instead of
 * GAM_1.ufm/GAM_1s.ufm), EG2IC1, EG3IC1, EG4ICR, the energy assembly and
 * XCfock_correction (:728-784).  G4 may be NULL: the 5-particle values are
 * then formed per entry from G2s and Gee (GAM4_DSCF).  comps as the C ABI's
 * neorx_xchf_fock_dev returns them. */
void orc_xchf_fock(i64 nebf, i64 npbf, i64 nelec, double enuc, const double *DE, const double *DP, const double *He,
                   const double *Hp, const double *Gep, const double *Gee, const double *G1, const double *G1s,
                   const double *G2, const double *G2s, const double *G3, const double *G4, double *FE, double *FP,
                   double *comps)
{
    i64 ne = nebf, np = npbf;
    double *se1 = calloc(ne * ne, 8), *se2 = calloc(ne * ne, 8), *sp1 = calloc(np * np, 8), *sp2 = calloc(np * np, 8);
    double *xfe = calloc(ne * ne, 8), *xfp = calloc(np * np, 8);
#define DEm(i, j) DE[(j) * ne + (i)]
#define DPm(i, j) DP[(j) * np + (i)]
    double e_ecore = 0, e_pcore = 0, e_ep = 0, e_ee = 0, e1 = 0, s1 = 0, e2 = 0, s2 = 0, e3 = 0, e4 = 0;
    for (i64 i = 0; i < ne; ++i)
        for (i64 j = 0; j < ne; ++j) {
            FE[j * ne + i] = He[i * ne + j];
            e_ecore += DEm(i, j) * He[i * ne + j];
        }
    for (i64 i = 0; i < np; ++i)
        for (i64 j = 0; j < np; ++j) {
            FP[j * np + i] = Hp[i * np + j];
            e_pcore += DPm(i, j) * Hp[i * np + j];
        }
    /* E_from_GAM_ep */
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i = 0; i < ne; ++i)
                for (i64 j = 0; j < ne; ++j) {
                    double v = Gep[((ip * np + jp) * ne + i) * ne + j];
                    xfe[j * ne + i] += DPm(ip, jp) * v;
                    xfp[jp * np + ip] += DEm(i, j) * v;
                    e_ep += DEm(i, j) * DPm(ip, jp) * v;
                }
    for (i64 t = 0; t < ne * ne; ++t) FE[t] += xfe[t];
    for (i64 t = 0; t < np * np; ++t) FP[t] += xfp[t];
    /* E_from_GAM_ee */
    if (nelec > 1) {
        memset(xfe, 0, 8 * ne * ne);
        for (i64 i2 = 0; i2 < ne; ++i2)
            for (i64 j2 = 0; j2 < ne; ++j2)
                for (i64 i1 = 0; i1 < ne; ++i1)
                    for (i64 j1 = 0; j1 < ne; ++j1) {
                        double v1 = Gee[((i1 * ne + j1) * ne + i2) * ne + j2];
                        double v2 = Gee[((i1 * ne + j2) * ne + i2) * ne + j1];
                        double val = (v1 - 0.5 * v2) * 0.5;
                        xfe[j1 * ne + i1] += 2.0 * DEm(i2, j2) * val;
                        e_ee += DEm(i1, j1) * DEm(i2, j2) * val;
                    }
        for (i64 t = 0; t < ne * ne; ++t) FE[t] += xfe[t];
    }
    /* S_from_GAM_1s, E_from_GAM_1 */
    memset(xfe, 0, 8 * ne * ne);
    memset(xfp, 0, 8 * np * np);
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i = 0; i < ne; ++i)
                for (i64 j = 0; j < ne; ++j) {
                    i64 ia = ((ip * np + jp) * ne + i) * ne + j;
                    se1[j * ne + i] += DPm(ip, jp) * G1s[ia];
                    sp1[jp * np + ip] += DEm(i, j) * G1s[ia];
                    s1 += DEm(i, j) * DPm(ip, jp) * G1s[ia];
                    xfe[j * ne + i] += DPm(ip, jp) * G1[ia];
                    xfp[jp * np + ip] += DEm(i, j) * G1[ia];
                    e1 += DEm(i, j) * DPm(ip, jp) * G1[ia];
                }
    for (i64 t = 0; t < ne * ne; ++t) FE[t] += xfe[t];
    for (i64 t = 0; t < np * np; ++t) FP[t] += xfp[t];
#undef DEm
#undef DPm
    if (nelec > 1) {
        eg2_impl(ne, np, DE, DP, G2, G2s, FE, se2, FP, sp2, &e2, &s2, 1);
        if (nelec > 2) {
            orc_eg3ic1(ne, np, DE, DP, G3, FE, FP, &e3);
            if (G4)
                orc_eg4icr(ne, np, DE, DP, G4, FE, FP, &e4);
            else
                orc_gam4_dscf(ne, np, G2s, Gee, DE, DP, FE, FP, &e4);
        }
    }
    double S = 1.0 + s1 + s2;
    double psi = (e_ecore + e_pcore + e_ep + e_ee + e1 + e2 + e3 + e4) / S;
    double H = psi / S;
    /* XCfock_correction */
    for (i64 t = 0; t < ne * ne; ++t) FE[t] = (1.0 / S) * FE[t] - H * (se1[t] + 2.0 * se2[t]);
    for (i64 t = 0; t < np * np; ++t) FP[t] = (1.0 / S) * FP[t] - H * (sp1[t] + sp2[t]);
    double c[14] = {psi + enuc, enuc, e_ecore, e_pcore, e_ep, e_ee, e1, e2, e3, e4, S, s1, s2, H};
    memcpy(comps, c, sizeof(c));
    free(se1); free(se2); free(sp1); free(sp2); free(xfe); free(xfp);
}

/* cyclic Jacobi eigen-decomposition of the symmetric n x n column-major a
 * (destroyed: eigenvalues on the diagonal), eigenvectors in the columns of v */
static void jacobi_cpu(i64 n, double *a, double *v)
{
    for (i64 t = 0; t < n * n; ++t) v[t] = (t % n == t / n) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (i64 j = 0; j < n; ++j)
            for (i64 i = 0; i < n; ++i) {
                if (i == j)
                    dia += a[j * n + i] * a[j * n + i];
                else
                    off += a[j * n + i] * a[j * n + i];
            }
        if (off <= 1e-30 * dia || off == 0.0) break;
        for (i64 p = 0; p < n - 1; ++p)
            for (i64 q = p + 1; q < n; ++q) {
                double apq = a[q * n + p];
                if (apq == 0.0) continue;
                double th = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (i64 k = 0; k < n; ++k) { /* rows p, q */
                    double x = a[k * n + p], y = a[k * n + q];
                    a[k * n + p] = c * x - s * y;
                    a[k * n + q] = s * x + c * y;
                }
                for (i64 k = 0; k < n; ++k) { /* columns p, q */
                    double x = a[p * n + k], y = a[q * n + k];
                    a[p * n + k] = c * x - s * y;
                    a[q * n + k] = s * x + c * y;
                    x = v[p * n + k];
                    y = v[q * n + k];
                    v[p * n + k] = c * x - s * y;
                    v[q * n + k] = s * x + c * y;
                }
            }
    }
}

/* ROOTHAN (xcscf.f:766-952): X = S^-1/2 (symmetric orthogonalisation),
 * F' = X^T F X, its eigenvectors C' with eigenvalues ascending, C = X C',
 * D = prefac sum_{k = kstart..klast} C(:,k) C(:,k)^T (1-based k). */
void orc_roothan(i64 n, const double *S, const double *F, i64 kstart, i64 klast, double prefac, double *C,
                 double *ev, double *D)
{
    double *a = malloc(8 * n * n), *v = malloc(8 * n * n), *X = malloc(8 * n * n), *t = malloc(8 * n * n);
    i64 *ord = malloc(sizeof(i64) * n);
    memcpy(a, S, 8 * n * n);
    jacobi_cpu(n, a, v);
    for (i64 i = 0; i < n; ++i)
        for (i64 j = 0; j < n; ++j) {
            double x = 0.0;
            for (i64 k = 0; k < n; ++k) x += v[k * n + i] * v[k * n + j] / sqrt(a[k * n + k]);
            X[j * n + i] = x;
        }
    for (i64 i = 0; i < n; ++i)
        for (i64 j = 0; j < n; ++j) {
            double x = 0.0;
            for (i64 k = 0; k < n; ++k) x += F[k * n + i] * X[j * n + k];
            t[j * n + i] = x;
        }
    for (i64 i = 0; i < n; ++i)
        for (i64 j = 0; j < n; ++j) {
            double x = 0.0;
            for (i64 k = 0; k < n; ++k) x += X[i * n + k] * t[j * n + k];
            a[j * n + i] = x;
        }
    jacobi_cpu(n, a, v);
    for (i64 i = 0; i < n; ++i) ord[i] = i;
    for (i64 i = 1; i < n; ++i) /* insertion sort by eigenvalue */
        for (i64 j = i; j > 0 && a[ord[j] * n + ord[j]] < a[ord[j - 1] * n + ord[j - 1]]; --j) {
            i64 x = ord[j];
            ord[j] = ord[j - 1];
            ord[j - 1] = x;
        }
    for (i64 k = 0; k < n; ++k) {
        ev[k] = a[ord[k] * n + ord[k]];
        for (i64 i = 0; i < n; ++i) {
            double x = 0.0;
            for (i64 l = 0; l < n; ++l) x += X[l * n + i] * v[ord[k] * n + l];
            C[k * n + i] = x;
        }
    }
    for (i64 i = 0; i < n; ++i)
        for (i64 j = 0; j < n; ++j) {
            double x = 0.0;
            for (i64 k = kstart - 1; k < klast; ++k) x += C[k * n + i] * C[k * n + j];
            D[j * n + i] = prefac * x;
        }
    free(a); free(v); free(X); free(t); free(ord);
}

/* The SCF loop of xcscf (xcscf.f:2-560) with ROOTHAN for both particles
 * (no SOSCF step): per iteration the nuclear Fock build with (DE, DP),
 * ROOTHAN for DP, the electronic build with (DE, new DP), ROOTHAN for DE,
 * then DENDIF on both; converged when max|dDE| < tole (and max|dDP| < tolp
 * if npbf > 1).  ce_guess (column-major orbitals, NULL: the core guess of
 * guess_elec) gives DE; guess_prot gives DP.  trace: per iteration
 * [E_total, dE, max|dDE|, max|dDP|].  Returns the number of iterations
 * (negative if not converged). */
i64 orc_xcscf(i64 nebf, i64 npbf, i64 nelec, i64 nucst, double enuc, const double *Se, const double *Sp,
              const double *He, const double *Hp, const double *Gep, const double *Gee, const double *G1,
              const double *G1s, const double *G2, const double *G2s, const double *G3, const double *G4,
              const double *ce_guess, i64 maxit, double tole, double tolp, double *comps_init, double *comps,
              double *trace)
{
    i64 ne = nebf, np = npbf, nocc = nelec > 1 ? nelec / 2 : 1;
    double pre = nelec > 1 ? 2.0 : 1.0;
    double *DE = calloc(ne * ne, 8), *DP = calloc(np * np, 8), *DE0 = calloc(ne * ne, 8), *DP0 = calloc(np * np, 8);
    double *FE = calloc(ne * ne, 8), *FP = calloc(np * np, 8), *FEd = calloc(ne * ne, 8), *FPd = calloc(np * np, 8);
    double *CE = calloc(ne * ne, 8), *CP = calloc(np * np, 8), *evE = calloc(ne, 8), *evP = calloc(np, 8);
    double *h = calloc(ne * ne + np * np, 8);
    if (ce_guess) {
        for (i64 i = 0; i < ne; ++i)
            for (i64 j = 0; j < ne; ++j) {
                double x = 0.0;
                for (i64 k = 0; k < nocc; ++k) x += pre * ce_guess[k * ne + i] * ce_guess[k * ne + j];
                DE[j * ne + i] = x;
            }
    } else {
        for (i64 i = 0; i < ne; ++i)
            for (i64 j = 0; j < ne; ++j) h[j * ne + i] = He[i * ne + j];
        orc_roothan(ne, Se, h, 1, nocc, pre, CE, evE, DE);
    }
    for (i64 i = 0; i < np; ++i)
        for (i64 j = 0; j < np; ++j) h[j * np + i] = Hp[i * np + j];
    orc_roothan(np, Sp, h, nucst, nucst, 1.0, CP, evP, DP);
    double eold = 0.0;
    i64 it, done = 0;
    for (it = 1; it <= maxit && !done; ++it) {
        double c[14];
        orc_xchf_fock(ne, np, nelec, enuc, DE, DP, He, Hp, Gep, Gee, G1, G1s, G2, G2s, G3, G4, FEd, FP, c);
        if (it == 1 && comps_init) memcpy(comps_init, c, sizeof(c));
        orc_roothan(np, Sp, FP, nucst, nucst, 1.0, CP, evP, DP);
        orc_xchf_fock(ne, np, nelec, enuc, DE, DP, He, Hp, Gep, Gee, G1, G1s, G2, G2s, G3, G4, FE, FPd, c);
        orc_roothan(ne, Se, FE, 1, nocc, pre, CE, evE, DE);
        double de = 0.0, dp = 0.0;
        for (i64 t = 0; t < ne * ne; ++t) {
            de = fmax(de, fabs(DE0[t] - DE[t]));
            DE0[t] = DE[t];
        }
        for (i64 t = 0; t < np * np; ++t) {
            dp = fmax(dp, fabs(DP0[t] - DP[t]));
            DP0[t] = DP[t];
        }
        if (trace) {
            trace[4 * (it - 1)] = c[0];
            trace[4 * (it - 1) + 1] = c[0] - eold;
            trace[4 * (it - 1) + 2] = de;
            trace[4 * (it - 1) + 3] = dp;
        }
        eold = c[0];
        if (comps) memcpy(comps, c, sizeof(c));
        done = np > 1 ? (de < tole && dp < tolp) : (de < tole);
    }
    free(DE); free(DP); free(DE0); free(DP0); free(FE); free(FP); free(FEd); free(FPd);
    free(CE); free(CP); free(evE); free(evP); free(h);
    return done ? it - 1 : -(it - 1);
}

/* ------------------------------------------------------------------------
 * RXCHF (source/rxchf/, one special electron): the e-e-p tensors and their
 * Fock contractions.
 * ------------------------------------------------------------------------ */
/* RXCHF_GAM2_IC1 (rxchf_gam2_IC1_OMP.f:2-250) with RXCHF_contract_omega2_conv
 * (rxchf_gam2_CONV_OMP.f:388-580): the contracted ansE1 as is (GM2_1), the
 * pair-swap averages (x12 + x21)/2 of the contracted ansE2 (GM2_2) and ansS
 * (GM2s) */
void orc_rxchf_gam2_ic1(BAS_ARGS, SYS_ARGS, double *GM2_1, double *GM2_2, double *GM2s)
{
    BAS_INIT;
    SYS_INIT;
    i64 ne = nebf, np = npbf, ng2 = ne * ne * ne * ne * np * np;
    double *X2 = (double *)malloc(sizeof(double) * ng2), *XS = (double *)malloc(sizeof(double) * ng2);
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            double o1 = 0.0, o2 = 0.0, os = 0.0;
                            for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                    for (i64 a2 = kpestr[i2] - 1; a2 < kpeend[i2]; ++a2)
                                        for (i64 c2 = kpestr[j2] - 1; c2 < kpeend[j2]; ++c2) {
                                            prim_t b[3] = {eprim(&B, a1), eprim(&B, a2), pprim(&B, ip)};
                                            prim_t k[3] = {eprim(&B, c1), eprim(&B, c2), pprim(&B, jp)};
                                            double e1, e2, sv;
                                            rxchf2_prim(&S, b, k, &e1, &e2, &sv);
                                            double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] * agebfcc[c1] *
                                                        agebfcc[a2] * agebfcc[c2];
                                            o1 += e1 * cf;
                                            o2 += e2 * cf;
                                            os += sv * cf;
                                        }
                            i64 ia = ix2(ne, np, ip, jp, i1, j1, i2, j2);
                            GM2_1[ia] = o1;
                            X2[ia] = o2;
                            XS[ia] = os;
                        }
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2) {
                            i64 a12 = ix2(ne, np, ip, jp, i1, j1, i2, j2), a21 = ix2(ne, np, ip, jp, i2, j2, i1, j1);
                            GM2_2[a12] = (X2[a12] + X2[a21]) / 2.0;
                            GM2s[a12] = (XS[a12] + XS[a21]) / 2.0;
                        }
    free(X2);
    free(XS);
}

#define RXA(i, j) DAE[(j) * ne + (i)]
#define RXB(i, j) DBE[(j) * ne + (i)]
#define RXP(i, j) DP[(j) * np + (i)]
/* RXCHF_FAE_OMG2 / _FP_OMG2 / _FBE_OMG2 (rxchf_omg2.f:2-660); which = 0, 1,
 * 2 selects the routine.  Per entry (rxchf_omg2.f:193-227, :411-445,
 * :629-663): val1 = GM2_1(i1 j1 i2 j2), val2 = GM2_2(i1 j2 i2 j1),
 * Sval = GM2s(i1 j2 i2 j1).  F and SF accumulate; E, S overwritten. */
void orc_rxchf_omg2(i64 which, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                    const double *GM2_1, const double *GM2_2, const double *GM2s, double *F, double *SF, double *E,
                    double *Sv)
{
    i64 ne = nebf, np = npbf;
    i64 nf = which == 1 ? np * np : ne * ne;
    double *xf = (double *)calloc(nf, sizeof(double)), *xs = (double *)calloc(nf, sizeof(double));
    double e = 0.0, sv = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
    for (i64 jp = 0; jp < np; ++jp)
    for (i64 i1 = 0; i1 < ne; ++i1)
    for (i64 j1 = 0; j1 < ne; ++j1)
    for (i64 i2 = 0; i2 < ne; ++i2)
    for (i64 j2 = 0; j2 < ne; ++j2) {
        double val1 = GM2_1[ix2(ne, np, ip, jp, i1, j1, i2, j2)];
        double val2 = GM2_2[ix2(ne, np, ip, jp, i1, j2, i2, j1)];
        double sval = GM2s[ix2(ne, np, ip, jp, i1, j2, i2, j1)];
        double w = RXP(ip, jp) * RXA(i2, j2) * RXB(i1, j1);
        if (which == 0) {
            xf[j2 * ne + i2] += RXP(ip, jp) * RXB(i1, j1) * (2.0 * val1 - val2);
            xs[j2 * ne + i2] += RXP(ip, jp) * RXB(i1, j1) * sval;
        } else if (which == 1) {
            xf[jp * np + ip] += RXA(i2, j2) * RXB(i1, j1) * (val1 - 0.5 * val2);
            xs[jp * np + ip] -= 0.5 * RXA(i2, j2) * RXB(i1, j1) * sval;
        } else {
            xf[j1 * ne + i1] += RXA(i2, j2) * RXP(ip, jp) * (val1 - 0.5 * val2);
            xs[j1 * ne + i1] -= 0.5 * RXA(i2, j2) * RXP(ip, jp) * sval;
        }
        e += w * (val1 - 0.5 * val2);
        sv += w * (-0.5) * sval;
    }
    for (i64 k = 0; k < nf; ++k) {
        F[k] += xf[k];
        SF[k] += xs[k];
    }
    *E = e;
    *Sv = sv;
    free(xf);
    free(xs);
}
#undef RXA
#undef RXB
#undef RXP

/* RXCHF_GAM3_IC1 (rxchf_gam3_IC1_OMP.f:2-283) with
 * RXCHF_contract_omega3_conv (rxchf_gam3_CONV_OMP.f:479-713): contracted ans1
 * averaged over the exchange of pairs 2 and 3, (x123 + x132)/2 (GM3_1,
 * :211-217), and contracted ans2 as is (GM3_2, :235-238) */
void orc_rxchf_gam3_ic1(BAS_ARGS, SYS_ARGS, double *GM3_1, double *GM3_2)
{
    BAS_INIT;
    SYS_INIT;
    i64 ne = nebf, np = npbf, ng3 = ne * ne * ne * ne * ne * ne * np * np;
    double *X1 = (double *)malloc(sizeof(double) * ng3);
#pragma omp parallel for collapse(3) schedule(dynamic, 1)
    for (i64 ip = 0; ip < np; ++ip)
        for (i64 jp = 0; jp < np; ++jp)
            for (i64 i1 = 0; i1 < ne; ++i1)
                for (i64 j1 = 0; j1 < ne; ++j1)
                    for (i64 i2 = 0; i2 < ne; ++i2)
                        for (i64 j2 = 0; j2 < ne; ++j2)
                            for (i64 i3 = 0; i3 < ne; ++i3)
                                for (i64 j3 = 0; j3 < ne; ++j3) {
                                    double o1 = 0.0, o2 = 0.0;
                                    for (i64 a1 = kpestr[i1] - 1; a1 < kpeend[i1]; ++a1)
                                    for (i64 c1 = kpestr[j1] - 1; c1 < kpeend[j1]; ++c1)
                                    for (i64 a2 = kpestr[i2] - 1; a2 < kpeend[i2]; ++a2)
                                    for (i64 c2 = kpestr[j2] - 1; c2 < kpeend[j2]; ++c2)
                                    for (i64 a3 = kpestr[i3] - 1; a3 < kpeend[i3]; ++a3)
                                    for (i64 c3 = kpestr[j3] - 1; c3 < kpeend[j3]; ++c3) {
                                        prim_t b[4] = {eprim(&B, a1), eprim(&B, a2), eprim(&B, a3), pprim(&B, ip)};
                                        prim_t k[4] = {eprim(&B, c1), eprim(&B, c2), eprim(&B, c3), pprim(&B, jp)};
                                        double v1, v2;
                                        rxchf3_prim(&S, b, k, &v1, &v2);
                                        double cf = agnbfcc[ip] * agnbfcc[jp] * agebfcc[a1] * agebfcc[c1] *
                                                    agebfcc[a2] * agebfcc[c2] * agebfcc[a3] * agebfcc[c3];
                                        o1 += v1 * cf;
                                        o2 += v2 * cf;
                                    }
                                    i64 ia = ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3);
                                    X1[ia] = o1;
                                    GM3_2[ia] = o2;
                                }
    for (i64 ip = 0; ip < np; ++ip)
    for (i64 jp = 0; jp < np; ++jp)
    for (i64 i1 = 0; i1 < ne; ++i1)
    for (i64 j1 = 0; j1 < ne; ++j1)
    for (i64 i2 = 0; i2 < ne; ++i2)
    for (i64 j2 = 0; j2 < ne; ++j2)
    for (i64 i3 = 0; i3 < ne; ++i3)
    for (i64 j3 = 0; j3 < ne; ++j3) {
        i64 a = ix3(ne, np, ip, jp, i1, j1, i2, j2, i3, j3), c = ix3(ne, np, ip, jp, i1, j1, i3, j3, i2, j2);
        GM3_1[a] = (X1[a] + X1[c]) * 0.5;
    }
    free(X1);
}

#define RXA(i, j) DAE[(j) * ne + (i)]
#define RXB(i, j) DBE[(j) * ne + (i)]
#define RXP(i, j) DP[(j) * np + (i)]
#define G31(a1, b1, a2, b2, a3, b3) GM3_1[ix3(ne, np, ip, jp, a1, b1, a2, b2, a3, b3)]
#define G32(a1, b1, a2, b2, a3, b3) GM3_2[ix3(ne, np, ip, jp, a1, b1, a2, b2, a3, b3)]
/* RXCHF_FAE_OMG3 / _FP_OMG3 / _FBE_OMG3 (rxchf_omg3.f:2-708); which = 0, 1, 2
 * selects the routine.  Per entry (ip,jp,i1,j1,i2,j2,i3,j3):
 *   val1 = GM3_1(i1 j1 i2 j2 i3 j3) - GM3_1(i1 j1 i2 j3 i3 j2)/2
 * (:215, :465, :688).
 * FAE (:226-251): val2 = (GM3_2(i1 j2 i2 j1 i3 j3) + GM3_2(i1 i2 j2 j1 i3 j3))
 *                        - (GM3_2(i1 j3 i2 j1 i3 j2) + GM3_2(i1 j3 j2 j1 i3 i2))/2,
 *   val3 = (GM3_2(i1 j3 i3 j1 i2 j2) + GM3_2(i1 j3 i3 j1 j2 i2))
 *          - (GM3_2(i1 j2 i3 j1 i2 j3) + GM3_2(i1 i2 i3 j1 j2 j3))/2,
 *   FAE(i2,j2) += DAE(i3,j3) DBE(i1,j1) DP(ip,jp) (2 val1 - (val2 + val3)/2),
 *   E += DP DBE(i1,j1) DAE(i2,j2) DAE(i3,j3) (val1 - val2/2) / 2;
 * FP, FBE (:466-475, :689-698): val2 = GM3_2(i1 j2 i2 j1 i3 j3)
 *   - GM3_2(i1 j3 i2 j1 i3 j2)/2, w = DAE(i2,j2) DAE(i3,j3) (val1 - val2) / 2,
 *   FP(ip,jp) += DBE(i1,j1) w, FBE(i1,j1) += DP(ip,jp) w, E += DP DBE w.
 * F accumulates (add2fock); E is overwritten. */
void orc_rxchf_omg3(i64 which, i64 nebf, i64 npbf, const double *DAE, const double *DBE, const double *DP,
                    const double *GM3_1, const double *GM3_2, double *F, double *E)
{
    i64 ne = nebf, np = npbf;
    i64 nf = which == 1 ? np * np : ne * ne;
    double *xf = (double *)calloc(nf, sizeof(double));
    double e = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
    for (i64 jp = 0; jp < np; ++jp)
    for (i64 i1 = 0; i1 < ne; ++i1)
    for (i64 j1 = 0; j1 < ne; ++j1)
    for (i64 i2 = 0; i2 < ne; ++i2)
    for (i64 j2 = 0; j2 < ne; ++j2)
    for (i64 i3 = 0; i3 < ne; ++i3)
    for (i64 j3 = 0; j3 < ne; ++j3) {
        double val1 = G31(i1, j1, i2, j2, i3, j3) - 0.5 * G31(i1, j1, i2, j3, i3, j2);
        if (which == 0) {
            double val2 = (G32(i1, j2, i2, j1, i3, j3) + G32(i1, i2, j2, j1, i3, j3)) -
                          0.5 * (G32(i1, j3, i2, j1, i3, j2) + G32(i1, j3, j2, j1, i3, i2));
            double val3 = (G32(i1, j3, i3, j1, i2, j2) + G32(i1, j3, i3, j1, j2, i2)) -
                          0.5 * (G32(i1, j2, i3, j1, i2, j3) + G32(i1, i2, i3, j1, j2, j3));
            xf[j2 * ne + i2] += RXA(i3, j3) * RXB(i1, j1) * RXP(ip, jp) * (2.0 * val1 - 0.5 * (val2 + val3));
            e += 0.5 * RXP(ip, jp) * RXB(i1, j1) * RXA(i2, j2) * RXA(i3, j3) * (val1 - 0.5 * val2);
        } else {
            double val2 = G32(i1, j2, i2, j1, i3, j3) - 0.5 * G32(i1, j3, i2, j1, i3, j2);
            double w = 0.5 * RXA(i2, j2) * RXA(i3, j3) * (val1 - val2);
            if (which == 1)
                xf[jp * np + ip] += RXB(i1, j1) * w;
            else
                xf[j1 * ne + i1] += RXP(ip, jp) * w;
            e += RXP(ip, jp) * RXB(i1, j1) * w;
        }
    }
    for (i64 k = 0; k < nf; ++k) F[k] += xf[k];
    *E = e;
    free(xf);
}
#undef G31
#undef G32

/* RXCHF_FP_OMG1 / RXCHF_FBE_OMG1 (rxchf_omg1.f:2-165, per entry :60-66 and
 * :142-148) over the GAM_1 and
 * GAM_1s arrays (pack_4D(nebf,nebf,npbf, je1,ie1,jp,ip) order, as the .ufm
 * files hold them); which = 1 (FP) or 2 (FBE).  F accumulates (add2fock), SF
 * is overwritten with the GAM_1s contraction, E and S overwritten. */
void orc_rxchf_omg1(i64 which, i64 nebf, i64 npbf, const double *DBE, const double *DP, const double *G1,
                    const double *G1s, double *F, double *SF, double *E, double *Sv)
{
    i64 ne = nebf, np = npbf;
    i64 nf = which == 1 ? np * np : ne * ne;
    double *xf = (double *)calloc(nf, sizeof(double));
    double e = 0.0, sv = 0.0;
    for (i64 k = 0; k < nf; ++k) SF[k] = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
    for (i64 jp = 0; jp < np; ++jp)
    for (i64 i1 = 0; i1 < ne; ++i1)
    for (i64 j1 = 0; j1 < ne; ++j1) {
        i64 ia = ((ip * np + jp) * ne + i1) * ne + j1;
        if (which == 1) {
            xf[jp * np + ip] += RXB(i1, j1) * G1[ia];
            SF[jp * np + ip] += RXB(i1, j1) * G1s[ia];
        } else {
            xf[j1 * ne + i1] += RXP(ip, jp) * G1[ia];
            SF[j1 * ne + i1] += RXP(ip, jp) * G1s[ia];
        }
        e += RXP(ip, jp) * RXB(i1, j1) * G1[ia];
        sv += RXP(ip, jp) * RXB(i1, j1) * G1s[ia];
    }
    for (i64 k = 0; k < nf; ++k) F[k] += xf[k];
    *E = e;
    *Sv = sv;
    free(xf);
}
#undef RXA
#undef RXB
#undef RXP

/* RXCUHF (open-shell regular electrons, one special electron) */
#define RXL(i, j) Dal[(j) * ne + (i)]
#define RXBt(i, j) Dbt[(j) * ne + (i)]
#define RXT(i, j) Dto[(j) * ne + (i)]
#define RXB(i, j) DBE[(j) * ne + (i)]
#define RXP(i, j) DP[(j) * np + (i)]
/* RXCUHF_FE_OMG2 / _FP_OMG2 / _FBE_OMG2 (rxchf/rxcuhf_omg2.f:2-694); which =
 * 0, 1, 2.  Per entry (:219-243, :438-454, :664-680): val1 = GM2_1(i1 j1 i2 j2),
 * val2 = GM2_2(i1 j2 i2 j1), Sval = GM2s(i1 j2 i2 j1);
 *   FE:  FAalp(i2,j2) += DP DBE (val1 - val2), SAalp += DP DBE Sval,
 *        FAbet(i2,j2) += DP DBE val1;
 *   u = DAtot(i2,j2) val1 - DAalp(i2,j2) val2, us = -DAalp(i2,j2) Sval:
 *   FP(ip,jp) += DBE u, SP += DBE us;  FBE(i1,j1) += DP u, SBE += DP us;
 *   E = sum DP DBE u, S = sum DP DBE us.
 * F, F2 (FAbet, which 0), SF accumulate (add2fock); E, S overwritten. */
void orc_rxcuhf_omg2(i64 which, i64 nebf, i64 npbf, const double *Dal, const double *Dto, const double *DBE,
                     const double *DP, const double *GM2_1, const double *GM2_2, const double *GM2s, double *F,
                     double *F2, double *SF, double *E, double *Sv)
{
    i64 ne = nebf, np = npbf;
    i64 nf = which == 1 ? np * np : ne * ne;
    double *xf = (double *)calloc(nf, sizeof(double)), *xs = (double *)calloc(nf, sizeof(double));
    double *xb = (double *)calloc(nf, sizeof(double));
    double e = 0.0, sv = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
    for (i64 jp = 0; jp < np; ++jp)
    for (i64 i1 = 0; i1 < ne; ++i1)
    for (i64 j1 = 0; j1 < ne; ++j1)
    for (i64 i2 = 0; i2 < ne; ++i2)
    for (i64 j2 = 0; j2 < ne; ++j2) {
        double val1 = GM2_1[ix2(ne, np, ip, jp, i1, j1, i2, j2)];
        double val2 = GM2_2[ix2(ne, np, ip, jp, i1, j2, i2, j1)];
        double sval = GM2s[ix2(ne, np, ip, jp, i1, j2, i2, j1)];
        double u = RXT(i2, j2) * val1 - RXL(i2, j2) * val2, us = -RXL(i2, j2) * sval;
        if (which == 0) {
            xf[j2 * ne + i2] += RXP(ip, jp) * RXB(i1, j1) * (val1 - val2);
            xs[j2 * ne + i2] += RXP(ip, jp) * RXB(i1, j1) * sval;
            xb[j2 * ne + i2] += RXP(ip, jp) * RXB(i1, j1) * val1;
        } else if (which == 1) {
            xf[jp * np + ip] += RXB(i1, j1) * u;
            xs[jp * np + ip] += RXB(i1, j1) * us;
        } else {
            xf[j1 * ne + i1] += RXP(ip, jp) * u;
            xs[j1 * ne + i1] += RXP(ip, jp) * us;
        }
        e += RXP(ip, jp) * RXB(i1, j1) * u;
        sv += RXP(ip, jp) * RXB(i1, j1) * us;
    }
    for (i64 k = 0; k < nf; ++k) {
        F[k] += xf[k];
        SF[k] += xs[k];
        if (which == 0) F2[k] += xb[k];
    }
    *E = e;
    *Sv = sv;
    free(xf);
    free(xs);
    free(xb);
}

#define G31(a1, b1, a2, b2, a3, b3) GM3_1[ix3(ne, np, ip, jp, a1, b1, a2, b2, a3, b3)]
#define G32(a1, b1, a2, b2, a3, b3) GM3_2[ix3(ne, np, ip, jp, a1, b1, a2, b2, a3, b3)]
/* RXCUHF_FE_OMG3 / _FP_OMG3 / _FBE_OMG3 (rxchf/rxcuhf_omg3.f:2-728); which =
 * 0, 1, 2.  Per entry (ip,jp,i1,j1,i2,j2,i3,j3), with
 *   val1a = GM3_1(i1 j1 i2 j2 i3 j3), val1b = GM3_1(i1 j1 i2 j3 i3 j2),
 *   g213 = GM3_2(i1 j2 i2 j1 i3 j3), g312 = GM3_2(i1 j3 i2 j1 i3 j2),
 *   w = DAtot2 DAtot3 val1a / 2 - (DAalp2 DAalp3 + DAbet2 DAbet3) val1b / 2
 *       - DAalp2 DAtot3 g213 + DAalp2 DAalp3 g312        (:253-259, :476-489, :706-719)
 * (densities at (i2,j2) and (i3,j3)): FP(ip,jp) += DBE w, FBE(i1,j1) += DP w,
 * E = sum DP DBE w for all three routines; FE (:233-251):
 *   val2a = (g213 + GM3_2(i1 i2 j2 j1 i3 j3))/2, val2b = (g312 + GM3_2(i1 j3 j2 j1 i3 i2))/2,
 *   val2c = (GM3_2(i1 j3 i3 j1 i2 j2) + GM3_2(i1 j3 i3 j1 j2 i2))/2,
 *   val2d = (GM3_2(i1 j2 i3 j1 i2 j3) + GM3_2(i1 i2 i3 j1 j2 j3))/2,
 *   FAalp(i2,j2) += DBE DP (DAtot3 val1a - DAalp3 val1b - DAtot3 val2a + DAalp3 val2b
 *                           - DAalp3 val2c + DAalp3 val2d),
 *   FAbet(i2,j2) += DBE DP (DAtot3 val1a - DAbet3 val1b - DAalp3 val2c).
 * F, F2 (FAbet, which 0) accumulate; E overwritten. */
void orc_rxcuhf_omg3(i64 which, i64 nebf, i64 npbf, const double *Dal, const double *Dbt, const double *Dto,
                     const double *DBE, const double *DP, const double *GM3_1, const double *GM3_2, double *F,
                     double *F2, double *E)
{
    i64 ne = nebf, np = npbf;
    i64 nf = which == 1 ? np * np : ne * ne;
    double *xf = (double *)calloc(nf, sizeof(double)), *xb = (double *)calloc(nf, sizeof(double));
    double e = 0.0;
    for (i64 ip = 0; ip < np; ++ip)
    for (i64 jp = 0; jp < np; ++jp)
    for (i64 i1 = 0; i1 < ne; ++i1)
    for (i64 j1 = 0; j1 < ne; ++j1)
    for (i64 i2 = 0; i2 < ne; ++i2)
    for (i64 j2 = 0; j2 < ne; ++j2)
    for (i64 i3 = 0; i3 < ne; ++i3)
    for (i64 j3 = 0; j3 < ne; ++j3) {
        double val1a = G31(i1, j1, i2, j2, i3, j3), val1b = G31(i1, j1, i2, j3, i3, j2);
        double g213 = G32(i1, j2, i2, j1, i3, j3), g312 = G32(i1, j3, i2, j1, i3, j2);
        double w = 0.5 * RXT(i2, j2) * RXT(i3, j3) * val1a -
                   0.5 * (RXL(i2, j2) * RXL(i3, j3) + RXBt(i2, j2) * RXBt(i3, j3)) * val1b -
                   RXL(i2, j2) * RXT(i3, j3) * g213 + RXL(i2, j2) * RXL(i3, j3) * g312;
        if (which == 0) {
            double val2a = (g213 + G32(i1, i2, j2, j1, i3, j3)) * 0.5;
            double val2b = (g312 + G32(i1, j3, j2, j1, i3, i2)) * 0.5;
            double val2c = (G32(i1, j3, i3, j1, i2, j2) + G32(i1, j3, i3, j1, j2, i2)) * 0.5;
            double val2d = (G32(i1, j2, i3, j1, i2, j3) + G32(i1, i2, i3, j1, j2, j3)) * 0.5;
            double f = RXB(i1, j1) * RXP(ip, jp);
            xf[j2 * ne + i2] += f * (RXT(i3, j3) * val1a - RXL(i3, j3) * val1b - RXT(i3, j3) * val2a +
                                     RXL(i3, j3) * val2b - RXL(i3, j3) * val2c + RXL(i3, j3) * val2d);
            xb[j2 * ne + i2] += f * (RXT(i3, j3) * val1a - RXBt(i3, j3) * val1b - RXL(i3, j3) * val2c);
        } else if (which == 1) {
            xf[jp * np + ip] += RXB(i1, j1) * w;
        } else {
            xf[j1 * ne + i1] += RXP(ip, jp) * w;
        }
        e += RXP(ip, jp) * RXB(i1, j1) * w;
    }
    for (i64 k = 0; k < nf; ++k) {
        F[k] += xf[k];
        if (which == 0) F2[k] += xb[k];
    }
    *E = e;
    free(xf);
    free(xb);
}
#undef G31
#undef G32
#undef RXL
#undef RXBt
#undef RXT
#undef RXB
#undef RXP
