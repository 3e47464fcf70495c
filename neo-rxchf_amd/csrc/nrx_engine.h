// nrx_engine.h -- per-tuple geminal integral engine (device code, also
// compiled for the host by the unit tests of the engine itself).
//
// One primitive tuple of the NEO-XCHF geminal path is evaluated by a single
// lane.  Every elementary integral is computed per Cartesian direction with
// univariate polynomials only:
//
//   * each particle i carries its bra*ket product as a Gaussian
//     exp(-p_i y^2) (y = x - P_i) times a polynomial f_i(y);
//   * geminals form a star around the quantum proton (the hub), so the
//     electrons that the operator does not touch are integrated out one at a
//     time: a leaf's conditional expectation is a polynomial in the hub
//     coordinate (Gaussian smoothing + linear substitution) and its Gaussian
//     folds into the hub's Gaussian;
//   * a Coulomb-type operator leaves a one- or two-variable core; both sides
//     are written in the Hermite-Gaussian basis, convolved, and combined with
//     the derivative polynomials R_k(u^2) of exp(-lambda(u) D^2), lambda(u) =
//     lambda0 + lambda1 u^2, so that the whole integral is
//        const * exp(-lambda0 D^2) * sum_n c_n F_n(lambda1 D^2);
//     a geminal acting on the same pair as the Coulomb operator is folded into
//     the kernel (lambda0 != 0).  When the hub is linked to both Coulomb
//     particles its conditional expectation is expanded as a short sum of
//     separable products.
//   * F_n follows the reference's two-branch Boys semantics: for T > 20 the
//     asymptotic form sqrt(pi/T)/2 * (2m-1)!!/(2T)^m exactly as
//     AC_Integrals/common.f:317-341 of the reference; for T <= 20 (where the
//     reference's truncated series is within ~1e-13 of the exact function)
//     a 7-term Taylor expansion from a 0.05-spaced table of exact values,
//     then downward recursion.
//
// The kinetic operator is the second derivative of the ket along each
// direction, expanded into overlaps with modified polynomials of the particle
// and of its ket-geminal partner (same operator as G3_MD_T1d,
// CS_Integrals/XC1_G3_MD.f:2680-2868, and pgi_Lapx1, AC_Integrals/g2.f:655).
//
// Work shared by several integrals of one tuple is computed once: the
// Hermite expansion of the core is reused for every classical nucleus, the
// overlaps of the kinetic operator are shared by its variants, and the terms
// that depend only on gamma_k + gamma_l are evaluated for k <= l.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NRX_HD __host__ __device__ inline
#else
#include <cmath>
#define NRX_HD inline
#endif

// Host builds used for accounting may define NRX_COUNT_FLOPS to tally the
// floating-point operations the algorithm performs (FMA = 2).
#if defined(NRX_COUNT_FLOPS) && !defined(__HIP_DEVICE_COMPILE__)
extern thread_local double nrx_flops;
#define NRX_CNT(x) (nrx_flops += (double)(x))
#else
#define NRX_CNT(x) ((void)0)
#endif

namespace nrx {

// polynomial capacities: per-particle polynomials have degree <= 4 (d on both
// sides) or <= 6 inside the kinetic operator; absorbed/core polynomials have
// degree <= 16 (four particles with d functions).
constexpr int NS = 7;
constexpr int NP = 18;
constexpr int NPART = 4;
constexpr double PI = 3.141592653589793238462643383279502884;
// pgike1 (AC_Integrals/g2.f:216) normalises its overlaps with this value
constexpr double PI_G2 = 3.141592650;

template <int N> struct PolyT {
    double c[N];
    int n; // number of coefficients (degree + 1)
};
using Poly = PolyT<NP>;
using PolyS = PolyT<NS>;

template <int N> NRX_HD void p_const(PolyT<N> &p, double v)
{
    p.n = 1;
    p.c[0] = v;
}

template <int NA, int NB> NRX_HD void p_copy(const PolyT<NA> &a, PolyT<NB> &b)
{
    int n = a.n < NB ? a.n : NB;
    for (int k = 0; k < n; ++k) b.c[k] = a.c[k];
    b.n = n;
}

// r = a * b (truncated to NR coefficients; degrees are sized to fit)
template <int NA, int NB, int NR> NRX_HD void p_mul(const PolyT<NA> &a, const PolyT<NB> &b, PolyT<NR> &r)
{
    int n = a.n + b.n - 1;
    if (n > NR) n = NR;
    for (int k = 0; k < n; ++k) r.c[k] = 0.0;
    for (int i = 0; i < a.n; ++i) {
        double ai = a.c[i];
        int jm = b.n;
        if (i + jm > n) jm = n - i;
        for (int j = 0; j < jm; ++j) r.c[i + j] += ai * b.c[j];
        NRX_CNT(2 * jm);
    }
    r.n = n;
}

// a *= (y + s)
template <int N> NRX_HD void p_mul_lin(PolyT<N> &a, double s)
{
    int n = a.n;
    if (n < N) {
        a.c[n] = a.c[n - 1];
        for (int k = n - 1; k > 0; --k) a.c[k] = a.c[k - 1] + s * a.c[k];
        a.c[0] = s * a.c[0];
        a.n = n + 1;
        NRX_CNT(2 * n - 1);
    }
}

// (y + sa)^la (y + sb)^lb
template <int N> NRX_HD void p_pair(int la, double sa, int lb, double sb, PolyT<N> &r)
{
    p_const(r, 1.0);
    for (int i = 0; i < la; ++i) p_mul_lin(r, sa);
    for (int i = 0; i < lb; ++i) p_mul_lin(r, sb);
}

// h(s) = E[f(s + e)], e ~ N(0, var)
template <int N> NRX_HD void p_smooth(const PolyT<N> &f, double var, PolyT<N> &h)
{
    double mom[N];
    mom[0] = 1.0;
    if (N > 1) mom[1] = 0.0;
    for (int k = 2; k < f.n; ++k) mom[k] = (k & 1) ? 0.0 : mom[k - 2] * (k - 1) * var;
    for (int j = 0; j < f.n; ++j) {
        double acc = f.c[j], bin = 1.0; // C(k, j) for k = j + 2i
        for (int k = j + 2; k < f.n; k += 2) {
            bin = bin * (double)(k * (k - 1)) / (double)((k - j) * (k - j - 1));
            acc += f.c[k] * bin * mom[k - j];
        }
        h.c[j] = acc;
        NRX_CNT(6 * ((f.n - j) / 2));
    }
    NRX_CNT(2 * f.n);
    h.n = f.n;
}

template <int N> NRX_HD double p_eval(const PolyT<N> &h, double s)
{
    double v = 0.0;
    for (int k = h.n - 1; k >= 0; --k) v = v * s + h.c[k];
    NRX_CNT(2 * h.n);
    return v;
}

// g(y) = h(m0 + beta y)  (g may alias h)
template <int N> NRX_HD void p_compose(const PolyT<N> &h, double m0, double beta, PolyT<N> &g)
{
    PolyT<N> t = h;
    for (int i = 0; i < t.n; ++i)
        for (int k = t.n - 2; k >= i; --k) t.c[k] += m0 * t.c[k + 1];
    double bp = 1.0;
    for (int k = 0; k < t.n; ++k) {
        g.c[k] = t.c[k] * bp;
        bp *= beta;
    }
    g.n = t.n;
    NRX_CNT(t.n * (t.n - 1) + 2 * t.n);
}

// Hermite-Gaussian expansion: f(s) exp(-w s^2) = sum_t e_t Lambda_t(s),
// Lambda_t = d^t/dc^t exp(-w (s - c)^2) at c = 0;
// s Lambda_t = Lambda_{t+1}/(2w) + t Lambda_{t-1}
template <int N> NRX_HD void p_hermite(const PolyT<N> &f, double w, PolyT<N> &e)
{
    double i2w = 0.5 / w;
    int n = f.n;
    double cur[N];
    cur[0] = f.c[n - 1];
    int m = 1;
    for (int k = n - 2; k >= 0; --k) {
        double prev = 0.0; // cur[t-1] before update
        for (int t = 0; t <= m; ++t) {
            double old = (t < m) ? cur[t] : 0.0;
            double v = prev * i2w;
            if (t + 1 < m) v += (t + 1) * cur[t + 1];
            prev = old;
            cur[t] = v;
        }
        ++m;
        cur[0] += f.c[k];
        NRX_CNT(4 * m);
    }
    for (int t = 0; t < m; ++t) e.c[t] = cur[t];
    e.n = m;
}

// ------------------------------------------------------------------
// Boys function, reference semantics, F[0..M]
// ------------------------------------------------------------------
constexpr int BT_N = 401;      // grid T = 0, 0.05, ..., 20
constexpr double BT_H = 0.05;
constexpr int BT_M = 32;       // orders stored per grid point
constexpr int BT_K = 7;        // Taylor terms
constexpr int BT_MAXM = BT_M - BT_K; // largest order evaluated from the table

NRX_HD void boys_ref(int M, double T, double *F, const double *tab)
{
    if (T > 20.0) {
        // common.f:326-328
        F[0] = 0.88622692545275801365 / sqrt(T);
        for (int l = 1; l <= M; ++l) F[l] = F[l - 1] * (1.0 * l - 0.5) / T;
        NRX_CNT(3 * M + 2);
        return;
    }
    int i = (int)(T * (1.0 / BT_H) + 0.5);
    double dt = T - i * BT_H;
    const double *row = tab + i * BT_M + M;
    // F_M(T) = sum_k F_{M+k}(T_i) (-dt)^k / k!
    double v = row[BT_K - 1];
    for (int k = BT_K - 1; k >= 1; --k) v = row[k - 1] - dt * v * (1.0 / k);
    F[M] = v;
    double ex = exp(-T), t2 = 2.0 * T;
    for (int l = M - 1; l >= 0; --l) F[l] = (t2 * F[l + 1] + ex) * (1.0 / (2 * l + 1));
    NRX_CNT(4 * BT_K + 4 * M + 10);
}

// host-side construction of the table (exact values, extended precision)
inline void boys_table_fill(double *tab)
{
    for (int i = 0; i < BT_N; ++i) {
        long double T = (long double)i * (long double)BT_H;
        for (int m = 0; m < BT_M; ++m) {
            // F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
            long double term = 1.0L / (2 * m + 1), sum = term;
            for (int k = 1; k < 5000; ++k) {
                term = term * 2.0L * T / (2 * m + 2 * k + 1);
                sum += term;
                if (term < 1e-24L * sum) break;
            }
            tab[i * BT_M + m] = (double)(expl(-T) * sum);
        }
    }
}

// ------------------------------------------------------------------
// per-particle bra*ket product
// ------------------------------------------------------------------
struct Prim {
    double a;
    double A[3];
    int l[3];
};

struct Pair {
    double p, P[3], K; // exponent, centre, prod_d exp(-ab/p (A-B)^2)
    PolyS f[3];        // per-direction polynomial in y = x - P
    double beta, B[3], sA[3], sB[3];
    int la[3], lb[3];
};

NRX_HD void make_pair(const Prim &bra, const Prim &ket, Pair &q)
{
    double a = bra.a, b = ket.a, p = a + b;
    q.p = p;
    q.beta = b;
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        double P = (a * bra.A[d] + b * ket.A[d]) / p;
        double AB = bra.A[d] - ket.A[d];
        r2 += AB * AB;
        q.P[d] = P;
        q.B[d] = ket.A[d];
        q.sA[d] = P - bra.A[d];
        q.sB[d] = P - ket.A[d];
        q.la[d] = bra.l[d];
        q.lb[d] = ket.l[d];
        p_pair(bra.l[d], q.sA[d], ket.l[d], q.sB[d], q.f[d]);
    }
    q.K = exp(-a * b / p * r2);
    NRX_CNT(3 * 12 + 6);
}

// ------------------------------------------------------------------
// Star elimination per direction
// ------------------------------------------------------------------
// variable being reduced: polynomial F in the particle coordinate y and
// Gaussian exp(-w (y - c)^2)
template <int N> struct VarT {
    PolyT<N> F;
    double w, c;
};
using Var = VarT<NP>;

// Integrate leaf L (edge g, d = P_leaf - P_parent) into parent Q; returns the
// scalar factor sqrt(pi/(wL+g)) * exp(...).
template <int NL, int NQ>
NRX_HD double absorb_leaf(const PolyT<NL> &LF, double Lw, double Lc, double g, double d, VarT<NQ> &Q, double pic)
{
    double W = Lw + g;
    PolyT<NL> h;
    p_smooth(LF, 0.5 / W, h);
    double fac = sqrt(pic / W);
    if (g == 0.0) {
        double v = p_eval(h, Lc);
        for (int k = 0; k < Q.F.n; ++k) Q.F.c[k] *= v;
        NRX_CNT(Q.F.n + 4);
        return fac;
    }
    double beta = g / W, m0 = (Lw * Lc - g * d) / W;
    p_compose(h, m0, beta, h);
    PolyT<NQ> r;
    p_mul(Q.F, h, r);
    Q.F = r;
    double kap = Lw * g / W;
    double cl = d + Lc;
    double Wq = Q.w + kap;
    double diff = Q.c - cl;
    fac *= exp(-(Q.w * kap / Wq) * diff * diff);
    Q.c = (Q.w * Q.c + kap * cl) / Wq;
    Q.w = Wq;
    NRX_CNT(26);
    return fac;
}

template <int N> NRX_HD double close_var(const VarT<N> &V, double pic)
{
    PolyT<N> h;
    p_smooth(V.F, 0.5 / V.w, h);
    return sqrt(pic / V.w) * p_eval(h, V.c);
}

// Integral description: particles 0..n-1, hub = n-1, electron i linked to the
// hub by gam[i] (bra+ket).
struct Sys {
    int n;
    const Pair *q[NPART];
    double gam[NPART];
    const double *bt; // Boys table
};

NRX_HD double Kprod(const Sys &s)
{
    double k = 1.0;
    for (int i = 0; i < s.n; ++i) k *= s.q[i]->K;
    return k;
}

// hub state along d after absorbing every electron whose skip flag is unset
NRX_HD double hub_absorb(const Sys &s, int d, unsigned skip, Var &H, double pic)
{
    int h = s.n - 1;
    p_copy(s.q[h]->f[d], H.F);
    H.w = s.q[h]->p;
    H.c = 0.0;
    double fac = 1.0;
    for (int i = 0; i < h; ++i)
        if (!((skip >> i) & 1u))
            fac *= absorb_leaf(s.q[i]->f[d], s.q[i]->p, 0.0, s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], H, pic);
    return fac;
}

// 1-D overlap along d with the given per-particle polynomials
NRX_HD double ovl_dim(const Sys &s, int d, const PolyS *const *F, double pic)
{
    int h = s.n - 1;
    Var H;
    p_copy(*F[h], H.F);
    H.w = s.q[h]->p;
    H.c = 0.0;
    double fac = 1.0;
    for (int i = 0; i < h; ++i) fac *= absorb_leaf(*F[i], s.q[i]->p, 0.0, s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], H, pic);
    return fac * close_var(H, pic);
}

NRX_HD double ovl(const Sys &s, double pic = PI)
{
    double r = Kprod(s);
    for (int d = 0; d < 3; ++d) {
        const PolyS *F[NPART];
        for (int i = 0; i < s.n; ++i) F[i] = &s.q[i]->f[d];
        r *= ovl_dim(s, d, F, pic);
    }
    return r;
}

// Kinetic energy of particle a (-1/(2m) <bra| d^2/dx_a^2 |ket>) for nv <= 2
// variants sharing one geminal configuration: variant v has a ket geminal of
// exponent gk[v] between a and partner c (gk[v] == 0 or c < 0: none).
NRX_HD void kin_multi(const Sys &s, int a, int c, const double *gk, int nv, double mass, double pic, double *out)
{
    double S[3], T[3][2];
    const Pair &A = *s.q[a];
    for (int d = 0; d < 3; ++d) {
        const PolyS *F[NPART];
        for (int i = 0; i < s.n; ++i) F[i] = &s.q[i]->f[d];
        S[d] = ovl_dim(s, d, F, pic);
        int L = A.lb[d];
        PolyS xm[5]; // u_A(y) X^{L-2+j}, j = 0..4
        for (int j = 0; j < 5; ++j)
            if (L - 2 + j >= 0) p_pair(A.la[d], A.sA[d], L - 2 + j, A.sB[d], xm[j]);
        for (int v = 0; v < nv; ++v) {
            double g = (c >= 0) ? gk[v] : 0.0;
            double al = A.beta + g;
            double dl = (g != 0.0) ? g * (A.B[d] - s.q[c]->B[d]) : 0.0;
            // F0 = L(L-1)X^{L-2} - 4L dl X^{L-1} + (-4L al + 4dl^2 - 2al) X^L + 8 al dl X^{L+1} + 4al^2 X^{L+2}
            // F1 = 4 L g X^{L-1} - 8 dl g X^L - 8 al g X^{L+1}      (times X_c)
            // F2 = 4 g^2 X^L                                        (times X_c^2)
            double cf0[5] = {(double)(L * (L - 1)), -4.0 * L * dl, -4.0 * L * al + 4.0 * dl * dl - 2.0 * al,
                             8.0 * al * dl, 4.0 * al * al};
            PolyS F0;
            F0.n = 1;
            F0.c[0] = 0.0;
            for (int j = 0; j < 5; ++j) {
                if (L - 2 + j < 0 || cf0[j] == 0.0) continue;
                for (int k = F0.n; k < xm[j].n; ++k) F0.c[k] = 0.0;
                if (xm[j].n > F0.n) F0.n = xm[j].n;
                for (int k = 0; k < xm[j].n; ++k) F0.c[k] += cf0[j] * xm[j].c[k];
                NRX_CNT(2 * xm[j].n);
            }
            const PolyS *F[NPART];
            for (int i = 0; i < s.n; ++i) F[i] = &s.q[i]->f[d];
            F[a] = &F0;
            double t = ovl_dim(s, d, F, pic);
            if (g != 0.0) {
                double cf1[5] = {0.0, 4.0 * L * g, -8.0 * dl * g, -8.0 * al * g, 0.0};
                PolyS F1, F2;
                F1.n = 1;
                F1.c[0] = 0.0;
                for (int j = 1; j < 4; ++j) {
                    if (L - 2 + j < 0 || cf1[j] == 0.0) continue;
                    for (int k = F1.n; k < xm[j].n; ++k) F1.c[k] = 0.0;
                    if (xm[j].n > F1.n) F1.n = xm[j].n;
                    for (int k = 0; k < xm[j].n; ++k) F1.c[k] += cf1[j] * xm[j].c[k];
                }
                F2 = xm[2];
                for (int k = 0; k < F2.n; ++k) F2.c[k] *= 4.0 * g * g;
                PolyS fc1 = s.q[c]->f[d];
                double xc = s.q[c]->P[d] - s.q[c]->B[d];
                p_mul_lin(fc1, xc);
                PolyS fc2 = fc1;
                p_mul_lin(fc2, xc);
                F[a] = &F1;
                F[c] = &fc1;
                t += ovl_dim(s, d, F, pic);
                F[a] = &F2;
                F[c] = &fc2;
                t += ovl_dim(s, d, F, pic);
            }
            T[d][v] = t;
        }
    }
    double kp = Kprod(s);
    for (int v = 0; v < nv; ++v) {
        double r = T[0][v] * S[1] * S[2] + S[0] * T[1][v] * S[2] + S[0] * S[1] * T[2][v];
        out[v] = -0.5 * r / mass * kp;
    }
}

NRX_HD double kin(const Sys &s, int a, int c, double gk, double mass, double pic = PI)
{
    double o;
    kin_multi(s, a, c, &gk, 1, mass, pic, &o);
    return o;
}

// P(u^2) = sum_k hk R_k(u^2), R_0 = 1, R_1 = -2 lam D,
// R_{k+1} = -2 lam (D R_k + k R_{k-1}), lam = l0 + l1 u^2
NRX_HD void hermite_coulomb(const Poly &hk, double D, double l0, double l1, Poly &P)
{
    double Rm[NP], R[NP];
    int rn = 1;
    R[0] = 1.0;
    Rm[0] = 0.0;
    P.n = hk.n;
    for (int j = 0; j < hk.n; ++j) P.c[j] = 0.0;
    P.c[0] = hk.c[0];
    for (int k = 0; k + 1 < hk.n; ++k) {
        // Rn = -2 (l0 + l1 u) (D R + k Rm); Rm has rn-1 coefficients
        double prevt = 0.0;
        for (int j = 0; j <= rn; ++j) {
            double tj = (j < rn) ? D * R[j] + ((j < rn - 1) ? k * Rm[j] : 0.0) : 0.0;
            double v = l0 * tj + l1 * prevt;
            prevt = tj;
            if (j < rn) Rm[j] = R[j];
            R[j] = -2.0 * v;
        }
        ++rn;
        double h = hk.c[k + 1];
        for (int j = 0; j < rn; ++j) P.c[j] += h * R[j];
        NRX_CNT(9 * rn);
    }
}

NRX_HD double boys_combine(const Poly *P3, double T, const double *bt)
{
    Poly xy, xyz;
    p_mul(P3[0], P3[1], xy);
    p_mul(xy, P3[2], xyz);
    double F[NP];
    int M = xyz.n - 1;
    boys_ref(M, T, F, bt);
    double s = 0.0;
    for (int k = 0; k <= M; ++k) s += xyz.c[k] * F[k];
    NRX_CNT(2 * (M + 1));
    return s;
}

NRX_HD void underflow(double &x)
{
    if (fabs(x) < 1.0e-100) x = 0.0;
}

// Hermite coefficients and absolute centre of the core variable of a
// nuclear-attraction integral on particle a (everything else absorbed)
NRX_HD double nuc_core(const Sys &s, int a, int d, Poly &e, double &X, double &w, double pic)
{
    int h = s.n - 1;
    Var H;
    double fac = hub_absorb(s, d, (a == h) ? 0u : (1u << a), H, pic);
    Var A;
    int core = h;
    if (a != h) {
        p_copy(s.q[a]->f[d], A.F);
        A.w = s.q[a]->p;
        A.c = 0.0;
        fac *= absorb_leaf(H.F, H.w, H.c, s.gam[a], s.q[h]->P[d] - s.q[a]->P[d], A, pic);
        core = a;
    } else {
        A = H;
    }
    Poly sh;
    p_compose(A.F, A.c, 1.0, sh);
    p_hermite(sh, A.w, e);
    X = s.q[core]->P[d] + A.c;
    w = A.w;
    return fac;
}

// sum_c zan[c] <1/|r_a - C_c|>, each centre's value flushed below 1e-100 when
// flush is set (the reference applies underflow() per centre)
NRX_HD double nuc_sum(const Sys &s, int a, int nat, const double *cat, const double *zan, bool flush, double pic = PI)
{
    Poly e[3];
    double X[3], w = 1.0, fac = Kprod(s);
    for (int d = 0; d < 3; ++d) fac *= nuc_core(s, a, d, e[d], X[d], w, pic);
    double pre = fac * 2.0 * pic / w;
    double tot = 0.0;
    for (int c = 0; c < nat; ++c) {
        Poly P3[3];
        double D[3];
        for (int d = 0; d < 3; ++d) {
            D[d] = X[d] - cat[3 * c + d];
            hermite_coulomb(e[d], D[d], 0.0, w, P3[d]);
        }
        double T = w * (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
        double v = pre * boys_combine(P3, T, s.bt);
        if (flush) underflow(v);
        tot += zan[c] * v;
    }
    return tot;
}

NRX_HD double nuc(const Sys &s, int a, const double *C, double pic = PI)
{
    double z = 1.0;
    return nuc_sum(s, a, 1, C, &z, false, pic);
}

// Hermite convolution of the two Coulomb sides: hk[t+u] += ea[t] eb[u] (-1)^u
NRX_HD void conv_add(const Poly &ea, const Poly &eb, Poly &hk)
{
    int n = ea.n + eb.n - 1;
    for (int k = hk.n; k < n; ++k) hk.c[k] = 0.0;
    if (n > hk.n) hk.n = n;
    for (int t = 0; t < ea.n; ++t)
        for (int u = 0; u < eb.n; ++u) hk.c[t + u] += ea.c[t] * ((u & 1) ? -eb.c[u] : eb.c[u]);
    NRX_CNT(2 * ea.n * eb.n);
}

// Coulomb <1/|r_a - r_b|>, a < b
NRX_HD double coul(const Sys &s, int a, int b, double pic = PI)
{
    int h = s.n - 1;
    Poly P3[3];
    double Dv[3], fac = Kprod(s);
    double wa = 1.0, wb = 1.0, gab = 0.0;
    for (int d = 0; d < 3; ++d) {
        unsigned skip = (1u << a) | ((b < h) ? (1u << b) : 0u);
        Var H;
        fac *= hub_absorb(s, d, skip, H, pic);
        Var A, B;
        p_copy(s.q[a]->f[d], A.F);
        A.w = s.q[a]->p;
        A.c = 0.0;
        Poly hk;
        hk.n = 1;
        hk.c[0] = 0.0;
        if (b == h) {
            gab = s.gam[a];
            B = H;
        } else {
            p_copy(s.q[b]->f[d], B.F);
            B.w = s.q[b]->p;
            B.c = 0.0;
            double ga = s.gam[a], gb = s.gam[b];
            if (ga == 0.0 || gb == 0.0) {
                gab = 0.0;
                if (ga != 0.0)
                    fac *= absorb_leaf(H.F, H.w, H.c, ga, s.q[h]->P[d] - s.q[a]->P[d], A, pic);
                else
                    fac *= absorb_leaf(H.F, H.w, H.c, gb, s.q[h]->P[d] - s.q[b]->P[d], B, pic);
            } else {
                // hub linked to both: integrate the hub conditionally on (y_a, y_b)
                double W = H.w + ga + gb;
                double dpa = s.q[h]->P[d] - s.q[a]->P[d], dpb = s.q[h]->P[d] - s.q[b]->P[d];
                fac *= sqrt(pic / W);
                double ka = H.w * ga / W, kb = H.w * gb / W;
                double ca = dpa + H.c, cb = dpb + H.c;
                {
                    double Wn = A.w + ka, df = A.c - ca;
                    fac *= exp(-(A.w * ka / Wn) * df * df);
                    A.c = (A.w * A.c + ka * ca) / Wn;
                    A.w = Wn;
                }
                {
                    double Wn = B.w + kb, df = B.c - cb;
                    fac *= exp(-(B.w * kb / Wn) * df * df);
                    B.c = (B.w * B.c + kb * cb) / Wn;
                    B.w = Wn;
                }
                gab = ga * gb / W;
                // h(m0 + ba ya + bb yb) = sum_j (bb yb)^j / j! h^{(j)}(m0 + ba ya)
                Poly hs;
                p_smooth(H.F, 0.5 / W, hs);
                double m0 = (H.w * H.c - ga * dpa - gb * dpb) / W, ba = ga / W, bb = gb / W;
                double bbj = 1.0;
                int hn = hs.n;
                for (int j = 0; j < hn; ++j) {
                    if (j > 0) {
                        for (int k = 0; k + 1 < hn - (j - 1); ++k) hs.c[k] = hs.c[k + 1] * (k + 1) / (double)j;
                        bbj *= bb;
                    }
                    Poly der;
                    der.n = hn - j;
                    for (int k = 0; k < der.n; ++k) der.c[k] = hs.c[k];
                    Poly ga_, Fa, Fb, ea, eb;
                    p_compose(der, m0, ba, ga_);
                    p_mul(A.F, ga_, Fa);
                    Fb.n = B.F.n + j;
                    for (int k = 0; k < j; ++k) Fb.c[k] = 0.0;
                    for (int k = 0; k < B.F.n; ++k) Fb.c[k + j] = B.F.c[k] * bbj;
                    p_compose(Fa, A.c, 1.0, Fa);
                    p_compose(Fb, B.c, 1.0, Fb);
                    p_hermite(Fa, A.w, ea);
                    p_hermite(Fb, B.w, eb);
                    conv_add(ea, eb, hk);
                }
                goto have_hk;
            }
        }
        {
            Poly sa, sb, ea, eb;
            p_compose(A.F, A.c, 1.0, sa);
            p_compose(B.F, B.c, 1.0, sb);
            p_hermite(sa, A.w, ea);
            p_hermite(sb, B.w, eb);
            conv_add(ea, eb, hk);
        }
    have_hk:
        wa = A.w;
        wb = B.w;
        double rho = wa * wb / (wa + wb);
        double l0 = rho * gab / (rho + gab), l1 = rho * rho / (rho + gab);
        Dv[d] = (s.q[a]->P[d] + A.c) - (s.q[b]->P[d] + B.c);
        hermite_coulomb(hk, Dv[d], l0, l1, P3[d]);
    }
    double rho = wa * wb / (wa + wb);
    double l0 = rho * gab / (rho + gab), l1 = rho * rho / (rho + gab);
    double D2 = Dv[0] * Dv[0] + Dv[1] * Dv[1] + Dv[2] * Dv[2];
    double pre = pic / (wa + wb);
    pre = pre * sqrt(pre) * 2.0 * pic / (rho + gab);
    return fac * pre * exp(-l0 * D2) * boys_combine(P3, l1 * D2, s.bt);
}

// ------------------------------------------------------------------
// Term assembly of one primitive tuple
// ------------------------------------------------------------------
struct SysParams {
    int nat, ngtg;
    double pmass;
    const double *cat, *zan, *bcoef, *gamma;
    const double *bt; // Boys table
};

NRX_HD Sys mksys(const SysParams &S, int n, const Pair *p0, const Pair *p1, const Pair *p2, const Pair *p3, double g0,
                 double g1, double g2)
{
    Sys s;
    s.n = n;
    s.q[0] = p0;
    s.q[1] = p1;
    s.q[2] = p2;
    s.q[3] = p3;
    s.gam[0] = g0;
    s.gam[1] = g1;
    s.gam[2] = g2;
    s.gam[3] = 0.0;
    s.bt = S.bt;
    return s;
}

// xcalc_GAM1_MD (gam_1_OMP.f:398-839).  q[0] = electron, q[1] = proton.
// Terms that depend on gamma_k + gamma_l only are summed over k <= l.
NRX_HD void gam1_tuple(const SysParams &S, const Pair *q, double &ansE, double &ansS)
{
    const Pair *e = &q[0], *p = &q[1];
    double tegx = 0, tpgx = 0, texg = 0, tpxg = 0, tegg = 0, tpgg = 0;
    double vcegx = 0, vcegg = 0, vcpgx = 0, vcpgg = 0, vmgx = 0, vmgg = 0;
    double ogx = 0, ogg = 0;
    for (int k = 0; k < S.ngtg; ++k) {
        double gv = S.gamma[k], bv = S.bcoef[k];
        Sys s1 = mksys(S, 2, e, p, 0, 0, gv, 0, 0);  // (e, p): hub = p
        Sys s1r = mksys(S, 2, p, e, 0, 0, gv, 0, 0); // (p, e): hub = e
        double gks[2] = {0.0, gv}, kk[2];
        kin_multi(s1, 0, 1, gks, 2, 1.0, PI_G2, kk);
        tegx += bv * kk[0];
        texg += bv * kk[1];
        kin_multi(s1r, 0, 1, gks, 2, S.pmass, PI_G2, kk);
        tpgx += bv * kk[0];
        tpxg += bv * kk[1];
        vcegx += bv * nuc_sum(s1, 0, S.nat, S.cat, S.zan, false);
        vcpgx += bv * nuc_sum(s1, 1, S.nat, S.cat, S.zan, false);
        vmgx += bv * coul(s1, 0, 1);
        ogx += bv * ovl(s1);
        for (int l = 0; l < S.ngtg; ++l) {
            double gw = S.gamma[l], bb = bv * S.bcoef[l];
            Sys s2 = mksys(S, 2, e, p, 0, 0, gv + gw, 0, 0);
            Sys s2r = mksys(S, 2, p, e, 0, 0, gv + gw, 0, 0);
            tegg += bb * kin(s2, 0, 1, gw, 1.0, PI_G2);
            tpgg += bb * kin(s2r, 0, 1, gw, S.pmass, PI_G2);
            if (l < k) continue;
            double w = (l == k) ? bb : 2.0 * bb;
            vcegg += w * nuc_sum(s2, 0, S.nat, S.cat, S.zan, false);
            vcpgg += w * nuc_sum(s2, 1, S.nat, S.cat, S.zan, false);
            vmgg += w * coul(s2, 0, 1);
            ogg += w * ovl(s2);
        }
    }
    double te = tegx + texg + tegg, tp = tpgx + tpxg + tpgg;
    double vce = 2.0 * vcegx + vcegg, vcp = 2.0 * vcpgx + vcpgg, vmx = 2.0 * vmgx + vmgg;
    ansE = te + tp - vce + vcp - vmx;
    ansS = 2.0 * ogx + ogg;
}

// xcalc_GAM2_MD (gam_2_OMP.f:546-1077).  q = (e1, e2, p).
NRX_HD void gam2_tuple(const SysParams &S, const Pair *q, double &ansE, double &ansS)
{
    const Pair *e1 = &q[0], *e2 = &q[1], *p = &q[2];
    double gTE = 0, gVEC = 0, gVEE = 0, gVEP = 0;
    for (int ik = 0; ik < S.ngtg; ++ik) {
        double gk = S.gamma[ik], bk = S.bcoef[ik];
        Sys s = mksys(S, 3, e1, e2, p, 0, 0.0, gk, 0); // g(e2,p)
        double xgTE = kin(s, 0, -1, 0.0, 1.0);
        double xgVEC = -nuc_sum(s, 0, S.nat, S.cat, S.zan, true);
        double xgVEP = coul(s, 0, 2);
        underflow(xgVEP);
        xgVEP = -xgVEP;
        double xgVEE = coul(s, 0, 1);
        underflow(xgVEE);
        gTE += bk * xgTE;
        gVEC += bk * xgVEC;
        gVEP += bk * xgVEP;
        gVEE += bk * xgVEE;
    }
    double gTPg = 0, gTEg1 = 0, gTEg2 = 0, gTEg3 = 0, gVPCg = 0, gVECg1 = 0, gVECg2 = 0;
    double gVEPg1 = 0, gVEPg2 = 0, gVEEg1 = 0, gVEEg2 = 0, gsg = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            double gk = S.gamma[ik], gl = S.gamma[il];
            double bb = S.bcoef[ik] * S.bcoef[il];
            // sb: g(e1,p)=gk, g(e2,p)=gl.  The reference's xgTEg2 of the pair
            // (il,ik) (bra g(e2,p)=gl, ket g(e1,p)=gk) is the kinetic operator
            // on this configuration with ket partner exponent gk.
            Sys sb = mksys(S, 3, e1, e2, p, 0, gk, gl, 0);
            double gks[2] = {0.0, gk}, kk[2];
            kin_multi(sb, 0, 2, gks, 2, 1.0, PI, kk);
            gTEg1 += bb * kk[0];
            gTEg2 += bb * kk[1];
            gTPg += bb * kin(sb, 2, 1, gl, S.pmass);
            gVPCg += bb * nuc_sum(sb, 2, S.nat, S.cat, S.zan, true);
            gVECg1 -= bb * nuc_sum(sb, 0, S.nat, S.cat, S.zan, true);
            double v = coul(sb, 0, 2);
            underflow(v);
            gVEPg1 -= bb * v;
            v = coul(sb, 0, 1);
            underflow(v);
            gVEEg2 += bb * v;
            gsg += bb * ovl(sb);
            if (il < ik) continue;
            double w = (il == ik) ? bb : 2.0 * bb;
            Sys s2 = mksys(S, 3, e1, e2, p, 0, 0.0, gk + gl, 0); // g(e2,p)=gk+gl
            Sys s1 = mksys(S, 3, e1, e2, p, 0, gk + gl, 0.0, 0); // g(e1,p)=gk+gl
            gTEg3 += w * kin(s2, 0, -1, 0.0, 1.0);
            gVECg2 -= w * nuc_sum(s2, 0, S.nat, S.cat, S.zan, true);
            v = coul(s2, 0, 2);
            underflow(v);
            gVEPg2 -= w * v;
            v = coul(s1, 0, 1);
            underflow(v);
            gVEEg1 += w * v;
        }
    ansE = 2.0 * (gTE + gVEC + gVEP) + 4.0 * gVEE * 0.5 + gTPg + gVPCg + gTEg1 + gTEg3 + 2.0 * (gVEPg1 + gVECg1) +
           gTEg2 + gVEPg2 + gVECg2 + 2.0 * gVEEg1 * 0.5 + 2.0 * gVEEg2 * 0.5;
    ansS = gsg;
}

// xcalc_GAM3_MD (gam_3_OMP.f:617-983).  q = (e1, e2, e3, p).
NRX_HD void gam3_tuple(const SysParams &S, const Pair *q, double &ans)
{
    const Pair *e1 = &q[0], *e2 = &q[1], *e3 = &q[2], *p = &q[3];
    // geminal-independent factors
    Sys s23 = mksys(S, 2, e2, e3, 0, 0, 0.0, 0, 0);
    double vee23 = coul(s23, 0, 1);
    underflow(vee23);
    Sys s3 = mksys(S, 1, e3, 0, 0, 0, 0, 0, 0);
    double hcore = kin(s3, 0, -1, 0.0, 1.0) - nuc_sum(s3, 0, S.nat, S.cat, S.zan, false);
    underflow(hcore);
    double gVEE = 0.0;
    for (int ik = 0; ik < S.ngtg; ++ik) {
        Sys s = mksys(S, 2, e1, p, 0, 0, S.gamma[ik], 0, 0);
        double xx = ovl(s);
        underflow(xx);
        gVEE += S.bcoef[ik] * (xx * vee23);
    }
    double gHEg = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            double gk = S.gamma[ik], gl = S.gamma[il];
            double bb = S.bcoef[ik] * S.bcoef[il];
            Sys s3p = mksys(S, 3, e1, e2, p, 0, gk, gl, 0);
            double yy = ovl(s3p);
            Sys s4a = mksys(S, 4, e1, e2, e3, p, gk, gl, 0.0); // 1/r34, g14 g24
            double zz = coul(s4a, 2, 3);
            underflow(yy);
            underflow(zz);
            gHEg += bb * ((hcore * yy) + (zz * (-1.0)));
            Sys s4b = mksys(S, 4, e1, e2, e3, p, gk, 0.0, gl); // 1/r12, g14 g34
            gVEEg2 += bb * coul(s4b, 0, 1);
            if (il < ik) continue;
            double w = (il == ik) ? bb : 2.0 * bb;
            Sys s2 = mksys(S, 2, e1, p, 0, 0, gk + gl, 0, 0);
            double xx = ovl(s2);
            underflow(xx);
            gVEEg1 += w * (xx * vee23);
        }
    ans = 2.0 * gVEE * 0.5 + gHEg + gVEEg1 * 0.5 + 4.0 * gVEEg2 * 0.5;
}

} // namespace nrx
