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
//     the kernel (lambda0 != 0).  The only time the two-sided elimination
//     (hub linked to both Coulomb particles) appears, the hub's conditional
//     expectation is expanded as a short sum of separable products.
//   * F_n uses the reference's two-branch Boys semantics (asymptotic form for
//     T > 20, truncated series with 1e-12 cut otherwise), see
//     AC_Integrals/common.f:317-341 of the reference.
//
// The kinetic operator is the second derivative of the ket along each
// direction, expanded into overlaps with modified polynomials of the particle
// and of its ket-geminal partner (same operator as G3_MD_T1d,
// CS_Integrals/XC1_G3_MD.f:2680-2868, and pgi_Lapx1, AC_Integrals/g2.f:655).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NRX_HD __host__ __device__ inline
#else
#include <cmath>
#define NRX_HD inline
#endif

namespace nrx {

// polynomial capacity: degree <= NP-1.  Four particles with d functions on
// both sides give per-direction degree <= 16 (+2 for the kinetic operator).
constexpr int NP = 20;
constexpr int NPART = 4;
constexpr double PI = 3.141592653589793238462643383279502884;
// pgike1 (AC_Integrals/g2.f:216) normalises its overlaps with this value
constexpr double PI_G2 = 3.141592650;

struct Poly {
    double c[NP];
    int n; // number of coefficients (degree + 1)
};

NRX_HD void p_zero(Poly &p) { p.n = 1; p.c[0] = 0.0; }
NRX_HD void p_const(Poly &p, double v) { p.n = 1; p.c[0] = v; }

// r = a * b (truncated to NP coefficients; callers size degrees to fit)
NRX_HD void p_mul(const Poly &a, const Poly &b, Poly &r)
{
    int n = a.n + b.n - 1;
    if (n > NP) n = NP;
    for (int k = 0; k < n; ++k) r.c[k] = 0.0;
    for (int i = 0; i < a.n; ++i) {
        double ai = a.c[i];
        if (ai == 0.0) continue;
        int jm = b.n;
        if (i + jm > n) jm = n - i;
        for (int j = 0; j < jm; ++j) r.c[i + j] += ai * b.c[j];
    }
    r.n = n;
}

// a *= (y + s)
NRX_HD void p_mul_lin(Poly &a, double s)
{
    int n = a.n;
    if (n < NP) {
        a.c[n] = a.c[n - 1];
        for (int k = n - 1; k > 0; --k) a.c[k] = a.c[k - 1] + s * a.c[k];
        a.c[0] = s * a.c[0];
        a.n = n + 1;
    }
}

// (y + sa)^la (y + sb)^lb
NRX_HD void p_pair(int la, double sa, int lb, double sb, Poly &r)
{
    p_const(r, 1.0);
    for (int i = 0; i < la; ++i) p_mul_lin(r, sa);
    for (int i = 0; i < lb; ++i) p_mul_lin(r, sb);
}

// h(s) = E[f(s + e)], e ~ N(0, var)
NRX_HD void p_smooth(const Poly &f, double var, Poly &h)
{
    // moments of e: m[2j] = (2j-1)!! var^j
    double mom[NP];
    mom[0] = 1.0;
    if (NP > 1) mom[1] = 0.0;
    for (int k = 2; k < f.n; ++k) mom[k] = (k & 1) ? 0.0 : mom[k - 2] * (k - 1) * var;
    for (int j = 0; j < f.n; ++j) {
        double acc = 0.0, bin = 1.0; // bin = C(k, j) starting at k = j
        for (int k = j; k < f.n; k += 1) {
            if (((k - j) & 1) == 0) acc += f.c[k] * bin * mom[k - j];
            bin = bin * (k + 1) / (k + 1 - j);
        }
        h.c[j] = acc;
    }
    h.n = f.n;
}

// evaluate h at s
NRX_HD double p_eval(const Poly &h, double s)
{
    double v = 0.0;
    for (int k = h.n - 1; k >= 0; --k) v = v * s + h.c[k];
    return v;
}

// g(y) = h(m0 + beta y)
NRX_HD void p_compose(const Poly &h, double m0, double beta, Poly &g)
{
    // Taylor shift to m0 (Horner synthetic division), then scale
    Poly t = h;
    for (int i = 0; i < t.n; ++i)
        for (int k = t.n - 2; k >= i; --k) t.c[k] += m0 * t.c[k + 1];
    double bp = 1.0;
    for (int k = 0; k < t.n; ++k) {
        g.c[k] = t.c[k] * bp;
        bp *= beta;
    }
    g.n = t.n;
}

// Hermite-Gaussian expansion: f(s) exp(-w s^2) = sum_t e_t Lambda_t(s),
// Lambda_t = d^t/dc^t exp(-w (s - c)^2) at c = 0;  s Lambda_t = Lambda_{t+1}/(2w) + t Lambda_{t-1}
NRX_HD void p_hermite(const Poly &f, double w, Poly &e)
{
    double i2w = 0.5 / w;
    int n = f.n;
    e.n = 1;
    e.c[0] = f.c[n - 1];
    for (int k = n - 2; k >= 0; --k) {
        // e <- s*e + f_k
        int m = e.n;
        double nw[NP];
        for (int t = 0; t <= m; ++t) {
            double v = 0.0;
            if (t >= 1) v += e.c[t - 1] * i2w;
            if (t + 1 < m) v += (t + 1) * e.c[t + 1];
            nw[t] = v;
        }
        for (int t = 0; t <= m; ++t) e.c[t] = nw[t];
        e.n = m + 1;
        e.c[0] += f.c[k];
    }
}

// ------------------------------------------------------------------
// Boys function, reference semantics (XGAMMAF): F[0..M]
// ------------------------------------------------------------------
NRX_HD void boys_ref(int M, double T, double *F)
{
    const double SQRPI = 0.88622692545275801365;
    if (T > 20.0) {
        F[0] = SQRPI / sqrt(T);
        for (int l = 1; l <= M; ++l) F[l] = F[l - 1] * (1.0 * l - 0.5) / T;
    } else {
        double pt5pm = 0.5 + M;
        double expti = exp(-T), ti2 = 2.0 * T;
        double term = 0.5 / pt5pm, sumk = term;
        for (int k = 1; k <= 1000; ++k) {
            term = term * T / (pt5pm + k);
            sumk += term;
            if (term < 1.0e-12) break;
        }
        F[M] = expti * sumk;
        for (int l = M - 1; l >= 0; --l) F[l] = (F[l + 1] * ti2 + expti) / (2.0 * l + 1.0);
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
    Poly f[3];         // per-direction polynomial in y = x - P
    double beta, B[3], sA[3], sB[3];
    int la[3], lb[3];
};

NRX_HD void make_pair(const Prim &bra, const Prim &ket, Pair &q)
{
    double a = bra.a, b = ket.a, p = a + b;
    q.p = p;
    q.beta = b;
    double K = 1.0;
    for (int d = 0; d < 3; ++d) {
        double P = (a * bra.A[d] + b * ket.A[d]) / p;
        double AB = bra.A[d] - ket.A[d];
        K *= exp(-a * b / p * AB * AB);
        q.P[d] = P;
        q.B[d] = ket.A[d];
        q.sA[d] = P - bra.A[d];
        q.sB[d] = P - ket.A[d];
        q.la[d] = bra.l[d];
        q.lb[d] = ket.l[d];
        p_pair(bra.l[d], q.sA[d], ket.l[d], q.sB[d], q.f[d]);
    }
    q.K = K;
}

// ------------------------------------------------------------------
// Star elimination per direction
// ------------------------------------------------------------------
// One variable being reduced: polynomial F in the particle coordinate y and
// Gaussian exp(-w (y - c)^2).
struct Var {
    Poly F;
    double w, c;
};

// Integrate leaf L (edge g, d = P_leaf - P_parent) into parent Q; returns the
// scalar factor sqrt(pi/(wL+g)) * exp(...).
NRX_HD double absorb_leaf(const Var &L, double g, double d, Var &Q, double pic)
{
    double W = L.w + g;
    Poly h;
    p_smooth(L.F, 0.5 / W, h);
    double fac = sqrt(pic / W);
    if (g == 0.0) {
        double v = p_eval(h, L.c);
        Poly t = Q.F;
        for (int k = 0; k < t.n; ++k) t.c[k] *= v;
        Q.F = t;
        return fac;
    }
    double beta = g / W, m0 = (L.w * L.c - g * d) / W;
    Poly gq, r;
    p_compose(h, m0, beta, gq);
    p_mul(Q.F, gq, r);
    Q.F = r;
    double kap = L.w * g / W;
    double cl = d + L.c;
    double Wq = Q.w + kap;
    double diff = Q.c - cl;
    fac *= exp(-(Q.w * kap / Wq) * diff * diff);
    Q.c = (Q.w * Q.c + kap * cl) / Wq;
    Q.w = Wq;
    return fac;
}

// integrate a single remaining variable
NRX_HD double close_var(const Var &V, double pic)
{
    Poly h;
    p_smooth(V.F, 0.5 / V.w, h);
    return sqrt(pic / V.w) * p_eval(h, V.c);
}

// Integral description: particles 0..n-1, hub = n-1, electron i linked to the
// hub by gam[i] (bra+ket).
struct Sys {
    int n;
    const Pair *q[NPART];
    double gam[NPART];
};

// 1-D overlap along direction d with the given per-particle polynomials
NRX_HD double ovl_dim(const Sys &s, int d, const Poly *const *F, double pic)
{
    int h = s.n - 1;
    Var H;
    H.F = *F[h];
    H.w = s.q[h]->p;
    H.c = 0.0;
    double fac = 1.0;
    for (int i = 0; i < h; ++i) {
        Var L;
        L.F = *F[i];
        L.w = s.q[i]->p;
        L.c = 0.0;
        fac *= absorb_leaf(L, s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], H, pic);
    }
    return fac * close_var(H, pic);
}

NRX_HD double Kprod(const Sys &s)
{
    double k = 1.0;
    for (int i = 0; i < s.n; ++i) k *= s.q[i]->K;
    return k;
}

NRX_HD double ovl(const Sys &s, double pic = PI)
{
    double r = Kprod(s);
    for (int d = 0; d < 3; ++d) {
        const Poly *F[NPART];
        for (int i = 0; i < s.n; ++i) F[i] = &s.q[i]->f[d];
        r *= ovl_dim(s, d, F, pic);
    }
    return r;
}

// Kinetic energy of particle a (-1/(2m) <bra| d^2/dx_a^2 |ket>), ket geminal
// gk between a and partner c (c < 0: none).
NRX_HD double kin(const Sys &s, int a, int c, double gk, double mass, double pic = PI)
{
    double S[3], T[3];
    const Pair &A = *s.q[a];
    for (int d = 0; d < 3; ++d) {
        const Poly *F[NPART];
        for (int i = 0; i < s.n; ++i) F[i] = &s.q[i]->f[d];
        S[d] = ovl_dim(s, d, F, pic);
        int L = A.lb[d];
        double al = A.beta + gk;
        double dl = (c >= 0) ? gk * (A.B[d] - s.q[c]->B[d]) : 0.0;
        // u_A(y) (y + sB)^m for the needed m, combined with coefficients
        // F0 = L(L-1)X^{L-2} - 4L al X^L - 4L dl X^{L-1} + 4al^2 X^{L+2} + 4dl^2 X^L + 8 al dl X^{L+1} - 2 al X^L
        // F1 = 4 L gk X^{L-1} - 8 al gk X^{L+1} - 8 dl gk X^L
        // F2 = 4 gk^2 X^L
        double cf0[5], cf1[5], cf2[5]; // coefficients of X^{L-2+j}, j=0..4
        for (int j = 0; j < 5; ++j) cf0[j] = cf1[j] = cf2[j] = 0.0;
        cf0[0] = (double)(L * (L - 1));
        cf0[1] = -4.0 * L * dl;
        cf0[2] = -4.0 * L * al + 4.0 * dl * dl - 2.0 * al;
        cf0[3] = 8.0 * al * dl;
        cf0[4] = 4.0 * al * al;
        cf1[1] = 4.0 * L * gk;
        cf1[2] = -8.0 * dl * gk;
        cf1[3] = -8.0 * al * gk;
        cf2[2] = 4.0 * gk * gk;
        Poly F0, F1, F2;
        // build sum_j cf[j] X^{L-2+j} * u_A
        Poly acc0, acc1, acc2;
        for (int k = 0; k < NP; ++k) { acc0.c[k] = acc1.c[k] = acc2.c[k] = 0.0; }
        acc0.n = acc1.n = acc2.n = 1;
        for (int j = 0; j < 5; ++j) {
            int m = L - 2 + j;
            if (m < 0) continue;
            if (cf0[j] == 0.0 && cf1[j] == 0.0 && cf2[j] == 0.0) continue;
            Poly xm;
            p_pair(A.la[d], A.sA[d], m, A.sB[d], xm);
            for (int k = 0; k < xm.n; ++k) {
                acc0.c[k] += cf0[j] * xm.c[k];
                acc1.c[k] += cf1[j] * xm.c[k];
                acc2.c[k] += cf2[j] * xm.c[k];
            }
            if (xm.n > acc0.n) acc0.n = acc1.n = acc2.n = xm.n;
        }
        F0 = acc0;
        F1 = acc1;
        F2 = acc2;
        F[a] = &F0;
        double t = ovl_dim(s, d, F, pic);
        if (c >= 0 && gk != 0.0) {
            Poly fc1 = s.q[c]->f[d];
            double xc = s.q[c]->P[d] - s.q[c]->B[d];
            p_mul_lin(fc1, xc);
            Poly fc2 = fc1;
            p_mul_lin(fc2, xc);
            F[a] = &F1;
            F[c] = &fc1;
            t += ovl_dim(s, d, F, pic);
            F[a] = &F2;
            F[c] = &fc2;
            t += ovl_dim(s, d, F, pic);
        }
        T[d] = t;
    }
    double r = T[0] * S[1] * S[2] + S[0] * T[1] * S[2] + S[0] * S[1] * T[2];
    return -0.5 * r / mass * Kprod(s);
}

// accumulate into P(u^2) the Hermite-Coulomb combination sum_k hk R_k(u^2),
// R_0 = 1, R_1 = -2 lam D, R_{k+1} = -2 lam (D R_k + k R_{k-1}),
// lam = l0 + l1 u^2.
NRX_HD void hermite_coulomb(const Poly &hk, double D, double l0, double l1, Poly &P)
{
    Poly Rm, R, Rn;
    p_const(R, 1.0);
    Rm.n = 1;
    Rm.c[0] = 0.0;
    P.n = 1;
    P.c[0] = hk.c[0];
    for (int k = 0; k + 1 < hk.n; ++k) {
        // Rn = -2 (l0 + l1 u) (D R + k Rm)
        int m = R.n;
        double tmp[NP];
        for (int j = 0; j < m; ++j) tmp[j] = D * R.c[j] + ((j < Rm.n) ? k * Rm.c[j] : 0.0);
        Rn.n = m + 1;
        for (int j = 0; j <= m; ++j) {
            double v = 0.0;
            if (j < m) v += l0 * tmp[j];
            if (j >= 1) v += l1 * tmp[j - 1];
            Rn.c[j] = -2.0 * v;
        }
        Rm = R;
        R = Rn;
        double h = hk.c[k + 1];
        if (h != 0.0) {
            for (int j = P.n; j < R.n; ++j) P.c[j] = 0.0;
            if (R.n > P.n) P.n = R.n;
            for (int j = 0; j < R.n; ++j) P.c[j] += h * R.c[j];
        }
    }
}

NRX_HD double boys_combine(const Poly *P3, double T)
{
    Poly xy, xyz;
    p_mul(P3[0], P3[1], xy);
    p_mul(xy, P3[2], xyz);
    double F[3 * NP];
    int M = xyz.n - 1;
    boys_ref(M, T, F);
    double s = 0.0;
    for (int k = 0; k <= M; ++k) s += xyz.c[k] * F[k];
    return s;
}

// nuclear attraction <1/|r_a - C|>
NRX_HD double nuc(const Sys &s, int a, const double *C, double pic = PI)
{
    int h = s.n - 1;
    Poly P3[3];
    double Dv[3], fac = Kprod(s), wc = 0.0;
    for (int d = 0; d < 3; ++d) {
        Var V[NPART];
        for (int i = 0; i < s.n; ++i) {
            V[i].F = s.q[i]->f[d];
            V[i].w = s.q[i]->p;
            V[i].c = 0.0;
        }
        for (int i = 0; i < h; ++i)
            if (i != a) fac *= absorb_leaf(V[i], s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], V[h], pic);
        int core = h;
        if (a != h) {
            fac *= absorb_leaf(V[h], s.gam[a], s.q[h]->P[d] - s.q[a]->P[d], V[a], pic);
            core = a;
        }
        Var &X = V[core];
        Poly sh, e;
        p_compose(X.F, X.c, 1.0, sh);
        p_hermite(sh, X.w, e);
        Dv[d] = s.q[core]->P[d] + X.c - C[d];
        wc = X.w;
        hermite_coulomb(e, Dv[d], 0.0, X.w, P3[d]);
    }
    double T = wc * (Dv[0] * Dv[0] + Dv[1] * Dv[1] + Dv[2] * Dv[2]);
    return fac * (2.0 * pic / wc) * boys_combine(P3, T);
}

// Coulomb <1/|r_a - r_b|>, a < b
NRX_HD double coul(const Sys &s, int a, int b, double pic = PI)
{
    int h = s.n - 1;
    Poly P3[3];
    double Dv[3], fac = Kprod(s);
    double wa = 0, wb = 0, gab = 0.0;
    for (int d = 0; d < 3; ++d) {
        Var V[NPART];
        for (int i = 0; i < s.n; ++i) {
            V[i].F = s.q[i]->f[d];
            V[i].w = s.q[i]->p;
            V[i].c = 0.0;
        }
        for (int i = 0; i < h; ++i)
            if (i != a && i != b) fac *= absorb_leaf(V[i], s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], V[h], pic);
        Poly hk;
        hk.n = 1;
        hk.c[0] = 0.0;
        if (b == h) {
            gab = s.gam[a];
        } else {
            double ga = s.gam[a], gb = s.gam[b];
            if (ga == 0.0 || gb == 0.0) {
                gab = 0.0;
                int tgt = (ga != 0.0) ? a : b;
                fac *= absorb_leaf(V[h], (ga != 0.0) ? ga : gb, s.q[h]->P[d] - s.q[tgt]->P[d], V[tgt], pic);
            } else {
                // hub linked to both: integrate the hub conditionally on (y_a, y_b)
                Var &Hv = V[h];
                double W = Hv.w + ga + gb;
                double dpa = s.q[h]->P[d] - s.q[a]->P[d], dpb = s.q[h]->P[d] - s.q[b]->P[d];
                fac *= sqrt(pic / W);
                // induced Gaussians on a and b
                double ka = Hv.w * ga / W, kb = Hv.w * gb / W;
                double ca = dpa + Hv.c, cb = dpb + Hv.c;
                {
                    double Wn = V[a].w + ka, df = V[a].c - ca;
                    fac *= exp(-(V[a].w * ka / Wn) * df * df);
                    V[a].c = (V[a].w * V[a].c + ka * ca) / Wn;
                    V[a].w = Wn;
                }
                {
                    double Wn = V[b].w + kb, df = V[b].c - cb;
                    fac *= exp(-(V[b].w * kb / Wn) * df * df);
                    V[b].c = (V[b].w * V[b].c + kb * cb) / Wn;
                    V[b].w = Wn;
                }
                gab = ga * gb / W;
                // conditional expectation of the hub polynomial:
                // h(m0 + ba ya + bb yb) = sum_j (bb yb)^j / j! h^{(j)}(m0 + ba ya)
                Poly hs;
                p_smooth(Hv.F, 0.5 / W, hs);
                double m0 = (Hv.w * Hv.c - ga * dpa - gb * dpb) / W, ba = ga / W, bb = gb / W;
                Poly der = hs; // h^{(j)}/j!
                double bbj = 1.0;
                // Gaussian centres/weights of the core are final now
                for (int j = 0; j < hs.n; ++j) {
                    if (j > 0) {
                        // der <- derivative / j
                        for (int k = 0; k + 1 < der.n; ++k) der.c[k] = der.c[k + 1] * (k + 1) / (double)j;
                        der.n = der.n - 1;
                        if (der.n < 1) break;
                        bbj *= bb;
                    }
                    Poly ga_, Fa, Fb, sa, sb, ea, eb;
                    p_compose(der, m0, ba, ga_);
                    p_mul(V[a].F, ga_, Fa);
                    Fb = V[b].F;
                    for (int k = 0; k < j; ++k) p_mul_lin(Fb, 0.0);
                    for (int k = 0; k < Fb.n; ++k) Fb.c[k] *= bbj;
                    p_compose(Fa, V[a].c, 1.0, sa);
                    p_compose(Fb, V[b].c, 1.0, sb);
                    p_hermite(sa, V[a].w, ea);
                    p_hermite(sb, V[b].w, eb);
                    int n = ea.n + eb.n - 1;
                    for (int k = hk.n; k < n; ++k) hk.c[k] = 0.0;
                    if (n > hk.n) hk.n = n;
                    for (int t = 0; t < ea.n; ++t)
                        for (int u = 0; u < eb.n; ++u) hk.c[t + u] += ea.c[t] * ((u & 1) ? -eb.c[u] : eb.c[u]);
                }
                goto have_hk;
            }
        }
        {
            Poly sa, sb, ea, eb;
            p_compose(V[a].F, V[a].c, 1.0, sa);
            p_compose(V[b].F, V[b].c, 1.0, sb);
            p_hermite(sa, V[a].w, ea);
            p_hermite(sb, V[b].w, eb);
            hk.n = ea.n + eb.n - 1;
            for (int k = 0; k < hk.n; ++k) hk.c[k] = 0.0;
            for (int t = 0; t < ea.n; ++t)
                for (int u = 0; u < eb.n; ++u) hk.c[t + u] += ea.c[t] * ((u & 1) ? -eb.c[u] : eb.c[u]);
        }
    have_hk:
        wa = V[a].w;
        wb = V[b].w;
        double rho = wa * wb / (wa + wb);
        double l0 = rho * gab / (rho + gab), l1 = rho * rho / (rho + gab);
        Dv[d] = (s.q[a]->P[d] + V[a].c) - (s.q[b]->P[d] + V[b].c);
        hermite_coulomb(hk, Dv[d], l0, l1, P3[d]);
    }
    double rho = wa * wb / (wa + wb);
    double l0 = rho * gab / (rho + gab), l1 = rho * rho / (rho + gab);
    double D2 = Dv[0] * Dv[0] + Dv[1] * Dv[1] + Dv[2] * Dv[2];
    double pre = pic / (wa + wb);
    pre = pre * sqrt(pre) * 2.0 * pic / (rho + gab);
    return fac * pre * exp(-l0 * D2) * boys_combine(P3, l1 * D2);
}

NRX_HD void underflow(double &x)
{
    if (fabs(x) < 1.0e-100) x = 0.0;
}

// ------------------------------------------------------------------
// Term assembly of one primitive tuple
// ------------------------------------------------------------------
struct SysParams {
    int nat, ngtg;
    double pmass;
    const double *cat, *zan, *bcoef, *gamma;
};

NRX_HD Sys mksys(int n, const Pair *p0, const Pair *p1, const Pair *p2, const Pair *p3, double g0, double g1,
                 double g2)
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
    return s;
}

// xcalc_GAM1_MD (gam_1_OMP.f:398-839).  q[0] = electron, q[1] = proton.
NRX_HD void gam1_tuple(const SysParams &S, const Pair *q, double &ansE, double &ansS)
{
    const Pair *e = &q[0], *p = &q[1];
    double tegx = 0, tpgx = 0, texg = 0, tpxg = 0, tegg = 0, tpgg = 0;
    double vcegx = 0, vcegg = 0, vcpgx = 0, vcpgg = 0, vmgx = 0, vmgg = 0;
    double ogx = 0, ogg = 0;
    for (int k = 0; k < S.ngtg; ++k) {
        double gv = S.gamma[k], bv = S.bcoef[k];
        Sys s1 = mksys(2, e, p, 0, 0, gv, 0, 0);   // particle order (e, p): hub = p
        Sys s1r = mksys(2, p, e, 0, 0, gv, 0, 0);  // particle order (p, e): hub = e
        tegx += bv * kin(s1, 0, -1, 0.0, 1.0, PI_G2);
        tpgx += bv * kin(s1r, 0, -1, 0.0, S.pmass, PI_G2);
        double vce = 0, vcp = 0;
        for (int c = 0; c < S.nat; ++c) {
            vce += nuc(s1, 0, S.cat + 3 * c) * S.zan[c];
            vcp += nuc(s1, 1, S.cat + 3 * c) * S.zan[c];
        }
        vcegx += bv * vce;
        vcpgx += bv * vcp;
        vmgx += bv * coul(s1, 0, 1);
        texg += bv * kin(s1, 0, 1, gv, 1.0, PI_G2);
        tpxg += bv * kin(s1r, 0, 1, gv, S.pmass, PI_G2);
        ogx += bv * ovl(s1);
        for (int l = 0; l < S.ngtg; ++l) {
            double gw = S.gamma[l], bw = S.bcoef[l];
            Sys s2 = mksys(2, e, p, 0, 0, gv + gw, 0, 0);
            Sys s2r = mksys(2, p, e, 0, 0, gv + gw, 0, 0);
            tegg += bv * bw * kin(s2, 0, 1, gw, 1.0, PI_G2);
            tpgg += bv * bw * kin(s2r, 0, 1, gw, S.pmass, PI_G2);
            double vce2 = 0, vcp2 = 0;
            for (int c = 0; c < S.nat; ++c) {
                vce2 += nuc(s2, 0, S.cat + 3 * c) * S.zan[c];
                vcp2 += nuc(s2, 1, S.cat + 3 * c) * S.zan[c];
            }
            vcegg += bv * bw * vce2;
            vcpgg += bv * bw * vcp2;
            vmgg += bv * bw * coul(s2, 0, 1);
            ogg += bv * bw * ovl(s2);
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
        Sys s = mksys(3, e1, e2, p, 0, 0.0, gk, 0); // g(e2,p)
        double xgTE = kin(s, 0, -1, 0.0, 1.0);
        double xgVEC = 0.0;
        for (int c = 0; c < S.nat; ++c) {
            double v = nuc(s, 0, S.cat + 3 * c);
            underflow(v);
            xgVEC -= S.zan[c] * v;
        }
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
            Sys sb = mksys(3, e1, e2, p, 0, gk, gl, 0);      // g(e1,p)=gk, g(e2,p)=gl
            Sys sc = mksys(3, e1, e2, p, 0, gl, gk, 0);      // g(e2,p)=gk, g(e1,p)=gl
            Sys s2 = mksys(3, e1, e2, p, 0, 0.0, gk + gl, 0); // g(e2,p)=gk+gl
            Sys s1 = mksys(3, e1, e2, p, 0, gk + gl, 0.0, 0); // g(e1,p)=gk+gl
            // kinetic of the proton: bra g(p,e1)=gk, ket g(p,e2)=gl
            Sys sp = mksys(3, e1, e2, p, 0, gk, gl, 0);
            double xgTPg = kin(sp, 2, 1, gl, S.pmass);
            double xgTEg1 = kin(sb, 0, -1, 0.0, 1.0);
            double xgTEg2 = kin(sc, 0, 2, gl, 1.0);
            double xgTEg3 = kin(s2, 0, -1, 0.0, 1.0);
            double xgVPCg = 0, xgVECg1 = 0, xgVECg2 = 0;
            for (int c = 0; c < S.nat; ++c) {
                const double *C = S.cat + 3 * c;
                double z = S.zan[c], v;
                v = nuc(sb, 2, C);
                underflow(v);
                xgVPCg += z * v;
                v = nuc(sb, 0, C);
                underflow(v);
                xgVECg1 -= z * v;
                v = nuc(s2, 0, C);
                underflow(v);
                xgVECg2 -= z * v;
            }
            double xgVEPg1 = coul(sb, 0, 2);
            underflow(xgVEPg1);
            xgVEPg1 = -xgVEPg1;
            double xgVEPg2 = coul(s2, 0, 2);
            underflow(xgVEPg2);
            xgVEPg2 = -xgVEPg2;
            double xgVEEg1 = coul(s1, 0, 1);
            underflow(xgVEEg1);
            double xgVEEg2 = coul(sb, 0, 1);
            underflow(xgVEEg2);
            double xgsg = ovl(sb);
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
    ansE = 2.0 * (gTE + gVEC + gVEP) + 4.0 * gVEE * 0.5 + gTPg + gVPCg + gTEg1 + gTEg3 + 2.0 * (gVEPg1 + gVECg1) +
           gTEg2 + gVEPg2 + gVECg2 + 2.0 * gVEEg1 * 0.5 + 2.0 * gVEEg2 * 0.5;
    ansS = gsg;
}

// xcalc_GAM3_MD (gam_3_OMP.f:617-983).  q = (e1, e2, e3, p).
NRX_HD void gam3_tuple(const SysParams &S, const Pair *q, double &ans)
{
    const Pair *e1 = &q[0], *e2 = &q[1], *e3 = &q[2], *p = &q[3];
    // geminal-independent factors
    Sys s23 = mksys(2, e2, e3, 0, 0, 0.0, 0, 0);
    double vee23 = coul(s23, 0, 1);
    underflow(vee23);
    Sys s3 = mksys(1, e3, 0, 0, 0, 0, 0, 0);
    double xke = kin(s3, 0, -1, 0.0, 1.0);
    double vc = 0.0;
    for (int c = 0; c < S.nat; ++c) vc += S.zan[c] * nuc(s3, 0, S.cat + 3 * c);
    double hcore = xke + vc * (-1.0);
    underflow(hcore);
    double gVEE = 0.0;
    for (int ik = 0; ik < S.ngtg; ++ik) {
        Sys s = mksys(2, e1, p, 0, 0, S.gamma[ik], 0, 0);
        double xx = ovl(s);
        underflow(xx);
        gVEE += S.bcoef[ik] * (xx * vee23);
    }
    double gHEg = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            double gk = S.gamma[ik], gl = S.gamma[il];
            Sys s3p = mksys(3, e1, e2, p, 0, gk, gl, 0);
            double yy = ovl(s3p);
            Sys s4a = mksys(4, e1, e2, e3, p, gk, gl, 0.0); // 1/r34, g14 g24
            double zz = coul(s4a, 2, 3);
            underflow(yy);
            underflow(zz);
            double xgHEg = (hcore * yy) + (zz * (-1.0));
            Sys s2 = mksys(2, e1, p, 0, 0, gk + gl, 0, 0);
            double xx = ovl(s2);
            underflow(xx);
            double xgVEEg1 = xx * vee23;
            Sys s4b = mksys(4, e1, e2, e3, p, gk, 0.0, gl); // 1/r12, g14 g34
            double xgVEEg2 = coul(s4b, 0, 1);
            double bb = S.bcoef[ik] * S.bcoef[il];
            gHEg += bb * xgHEg;
            gVEEg1 += bb * xgVEEg1;
            gVEEg2 += bb * xgVEEg2;
        }
    ans = 2.0 * gVEE * 0.5 + gHEg + gVEEg1 * 0.5 + 4.0 * gVEEg2 * 0.5;
}

} // namespace nrx
