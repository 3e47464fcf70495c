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
//     reference's truncated series is within ~5e-12 of the exact function)
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
//
// Register residency: every polynomial has a compile-time capacity and every
// loop over coefficients has a compile-time trip count (coefficients past the
// degree are kept at zero), so after inlining all coefficient arrays are
// indexed by constants and live in VGPRs.  The capacity NL of the absorbed
// polynomials is a template parameter: kernels are instantiated for a small
// capacity that covers almost every tuple and for a general one.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NRX_HD __host__ __device__ inline __attribute__((always_inline))
#else
#include <cmath>
#define NRX_HD inline
#endif

#define NRX_UNROLL _Pragma("unroll")
// loops over the three Cartesian directions are kept rolled so that only one
// direction's polynomials are live at a time
#define NRX_DIRS _Pragma("nounroll")

// Compiler fence between elementary integrals of one tuple: the pair records
// are re-read after it, so values of one integral are not kept live into the
// next one (they are cheap to reload and expensive to hold in registers).
#if defined(__HIP_DEVICE_COMPILE__)
#define NRX_FENCE() asm volatile("" ::: "memory")
#else
#define NRX_FENCE() ((void)0)
#endif

// Host builds used for accounting may define NRX_COUNT_FLOPS to tally the
// floating-point operations the algorithm needs (FMA = 2): each polynomial
// operation is counted over the coefficients its operands' degrees make
// nonzero, not over the zero-padded capacity the kernels execute.
#if defined(NRX_COUNT_FLOPS) && !defined(__HIP_DEVICE_COMPILE__)
extern thread_local double nrx_flops;
#define NRX_CNT(x) (nrx_flops += (double)(x))
#else
#define NRX_CNT(x) ((void)0)
#endif

// Largest angular momentum per primitive the engine is compiled for.  The
// library compiles it twice: for l <= 2 (the tier kernels of nrx_lib.hip)
// and for l <= 3 (nrx_wide.hip, the tuples that need f functions).  The
// capacity-dependent part lives in an inline namespace per build, so the two
// builds never share a symbol.
#ifndef NRX_LMAX
#define NRX_LMAX 2
#endif
#if NRX_LMAX == 2
#define NRX_CAP_NS cap2
#elif NRX_LMAX == 3
#define NRX_CAP_NS cap3
#else
#error "NRX_LMAX must be 2 or 3"
#endif

namespace nrx {

constexpr double PI = 3.141592653589793238462643383279502884;
// pgike1 (AC_Integrals/g2.f:216) normalises its overlaps with this value
constexpr double PI_G2 = 3.141592650;
// a two-particle overlap carries pi^3 (pi^(1/2) per Gaussian integration, two
// per direction): this factor turns one normalised with PI_G2 into one with PI
constexpr double G2_TO_PI3 = (PI / PI_G2) * (PI / PI_G2) * (PI / PI_G2);

template <int N> struct PolyT {
    double c[N]; // c[k] = 0 for k >= n
    int n;       // number of coefficients (degree + 1)
};

NRX_HD constexpr double binomc(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
NRX_HD constexpr double dfact(int n) // (n)!! for odd n, 1 for n <= 0
{
    double r = 1.0;
    for (int i = n; i > 1; i -= 2) r *= i;
    return r;
}
NRX_HD int imin(int a, int b) { return a < b ? a : b; }

// 1/x and 1/sqrt(x) for the engine's positive, well-scaled exponents (no
// zeros, infinities or denormals): the hardware estimate refined by two
// Newton-Raphson steps, r <- r + r (1 - x r) and y <- y + y (1 - x y^2) / 2,
// which leaves about an ulp; the host build uses the IEEE operations.
NRX_HD double rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}
NRX_HD double rsqrt(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
    y = __builtin_fma(0.5 * y, __builtin_fma(-x * y, y, 1.0), y);
    return __builtin_fma(0.5 * y, __builtin_fma(-x * y, y, 1.0), y);
#else
    return 1.0 / sqrt(x);
#endif
}

template <int N> NRX_HD void p_const(PolyT<N> &p, double v)
{
    NRX_UNROLL for (int k = 0; k < N; ++k) p.c[k] = 0.0;
    p.c[0] = v;
    p.n = 1;
}

template <int NA, int NR> NRX_HD void p_copy(const PolyT<NA> &a, PolyT<NR> &r)
{
    NRX_UNROLL for (int k = 0; k < NR; ++k) r.c[k] = (k < NA) ? a.c[k] : 0.0;
    r.n = imin(a.n, NR);
}

// r = a * b (capacity NR; degrees are sized to fit)
template <int NA, int NBB, int NR> NRX_HD void p_mul(const PolyT<NA> &a, const PolyT<NBB> &b, PolyT<NR> &r)
{
    double t[NR];
    NRX_UNROLL for (int k = 0; k < NR; ++k) t[k] = 0.0;
    NRX_UNROLL for (int i = 0; i < NA; ++i) {
        NRX_UNROLL for (int j = 0; j < NBB; ++j) {
            if (i + j < NR) t[i + j] += a.c[i] * b.c[j];
        }
    }
    int n = imin(a.n + b.n - 1, NR);
    NRX_UNROLL for (int k = 0; k < NR; ++k) r.c[k] = t[k];
    r.n = n;
    NRX_CNT(2 * a.n * b.n);
}

// a *= (y + s) when on, unchanged otherwise (branch-free)
template <int N> NRX_HD void p_mul_lin(PolyT<N> &a, double s, bool on = true)
{
    double g = on ? 1.0 : 0.0, sl = on ? s : 1.0;
    double t[N];
    NRX_UNROLL for (int k = 0; k < N; ++k) t[k] = ((k > 0) ? g * a.c[k - 1] : 0.0) + sl * a.c[k];
    NRX_UNROLL for (int k = 0; k < N; ++k) a.c[k] = t[k];
    NRX_CNT(on ? 2 * (a.n + 1) : 0);
    a.n = imin(a.n + (on ? 1 : 0), N);
}

// (y + sa)^la (y + sb)^lb with la, lb <= NRX_LMAX (the largest angular
// momentum of a system is checked when it is created)
template <int N> NRX_HD void p_pair(int la, double sa, int lb, double sb, PolyT<N> &r)
{
    p_const(r, 1.0);
    NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) p_mul_lin(r, sa, i < la);
    NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) p_mul_lin(r, sb, i < lb);
}

// h(s) = E[f(s + e)], e ~ N(0, var):  h_j = sum_i f_{j+2i} C(j+2i, j) (2i-1)!! var^i
template <int N> NRX_HD void p_smooth(const PolyT<N> &f, double var, PolyT<N> &h)
{
    double vp[N / 2 + 1];
    vp[0] = 1.0;
    NRX_UNROLL for (int i = 1; i <= N / 2; ++i) vp[i] = vp[i - 1] * var;
    double t[N];
    NRX_UNROLL for (int j = 0; j < N; ++j) {
        double acc = f.c[j];
        NRX_UNROLL for (int i = 1; j + 2 * i < N; ++i) acc += f.c[j + 2 * i] * (binomc(j + 2 * i, j) * dfact(2 * i - 1)) * vp[i];
        t[j] = acc;
    }
    NRX_UNROLL for (int j = 0; j < N; ++j) h.c[j] = t[j];
    h.n = f.n;
    NRX_CNT(f.n * f.n / 2 + f.n);
}

template <int N> NRX_HD double p_eval(const PolyT<N> &h, double s)
{
    double v = 0.0;
    NRX_UNROLL for (int k = N - 1; k >= 0; --k) v = v * s + h.c[k];
    NRX_CNT(2 * h.n);
    return v;
}

// g(y) = h(m0 + beta y):  g_j = beta^j sum_{k>=j} C(k,j) m0^{k-j} h_k   (g may alias h)
template <int N> NRX_HD void p_compose(const PolyT<N> &h, double m0, double beta, PolyT<N> &g)
{
    double mp[N], t[N];
    mp[0] = 1.0;
    NRX_UNROLL for (int i = 1; i < N; ++i) mp[i] = mp[i - 1] * m0;
    double bp = 1.0;
    NRX_UNROLL for (int j = 0; j < N; ++j) {
        double acc = h.c[j];
        NRX_UNROLL for (int k = j + 1; k < N; ++k) acc += binomc(k, j) * mp[k - j] * h.c[k];
        t[j] = acc * bp;
        bp *= beta;
    }
    NRX_UNROLL for (int j = 0; j < N; ++j) g.c[j] = t[j];
    g.n = h.n;
    NRX_CNT(h.n * h.n + 3 * h.n);
}

// Hermite-Gaussian expansion: f(s) exp(-w s^2) = sum_t e_t Lambda_t(s),
// Lambda_t = d^t/dc^t exp(-w (s - c)^2) at c = 0.  From
// s Lambda_t = Lambda_{t+1}/(2w) + t Lambda_{t-1}, in closed form
//   s^k exp(-w s^2) = sum_{t+2m=k} k! / (t! m! 2^m) a^(t+m) Lambda_t,  a = 1/(2w),
// so e_t = a^t sum_m f_{t+2m} (t+2m)! / (t! m! 2^m) a^m.
constexpr double hermite_coef(int t, int m)
{
    double r = 1.0;
    for (int i = t + 1; i <= t + 2 * m; ++i) r *= i; // (t+2m)!/t!
    for (int i = 2; i <= m; ++i) r /= i;             // / m!
    for (int i = 0; i < m; ++i) r *= 0.5;            // / 2^m
    return r;
}

template <int N> NRX_HD void p_hermite(const PolyT<N> &f, double w, PolyT<N> &e)
{
    double ap[N];
    ap[0] = 1.0;
    const double a = 0.5 * rcp(w);
    NRX_UNROLL for (int i = 1; i < N; ++i) ap[i] = ap[i - 1] * a;
    double t_[N];
    NRX_UNROLL for (int t = 0; t < N; ++t) {
        double acc = 0.0;
        NRX_UNROLL for (int m = 0; t + 2 * m < N; ++m) acc += f.c[t + 2 * m] * hermite_coef(t, m) * ap[m];
        t_[t] = acc * ap[t];
    }
    NRX_UNROLL for (int t = 0; t < N; ++t) e.c[t] = t_[t];
    e.n = f.n;
    NRX_CNT(f.n * f.n / 2 + 2 * f.n);
}

// ------------------------------------------------------------------
// Boys function, reference semantics
// ------------------------------------------------------------------
constexpr int BT_N = 401;      // grid T = 0, 0.05, ..., 20
constexpr double BT_H = 0.05;
constexpr int BT_M = 32;       // orders stored per grid point
constexpr int BT_K = 7;        // Taylor terms
constexpr int BT_MAXM = BT_M - BT_K; // largest order evaluated from the table

// F_M(T) for T <= 20 from the table: sum_k F_{M+k}(T_i) (-dt)^k / k!
NRX_HD double boys_top(int M, double T, const double *tab)
{
    int i = (int)(T * (1.0 / BT_H) + 0.5);
    double dt = T - i * BT_H;
    const double *row = tab + i * BT_M + M;
    double v = row[BT_K - 1];
    NRX_UNROLL for (int k = BT_K - 1; k >= 1; --k) v = row[k - 1] - dt * v * (1.0 / k);
    return v;
}

// F[0..M] (host-side helper and tests)
NRX_HD void boys_ref(int M, double T, double *F, const double *tab)
{
    if (T > 20.0) {
        // common.f:326-328
        F[0] = 0.88622692545275801365 / sqrt(T);
        for (int l = 1; l <= M; ++l) F[l] = F[l - 1] * (1.0 * l - 0.5) / T;
        return;
    }
    F[M] = boys_top(M, T, tab);
    double ex = exp(-T), t2 = 2.0 * T;
    for (int l = M - 1; l >= 0; --l) F[l] = (t2 * F[l + 1] + ex) * (1.0 / (2 * l + 1));
}

// sum_k c_k F_k(T) for a polynomial c of degree M = c.n - 1
template <int N> NRX_HD double boys_dot(const PolyT<N> &c, double T, const double *tab)
{
    int M = c.n - 1;
    double s = 0.0;
    if (T > 20.0) {
        double rT = rsqrt(T), f = 0.88622692545275801365 * rT, it = rT * rT;
        s = c.c[0] * f;
        NRX_UNROLL for (int l = 1; l < N; ++l) {
            f = f * (l - 0.5) * it;
            s += c.c[l] * f;
        }
    } else {
        double f = boys_top(M, T, tab);
        double ex = exp(-T), t2 = 2.0 * T;
        NRX_UNROLL for (int k = N - 1; k >= 0; --k) {
            double fn = (t2 * f + ex) * (1.0 / (2 * k + 1));
            f = (k < M) ? fn : f;
            s += c.c[k] * f; // c.c[k] = 0 for k > M
        }
    }
    NRX_CNT(5 * c.n + 2 * BT_K + 10);
    return s;
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
// bra*ket product of one particle (precomputed per primitive pair)
// ------------------------------------------------------------------
struct Prim {
    double a;
    double A[3];
    int l[3];
};

struct PairRec {
    double p, K, beta;       // exponent a+b, prod_d exp(-ab/p (A-B)^2), ket exponent
    double P[3], B[3];       // product centre, ket centre
    double sA[3], sB[3];     // P - A, P - B
    int la[3], lb[3];        // bra / ket powers
};

NRX_HD void make_pair(const Prim &bra, const Prim &ket, PairRec &q)
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
    }
    q.K = exp(-a * b / p * r2);
}

// per-direction polynomial of a particle in y = x - P
template <int N> NRX_HD void fpoly(const PairRec &q, int d, PolyT<N> &f)
{
    p_pair(q.la[d], q.sA[d], q.lb[d], q.sB[d], f);
}

// ------------------------------------------------------------------
// Star elimination per direction
// ------------------------------------------------------------------
template <int N> struct VarT {
    PolyT<N> F; // polynomial in the particle coordinate y
    double w, c; // Gaussian exp(-w (y - c)^2)
};

// Integrate leaf L (edge g, d = P_leaf - P_parent) into parent Q; returns the
// scalar factor sqrt(pi/(wL+g)) * exp(...).
template <int NLF, int NQ>
NRX_HD double absorb_leaf(const PolyT<NLF> &LF, double Lw, double Lc, double g, double d, VarT<NQ> &Q, double pic)
{
    // one reciprocal per new exponent, the quotients as products
    double W = Lw + g, rW = rsqrt(W), iW = rW * rW;
    PolyT<NLF> h;
    p_smooth(LF, 0.5 * iW, h);
    double fac = sqrt(pic) * rW;
    if (g == 0.0) {
        double v = p_eval(h, Lc);
        NRX_UNROLL for (int k = 0; k < NQ; ++k) Q.F.c[k] *= v;
        return fac;
    }
    double beta = g * iW, m0 = (Lw * Lc - g * d) * iW;
    p_compose(h, m0, beta, h);
    p_mul(Q.F, h, Q.F);
    double kap = Lw * beta;
    double cl = d + Lc;
    double Wq = Q.w + kap, iWq = rcp(Wq);
    double diff = Q.c - cl;
    fac *= exp(-(Q.w * kap * iWq) * diff * diff);
    Q.c = (Q.w * Q.c + kap * cl) * iWq;
    Q.w = Wq;
    NRX_CNT(26);
    return fac;
}

// E[F(y)] over exp(-w (y - c)^2) / sqrt(pi/w), given iw = 1/w
template <int N> NRX_HD double close_expect(const PolyT<N> &F, double iw, double c)
{
    PolyT<N> h;
    p_smooth(F, 0.5 * iw, h);
    return p_eval(h, c);
}

template <int N> NRX_HD double close_var(const VarT<N> &V, double pic)
{
    PolyT<N> h;
    double rw = rsqrt(V.w), iw = rw * rw;
    p_smooth(V.F, 0.5 * iw, h);
    return sqrt(pic) * rw * p_eval(h, V.c);
}

// system-wide inputs of a tuple evaluation
struct SysParams {
    int nat, ngtg;
    double pmass;
    const double *cat, *zan, *bcoef, *gamma;
    const double *bt; // Boys table
};

inline namespace NRX_CAP_NS {

constexpr int NS = 2 * NRX_LMAX + 1; // per-particle polynomial capacity (bra + ket powers <= 2 lmax)
constexpr int NK = NS + 2;           // per-particle polynomial inside the kinetic operator
constexpr int NB = 8 * NRX_LMAX + 1; // Boys-order polynomial (total angular momentum <= 8 lmax)
constexpr int NPART = 4;
// trip count of the kinetic operator's X^(L-2) factor loop (at most
// NRX_LMAX - 2 factors; the l <= 2 build keeps its historical two predicated
// steps, which never fire there)
constexpr int KIN_LM2 = NRX_LMAX > 2 ? NRX_LMAX - 2 : 2;

// Integral description: particles 0..n-1, hub = n-1, electron i linked to the
// hub by gam[i] (bra+ket).
struct Sys {
    int n;
    const PairRec *q[NPART];
    double gam[NPART];
    const double *bt; // Boys table
};

NRX_HD double Kprod(const Sys &s)
{
    double k = 1.0;
    NRX_UNROLL for (int i = 0; i < NPART; ++i)
        if (i < s.n) k *= s.q[i]->K;
    return k;
}

// 1-D overlap along d; particles ia and ic (when >= 0) use the given
// polynomials instead of their own
template <int NL>
NRX_HD double ovl_dim(const Sys &s, int d, int ia, const PolyT<NK> &Fa, int ic, const PolyT<NK> &Fc, double pic)
{
    int h = s.n - 1;
    VarT<NL> H;
    if (h == ia)
        p_copy(Fa, H.F);
    else if (h == ic)
        p_copy(Fc, H.F);
    else
        fpoly(*s.q[h], d, H.F);
    H.w = s.q[h]->p;
    H.c = 0.0;
    double fac = 1.0;
    NRX_UNROLL for (int i = 0; i < NPART - 1; ++i) {
        if (i >= h) continue;
        PolyT<NK> f;
        if (i == ia)
            f = Fa;
        else if (i == ic)
            f = Fc;
        else
            fpoly(*s.q[i], d, f);
        fac *= absorb_leaf(f, s.q[i]->p, 0.0, s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], H, pic);
    }
    return fac * close_var(H, pic);
}

template <int NL> NRX_HD double ovl(const Sys &s, double pic = PI)
{
    NRX_FENCE();
    double r = Kprod(s);
    PolyT<NK> dummy;
    p_const(dummy, 1.0);
    NRX_DIRS for (int d = 0; d < 3; ++d) r *= ovl_dim<NL>(s, d, -1, dummy, -1, dummy, pic);
    return r;
}

// Particle i's polynomial f (in its own coordinate) as a polynomial in the
// hub coordinate y after i is integrated out (the polynomial part of
// absorb_leaf): smooth(f, 1/(2W))(m0 + beta y), W = p_i + g_i, beta = g_i / W,
// m0 = -beta (P_i - P_hub).  Linear in f; the hub's own polynomial maps to
// itself.
template <int NF, int NR> NRX_HD void hub_map(const Sys &s, int i, int d, const PolyT<NF> &f, PolyT<NR> &r)
{
    const int h = s.n - 1;
    if (i == h) {
        p_copy(f, r);
        return;
    }
    const PairRec &Q = *s.q[i];
    double g = s.gam[i], iW = rcp(Q.p + g);
    PolyT<NF> t;
    p_smooth(f, 0.5 * iW, t);
    if (g == 0.0) {
        p_const(r, p_eval(t, 0.0));
        return;
    }
    double beta = g * iW;
    p_compose(t, -beta * (Q.P[d] - s.q[h]->P[d]), beta, t);
    p_copy(t, r);
}

template <int N> NRX_HD void p_add(const PolyT<N> &a, const PolyT<N> &b, PolyT<N> &r)
{
    NRX_UNROLL for (int k = 0; k < N; ++k) r.c[k] = a.c[k] + b.c[k];
    r.n = a.n > b.n ? a.n : b.n;
    NRX_CNT(r.n);
}

// Kinetic energy of particle a (-1/(2m) <bra| d^2/dx_a^2 |ket>) for nv <= 2
// variants sharing one geminal configuration: variant v has a ket geminal of
// exponent gk[v] between a and partner c (gk[v] == 0 or c < 0: none).
// Per direction every 1-D integral of the product rule is
// fac * sqrt(pi/w) * E[Q(y)] over one Gaussian exp(-w (y - c)^2) in the hub
// coordinate (the Gaussian factors of the star do not depend on the
// polynomials), with Q the product of the particles' polynomials mapped by
// hub_map; the factors are computed once and the terms of a variant are
// summed into one Q before the closing expectation.
template <int NL>
NRX_HD void kin_multi(const Sys &s, int a, int c, const double *gk, int nv, double mass, double pic, double *out,
                      double *ovl_out = nullptr)
{
    NRX_FENCE();
    // the kinetic polynomials (degree <= tuple degree < NL) at capacity min(NK, NL)
    constexpr int FK = NK < NL ? NK : NL;
    // product rule over directions, accumulated direction by direction:
    // R_v = sum_d T_d prod_{d' != d} S_d',  Pr = prod_d S_d
    double Pr = 1.0, R[2] = {0.0, 0.0};
    const PairRec &A = *s.q[a];
    const int h = s.n - 1;
    NRX_DIRS for (int d = 0; d < 3; ++d) {
        // Gaussian factors, leaf by leaf as absorb_leaf forms them
        double w = s.q[h]->p, cq = 0.0, fac = 1.0;
        NRX_UNROLL for (int i = 0; i < NPART - 1; ++i) {
            if (i >= h) continue;
            const PairRec &Q = *s.q[i];
            double g = s.gam[i], rW = rsqrt(Q.p + g), iW = rW * rW, fl = sqrt(pic) * rW;
            if (g != 0.0) {
                double kap = Q.p * (g * iW), cl = Q.P[d] - s.q[h]->P[d], Wq = w + kap, iWq = rcp(Wq);
                double diff = cq - cl;
                fl *= exp(-(w * kap * iWq) * diff * diff);
                cq = (w * cq + kap * cl) * iWq;
                w = Wq;
            }
            fac *= fl;
        }
        const double rw = rsqrt(w), iw = rw * rw, norm = fac * (sqrt(pic) * rw);
        // the particles other than a and c
        PolyT<NL> env;
        p_const(env, 1.0);
        NRX_UNROLL for (int i = 0; i < NPART; ++i) {
            if (i > h || i == a || i == c) continue;
            PolyT<NS> f;
            fpoly(*s.q[i], d, f);
            PolyT<NL> m;
            hub_map(s, i, d, f, m);
            p_mul(env, m, env);
        }
        PolyT<NL> base, q;
        if (c >= 0) {
            PolyT<NS> f;
            fpoly(*s.q[c], d, f);
            PolyT<NL> m;
            hub_map(s, c, d, f, m);
            p_mul(env, m, base);
        } else {
            base = env;
        }
        {
            PolyT<NS> f;
            fpoly(A, d, f);
            PolyT<NL> m;
            hub_map(s, a, d, f, m);
            p_mul(base, m, q);
        }
        double Sd = norm * close_expect(q, iw, cq);
        int L = A.lb[d];
        double sA = A.sA[d], sB = A.sB[d];
        int la = A.la[d];
        NRX_UNROLL for (int v = 0; v < 2; ++v) {
            if (v >= nv) continue;
            double g = (c >= 0) ? gk[v] : 0.0;
            double al = A.beta + g;
            double dl = (g != 0.0) ? g * (A.B[d] - s.q[c]->B[d]) : 0.0;
            // d^2/dx^2 of the ket X^L exp(-beta X^2) exp(-g (x - x_c)^2), X = y + sB:
            // F0 = L(L-1)X^{L-2} - 4L dl X^{L-1} + (-4L al + 4dl^2 - 2al) X^L + 8 al dl X^{L+1} + 4al^2 X^{L+2}
            // F1 = 4 L g X^{L-1} - 8 dl g X^L - 8 al g X^{L+1}      (times X_c)
            // F2 = 4 g^2 X^L                                        (times X_c^2)
            // each times the bra factor u_A = (y + sA)^la
            double cf0[5] = {(double)(L * (L - 1)), -4.0 * L * dl, -4.0 * L * al + 4.0 * dl * dl - 2.0 * al,
                             8.0 * al * dl, 4.0 * al * al};
            double cf1[5] = {0.0, 4.0 * L * g, -8.0 * dl * g, -8.0 * al * g, 0.0};
            // sum_j cf[j] X^{L-2+j} over L-2+j >= 0 by Horner in X, then X^{max(L-2,0)} u_A
            int j0 = (L >= 2) ? 0 : 2 - L;
            PolyT<FK> F0, F1;
            p_const(F0, cf0[4]);
            p_const(F1, cf1[4]);
            NRX_UNROLL for (int j = 3; j >= 0; --j) {
                bool on = j >= j0;
                p_mul_lin(F0, sB, on);
                p_mul_lin(F1, sB, on);
                F0.c[0] += on ? cf0[j] : 0.0;
                F1.c[0] += on ? cf1[j] : 0.0;
            }
            NRX_UNROLL for (int i = 0; i < KIN_LM2; ++i) {
                p_mul_lin(F0, sB, i < L - 2);
                p_mul_lin(F1, sB, i < L - 2);
            }
            NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) {
                p_mul_lin(F0, sA, i < la);
                p_mul_lin(F1, sA, i < la);
            }
            PolyT<NL> m0;
            hub_map(s, a, d, F0, m0);
            p_mul(base, m0, q);
            if (g != 0.0) {
                // partner c's polynomial times X_c and X_c^2, X_c = y_c + sB_c
                PolyT<FK> F2;
                p_const(F2, 4.0 * g * g);
                NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) p_mul_lin(F2, sB, i < L);
                NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) p_mul_lin(F2, sA, i < la);
                PolyT<FK> fc;
                fpoly(*s.q[c], d, fc);
                double xc = s.q[c]->sB[d];
                PolyT<NL> m1, m2, c1, c2, u;
                p_mul_lin(fc, xc);
                hub_map(s, c, d, fc, c1);
                p_mul_lin(fc, xc);
                hub_map(s, c, d, fc, c2);
                hub_map(s, a, d, F1, m1);
                hub_map(s, a, d, F2, m2);
                p_mul(m1, c1, u);
                p_mul(m2, c2, m2);
                p_add(u, m2, u);
                p_mul(env, u, u);
                p_add(q, u, q);
            }
            double t = norm * close_expect(q, iw, cq);
            R[v] = R[v] * Sd + Pr * t;
        }
        Pr *= Sd;
    }
    double kp = Kprod(s);
    NRX_UNROLL for (int v = 0; v < 2; ++v) {
        if (v >= nv) continue;
        out[v] = -0.5 * R[v] / mass * kp;
    }
    if (ovl_out) *ovl_out = kp * Pr; // the configuration's overlap, normalised with pic
}

template <int NL> NRX_HD double kin(const Sys &s, int a, int c, double gk, double mass, double pic = PI)
{
    double o[2];
    double g2[2] = {gk, gk};
    kin_multi<NL>(s, a, c, g2, 1, mass, pic, o);
    return o[0];
}

constexpr double factc(int n)
{
    double r = 1.0;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}

// The l0 = 0 case of hermite_coulomb (nuclear attraction) in closed form:
// R_t is the scaled Hermite polynomial
//   R_t = sum_{j=ceil(t/2)}^{t} (-1)^j t! / ((t-j)! (2j-t)!) (2D)^{2j-t} l1^j u^{2j},
// so P_j = (-l1)^j sum_{t=j}^{min(2j, N-1)} hk_t t! / ((t-j)! (2j-t)!) (2D)^{2j-t}.
template <int N> NRX_HD void hermite_coulomb0(const PolyT<N> &hk, double D, double l1, PolyT<N> &P)
{
    double dp[N];
    dp[0] = 1.0;
    NRX_UNROLL for (int i = 1; i < N; ++i) dp[i] = dp[i - 1] * (2.0 * D);
    double lp = 1.0;
    NRX_UNROLL for (int j = 0; j < N; ++j) {
        double acc = 0.0;
        NRX_UNROLL for (int t = j; t < N && t <= 2 * j; ++t)
            acc += hk.c[t] * (factc(t) / (factc(t - j) * factc(2 * j - t))) * dp[2 * j - t];
        P.c[j] = acc * lp;
        lp *= -l1;
    }
    P.n = hk.n;
    NRX_CNT(hk.n * hk.n / 2 + 3 * hk.n);
}

// P(u^2) = sum_k hk R_k(u^2), R_0 = 1, R_1 = -2 lam D,
// R_{k+1} = -2 lam (D R_k + k R_{k-1}), lam = l0 + l1 u^2.  R_t is the same
// scaled Hermite polynomial as above, in lam: P is hermite_coulomb0 with
// l1 = 1 (a polynomial in lam) composed with lam = l0 + l1 u^2 (a Taylor
// shift by l0, then the scaling by l1).
template <int N> NRX_HD void hermite_coulomb(const PolyT<N> &hk, double D, double l0, double l1, PolyT<N> &P)
{
    PolyT<N> q;
    hermite_coulomb0(hk, D, 1.0, q);
    p_compose(q, l0, l1, P);
}

NRX_HD void underflow(double &x)
{
    if (fabs(x) < 1.0e-100) x = 0.0;
}

// Hermite coefficients and absolute centre of the core variable of a
// nuclear-attraction integral on particle a (everything else absorbed)
template <int NL> NRX_HD double nuc_core(const Sys &s, int a, int d, PolyT<NL> &e, double &X, double &w, double pic)
{
    int h = s.n - 1;
    VarT<NL> H;
    fpoly(*s.q[h], d, H.F);
    H.w = s.q[h]->p;
    H.c = 0.0;
    double fac = 1.0;
    NRX_UNROLL for (int i = 0; i < NPART - 1; ++i) {
        if (i >= h || i == a) continue;
        PolyT<NS> f;
        fpoly(*s.q[i], d, f);
        fac *= absorb_leaf(f, s.q[i]->p, 0.0, s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], H, pic);
    }
    int core = h;
    VarT<NL> A;
    if (a != h) {
        fpoly(*s.q[a], d, A.F);
        A.w = s.q[a]->p;
        A.c = 0.0;
        fac *= absorb_leaf(H.F, H.w, H.c, s.gam[a], s.q[h]->P[d] - s.q[a]->P[d], A, pic);
        core = a;
    } else {
        A = H;
    }
    p_compose(A.F, A.c, 1.0, A.F);
    p_hermite(A.F, A.w, e);
    X = s.q[core]->P[d] + A.c;
    w = A.w;
    return fac;
}

// sum_c zan[c] <1/|r_a - C_c|>, each centre's value flushed below 1e-100 when
// flush is set (the reference applies underflow() per centre)
// sum_c zan[c] pre <1/|r - C_c|> over the core variable's Hermite
// coefficients e[d], centre X[d] and exponent w (the centre loop of nuc_sum)
template <int NL, int NBC>
NRX_HD double nuc_centres(const PolyT<NL> *e, const double *X, double w, double pre, int nat, const double *cat,
                          const double *zan, bool flush, const double *bt)
{
    double tot = 0.0;
    for (int c = 0; c < nat; ++c) {
        // product over directions, sized as it grows: x*y at capacity
        // 2NL-1 (capped at NBC), then *z at NBC
        constexpr int N2 = (2 * NL - 1 < NBC) ? 2 * NL - 1 : NBC;
        double Dx = X[0] - cat[3 * c], Dy = X[1] - cat[3 * c + 1], Dz = X[2] - cat[3 * c + 2];
        PolyT<NL> Px, Py;
        hermite_coulomb0(e[0], Dx, w, Px);
        hermite_coulomb0(e[1], Dy, w, Py);
        PolyT<N2> cxy;
        p_mul(Px, Py, cxy);
        hermite_coulomb0(e[2], Dz, w, Px);
        PolyT<NBC> cc;
        p_mul(cxy, Px, cc);
        double T = w * (((0.0 + Dx * Dx) + Dy * Dy) + Dz * Dz);
        double v = pre * boys_dot(cc, T, bt);
        if (flush) underflow(v);
        tot += zan[c] * v;
    }
    return tot;
}

template <int NL, int NBC = NB>
NRX_HD double nuc_sum(const Sys &s, int a, int nat, const double *cat, const double *zan, bool flush, double pic = PI)
{
    NRX_FENCE();
    PolyT<NL> e[3];
    double X[3], w = 1.0, fac = Kprod(s);
    NRX_UNROLL for (int d = 0; d < 3; ++d) fac *= nuc_core<NL>(s, a, d, e[d], X[d], w, pic);
    return nuc_centres<NL, NBC>(e, X, w, fac * 2.0 * pic * rcp(w), nat, cat, zan, flush, s.bt);
}

// One kinetic operator of kin_pair: particle a, partner c (c < 0: none),
// nv <= 2 variants with ket-geminal exponents gk[v], mass.
struct KinOp {
    int a, c, nv;
    double gk[2], mass;
};

// Two kinetic operators on one configuration (for instance the electron's
// and the proton's of an e-p star): the Gaussian factors, the particles'
// polynomials mapped to the hub coordinate and the overlap of each direction
// are computed once and shared, as are the product-rule overlaps S_d.
// out1[v], out2[v] as kin_multi's out; ovl_out (if set) the overlap.  With
// NUC set, nuc_out = nuc_sum(s, hub, ...) at capacity NC: the hub's core
// variable is the product polynomial of the overlap, on the same Gaussian.
template <int NL, int NC = NL, int NBC = NB, bool NUC = false>
NRX_HD void kin_pair(const Sys &s, const KinOp &o1, const KinOp &o2, double pic, double *out1, double *out2,
                     double *ovl_out = nullptr, double *nuc_out = nullptr, int nat = 0, const double *cat = nullptr,
                     const double *zan = nullptr, bool flush = false)
{
    NRX_FENCE();
    // the kinetic polynomials (degree <= tuple degree < NL) at capacity min(NK, NL)
    constexpr int FK = NK < NL ? NK : NL;
    const int h = s.n - 1;
    double Pr = 1.0, R[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    PolyT<NC> ne[3]; // the hub's Hermite coefficients per direction (NUC)
    double nX[3] = {0.0, 0.0, 0.0}, nfac = 1.0, nw = 1.0;
    NRX_DIRS for (int d = 0; d < 3; ++d) {
        double w = s.q[h]->p, cq = 0.0, fac = 1.0;
        NRX_UNROLL for (int i = 0; i < NPART - 1; ++i) {
            if (i >= h) continue;
            const PairRec &Q = *s.q[i];
            double g = s.gam[i], rW = rsqrt(Q.p + g), iW = rW * rW, fl = sqrt(pic) * rW;
            if (g != 0.0) {
                double kap = Q.p * (g * iW), cl = Q.P[d] - s.q[h]->P[d], Wq = w + kap, iWq = rcp(Wq);
                double diff = cq - cl;
                fl *= exp(-(w * kap * iWq) * diff * diff);
                cq = (w * cq + kap * cl) * iWq;
                w = Wq;
            }
            fac *= fl;
        }
        const double rw = rsqrt(w), iw = rw * rw, norm = fac * (sqrt(pic) * rw);
        // every particle's own polynomial in the hub coordinate
        PolyT<NL> m[NPART];
        NRX_UNROLL for (int i = 0; i < NPART; ++i) {
            if (i > h) {
                p_const(m[i], 1.0);
                continue;
            }
            PolyT<NS> f;
            fpoly(*s.q[i], d, f);
            hub_map(s, i, d, f, m[i]);
        }
        PolyT<NL> q;
        p_const(q, 1.0);
        NRX_UNROLL for (int i = 0; i < NPART; ++i)
            if (i <= h) p_mul(q, m[i], q);
        const double Sd = norm * close_expect(q, iw, cq);
        if constexpr (NUC) {
            PolyT<NC> qc, ed;
            p_copy(q, qc);
            p_compose(qc, cq, 1.0, qc);
            p_hermite(qc, w, ed);
            // the direction loop stays rolled: store by uniform branches
            if (d == 0)
                ne[0] = ed;
            else if (d == 1)
                ne[1] = ed;
            else
                ne[2] = ed;
            double x = s.q[h]->P[d] + cq;
            nX[0] = (d == 0) ? x : nX[0];
            nX[1] = (d == 1) ? x : nX[1];
            nX[2] = (d == 2) ? x : nX[2];
            nfac *= fac;
            nw = w;
        }
        NRX_UNROLL for (int k = 0; k < 2; ++k) {
            const KinOp &o = (k == 0) ? o1 : o2;
            const int a = o.a, c = o.c;
            const PairRec &A = *s.q[a];
            PolyT<NL> env, base;
            p_const(env, 1.0);
            NRX_UNROLL for (int i = 0; i < NPART; ++i)
                if (i <= h && i != a && i != c) p_mul(env, m[i], env);
            if (c >= 0)
                p_mul(env, m[c], base);
            else
                base = env;
            int L = A.lb[d];
            double sA = A.sA[d], sB = A.sB[d];
            int la = A.la[d];
            NRX_UNROLL for (int v = 0; v < 2; ++v) {
                if (v >= o.nv) continue;
                double g = (c >= 0) ? o.gk[v] : 0.0;
                double al = A.beta + g;
                double dl = (g != 0.0) ? g * (A.B[d] - s.q[c]->B[d]) : 0.0;
                // the F0, F1, F2 of kin_multi
                double cf0[5] = {(double)(L * (L - 1)), -4.0 * L * dl, -4.0 * L * al + 4.0 * dl * dl - 2.0 * al,
                                 8.0 * al * dl, 4.0 * al * al};
                double cf1[5] = {0.0, 4.0 * L * g, -8.0 * dl * g, -8.0 * al * g, 0.0};
                int j0 = (L >= 2) ? 0 : 2 - L;
                PolyT<FK> F0, F1;
                p_const(F0, cf0[4]);
                p_const(F1, cf1[4]);
                NRX_UNROLL for (int j = 3; j >= 0; --j) {
                    bool on = j >= j0;
                    p_mul_lin(F0, sB, on);
                    p_mul_lin(F1, sB, on);
                    F0.c[0] += on ? cf0[j] : 0.0;
                    F1.c[0] += on ? cf1[j] : 0.0;
                }
                NRX_UNROLL for (int i = 0; i < KIN_LM2; ++i) {
                    p_mul_lin(F0, sB, i < L - 2);
                    p_mul_lin(F1, sB, i < L - 2);
                }
                NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) {
                    p_mul_lin(F0, sA, i < la);
                    p_mul_lin(F1, sA, i < la);
                }
                PolyT<NL> m0;
                hub_map(s, a, d, F0, m0);
                p_mul(base, m0, q);
                if (g != 0.0) {
                    PolyT<FK> F2;
                    p_const(F2, 4.0 * g * g);
                    NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) p_mul_lin(F2, sB, i < L);
                    NRX_UNROLL for (int i = 0; i < NRX_LMAX; ++i) p_mul_lin(F2, sA, i < la);
                    PolyT<FK> fc;
                    fpoly(*s.q[c], d, fc);
                    double xc = s.q[c]->sB[d];
                    PolyT<NL> m1, m2, c1, c2, u;
                    p_mul_lin(fc, xc);
                    hub_map(s, c, d, fc, c1);
                    p_mul_lin(fc, xc);
                    hub_map(s, c, d, fc, c2);
                    hub_map(s, a, d, F1, m1);
                    hub_map(s, a, d, F2, m2);
                    p_mul(m1, c1, u);
                    p_mul(m2, c2, m2);
                    p_add(u, m2, u);
                    p_mul(env, u, u);
                    p_add(q, u, q);
                }
                double t = norm * close_expect(q, iw, cq);
                R[k][v] = R[k][v] * Sd + Pr * t;
            }
        }
        Pr *= Sd;
    }
    double kp = Kprod(s);
    NRX_UNROLL for (int v = 0; v < 2; ++v) {
        if (v < o1.nv) out1[v] = -0.5 * R[0][v] / o1.mass * kp;
        if (v < o2.nv) out2[v] = -0.5 * R[1][v] / o2.mass * kp;
    }
    if (ovl_out) *ovl_out = kp * Pr;
    if constexpr (NUC) {
        // the Gaussian factors carry pic^(1/2) per leaf and direction; the
        // nuclear attraction is normalised with PI
        double r = PI / pic, sc = 1.0;
        NRX_UNROLL for (int i = 0; i < NPART - 1; ++i)
            if (i < h) sc *= r * sqrt(r);
        *nuc_out = nuc_centres<NC, NBC>(ne, nX, nw, kp * nfac * sc * 2.0 * PI * rcp(nw), nat, cat, zan, flush, s.bt);
    }
}

// hk[t+u] += ea[t] eb[u] (-1)^u
template <int N> NRX_HD void conv_add(const PolyT<N> &ea, const PolyT<N> &eb, PolyT<N> &hk)
{
    NRX_UNROLL for (int t = 0; t < N; ++t) {
        NRX_UNROLL for (int u = 0; t + u < N; ++u) hk.c[t + u] += ea.c[t] * ((u & 1) ? -eb.c[u] : eb.c[u]);
    }
    int n = imin(ea.n + eb.n - 1, N);
    hk.n = hk.n > n ? hk.n : n;
    NRX_CNT(2 * ea.n * eb.n);
}

// Coulomb <1/|r_a - r_b|>, a < b.  With b the hub and hub_ovl set, also the
// overlap of the star without particle a (the hub with every leaf but a
// absorbed, which the Coulomb integral forms anyway).
template <int NL, int NBC = NB>
NRX_HD double coul(const Sys &s, int a, int b, double pic = PI, double *hub_ovl = nullptr)
{
    NRX_FENCE();
    int h = s.n - 1;
    PolyT<NBC> cc; // product over directions of the polynomials in u^2
    p_const(cc, 1.0);
    double D2 = 0.0, fac = Kprod(s), ho = 1.0;
    double wa = 1.0, wb = 1.0, gab = 0.0;
    NRX_DIRS for (int d = 0; d < 3; ++d) {
        VarT<NL> H;
        fpoly(*s.q[h], d, H.F);
        H.w = s.q[h]->p;
        H.c = 0.0;
        double fh = 1.0;
        NRX_UNROLL for (int i = 0; i < NPART - 1; ++i) {
            if (i >= h || i == a || i == b) continue;
            PolyT<NS> f;
            fpoly(*s.q[i], d, f);
            fh *= absorb_leaf(f, s.q[i]->p, 0.0, s.gam[i], s.q[i]->P[d] - s.q[h]->P[d], H, pic);
        }
        fac *= fh;
        if (hub_ovl) ho *= fh * close_var(H, pic);
        VarT<NL> A, B;
        fpoly(*s.q[a], d, A.F);
        A.w = s.q[a]->p;
        A.c = 0.0;
        PolyT<NL> hk;
        p_const(hk, 0.0);
        bool cycle = false;
        if (b == h) {
            gab = s.gam[a];
            B = H;
        } else {
            fpoly(*s.q[b], d, B.F);
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
                cycle = true;
                double W = H.w + ga + gb, rW = rsqrt(W), iW = rW * rW;
                double dpa = s.q[h]->P[d] - s.q[a]->P[d], dpb = s.q[h]->P[d] - s.q[b]->P[d];
                fac *= sqrt(pic) * rW;
                double ka = H.w * ga * iW, kb = H.w * gb * iW;
                double ca = dpa + H.c, cb = dpb + H.c;
                {
                    double Wn = A.w + ka, iWn = rcp(Wn), df = A.c - ca;
                    fac *= exp(-(A.w * ka * iWn) * df * df);
                    A.c = (A.w * A.c + ka * ca) * iWn;
                    A.w = Wn;
                }
                {
                    double Wn = B.w + kb, iWn = rcp(Wn), df = B.c - cb;
                    fac *= exp(-(B.w * kb * iWn) * df * df);
                    B.c = (B.w * B.c + kb * cb) * iWn;
                    B.w = Wn;
                }
                gab = ga * gb * iW;
                // h(m0 + ba ya + bb yb) = sum_j (bb yb)^j / j! h^{(j)}(m0 + ba ya);
                // the hub polynomial has degree <= NS - 1 here (its own
                // polynomial only: every other electron is a or b)
                PolyT<NS> hs;
                p_copy(H.F, hs);
                p_smooth(hs, 0.5 * iW, hs);
                double m0 = (H.w * H.c - ga * dpa - gb * dpb) * iW, ba = ga * iW, bb = gb * iW;
                // D = h^{(j)} / j!  (D_{j+1,k} = (k+1)/(j+1) D_{j,k+1}),
                // Yb = B.F (bb yb)^j, kept as the Hermite coefficients eb of
                // Yb(s + B.c) exp(-B.w s^2): the next term multiplies by
                // bb (s + B.c), which s Lambda_t = Lambda_{t+1}/(2w) + t Lambda_{t-1}
                // turns into a three-term update.  The loop stays rolled, one
                // term live, and ends after the hub polynomial's degree (D = 0
                // beyond, and Yb's degree stays within capacity until then).
                PolyT<NS> D = hs;
                PolyT<NL> eb;
                {
                    PolyT<NL> Fb;
                    p_compose(B.F, B.c, 1.0, Fb);
                    p_hermite(Fb, B.w, eb);
                }
                const double i2wb = 0.5 * rcp(B.w);
                _Pragma("nounroll") for (int j = 0; j < hs.n; ++j) {
                    PolyT<NS> der;
                    p_compose(D, m0, ba, der);
                    PolyT<NL> Fa, ea;
                    p_mul(A.F, der, Fa);
                    p_compose(Fa, A.c, 1.0, Fa);
                    p_hermite(Fa, A.w, ea);
                    conv_add(ea, eb, hk);
                    double ij = rcp((double)(j + 1));
                    NRX_UNROLL for (int k = 0; k < NS; ++k) D.c[k] = (k + 1 < NS) ? D.c[k + 1] * ((k + 1) * ij) : 0.0;
                    double nb[NL];
                    NRX_UNROLL for (int t = 0; t < NL; ++t) {
                        double v = B.c * eb.c[t];
                        if (t >= 1) v += eb.c[t - 1] * i2wb;
                        if (t + 1 < NL) v += (t + 1) * eb.c[t + 1];
                        nb[t] = bb * v;
                    }
                    NRX_UNROLL for (int t = 0; t < NL; ++t) eb.c[t] = nb[t];
                    eb.n = imin(eb.n + 1, NL);
                    NRX_CNT(5 * eb.n);
                }
            }
        }
        if (!cycle) {
            PolyT<NL> ea, eb;
            p_compose(A.F, A.c, 1.0, A.F);
            p_compose(B.F, B.c, 1.0, B.F);
            p_hermite(A.F, A.w, ea);
            p_hermite(B.F, B.w, eb);
            conv_add(ea, eb, hk);
        }
        wa = A.w;
        wb = B.w;
        double rho = wa * wb * rcp(wa + wb), irg = rcp(rho + gab);
        double l0 = rho * gab * irg, l1 = rho * rho * irg;
        double Dd = (s.q[a]->P[d] + A.c) - (s.q[b]->P[d] + B.c);
        PolyT<NL> P;
        if (gab == 0.0) // l0 = 0: no geminal between a and b (uniform across a wave)
            hermite_coulomb0(hk, Dd, l1, P);
        else
            hermite_coulomb(hk, Dd, l0, l1, P);
        p_mul(cc, P, cc);
        D2 += Dd * Dd;
    }
    double rab = rsqrt(wa + wb), iab = rab * rab;
    double rho = wa * wb * iab, irg = rcp(rho + gab);
    double l0 = rho * gab * irg, l1 = rho * rho * irg;
    double pre = pic * iab;
    pre = pre * (sqrt(pic) * rab) * 2.0 * pic * irg;
    if (hub_ovl) {
        double k = 1.0;
        NRX_UNROLL for (int i = 0; i < NPART; ++i)
            if (i < s.n && i != a) k *= s.q[i]->K;
        *hub_ovl = k * ho;
    }
    return fac * pre * (gab == 0.0 ? 1.0 : exp(-l0 * D2)) * boys_dot(cc, l1 * D2, s.bt);
}

// ------------------------------------------------------------------
// Term assembly of one primitive tuple
// ------------------------------------------------------------------

NRX_HD Sys mksys(const SysParams &S, int n, const PairRec *p0, const PairRec *p1, const PairRec *p2,
                 const PairRec *p3, double g0, double g1, double g2)
{
    Sys s;
    s.n = n;
    s.q[0] = p0;
    s.q[1] = p1 ? p1 : p0;
    s.q[2] = p2 ? p2 : p0;
    s.q[3] = p3 ? p3 : p0;
    s.gam[0] = g0;
    s.gam[1] = g1;
    s.gam[2] = g2;
    s.gam[3] = 0.0;
    s.bt = S.bt;
    return s;
}

// polynomial capacity needed by a tuple: the largest per-direction sum of
// bra+ket powers over its particles, plus 2 where a multi-particle kinetic
// operator raises one polynomial (e-p and e-e-p classes).  Only the kinetic
// terms need the 2: the tuple functions of those classes evaluate overlap,
// nuclear attraction and Coulomb terms at capacity NL - 2.
NRX_HD int tuple_degree(const PairRec *const *q, int n, bool kinetic)
{
    int m = 0;
    for (int d = 0; d < 3; ++d) {
        int s = 0;
        for (int i = 0; i < n; ++i) s += q[i]->la[d] + q[i]->lb[d];
        m = s > m ? s : m;
    }
    return m + (kinetic ? 2 : 0);
}

// xcalc_GAM1_MD (gam_1_OMP.f:398-839).  e = electron, p = proton.
// Terms that depend on gamma_k + gamma_l only are summed over k <= l.
// With MULT set: RXCHFmult_xcalc_GAM1_MD (rxchfmult/rxchfmult_gam1.f:399-840),
// the same value without the one-geminal terms.
template <int NL, int NBC, bool MULT = false>
NRX_HD void gam1_tuple(const SysParams &S, const PairRec *e, const PairRec *p, double &ansE, double &ansS)
{
    constexpr int NC = NL - 2; // capacity of the terms without a kinetic operator
    double tegx = 0, tpgx = 0, texg = 0, tpxg = 0, tegg = 0, tpgg = 0;
    double vcegx = 0, vcegg = 0, vcpgx = 0, vcpgg = 0, vmgx = 0, vmgg = 0;
    double ogx = 0, ogg = 0;
    for (int k = 0; k < S.ngtg; ++k) {
        double gv = S.gamma[k], bv = S.bcoef[k];
        if constexpr (!MULT) {
            Sys s1 = mksys(S, 2, e, p, 0, 0, gv, 0, 0);  // (e, p): hub = p
            // the electron's and the proton's kinetic operators on the same
            // two-particle star: one pass serves both
            double ke[2], kp[2], ov, vp;
            kin_pair<NL, NC, NBC, true>(s1, KinOp{0, 1, 2, {0.0, gv}, 1.0}, KinOp{1, 0, 2, {0.0, gv}, S.pmass}, PI_G2,
                                        ke, kp, &ov, &vp, S.nat, S.cat, S.zan, false);
            tegx += bv * ke[0];
            texg += bv * ke[1];
            tpgx += bv * kp[0];
            tpxg += bv * kp[1];
            vcegx += bv * nuc_sum<NC, NBC>(s1, 0, S.nat, S.cat, S.zan, false);
            vcpgx += bv * vp;
            vmgx += bv * coul<NC, NBC>(s1, 0, 1);
            ogx += bv * (ov * G2_TO_PI3);
        }
        // (k,l) and (l,k) share the configuration gamma_k + gamma_l; their
        // kinetic terms differ only in the ket-geminal exponent (gamma_l,
        // gamma_k), two variants of one kinetic call
        for (int l = k; l < S.ngtg; ++l) {
            double gw = S.gamma[l], bb = bv * S.bcoef[l];
            Sys s2 = mksys(S, 2, e, p, 0, 0, gv + gw, 0, 0);
            int nv = (l == k) ? 1 : 2;
            double ke[2], kp[2], ov, vp;
            kin_pair<NL, NC, NBC, true>(s2, KinOp{0, 1, nv, {gw, gv}, 1.0}, KinOp{1, 0, nv, {gw, gv}, S.pmass}, PI_G2,
                                        ke, kp, &ov, &vp, S.nat, S.cat, S.zan, false);
            tegg += (l == k) ? bb * ke[0] : bb * ke[0] + bb * ke[1];
            tpgg += (l == k) ? bb * kp[0] : bb * kp[0] + bb * kp[1];
            double w = (l == k) ? bb : 2.0 * bb;
            vcegg += w * nuc_sum<NC, NBC>(s2, 0, S.nat, S.cat, S.zan, false);
            vcpgg += w * vp;
            vmgg += w * coul<NC, NBC>(s2, 0, 1);
            ogg += w * (ov * G2_TO_PI3);
        }
    }
    double te = tegx + texg + tegg, tp = tpgx + tpxg + tpgg;
    double vce = 2.0 * vcegx + vcegg, vcp = 2.0 * vcpgx + vcpgg, vmx = 2.0 * vmgx + vmgg;
    ansE = te + tp - vce + vcp - vmx;
    ansS = 2.0 * ogx + ogg;
}

// xcalc_GAM2_MD (gam_2_OMP.f:546-1077).  With MULT set: RXCHFmult_GAM2_MD
// (rxchfmult/rxchfmult_gam2.f:627-1162), where e1 and e2 come from different
// bases.  It drops the single-geminal terms, places the geminal of gVEEg1 on
// e2 (gamA23/gamB23, :1067-1068) instead of e1, and returns three values:
// ansE = the two-geminal energy, ansE2 = gVEPg2 + gVEEg1, ansS = gsg; its
// exchange form RXCHFmult_xcalc_GAM2_MDex (rxchfmult_gam2ex.f:559-1097) adds
// ansE3 = gVEEg2.
template <int NL, int NBC, bool MULT = false>
NRX_HD void gam2_core(const SysParams &S, const PairRec *e1, const PairRec *e2, const PairRec *p, double &ansE,
                      double &ansS, double &ansE2, double &ansE3)
{
    constexpr int NC = NL - 2; // capacity of the terms without a kinetic operator
    // Where e1 has no geminal (g(e1,p) = 0), its kinetic and nuclear-attraction
    // integrals factor into e1's one-particle value times the (e2, p) overlap
    // with the geminal: t1, v1 below, and ovl2(g).
    Sys s1e = mksys(S, 1, e1, 0, 0, 0, 0.0, 0, 0);
    const double t1 = kin<NL>(s1e, 0, -1, 0.0, 1.0);
    const double v1 = nuc_sum<NC, NBC>(s1e, 0, S.nat, S.cat, S.zan, true);
    auto ovl2 = [&](double g) {
        Sys s2p = mksys(S, 2, e2, p, 0, 0, g, 0, 0);
        return ovl<NC>(s2p);
    };
    double gTE = 0, gVEC = 0, gVEE = 0, gVEP = 0;
    if constexpr (!MULT)
    for (int ik = 0; ik < S.ngtg; ++ik) {
        double gk = S.gamma[ik], bk = S.bcoef[ik];
        Sys s = mksys(S, 3, e1, e2, p, 0, 0.0, gk, 0); // g(e2,p)
        double o2 = ovl2(gk);
        double xgTE = t1 * o2;
        double xgVEC = -(v1 * o2);
        double xgVEP = coul<NC, NBC>(s, 0, 2);
        underflow(xgVEP);
        xgVEP = -xgVEP;
        double xgVEE = coul<NC, NBC>(s, 0, 1);
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
            double kk[2], kq[2], ov, vp;
            kin_pair<NL, NC, NBC, true>(sb, KinOp{0, 2, 2, {0.0, gk}, 1.0}, KinOp{2, 1, 1, {gl, gl}, S.pmass}, PI, kk,
                                        kq, &ov, &vp, S.nat, S.cat, S.zan, true);
            gTEg1 += bb * kk[0];
            gTEg2 += bb * kk[1];
            gTPg += bb * kq[0];
            gVPCg += bb * vp;
            gVECg1 -= bb * nuc_sum<NC, NBC>(sb, 0, S.nat, S.cat, S.zan, true);
            double v = coul<NC, NBC>(sb, 0, 2);
            underflow(v);
            gVEPg1 -= bb * v;
            v = coul<NC, NBC>(sb, 0, 1);
            underflow(v);
            gVEEg2 += bb * v;
            gsg += bb * ov;
            if (il < ik) continue;
            double w = (il == ik) ? bb : 2.0 * bb;
            Sys s2 = mksys(S, 3, e1, e2, p, 0, 0.0, gk + gl, 0); // g(e2,p)=gk+gl
            Sys s1 = mksys(S, 3, e1, e2, p, 0, gk + gl, 0.0, 0); // g(e1,p)=gk+gl
            double o2 = ovl2(gk + gl);
            gTEg3 += w * (t1 * o2);
            gVECg2 -= w * (v1 * o2);
            v = coul<NC, NBC>(s2, 0, 2);
            underflow(v);
            gVEPg2 -= w * v;
            v = coul<NC, NBC>(MULT ? s2 : s1, 0, 1);
            underflow(v);
            gVEEg1 += w * v;
        }
    ansE = 2.0 * (gTE + gVEC + gVEP) + 4.0 * gVEE * 0.5 + gTPg + gVPCg + gTEg1 + gTEg3 + 2.0 * (gVEPg1 + gVECg1) +
           gTEg2 + gVEPg2 + gVECg2 + 2.0 * gVEEg1 * 0.5 + 2.0 * gVEEg2 * 0.5;
    ansS = gsg;
    ansE2 = gVEPg2 + gVEEg1;
    ansE3 = gVEEg2;
}

template <int NL, int NBC>
NRX_HD void gam2_tuple(const SysParams &S, const PairRec *e1, const PairRec *e2, const PairRec *p, double &ansE,
                       double &ansS)
{
    double unused, unused3;
    gam2_core<NL, NBC, false>(S, e1, e2, p, ansE, ansS, unused, unused3);
}

// xcalc_GAM3_MD (gam_3_OMP.f:617-983).  With MULT set:
// RXCHFmult_XCHF_GAM3_MD (rxchfmult/rxchfmult_gam3-xchf.f:630-1023), which
// drops the one-geminal gVEE term and relabels the electrons of the other
// terms; the relabelling is undone by the six-permutation average its
// builder applies, so the terms are evaluated here with the labels above.
template <int NL, int NBC, bool MULT = false>
NRX_HD void gam3_tuple(const SysParams &S, const PairRec *e1, const PairRec *e2, const PairRec *e3, const PairRec *p,
                       double &ans)
{
    // geminal-independent factors
    Sys s23 = mksys(S, 2, e2, e3, 0, 0, 0.0, 0, 0);
    double vee23 = coul<NL, NBC>(s23, 0, 1);
    underflow(vee23);
    Sys s3 = mksys(S, 1, e3, 0, 0, 0, 0, 0, 0);
    double hcore = kin<NL>(s3, 0, -1, 0.0, 1.0) - nuc_sum<NL, NBC>(s3, 0, S.nat, S.cat, S.zan, false);
    underflow(hcore);
    double gVEE = 0.0;
    for (int ik = 0; ik < (MULT ? 0 : S.ngtg); ++ik) {
        Sys s = mksys(S, 2, e1, p, 0, 0, S.gamma[ik], 0, 0);
        double xx = ovl<NL>(s);
        underflow(xx);
        gVEE += S.bcoef[ik] * (xx * vee23);
    }
    double gHEg = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            double gk = S.gamma[ik], gl = S.gamma[il];
            double bb = S.bcoef[ik] * S.bcoef[il];
            // 1/r34, g14 g24; its hub star (e1, e2, p) is the overlap yy
            Sys s4a = mksys(S, 4, e1, e2, e3, p, gk, gl, 0.0);
            double yy;
            double zz = coul<NL, NBC>(s4a, 2, 3, PI, &yy);
            underflow(yy);
            underflow(zz);
            gHEg += bb * ((hcore * yy) + (zz * (-1.0)));
            Sys s4b = mksys(S, 4, e1, e2, e3, p, gk, 0.0, gl); // 1/r12, g14 g34
            gVEEg2 += bb * coul<NL, NBC>(s4b, 0, 1);
            if (il < ik) continue;
            double w = (il == ik) ? bb : 2.0 * bb;
            Sys s2 = mksys(S, 2, e1, p, 0, 0, gk + gl, 0, 0);
            double xx = ovl<NL>(s2);
            underflow(xx);
            gVEEg1 += w * (xx * vee23);
        }
    if constexpr (MULT)
        ans = gHEg + gVEEg1 * 0.5 + 4.0 * gVEEg2 * 0.5;
    else
        ans = 2.0 * gVEE * 0.5 + gHEg + gVEEg1 * 0.5 + 4.0 * gVEEg2 * 0.5;
}

// RXCHF_xcalc_GAM3_MD (rxchf/rxchf_gam_3_OMP.f:617-1032): e1 the special
// electron, e2 and e3 regular, p the proton; only the two-geminal terms, two
// outputs.  One configuration per (k,l), g(e1,p) = gk, g(e2,p) = gl, e3
// unlinked, carries every 4-particle term:
//   ans1 = gVEEg1 = <(e1,p) overlap, g = gk + gl> x <1/r23>
//   ans2 = gHEg + gVEEg2 + gVEEg3 = hcore(e3) yy - <1/r(e3,p)> + <1/r13> + <1/r23>
// (the reference's gVEEg3 puts gk on e2 and gl on e1; summed over all (k,l)
// with the symmetric weights bcoef(k) bcoef(l) it equals <1/r23> of this
// configuration).
template <int NL, int NBC>
NRX_HD void rxchf3_tuple(const SysParams &S, const PairRec *e1, const PairRec *e2, const PairRec *e3,
                         const PairRec *p, double &ans1, double &ans2)
{
    Sys s23 = mksys(S, 2, e2, e3, 0, 0, 0.0, 0, 0);
    double vee23 = coul<NL, NBC>(s23, 0, 1);
    underflow(vee23);
    Sys s3 = mksys(S, 1, e3, 0, 0, 0, 0, 0, 0);
    double hcore = kin<NL>(s3, 0, -1, 0.0, 1.0) - nuc_sum<NL, NBC>(s3, 0, S.nat, S.cat, S.zan, false);
    underflow(hcore);
    double gHEg = 0, gVEEg1 = 0, gVEE23 = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            double gk = S.gamma[ik], gl = S.gamma[il];
            double bb = S.bcoef[ik] * S.bcoef[il];
            Sys s4 = mksys(S, 4, e1, e2, e3, p, gk, gl, 0.0);
            double yy;
            double zz = coul<NL, NBC>(s4, 2, 3, PI, &yy);
            underflow(yy);
            underflow(zz);
            gHEg += bb * ((hcore * yy) + (zz * (-1.0)));
            gVEE23 += bb * (coul<NL, NBC>(s4, 0, 2) + coul<NL, NBC>(s4, 1, 2));
            if (il < ik) continue;
            double w = (il == ik) ? bb : 2.0 * bb;
            Sys s2 = mksys(S, 2, e1, p, 0, 0, gk + gl, 0, 0);
            double xx = ovl<NL>(s2);
            underflow(xx);
            gVEEg1 += w * (xx * vee23);
        }
    ans1 = gVEEg1;
    ans2 = gHEg + gVEE23;
}

// RXCHFne_xcalc_GAM2_MD (rxchfne/rxchfne_gam_2_OMP.f:2-530): the RXCHF-ne
// e-e-p tuple, two outputs summed over all (k,l):
//   ansE1 = -<1/r(e2,p)> with g(e1,p) = gamma_k + gamma_l  +  <1/r12> with the same geminal
//   ansE2 =  <1/r12> with g(e1,p) = gamma_k, g(e2,p) = gamma_l
// RXCHF_xcalc_GAM2_MD (rxchf/rxchf_gam_2_OMP.f:546-1120): e1 the special
// electron, e2 a regular one, p the proton; only the two-geminal terms, in
// three outputs.  Configuration sb: g(e1,p) = gk on the bra, g(e2,p) = gl on
// the ket; configuration s1: g(e1,p) = gk + gl (e2 unlinked, so its kinetic
// and nuclear-attraction integrals factor into e2's one-particle value times
// the (e1, p) overlap):
//   ansE1 = gTEg3 + gVECg3 + gVEPg3 + gVEEg1       (s1: e2's one-body terms,
//           V(e2,p), V(e1,e2))
//   ansE2 = gTPg + gVPCg + gTEg1 + gVECg1 + gTEg2 + gVECg2 + gVEPg1 + gVEPg2
//           + gVEEg2                                (sb)
//   ansS  = gsg                                     (sb overlap)
template <int NL, int NBC>
NRX_HD void rxchf2_tuple(const SysParams &S, const PairRec *e1, const PairRec *e2, const PairRec *p, double &ansE1,
                         double &ansE2, double &ansS)
{
    constexpr int NC = NL - 2; // capacity of the terms without a kinetic operator
    Sys s2e = mksys(S, 1, e2, 0, 0, 0, 0.0, 0, 0);
    const double t2 = kin<NL>(s2e, 0, -1, 0.0, 1.0);
    const double v2 = nuc_sum<NC, NBC>(s2e, 0, S.nat, S.cat, S.zan, true);
    auto ovl1 = [&](double g) {
        Sys s1p = mksys(S, 2, e1, p, 0, 0, g, 0, 0);
        return ovl<NC>(s1p);
    };
    double gTPg = 0, gTEg1 = 0, gTEg2 = 0, gTEg3 = 0, gVPCg = 0, gVECg1 = 0, gVECg2 = 0, gVECg3 = 0;
    double gVEPg1 = 0, gVEPg2 = 0, gVEPg3 = 0, gVEEg1 = 0, gVEEg2 = 0, gsg = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            const double gk = S.gamma[ik], gl = S.gamma[il];
            const double bb = S.bcoef[ik] * S.bcoef[il];
            Sys sb = mksys(S, 3, e1, e2, p, 0, gk, gl, 0);
            // e1's kinetic energy (no ket geminal on e1) and the proton's (ket
            // geminal gl to e2) share the configuration's pass, with the
            // proton's nuclear attraction and the overlap
            double kk[2], kq[2], ov, vp;
            kin_pair<NL, NC, NBC, true>(sb, KinOp{0, 2, 1, {0.0, 0.0}, 1.0}, KinOp{2, 1, 1, {gl, gl}, S.pmass}, PI, kk,
                                        kq, &ov, &vp, S.nat, S.cat, S.zan, true);
            gTEg1 += bb * kk[0];
            gTPg += bb * kq[0];
            gVPCg += bb * vp;
            gsg += bb * ov;
            gTEg2 += bb * kin<NL>(sb, 1, 2, gl, 1.0); // e2, ket geminal gl to p
            gVECg1 -= bb * nuc_sum<NC, NBC>(sb, 0, S.nat, S.cat, S.zan, true);
            gVECg2 -= bb * nuc_sum<NC, NBC>(sb, 1, S.nat, S.cat, S.zan, true);
            double v = coul<NC, NBC>(sb, 0, 2);
            underflow(v);
            gVEPg1 -= bb * v;
            v = coul<NC, NBC>(sb, 1, 2);
            underflow(v);
            gVEPg2 -= bb * v;
            v = coul<NC, NBC>(sb, 0, 1);
            underflow(v);
            gVEEg2 += bb * v;
            if (il < ik) continue;
            const double w = (il == ik) ? bb : 2.0 * bb; // s1 depends on gk + gl only
            Sys s1 = mksys(S, 3, e1, e2, p, 0, gk + gl, 0.0, 0);
            const double o1 = ovl1(gk + gl);
            gTEg3 += w * (t2 * o1);
            gVECg3 -= w * (v2 * o1);
            v = coul<NC, NBC>(s1, 1, 2);
            underflow(v);
            gVEPg3 -= w * v;
            v = coul<NC, NBC>(s1, 0, 1);
            underflow(v);
            gVEEg1 += w * v;
        }
    ansE1 = gTEg3 + gVECg3 + gVEPg3 + gVEEg1;
    ansE2 = gTPg + gVPCg + gTEg1 + gVECg1 + gTEg2 + gVECg2 + gVEPg1 + gVEPg2 + gVEEg2;
    ansS = gsg;
}

template <int NL, int NBC>
NRX_HD void gam2ne_tuple(const SysParams &S, const PairRec *e1, const PairRec *e2, const PairRec *p, double &ansE1,
                         double &ansE2)
{
    double gVEPg2 = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            double gk = S.gamma[ik], gl = S.gamma[il];
            double bb = S.bcoef[ik] * S.bcoef[il];
            Sys s1 = mksys(S, 3, e1, e2, p, 0, gk + gl, 0.0, 0);
            double vep = coul<NL, NBC>(s1, 1, 2);
            underflow(vep);
            double vee1 = coul<NL, NBC>(s1, 0, 1);
            underflow(vee1);
            Sys s2 = mksys(S, 3, e1, e2, p, 0, gk, gl, 0);
            double vee2 = coul<NL, NBC>(s2, 0, 1);
            underflow(vee2);
            gVEPg2 += bb * (-vep);
            gVEEg1 += bb * vee1;
            gVEEg2 += bb * vee2;
        }
    ansE1 = gVEPg2 + gVEEg1;
    ansE2 = gVEEg2;
}

// RXCHFmult_INT_GAM3_MD (rxchfmult/rxchfmult_gam3-int.f:661-1054): e1 from
// the regular basis, e2 and e3 from the special-electron basis.
//   xgVEPg1 = -<1/r(e1,p)> with g(e2,p) = gamma_k, g(e3,p) = gamma_l
//   xgVEEg1 = <g(e2,p) = gamma_k + gamma_l> * <1/r(e1,e3)>
//   xgVEEg2 = <1/r(e1,e2)> with g(e2,p) = gamma_k, g(e3,p) = gamma_l
//   ans = gVEPg1 + gVEEg1 + 2 gVEEg2
// and, for its exchange form RXCHFmult_xcalc_GAM3_MDex (rxchfmult_gam3ex.f:712-1111),
// ans3 = 2 gVEEg2, ans4 = gVEEg1
template <int NL, int NBC>
NRX_HD void gam3int_tuple(const SysParams &S, const PairRec *e1, const PairRec *e2, const PairRec *e3,
                          const PairRec *p, double &ans, double &ans3, double &ans4)
{
    Sys s13 = mksys(S, 2, e1, e3, 0, 0, 0.0, 0, 0);
    double vee13 = coul<NL, NBC>(s13, 0, 1);
    underflow(vee13);
    double gVEPg1 = 0, gVEEg1 = 0, gVEEg2 = 0;
    for (int ik = 0; ik < S.ngtg; ++ik)
        for (int il = 0; il < S.ngtg; ++il) {
            double gk = S.gamma[ik], gl = S.gamma[il];
            double bb = S.bcoef[ik] * S.bcoef[il];
            Sys s4 = mksys(S, 4, e1, e2, e3, p, 0.0, gk, gl);
            double zz = coul<NL, NBC>(s4, 0, 3);
            underflow(zz);
            gVEPg1 += bb * (zz * (-1.0));
            gVEEg2 += bb * coul<NL, NBC>(s4, 0, 1);
            if (il < ik) continue;
            double w = (il == ik) ? bb : 2.0 * bb;
            Sys s2 = mksys(S, 2, e2, p, 0, 0, gk + gl, 0, 0);
            double xx = ovl<NL>(s2);
            underflow(xx);
            gVEEg1 += w * (xx * vee13);
        }
    ans = gVEPg1 + gVEEg1 + 2.0 * gVEEg2;
    ans3 = 2.0 * gVEEg2;
    ans4 = gVEEg1;
}

// Capacity tiers per tuple class: tier k evaluates with polynomial capacity
// NL[k] and Boys-order capacity NB[k].  A tuple runs in the first tier that
// covers its degree (tuple_degree <= NL-1) and its total angular momentum
// (tuple_L <= NB-1: the degree in u^2 of every Coulomb polynomial).
// Capacities past a tuple's need only add zero terms, so every tier gives
// the same values.
#if NRX_LMAX == 2
// The last tier covers every tuple allowed by l <= 2 (per particle and
// direction bra+ket powers <= 4): e-p 4+4+2 and L <= 8, e-e-p 3*4+2 and
// L <= 12, e-e-e-p 4*4 and L <= 16.  A tuple beyond it (an f function) has
// the wide tier N: the l <= 3 build evaluates it.
template <int CLS> struct Tier;
template <> struct Tier<1> {
    static constexpr int N = 4;
    static constexpr int NL[N] = {5, 6, 7, 11}, NB[N] = {7, 7, 9, 9};
    static constexpr bool KIN = true;
};
template <> struct Tier<2> {
    static constexpr int N = 5;
    static constexpr int NL[N] = {5, 6, 7, 9, 15}, NB[N] = {7, 9, 9, 13, 13};
    static constexpr bool KIN = true;
};
template <> struct Tier<3> {
    static constexpr int N = 4;
    static constexpr int NL[N] = {6, 7, 9, 17}, NB[N] = {9, 11, 17, 17};
    static constexpr bool KIN = false;
};
// class 4: the RXCHF-ne e-e-p tuple (Coulomb terms only: 3*4)
template <> struct Tier<4> {
    static constexpr int N = 3;
    static constexpr int NL[N] = {7, 9, 13}, NB[N] = {9, 13, 13};
    static constexpr bool KIN = false;
};
constexpr int NL_EE = 9; // two-electron repulsion only: 4+4
#else
// One tier covering every tuple allowed by l <= 3 (bra+ket powers <= 6 per
// particle and direction): e-p 6+6+2 and L <= 12, e-e-p 3*6+2 and L <= 18,
// e-e-e-p 4*6 and L <= 24, RXCHF-ne 3*6 and L <= 18.
template <int CLS> struct Tier;
template <> struct Tier<1> {
    static constexpr int N = 1;
    static constexpr int NL[N] = {15}, NB[N] = {13};
    static constexpr bool KIN = true;
};
template <> struct Tier<2> {
    static constexpr int N = 1;
    static constexpr int NL[N] = {21}, NB[N] = {19};
    static constexpr bool KIN = true;
};
template <> struct Tier<3> {
    static constexpr int N = 1;
    static constexpr int NL[N] = {25}, NB[N] = {25};
    static constexpr bool KIN = false;
};
template <> struct Tier<4> {
    static constexpr int N = 1;
    static constexpr int NL[N] = {19}, NB[N] = {19};
    static constexpr bool KIN = false;
};
constexpr int NL_EE = 13; // two-electron repulsion only: 6+6
#endif
// class 5: the RXCHFmult e-e-p tuple (two bases, kinetic terms as class 2);
// class 6: the RXCHFmult e-e-e-p interaction tuple (Coulomb and overlap: 4*4);
// class 7: the RXCHFmult XCHF e-e-e-p tuple (as class 3 without gVEE);
// class 8: the RXCHFmult e-p tuple (as class 1 without the one-geminal terms)
template <> struct Tier<5> : Tier<2> {};
template <> struct Tier<6> : Tier<3> {};
template <> struct Tier<7> : Tier<3> {};
template <> struct Tier<8> : Tier<1> {};
// class 9: the RXCHF (NBE = 1) e-e-p tuple (kinetic terms as class 2);
// class 10: the RXCHF e-e-e-p tuple (as class 3)
template <> struct Tier<9> : Tier<2> {};
template <> struct Tier<10> : Tier<3> {};

// electrons per tuple of a class
template <int CLS> constexpr int nel()
{
    return (CLS == 4 || CLS == 5 || CLS == 9) ? 2 : (CLS == 6 || CLS == 7 || CLS == 10) ? 3 : CLS == 8 ? 1 : CLS;
}

// q[0..CLS-1] electron pairs, q[CLS] proton pair
template <int CLS, int NL, int NBC = NB>
NRX_HD void tuple_eval(const SysParams &S, const PairRec *const *q, double &e, double &s)
{
    s = 0.0;
    if constexpr (CLS == 1)
        gam1_tuple<NL, NBC>(S, q[0], q[1], e, s);
    else if constexpr (CLS == 8)
        gam1_tuple<NL, NBC, true>(S, q[0], q[1], e, s);
    else if constexpr (CLS == 7)
        gam3_tuple<NL, NBC, true>(S, q[0], q[1], q[2], q[3], e);
    else if constexpr (CLS == 2)
        gam2_tuple<NL, NBC>(S, q[0], q[1], q[2], e, s);
    else if constexpr (CLS == 4)
        gam2ne_tuple<NL, NBC>(S, q[0], q[1], q[2], e, s);
    else if constexpr (CLS == 5) {
        double x, y;
        gam2_core<NL, NBC, true>(S, q[0], q[1], q[2], e, s, x, y);
    } else if constexpr (CLS == 9) {
        double x;
        rxchf2_tuple<NL, NBC>(S, q[0], q[1], q[2], e, x, s);
    } else if constexpr (CLS == 10) {
        rxchf3_tuple<NL, NBC>(S, q[0], q[1], q[2], q[3], e, s);
    } else if constexpr (CLS == 6) {
        double x;
        gam3int_tuple<NL, NBC>(S, q[0], q[1], q[2], q[3], e, s, x);
    }
    else
        gam3_tuple<NL, NBC>(S, q[0], q[1], q[2], q[3], e);
}

// four-output form: (ansE1, ansS, ansE2, ansE3) of class 5, (ans, ans3, ans4,
// 0) of class 6, (ansE1, ansS, ansE2, 0) of class 9; x and y are zero for the
// other classes
template <int CLS, int NL, int NBC = NB>
NRX_HD void tuple_eval(const SysParams &S, const PairRec *const *q, double &e, double &s, double &x, double &y)
{
    x = 0.0;
    y = 0.0;
    if constexpr (CLS == 5)
        gam2_core<NL, NBC, true>(S, q[0], q[1], q[2], e, s, x, y);
    else if constexpr (CLS == 9)
        rxchf2_tuple<NL, NBC>(S, q[0], q[1], q[2], e, x, s);
    else if constexpr (CLS == 6)
        gam3int_tuple<NL, NBC>(S, q[0], q[1], q[2], q[3], e, s, x);
    else
        tuple_eval<CLS, NL, NBC>(S, q, e, s);
}

template <int CLS, int NL, int NBC = NB>
NRX_HD void tuple_eval(const SysParams &S, const PairRec *const *q, double &e, double &s, double &x)
{
    double y;
    tuple_eval<CLS, NL, NBC>(S, q, e, s, x, y);
}

// total angular momentum of a tuple: bra+ket powers summed over particles
// and directions
NRX_HD int tuple_L(const PairRec *const *q, int n)
{
    int L = 0;
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) L += q[i]->la[d] + q[i]->lb[d];
    return L;
}

// the capacity tier a tuple runs in.  The e-e-e-p tuples of classes 3, 7 and
// 10 also apply a one-particle kinetic operator to e3 (their hcore term), whose
// polynomial needs e3's own largest per-direction degree + 2.
template <int CLS> NRX_HD int tuple_tier(const PairRec *const *q)
{
    using T = Tier<CLS>;
    // a Cartesian power above this build's l needs the other build
    int lc = 0;
    NRX_UNROLL for (int i = 0; i < nel<CLS>() + 1; ++i)
        NRX_UNROLL for (int d = 0; d < 3; ++d) {
            const int m = q[i]->la[d] > q[i]->lb[d] ? q[i]->la[d] : q[i]->lb[d];
            lc = m > lc ? m : lc;
        }
    if (lc > NRX_LMAX) return T::N;
    int deg = tuple_degree(q, nel<CLS>() + 1, T::KIN), L = tuple_L(q, nel<CLS>() + 1);
    if constexpr (CLS == 3 || CLS == 7 || CLS == 10) {
        NRX_UNROLL for (int d = 0; d < 3; ++d) {
            int k = q[2]->la[d] + q[2]->lb[d] + 2;
            deg = k > deg ? k : deg;
        }
    }
    for (int k = 0; k < T::N; ++k)
        if (deg <= T::NL[k] - 1 && L <= T::NB[k] - 1) return k;
    return T::N; // beyond this build's capacities
}

// the general tier of a class (covers every tuple)
template <int CLS> constexpr int tier_nl() { return Tier<CLS>::NL[Tier<CLS>::N - 1]; }
template <int CLS> constexpr int tier_nb() { return Tier<CLS>::NB[Tier<CLS>::N - 1]; }

} // inline namespace NRX_CAP_NS
} // namespace nrx
