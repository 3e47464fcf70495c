// nrx_tuples.h -- device helpers shared by the two builds of the integral
// engine in libneorx_hip.so (nrx_lib.hip: capacities for l <= 2;
// nrx_wide.hip: l <= 3): tuple sources, pair lookup, the sorted-list offsets
// of the tuple bins, and the contracted builders that evaluate the engine at
// a class's general capacity (templated on the build's NRX_LMAX so the two
// builds instantiate distinct kernels).
#pragma once
#include "nrx_internal.h"

// ------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------
NRX_DI SysParams d_params(const DevSys &S)
{
    SysParams P;
    P.nat = (int)S.nat;
    P.ngtg = (int)S.ngtg;
    P.pmass = S.pmass;
    P.cat = S.cat;
    P.zan = S.zan;
    P.bcoef = S.bcoef;
    P.gamma = S.gamma;
    P.bt = S.bt;
    return P;
}

// pair records of one tuple given original primitive indices
// (bra e..., ip, ket e..., jp)
template <int CLS> NRX_DI void tuple_pairs(const DevSys &S, const i64 *x, const PairRec **q)
{
    constexpr int E = nel<CLS>();
    for (int i = 0; i < E; ++i) q[i] = S.ep + x[i] * S.npebf + x[E + 1 + i];
    q[E] = S.pp + x[E] * S.npbf + x[2 * E + 1];
}

template <int CLS, int NL, int NBC> NRX_DI void eval_tuple(const DevSys &S, const i64 *x, double &e, double &s)
{
    const PairRec *q[nel<CLS>() + 1];
    tuple_pairs<CLS>(S, x, q);
    tuple_eval<CLS, NL, NBC>(d_params(S), q, e, s);
}

// index tuple t of a list, validated against the basis sizes
template <int CLS> NRX_DI bool load_tuple(const DevSys &S, const i64 *idx, i64 t, i64 *x)
{
    bool ok = true;
    constexpr int E = nel<CLS>();
    for (int k = 0; k < 2 * E + 2; ++k) {
        x[k] = idx[t * (2 * E + 2) + k];
        i64 lim = (k == E || k == 2 * E + 1) ? S.npbf : S.npebf;
        ok = ok && x[k] >= 0 && x[k] < lim;
    }
    return ok;
}

// the RXCHFmult classes whose electrons come from two bases
template <int CLS> constexpr bool two_basis() { return CLS == 5 || CLS == 6; }

// tuple r of the enumeration (ip,jp,ie1,je1,...,ieN,jeN), last index
// fastest, in the system's sorted primitive order (SORTED) or in the
// caller's original order.  The two-basis classes (5, 6) enumerate e1 over
// the first basis and the other electrons over the special-electron basis,
// in original order only.
template <int CLS, bool SORTED> NRX_DI void range_tuple(const DevSys &S, i64 r, i64 *x)
{
    static_assert(!two_basis<CLS>() || !SORTED, "two-basis classes are enumerated in original order");
    i64 np = S.npbf;
    constexpr int CLS_E = nel<CLS>();
    for (int i = CLS_E - 1; i >= 0; --i) {
        i64 ne = S.npebf, off = 0;
        if constexpr (two_basis<CLS>()) {
            ne = i == 0 ? S.npebf1 : S.npebf - S.npebf1;
            off = i == 0 ? 0 : S.npebf1;
        }
        i64 k = r % ne;
        x[CLS_E + 1 + i] = SORTED ? S.eord[k] : off + k;
        r /= ne;
        k = r % ne;
        x[i] = SORTED ? S.eord[k] : off + k;
        r /= ne;
    }
    i64 k = r % np;
    x[2 * CLS_E + 1] = SORTED ? S.pord[k] : k;
    r /= np;
    k = r % np;
    x[CLS_E] = SORTED ? S.pord[k] : k;
}

// Tuple sources: a caller's index list, or a range of the enumeration in
// sorted or in original primitive order.
enum { SRC_LIST = 0, SRC_SORTED = 1, SRC_ORIG = 2 };

template <int CLS, int SRC> NRX_DI void src_tuple(const DevSys &S, const i64 *idx, i64 start, i64 t, i64 *x)
{
    if constexpr (SRC == SRC_LIST)
        load_tuple<CLS>(S, idx, t, x);
    else
        range_tuple<CLS, SRC == SRC_SORTED>(S, start + t, x);
}

// offset of bin b in the sorted list: the counts of the bins before it
NRX_DI unsigned bin_offset(const unsigned *counts, int b)
{
    unsigned o = 0;
    for (int j = 0; j < b; ++j) o += counts[j];
    return o;
}

// GAM1_OMP_MD: one lane per contracted (ip,jp,iec,jec), primitive sums in
// the reference's loop order, at the general tier of the class.
// CLS 8: RXCHFmult_GAM1_MD (rxchfmult/rxchfmult_gam1.f:2-398), same loops.
template <int CLS, int LMAX>
__global__ void __launch_bounds__(256) k_gam1(DevSys S, double *__restrict__ G, double *__restrict__ Gs)
{
    static_assert(LMAX == NRX_LMAX, "instantiated by the build of its capacities");
    i64 ne = S.nebf, np = S.npbf;
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne * ne * np * np) return;
    i64 jc = t % ne, ic = (t / ne) % ne, jp = (t / (ne * ne)) % np, ip = t / (ne * ne * np);
    double acc = 0.0, accs = 0.0;
    for (i64 ie = S.kpestr[ic] - 1; ie < S.kpeend[ic]; ++ie)
        for (i64 je = S.kpestr[jc] - 1; je < S.kpeend[jc]; ++je) {
            i64 x[4] = {ie, ip, je, jp};
            double e, s;
            eval_tuple<CLS, tier_nl<CLS>(), tier_nb<CLS>()>(S, x, e, s);
            double cf = S.agnbfcc[ip] * S.agnbfcc[jp] * S.agebfcc[ie] * S.agebfcc[je];
            acc += e * cf;
            accs += s * cf;
        }
    G[t] = acc;
    Gs[t] = accs;
}

// calc_GAM_ee: lane per contracted quadruple, primitive sums in the
// reference's loop order (ie1, je1, ie2, je2)
template <int LMAX> __global__ void __launch_bounds__(256) k_gam_ee(DevSys S, double *__restrict__ G)
{
    static_assert(LMAX == NRX_LMAX, "instantiated by the build of its capacities");
    i64 ne = S.nebf;
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne * ne * ne * ne) return;
    i64 j2 = t % ne, i2 = (t / ne) % ne, j1 = (t / (ne * ne)) % ne, i1 = t / (ne * ne * ne);
    SysParams P = d_params(S);
    double acc = 0.0;
    for (i64 a1 = S.kpestr[i1] - 1; a1 < S.kpeend[i1]; ++a1)
        for (i64 c1 = S.kpestr[j1] - 1; c1 < S.kpeend[j1]; ++c1)
            for (i64 a2 = S.kpestr[i2] - 1; a2 < S.kpeend[i2]; ++a2)
                for (i64 c2 = S.kpestr[j2] - 1; c2 < S.kpeend[j2]; ++c2) {
                    const PairRec *q0 = S.ep + a1 * S.npebf + c1, *q1 = S.ep + a2 * S.npebf + c2;
                    Sys s = mksys(P, 2, q0, q1, 0, 0, 0.0, 0, 0);
                    acc += coul<NL_EE>(s, 0, 1) * S.agebfcc[a1] * S.agebfcc[c1] * S.agebfcc[a2] * S.agebfcc[c2];
                }
    G[t] = acc;
}

