// nrx_wide.hip -- the integral engine at the capacities of l <= 3 (f
// functions), for the tuples that the l <= 2 tier kernels of nrx_lib.hip
// cannot hold.
//
// The tier kernels bin every tuple; a tuple whose degree or total angular
// momentum exceeds their general tier (some particle carries an f function)
// goes to one extra bin, the wide bin, which k_prim_wide evaluates here with
// the engine compiled for NRX_LMAX = 3 (one capacity per class: e-p 15/13,
// e-e-p 21/19, e-e-e-p 25/25, RXCHF-ne 19/19).  Keeping these capacities in
// their own translation unit leaves the registers and code of the l <= 2
// kernels as they are.  The contracted builders that evaluate a class at its
// general capacity (GAM1_OMP_MD, calc_GAM_ee) have their l <= 3 instances
// here as well.
#define NRX_LMAX 3
#include "nrx_tuples.h"

namespace {

// lanes per block of the wide evaluation: one wavefront, so that blocks are
// scheduled per SIMD
constexpr int WIDE_BS = 64;
// most blocks of one wide launch; the blocks stride over the bin
constexpr i64 WIDE_MAX_BLOCKS = 4096;

// Every tuple of the wide bin: lane per listed tuple, grid-stride over the
// bin's count (read on the device; blocks past it exit at once).
template <int CLS, int SRC>
__global__ void __launch_bounds__(WIDE_BS) k_prim_wide(DevSys S, i64 start, const i64 *__restrict__ idx,
                                                       double *__restrict__ outE, double *__restrict__ outS,
                                                       double *__restrict__ outX, double *__restrict__ outY,
                                                       const unsigned *__restrict__ list,
                                                       const unsigned *__restrict__ counts, int wbin)
{
    constexpr int NP = nel<CLS>() + 1;
    const i64 n = counts[wbin];
    const unsigned off = bin_offset(counts, wbin);
    for (i64 k = (i64)blockIdx.x * WIDE_BS + threadIdx.x; k < n; k += (i64)gridDim.x * WIDE_BS) {
        const i64 t = (i64)list[off + k];
        i64 x[2 * NP];
        src_tuple<CLS, SRC>(S, idx, start, t, x); // a list was validated by the binning pass
        const PairRec *q[NP];
        tuple_pairs<CLS>(S, x, q);
        double e, s, xx, yy;
        tuple_eval<CLS, tier_nl<CLS>(), tier_nb<CLS>()>(d_params(S), q, e, s, xx, yy);
        outE[t] = e;
        if (outS) outS[t] = s;
        if (outX) outX[t] = xx;
        if (outY) outY[t] = yy;
    }
}

template <int CLS, int SRC>
int launch_wide(const DevSys &d, i64 start, const i64 *idx, double *oe, double *os, double *ox, double *oy,
                const unsigned *list, const unsigned *counts, int wbin, i64 n, hipStream_t st)
{
    const unsigned nb = (unsigned)std::min<i64>(WIDE_MAX_BLOCKS, std::max<i64>(1, (n + WIDE_BS - 1) / WIDE_BS));
    hipLaunchKernelGGL((k_prim_wide<CLS, SRC>), dim3(nb), dim3(WIDE_BS), 0, st, d, start, idx, oe, os, ox, oy, list,
                       counts, wbin);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

template <int CLS>
int launch_wide_src(int src, const DevSys &d, i64 start, const i64 *idx, double *oe, double *os, double *ox,
                    double *oy, const unsigned *list, const unsigned *counts, int wbin, i64 n, hipStream_t st)
{
    switch (src) {
    case SRC_LIST: return launch_wide<CLS, SRC_LIST>(d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case SRC_ORIG: return launch_wide<CLS, SRC_ORIG>(d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case SRC_SORTED:
        if constexpr (!two_basis<CLS>())
            return launch_wide<CLS, SRC_SORTED>(d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
        [[fallthrough]];
    default: return fail(NEORX_E_ARG, "wide tuples: class %d has no tuple source %d", CLS, src);
    }
}

} // namespace

int nrx_wide_prims(int cls, int src, const DevSys &d, i64 start, const i64 *idx, double *oe, double *os, double *ox,
                   double *oy, const unsigned *list, const unsigned *counts, int wbin, i64 n, hipStream_t st)
{
    switch (cls) {
    case 1: return launch_wide_src<1>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 2: return launch_wide_src<2>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 3: return launch_wide_src<3>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 4: return launch_wide_src<4>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 5: return launch_wide_src<5>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 6: return launch_wide_src<6>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 7: return launch_wide_src<7>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 8: return launch_wide_src<8>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 9: return launch_wide_src<9>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    case 10: return launch_wide_src<10>(src, d, start, idx, oe, os, ox, oy, list, counts, wbin, n, st);
    default: return fail(NEORX_E_ARG, "wide tuples: unknown class %d", cls);
    }
}

int nrx_wide_gam1(int cls, const DevSys &d, double *G, double *Gs, hipStream_t st)
{
    const i64 ne = d.nebf, np = d.npbf, n = ne * ne * np * np;
    if (cls == 1)
        hipLaunchKernelGGL((k_gam1<1, 3>), dim3(nblocks(n, 256)), dim3(256), 0, st, d, G, Gs);
    else if (cls == 8)
        hipLaunchKernelGGL((k_gam1<8, 3>), dim3(nblocks(n, 256)), dim3(256), 0, st, d, G, Gs);
    else
        return fail(NEORX_E_ARG, "GAM1 builder: class %d is not an e-p class", cls);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

int nrx_wide_gam_ee(const DevSys &d, double *G, hipStream_t st)
{
    const i64 ne = d.nebf, n = ne * ne * ne * ne;
    hipLaunchKernelGGL(k_gam_ee<3>, dim3(nblocks(n, 256)), dim3(256), 0, st, d, G);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}
