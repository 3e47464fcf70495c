// nrx_internal.h -- declarations shared by the translation units of
// libneorx_hip.so (nrx_lib.hip: integral kernels, builders and the
// reference-semantics contractions; nrx_scf.hip: the device-resident XCHF
// Fock build and SCF).  Nothing here is part of the C ABI: the functions are
// hidden from the shared library's symbol table.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/neorx.h"
#include "nrx_engine.h"

typedef int64_t i64;
#define NRX_DI __device__ inline __attribute__((always_inline))
#define NRX_HIDDEN __attribute__((visibility("hidden")))
using namespace nrx;

// error status of the calling thread (neorx_last_error); returns code
NRX_HIDDEN int fail(int code, const char *fmt, ...);

#define HIPCHK(x)                                                                                                  \
    do {                                                                                                           \
        hipError_t _e = (x);                                                                                       \
        if (_e != hipSuccess) return fail(NEORX_E_HIP, "%s failed: %s", #x, hipGetErrorString(_e));              \
    } while (0)

// NEORX_VERBOSE > 0: long builds report their progress on stderr
NRX_HIDDEN extern int g_verbose;

// neorx_init(0) unless a device is already selected
NRX_HIDDEN int ensure_init();

// ------------------------------------------------------------------
// system handle
// ------------------------------------------------------------------
struct DevSys {
    i64 nebf, npebf, npbf, nat, ngtg;
    i64 nebf1, npebf1;        // first electron basis; the rest is the special-electron basis (RXCHFmult)
    double pmass;
    const double *cat, *zan, *bcoef, *gamma, *agebfcc, *agnbfcc;
    const i64 *kpestr, *kpeend, *ampeb2c;
    const i64 *eord, *pord;   // enumeration order -> original primitive index
    const PairRec *ep, *pp;   // bra*ket tables [i][j] over primitives (npebf^2, npbf^2)
    const double *bt;         // Boys table (BT_N x BT_M)
    int lmax;                 // largest primitive angular momentum (3: f functions present)
};

struct System {
    DevSys d;
    std::vector<void *> allocs;
    std::vector<i64> eord, pord;
    std::vector<i64> kpestr; // host copy (1-based)
    int maxl;
};

NRX_HIDDEN System *get_sys(i64 h);

// ------------------------------------------------------------------
// the l <= 3 build of the engine (nrx_wide.hip)
// ------------------------------------------------------------------
// Tuples beyond the l <= 2 capacities (an f function on some particle) are
// binned by the tier kernels into the wide bin wbin of a launch's sorted
// list; this evaluates that bin (class cls, tuple source src) on stream st.
// n bounds the bin's size (the launch's tuple count).
NRX_HIDDEN int nrx_wide_prims(int cls, int src, const DevSys &d, i64 start, const i64 *idx, double *oe, double *os,
                              double *ox, double *oy, const unsigned *list, const unsigned *counts, int wbin, i64 n,
                              hipStream_t st);
// GAM1_OMP_MD (cls 1) / RXCHFmult_GAM1_MD (cls 8) and calc_GAM_ee at l <= 3
// capacities
NRX_HIDDEN int nrx_wide_gam1(int cls, const DevSys &d, double *G, double *Gs, hipStream_t st);
NRX_HIDDEN int nrx_wide_gam_ee(const DevSys &d, double *G, hipStream_t st);

// most blocks of a grid-stride launch (NEORX_GRID_CAP lowers it)
NRX_HIDDEN extern i64 g_grid_cap;

static inline unsigned nblocks(i64 n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Blocks for a grid-stride kernel over n items: at most g_grid_cap blocks, so
// that the grid's work-item count stays far below 2^32 whatever n is (the
// in-core GAM_3 of a 24-function basis with 8 nuclear functions has 1.2e10
// entries).
static inline unsigned grid_blocks(i64 n, int bs) { return (unsigned)std::min<i64>((n + bs - 1) / bs, g_grid_cap); }

// device buffer released when it goes out of scope
struct DevBuf {
    void *p = nullptr;
    bool own = true; // false: a resident tensor's memory, borrowed (not freed)
    ~DevBuf()
    {
        if (p && own) (void)hipFree(p);
    }
    int alloc(size_t bytes)
    {
        hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(NEORX_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        }
        return NEORX_OK;
    }
    double *d() { return (double *)p; }
};

// storage type of an in-core tensor read by the contractions
enum { NRX_F64 = 0, NRX_F32 = 1, NRX_BF16 = 2 };

// ------------------------------------------------------------------
// launchers defined in nrx_lib.hip, used by nrx_scf.hip.  They enqueue on st
// and neither allocate nor synchronise.
// ------------------------------------------------------------------
// GAM_ee (calc_GAM_ee) into d_gee (nebf^4)
NRX_HIDDEN int nrx_gam_ee_launch(System *s, double *d_gee, hipStream_t st);
// GAM_1, GAM_1s (GAM1_OMP_MD) into d_g1, d_g1s (nebf^2 npbf^2 each)
NRX_HIDDEN int nrx_gam1_launch(System *s, double *d_g1, double *d_g1s, hipStream_t st);
// EG3IC1 slab partials of nblk slabs (ip,jp,i1): part[blk][0..ne-1] = the FE
// row before the DP(ip,jp) weight, part[blk][ne] = its DE-weighted sum / 3;
// g of storage type dtype
NRX_HIDDEN int nrx_eg3_slabs(i64 ne, i64 nblk, const double *d_de, const void *g, int dtype, double *part,
                             hipStream_t st);
// the same rows as the symmetric bilinear form T(X, Y) of two densities (X on
// the (i2,j2) pair, Y on the (i3,j3) pair; T(DE, DE) are EG3IC1's rows), and
// part[blk][ne] = sum_x Z(i1,x) F[x] / 3, or 0 when d_z is null
NRX_HIDDEN int nrx_eg3_slabs2(i64 ne, i64 nblk, const double *d_x, const double *d_y, const double *d_z, const void *g,
                              int dtype, double *part, hipStream_t st);
// EG2IC1 slab partials of nblk slabs (ip,jp,i1): part[blk][4][ne] = dir, ex,
// dirS, exS (gs may be NULL: the S rows stay zero)
NRX_HIDDEN int nrx_eg2_slabs(i64 ne, i64 nblk, const double *d_de, const double *g, const double *gs, double *part,
                             hipStream_t st);
