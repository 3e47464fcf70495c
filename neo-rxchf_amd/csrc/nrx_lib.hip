// nrx_lib.hip -- libneorx_hip.so: HIP kernels for gfx950 and the C ABI of
// include/neorx.h.
//
// Layout in HBM:
//   * the basis and nuclei are copied once per system handle, in the
//     reference's own Fortran layouts (column-major, int64 indices);
//   * primitive tuples are enumerated over primitives re-ordered by
//     Cartesian power pattern and exponent, so that neighbouring lanes of a
//     wavefront share loop trip counts;
//   * contracted tensors use the reference packings (C row-major
//     [ip][jp][ie1][je1]...).
#include <atomic>
#include <chrono>
#include <map>
#include <mutex>
#include <numeric>

#include "nrx_internal.h"
#include "nrx_tuples.h"

// ------------------------------------------------------------------
// error handling
// ------------------------------------------------------------------
static thread_local std::string g_err;
static std::mutex g_mu;
static int g_device = -1;
static char g_arch[64] = "";
static double *g_boys_dev = nullptr; // Boys table in HBM, one per process

// Scratch of the allocation-free entry point neorx_eg3ic1_noalloc_dev: a
// file-scope device array of SCRATCH_DOUBLES doubles (16 MiB on every device
// the library runs on, held until the process ends).  In host code the
// array's name is not a device address: neorx_init looks its address up once
// per device (hipGetSymbolAddress, with that device current) and every host
// use goes through g_scratch[device].
static const i64 SCRATCH_DOUBLES = (i64)1 << 21;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
static const int MAX_DEVICES = 64;
static double *g_scratch[MAX_DEVICES] = {};

i64 g_grid_cap = (i64)1 << 20;       // most blocks of a grid-stride launch
int g_verbose = 0;                   // NEORX_VERBOSE: progress of long builds on stderr

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int ensure_init()
{
    if (g_device >= 0) return NEORX_OK;
    return neorx_init(0);
}

// ------------------------------------------------------------------
// system handle
// ------------------------------------------------------------------
static std::map<i64, System *> g_sys;
static i64 g_next = 1;

template <class T> static int upload(System *s, const T *h, size_t n, const T **out)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    s->allocs.push_back(p);
    if (h) HIPCHK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T *)p;
    return NEORX_OK;
}

System *get_sys(i64 h)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sys.find(h);
    return it == g_sys.end() ? nullptr : it->second;
}

extern "C" int neorx_init(int64_t device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(NEORX_E_NODEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(NEORX_E_ARG, "device %lld out of range (%d devices)", (long long)device, n);
    HIPCHK(hipSetDevice((int)device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, (int)device));
    snprintf(g_arch, sizeof(g_arch), "%s", prop.gcnArchName);
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(NEORX_E_NODEVICE, "device %lld is %s; this library is built for gfx950 only", (long long)device,
                    prop.gcnArchName);
    if (!g_boys_dev) {
        std::vector<double> tab(BT_N * BT_M);
        boys_table_fill(tab.data());
        HIPCHK(hipMalloc(&g_boys_dev, tab.size() * sizeof(double)));
        HIPCHK(hipMemcpy(g_boys_dev, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (device < MAX_DEVICES && !g_scratch[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NEORX_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch[device] = (double *)p;
    }
    g_verbose = getenv("NEORX_VERBOSE") ? atoi(getenv("NEORX_VERBOSE")) : 0;
    g_grid_cap = (i64)1 << 20;
    if (const char *cap = getenv("NEORX_GRID_CAP")) {
        long long v = atoll(cap);
        if (v < 1) return fail(NEORX_E_ARG, "NEORX_GRID_CAP=%s: expected a positive block count", cap);
        g_grid_cap = std::min<i64>(v, g_grid_cap);
    }
    g_device = (int)device;
    return NEORX_OK;
}

extern "C" int neorx_finalize(void)
{
    (void)neorx_release_resident();
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : g_sys) {
        for (void *p : kv.second->allocs) (void)hipFree(p);
        delete kv.second;
    }
    g_sys.clear();
    if (g_boys_dev) (void)hipFree(g_boys_dev);
    g_boys_dev = nullptr;
    std::fill(g_scratch, g_scratch + MAX_DEVICES, nullptr); // looked up again by the next neorx_init
    g_device = -1;
    return NEORX_OK;
}

extern "C" const char *neorx_last_error(void) { return g_err.c_str(); }
extern "C" const char *neorx_device_arch(void) { return g_arch; }

static Prim host_prim(i64 n, const double *ex, const i64 *am, const double *cen, i64 i)
{
    Prim p;
    p.a = ex[i];
    for (int d = 0; d < 3; ++d) {
        p.l[d] = (int)am[d * n + i];
        p.A[d] = cen[d * n + i];
    }
    return p;
}

static void pair_table(i64 n, const double *ex, const i64 *am, const double *cen, std::vector<PairRec> &tab)
{
    tab.resize(n * n);
    for (i64 i = 0; i < n; ++i)
        for (i64 j = 0; j < n; ++j) make_pair(host_prim(n, ex, am, cen, i), host_prim(n, ex, am, cen, j), tab[i * n + j]);
}

extern "C" int neorx_system_create(int64_t nebf, int64_t npebf, int64_t npbf, int64_t nat, int64_t ngtg, double pmass,
                                   const double *cat, const double *zan, const double *bcoef, const double *gamma,
                                   const int64_t *kpestr, const int64_t *kpeend, const int64_t *ampeb2c,
                                   const double *agebfcc, const double *agnbfcc, const double *elcex,
                                   const double *nucex, const int64_t *elcam, const int64_t *nucam,
                                   const double *elcbfc, const double *nucbfc, int64_t *handle)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (nebf <= 0 || npebf <= 0 || npbf <= 0 || nat < 0 || ngtg <= 0 || !handle)
        return fail(NEORX_E_ARG, "bad system sizes nebf=%lld npebf=%lld npbf=%lld nat=%lld ngtg=%lld",
                    (long long)nebf, (long long)npebf, (long long)npbf, (long long)nat, (long long)ngtg);
    int maxl = 0;
    for (i64 i = 0; i < npebf; ++i) {
        int l = (int)(elcam[i] + elcam[npebf + i] + elcam[2 * npebf + i]);
        for (int d = 0; d < 3; ++d)
            if (elcam[d * npebf + i] < 0) return fail(NEORX_E_ARG, "negative angular momentum");
        maxl = std::max(maxl, l);
    }
    for (i64 i = 0; i < npbf; ++i) {
        for (int d = 0; d < 3; ++d)
            if (nucam[d * npbf + i] < 0) return fail(NEORX_E_ARG, "negative angular momentum");
        maxl = std::max(maxl, (int)(nucam[i] + nucam[npbf + i] + nucam[2 * npbf + i]));
    }
    if (maxl > 3)
        return fail(NEORX_E_UNSUPPORTED, "angular momentum l=%d > 3 is not supported by this build", maxl);
    for (i64 c = 0; c < nebf; ++c)
        if (kpestr[c] < 1 || kpeend[c] < kpestr[c] || kpeend[c] > npebf)
            return fail(NEORX_E_ARG, "KPESTR/KPEEND out of range for contracted function %lld", (long long)c + 1);
    for (i64 i = 0; i < npebf; ++i)
        if (ampeb2c[i] < 1 || ampeb2c[i] > nebf) return fail(NEORX_E_ARG, "AMPEB2C out of range");
    // each contracted function owns a contiguous run of primitives, in order
    // (the layout make_KPE derives from AMPEB2C)
    for (i64 c = 0; c < nebf; ++c) {
        i64 expect = c == 0 ? 1 : kpeend[c - 1] + 1;
        if (kpestr[c] != expect) return fail(NEORX_E_ARG, "KPESTR(%lld) = %lld, expected %lld (primitives of contracted "
                                             "functions must be contiguous and in order)", (long long)c + 1,
                                             (long long)kpestr[c], (long long)expect);
        for (i64 i = kpestr[c] - 1; i < kpeend[c]; ++i)
            if (ampeb2c[i] != c + 1) return fail(NEORX_E_ARG, "AMPEB2C(%lld) disagrees with KPESTR/KPEEND", (long long)i + 1);
    }
    if (kpeend[nebf - 1] != npebf) return fail(NEORX_E_ARG, "KPEEND(NEBF) != NPEBF");

    System *s = new System();
    s->maxl = maxl;
    s->d.lmax = maxl;
    s->kpestr.assign(kpestr, kpestr + nebf);
    DevSys &d = s->d;
    d.nebf = nebf;
    d.npebf = npebf;
    d.nebf1 = nebf;
    d.npebf1 = npebf;
    d.npbf = npbf;
    d.nat = nat;
    d.ngtg = ngtg;
    d.pmass = pmass;
    d.bt = g_boys_dev;
    // enumeration order: by Cartesian power pattern, then exponent
    auto order = [](i64 n, const i64 *am, const double *ex, std::vector<i64> &ord) {
        ord.resize(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](i64 a, i64 b) {
            i64 la = am[a] + am[n + a] + am[2 * n + a], lb = am[b] + am[n + b] + am[2 * n + b];
            if (la != lb) return la < lb;
            for (int k = 0; k < 3; ++k)
                if (am[k * n + a] != am[k * n + b]) return am[k * n + a] > am[k * n + b];
            return ex[a] < ex[b];
        });
    };
    order(npebf, elcam, elcex, s->eord);
    order(npbf, nucam, nucex, s->pord);
    std::vector<PairRec> ept, ppt;
    pair_table(npebf, elcex, elcam, elcbfc, ept);
    pair_table(npbf, nucex, nucam, nucbfc, ppt);
    int r = 0;
    r |= upload(s, cat, 3 * nat, &d.cat);
    r |= upload(s, zan, nat, &d.zan);
    r |= upload(s, bcoef, ngtg, &d.bcoef);
    r |= upload(s, gamma, ngtg, &d.gamma);
    r |= upload(s, kpestr, nebf, &d.kpestr);
    r |= upload(s, kpeend, nebf, &d.kpeend);
    r |= upload(s, ampeb2c, npebf, &d.ampeb2c);
    r |= upload(s, agebfcc, npebf, &d.agebfcc);
    r |= upload(s, agnbfcc, npbf, &d.agnbfcc);
    r |= upload(s, s->eord.data(), npebf, &d.eord);
    r |= upload(s, s->pord.data(), npbf, &d.pord);
    r |= upload(s, ept.data(), ept.size(), &d.ep);
    r |= upload(s, ppt.data(), ppt.size(), &d.pp);
    if (r) {
        for (void *p : s->allocs) (void)hipFree(p);
        delete s;
        return NEORX_E_HIP;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    i64 h = g_next++;
    g_sys[h] = s;
    *handle = h;
    return NEORX_OK;
}

extern "C" int neorx_system_destroy(int64_t handle)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_sys.find(handle);
    if (it == g_sys.end()) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    for (void *p : it->second->allocs) (void)hipFree(p);
    delete it->second;
    g_sys.erase(it);
    return NEORX_OK;
}

extern "C" int neorx_system_order(int64_t handle, int64_t *eorder, int64_t *porder)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (eorder) std::copy(s->eord.begin(), s->eord.end(), eorder);
    if (porder) std::copy(s->pord.begin(), s->pord.end(), porder);
    return NEORX_OK;
}

// Tuples of a launch are sorted by bin, a counting sort in two passes:
// k_prim_bin gives each tuple its bin (capacity tier tuple_tier, and for the
// classes whose Coulomb terms take the hub-cycle path also the proton pair's
// largest per-direction degree, which bounds that path's series) and counts
// the bins, one atomic per wave and bin; k_prim_scatter writes the tuple
// positions bin by bin into one list; k_prim_tiers evaluates every bin with
// its tier's capacities, one lane per listed tuple, so that every wave runs
// at the capacity its tuples need.  Both passes count in LDS per block and
// touch each global counter once per block.  Out-of-range list entries are
// flagged (NaN results) and left out.
template <int CLS> constexpr int hub_bins() { return (CLS == 2 || CLS == 4 || CLS == 5 || CLS == 9) ? 5 : 1; }
template <int CLS> constexpr int n_bins() { return Tier<CLS>::N * hub_bins<CLS>(); }

template <int CLS> NRX_DI int tuple_bin(const PairRec *const *q)
{
    int k = tuple_tier<CLS>(q);
    if (k == Tier<CLS>::N) return n_bins<CLS>(); // the wide bin (l <= 3 build)
    if constexpr (hub_bins<CLS>() == 1) {
        return k;
    } else {
        const PairRec &P = *q[nel<CLS>()];
        int m = 0;
        NRX_UNROLL for (int d = 0; d < 3; ++d) m = max(m, P.la[d] + P.lb[d]);
        return k * hub_bins<CLS>() + min(m, hub_bins<CLS>() - 1);
    }
}

template <int CLS, int SRC>
__global__ void __launch_bounds__(256) k_prim_bin(DevSys S, i64 start, i64 ntup, const i64 *__restrict__ idx,
                                                  double *__restrict__ outE, double *__restrict__ outS,
                                                  double *__restrict__ outX, double *__restrict__ outY,
                                                  unsigned char *__restrict__ code, unsigned *__restrict__ counts)
{
    __shared__ unsigned lc[32];
    if (threadIdx.x < 32) lc[threadIdx.x] = 0u;
    __syncthreads();
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    int bin = -1;
    if (t < ntup) {
        i64 x[2 * nel<CLS>() + 2];
        bool ok = true;
        if constexpr (SRC == SRC_LIST)
            ok = load_tuple<CLS>(S, idx, t, x);
        else
            range_tuple<CLS, SRC == SRC_SORTED>(S, start + t, x);
        if (ok) {
            const PairRec *q[nel<CLS>() + 1];
            tuple_pairs<CLS>(S, x, q);
            bin = tuple_bin<CLS>(q);
        } else { // out-of-range primitive index: flag the entry
            outE[t] = __builtin_nan("");
            if (outS) outS[t] = __builtin_nan("");
            if (outX) outX[t] = __builtin_nan("");
            if (outY) outY[t] = __builtin_nan("");
        }
        code[t] = (unsigned char)(bin < 0 ? 255 : bin);
    }
    if (bin >= 0) atomicAdd(&lc[bin], 1u);
    __syncthreads();
    if (threadIdx.x < n_bins<CLS>() + 1 && lc[threadIdx.x]) atomicAdd(&counts[threadIdx.x], lc[threadIdx.x]);
}

template <int CLS>
__global__ void __launch_bounds__(256) k_prim_scatter(i64 ntup, const unsigned char *__restrict__ code,
                                                      const unsigned *__restrict__ counts,
                                                      unsigned *__restrict__ cursors, unsigned *__restrict__ list)
{
    __shared__ unsigned lc[32], lbase[32];
    if (threadIdx.x < 32) lc[threadIdx.x] = 0u;
    __syncthreads();
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    int bin = -1;
    if (t < ntup) {
        int c = code[t];
        bin = (c == 255) ? -1 : c;
    }
    unsigned loc = 0;
    if (bin >= 0) loc = atomicAdd(&lc[bin], 1u); // rank within the block's share of the bin
    __syncthreads();
    if (threadIdx.x < n_bins<CLS>() + 1)
        lbase[threadIdx.x] = bin_offset(counts, threadIdx.x) +
                             (lc[threadIdx.x] ? atomicAdd(&cursors[threadIdx.x], lc[threadIdx.x]) : 0u);
    __syncthreads();
    if (bin >= 0) list[lbase[bin] + loc] = (unsigned)t;
}

// Lanes per block of the tier kernel: one wavefront, so that blocks are
// scheduled per SIMD (the kernels run at one wave per SIMD), measured ~1 %
// faster than 256-lane blocks.
constexpr int TIER_BS = 64;

// One listed tuple of tier K: its pair records are staged in LDS (the
// evaluation reads them many times; at one wave per SIMD, the register use
// of these kernels, the LDS footprint costs no occupancy), then it is
// evaluated with the tier's capacities.
template <int CLS, int SRC, int K>
NRX_DI void tier_item(const DevSys &S, i64 start, const i64 *__restrict__ idx, double *__restrict__ outE,
                      double *__restrict__ outS, double *__restrict__ outX, double *__restrict__ outY, i64 t,
                      PairRec (*stage)[TIER_BS])
{
    constexpr int NP = nel<CLS>() + 1;
    i64 x[2 * NP];
    src_tuple<CLS, SRC>(S, idx, start, t, x); // a list was validated by k_prim_bin
    const PairRec *q[NP];
    tuple_pairs<CLS>(S, x, q);
    NRX_UNROLL for (int i = 0; i < NP; ++i) {
        stage[i][threadIdx.x] = *q[i];
        q[i] = &stage[i][threadIdx.x];
    }
    double e, s, xx, yy;
    tuple_eval<CLS, Tier<CLS>::NL[K], Tier<CLS>::NB[K]>(d_params(S), q, e, s, xx, yy);
    outE[t] = e;
    if (outS) outS[t] = s;
    if (outX) outX[t] = xx;
    if (outY) outY[t] = yy;
}

// block b of the fused launch: the bins are laid out from the last tier and
// the highest proton degree down (their tuples take longest, so they start
// first).  Locates the block's bin; -1 past the last bin.
template <int CLS> NRX_DI int block_bin(const unsigned *__restrict__ counts, i64 b, i64 &k0)
{
    for (int bin = n_bins<CLS>() - 1; bin >= 0; --bin) {
        i64 nb = ((i64)counts[bin] + TIER_BS - 1) / TIER_BS;
        if (b < nb) {
            k0 = b * TIER_BS;
            return bin;
        }
        b -= nb;
    }
    return -1;
}

// one inlined evaluation per tier: tier K of the bin's tuple
template <int CLS, int SRC, int K>
NRX_DI void tier_eval(int tier, const DevSys &S, i64 start, const i64 *__restrict__ idx, double *__restrict__ outE,
                      double *__restrict__ outS, double *__restrict__ outX, double *__restrict__ outY, i64 t,
                      PairRec (*stage)[TIER_BS])
{
    if constexpr (K >= 0) {
        if (tier == K) {
            tier_item<CLS, SRC, K>(S, start, idx, outE, outS, outX, outY, t, stage);
            return;
        }
        tier_eval<CLS, SRC, K - 1>(tier, S, start, idx, outE, outS, outX, outY, t, stage);
    }
}

// Every bin of a launch in one kernel, so that the bins' tails overlap:
// lane per listed tuple, blocks past the bins exit.
template <int CLS, int SRC>
__global__ void __launch_bounds__(TIER_BS) k_prim_tiers(DevSys S, i64 start, const i64 *__restrict__ idx,
                                                    double *__restrict__ outE, double *__restrict__ outS,
                                                    double *__restrict__ outX, double *__restrict__ outY,
                                                    const unsigned *__restrict__ list,
                                                    const unsigned *__restrict__ counts)
{
    __shared__ PairRec stage[nel<CLS>() + 1][TIER_BS];
    i64 k0 = 0;
    const int bin = block_bin<CLS>(counts, (i64)blockIdx.x, k0);
    if (bin < 0) return;
    const i64 k = k0 + threadIdx.x;
    if (k >= (i64)counts[bin]) return;
    const i64 t = (i64)list[bin_offset(counts, bin) + k];
    tier_eval<CLS, SRC, Tier<CLS>::N - 1>(bin / hub_bins<CLS>(), S, start, idx, outE, outS, outX, outY, t, stage);
}

__global__ void k_zero_u32n(unsigned *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}


// tuples per launch of the primitive kernels (bounds the sorted list, whose
// entries are 32-bit positions, and the bin counters)
static const i64 PRIM_CHUNK = (i64)1 << 24;

// stream-ordered scratch: the sorted tuple list, bin codes and counters
struct StreamBuf {
    void *p = nullptr;
    hipStream_t st = 0;
    ~StreamBuf()
    {
        if (p) (void)hipFreeAsync(p, st);
    }
    int alloc(size_t bytes, hipStream_t s)
    {
        st = s;
        hipError_t e = hipMallocAsync(&p, bytes, s);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(NEORX_E_NOMEM, "hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
        }
        return NEORX_OK;
    }
};

template <int CLS, int SRC>
static int launch_prims(const DevSys &d, i64 start, i64 ntup, const i64 *idx, double *oe, double *os, hipStream_t st,
                        double *ox = nullptr, double *oy = nullptr)
{
    constexpr int NBN = n_bins<CLS>();
    static_assert(NBN + 1 <= 32, "bin counters (with the wide bin) and cursors fit the 256-byte header");
    if constexpr (two_basis<CLS>())
        if (SRC != SRC_LIST && (d.npebf1 <= 0 || d.npebf1 >= d.npebf))
            return fail(NEORX_E_ARG, "class %d enumerates two electron bases; this handle has one", CLS);
    if (ntup <= 0) return NEORX_OK;
    i64 cap = std::min(ntup, PRIM_CHUNK);
    StreamBuf ws; // header: counts[32], cursors[32]; list: cap x u32; codes: cap x u8
    int rc = ws.alloc(256 + cap * sizeof(unsigned) + cap, st);
    if (rc) return rc;
    unsigned *counts = (unsigned *)ws.p, *cursors = counts + 32, *list = (unsigned *)((char *)ws.p + 256);
    unsigned char *code = (unsigned char *)(list + cap);
    for (i64 off = 0; off < ntup; off += PRIM_CHUNK) {
        i64 n = std::min(ntup - off, PRIM_CHUNK);
        const i64 *ix = SRC == SRC_LIST ? idx + off * (2 * nel<CLS>() + 2) : nullptr;
        double *e = oe + off, *s = os ? os + off : nullptr, *x = ox ? ox + off : nullptr;
        double *y = oy ? oy + off : nullptr;
        hipLaunchKernelGGL(k_zero_u32n, dim3(1), dim3(64), 0, st, counts, 64);
        hipLaunchKernelGGL((k_prim_bin<CLS, SRC>), dim3(nblocks(n, 256)), dim3(256), 0, st, d, start + off, n, ix, e, s,
                           x, y, code, counts);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL((k_prim_scatter<CLS>), dim3(nblocks(n, 256)), dim3(256), 0, st, n, code, counts, cursors,
                           list);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL((k_prim_tiers<CLS, SRC>), dim3(nblocks(n, TIER_BS) + NBN), dim3(TIER_BS), 0, st, d, start + off, ix,
                           e, s, x, y, list, counts);
        HIPCHK(hipGetLastError());
        if (d.lmax > 2) { // tuples with f functions: the wide bin
            rc = nrx_wide_prims(CLS, SRC, d, start + off, ix, e, s, x, y, list, counts, NBN, n, st);
            if (rc) return rc;
        }
    }
    return NEORX_OK;
}

static int prim_list(int64_t handle, int64_t cls, int64_t ntup, const int64_t *d_idx, double *d_ansE, double *d_ansS,
                     double *d_ansX, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (cls < 1 || cls > 10) return fail(NEORX_E_ARG, "class %lld not in 1..10", (long long)cls);
    if (ntup < 0 || (ntup > 0 && (!d_idx || !d_ansE))) return fail(NEORX_E_ARG, "bad tuple list arguments");
    if (ntup == 0) return NEORX_OK;
    hipStream_t st = (hipStream_t)stream;
    switch (cls) {
    case 1: return launch_prims<1, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 2: return launch_prims<2, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 3: return launch_prims<3, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 4: return launch_prims<4, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 5: return launch_prims<5, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 6: return launch_prims<6, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 7: return launch_prims<7, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 9: return launch_prims<9, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    case 10: return launch_prims<10, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    default: return launch_prims<8, SRC_LIST>(s->d, 0, ntup, d_idx, d_ansE, d_ansS, st, d_ansX);
    }
}

extern "C" int neorx_prim_list_dev(int64_t handle, int64_t cls, int64_t ntup, const int64_t *d_idx, double *d_ansE,
                                   double *d_ansS, void *stream)
{
    return prim_list(handle, cls, ntup, d_idx, d_ansE, d_ansS, nullptr, stream);
}

extern "C" int neorx_prim_list3_dev(int64_t handle, int64_t cls, int64_t ntup, const int64_t *d_idx, double *d_ansE,
                                    double *d_ansS, double *d_ansE2, void *stream)
{
    return prim_list(handle, cls, ntup, d_idx, d_ansE, d_ansS, d_ansE2, stream);
}

extern "C" int neorx_prim_range_dev(int64_t handle, int64_t cls, int64_t start, int64_t count, double *d_ansE,
                                    double *d_ansS, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (cls < 1 || cls > 4) return fail(NEORX_E_ARG, "class %lld not in 1..4", (long long)cls);
    double tot = 1.0;
    for (int i = 0; i < (cls == 4 ? 2 : cls); ++i) tot *= (double)s->d.npebf * (double)s->d.npebf;
    tot *= (double)s->d.npbf * (double)s->d.npbf;
    if (start < 0 || count < 0 || (double)start + (double)count > tot)
        return fail(NEORX_E_ARG, "range [%lld,%lld) outside the %.0f tuples of class %lld", (long long)start,
                    (long long)(start + count), tot, (long long)cls);
    if (count == 0) return NEORX_OK;
    if (!d_ansE) return fail(NEORX_E_ARG, "null output");
    hipStream_t st = (hipStream_t)stream;
    if (cls == 1) return launch_prims<1, SRC_SORTED>(s->d, start, count, nullptr, d_ansE, d_ansS, st);
    if (cls == 2) return launch_prims<2, SRC_SORTED>(s->d, start, count, nullptr, d_ansE, d_ansS, st);
    if (cls == 4) return launch_prims<4, SRC_SORTED>(s->d, start, count, nullptr, d_ansE, d_ansS, st);
    return launch_prims<3, SRC_SORTED>(s->d, start, count, nullptr, d_ansE, d_ansS, st);
}

// ------------------------------------------------------------------
// contracted builders
// ------------------------------------------------------------------
// Contracted tuples of GAM2_IC1 / GAM3_IC1 are built in two steps: the
// primitive kernels above fill a slab of primitive values in original order,
// then a contraction kernel sums each contracted tuple over its primitives
// (contract_omega2_conv / contract_omega3_conv loop order).  A slab is a run
// of consecutive leading indices L -- (ip,jp) for e-e-p, (ip,jp,i1) for
// e-e-e-p -- whose primitive tuples are contiguous in the enumeration
// because each contracted function owns a contiguous run of primitives.
//
// e-e-p: lane per contracted (L, i1, j1, i2, j2), L = ip*npbf + jp
// X holds leading indices from Xb on; with flush set every primitive value
// goes through underflow() first (contract_omega2, gam2_DSCF_OMP.f:514-515)
__global__ void __launch_bounds__(256) k_contract2(DevSys S, i64 L0, i64 n, i64 R0, const double *__restrict__ Y,
                                                   const double *__restrict__ Ys, double *__restrict__ X,
                                                   double *__restrict__ Xs, i64 Xb, int flush)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    i64 ne = S.nebf, np = S.npbf, pe = S.npebf;
    i64 ne4 = ne * ne * ne * ne, pe4 = pe * pe * pe * pe;
    i64 L = L0 + t / ne4, rem = t % ne4;
    i64 j2 = rem % ne, i2 = (rem / ne) % ne, j1 = (rem / (ne * ne)) % ne, i1 = rem / (ne * ne * ne);
    i64 ip = L / np, jp = L % np;
    double cp = S.agnbfcc[ip] * S.agnbfcc[jp];
    double om = 0.0, oms = 0.0;
    for (i64 a1 = S.kpestr[i1] - 1; a1 < S.kpeend[i1]; ++a1)
        for (i64 c1 = S.kpestr[j1] - 1; c1 < S.kpeend[j1]; ++c1)
            for (i64 a2 = S.kpestr[i2] - 1; a2 < S.kpeend[i2]; ++a2)
                for (i64 c2 = S.kpestr[j2] - 1; c2 < S.kpeend[j2]; ++c2) {
                    i64 r = L * pe4 + ((a1 * pe + c1) * pe + a2) * pe + c2 - R0;
                    double cf = cp * S.agebfcc[a1] * S.agebfcc[c1] * S.agebfcc[a2] * S.agebfcc[c2];
                    double y = Y[r], ys = Ys[r];
                    if (flush) {
                        underflow(y);
                        underflow(ys);
                    }
                    om += y * cf;
                    oms += ys * cf;
                }
    X[(L - Xb) * ne4 + rem] = om;
    Xs[(L - Xb) * ne4 + rem] = oms;
}

// GM2 = (x12 + x21)/2  (gam2_IC1_OMP.f:141-177)
// over n entries of whole (ip,jp) blocks
__global__ void __launch_bounds__(256) k_gam2_symm(i64 ne, i64 n, const double *__restrict__ X,
                                                   const double *__restrict__ Xs, double *__restrict__ G,
                                                   double *__restrict__ Gs)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 ne2 = ne * ne;
        i64 b = t / (ne2 * ne2), p1 = (t / ne2) % ne2, p2 = t % ne2;
        i64 t21 = (b * ne2 + p2) * ne2 + p1;
        G[t] = (X[t] + X[t21]) / 2.0;
        Gs[t] = (Xs[t] + Xs[t21]) / 2.0;
    }
}

// e-e-e-p: lane per contracted (L, j1, i2, j2, i3, j3), L = (ip*npbf + jp)*nebf + i1
__global__ void __launch_bounds__(256) k_contract3(DevSys S, i64 L0, i64 n, i64 R0, const double *__restrict__ Y,
                                                   double *__restrict__ X, i64 Xb)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    i64 ne = S.nebf, np = S.npbf, pe = S.npebf;
    i64 ne5 = ne * ne * ne * ne * ne, pe5 = pe * pe * pe * pe * pe;
    i64 L = L0 + t / ne5, rem = t % ne5;
    i64 j3 = rem % ne, i3 = (rem / ne) % ne, j2 = (rem / (ne * ne)) % ne, i2 = (rem / (ne * ne * ne)) % ne,
        j1 = rem / (ne * ne * ne * ne);
    i64 i1 = L % ne, pp = L / ne, ip = pp / np, jp = pp % np;
    double cp = S.agnbfcc[ip] * S.agnbfcc[jp];
    double om = 0.0;
    for (i64 a1 = S.kpestr[i1] - 1; a1 < S.kpeend[i1]; ++a1)
        for (i64 c1 = S.kpestr[j1] - 1; c1 < S.kpeend[j1]; ++c1)
            for (i64 a2 = S.kpestr[i2] - 1; a2 < S.kpeend[i2]; ++a2)
                for (i64 c2 = S.kpestr[j2] - 1; c2 < S.kpeend[j2]; ++c2)
                    for (i64 a3 = S.kpestr[i3] - 1; a3 < S.kpeend[i3]; ++a3)
                        for (i64 c3 = S.kpestr[j3] - 1; c3 < S.kpeend[j3]; ++c3) {
                            i64 r = (pp * pe + a1) * pe5 + (((c1 * pe + a2) * pe + c2) * pe + a3) * pe + c3 - R0;
                            double cf = cp * S.agebfcc[a1] * S.agebfcc[c1] * S.agebfcc[a2] * S.agebfcc[c2] *
                                        S.agebfcc[a3] * S.agebfcc[c3];
                            om += Y[r] * cf;
                        }
    X[(L - Xb) * ne5 + rem] = om;
}

// six-permutation average (gam3_IC1_OMP.f:149-263)
// over n entries of whole (ip,jp) blocks
__global__ void __launch_bounds__(256) k_gam3_symm(i64 ne, i64 n, const double *__restrict__ X,
                                                   double *__restrict__ G)
{
    i64 ne2 = ne * ne, blk = ne2 * ne2 * ne2;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 b = t / blk, r = t % blk;
        i64 p3 = r % ne2, p2 = (r / ne2) % ne2, p1 = r / (ne2 * ne2);
        const double *Xb = X + b * blk;
#define IX(a, c, e) (((a) * ne2 + (c)) * ne2 + (e))
        double s = Xb[IX(p1, p2, p3)] + Xb[IX(p1, p3, p2)] + Xb[IX(p2, p1, p3)] + Xb[IX(p2, p3, p1)] +
                   Xb[IX(p3, p1, p2)] + Xb[IX(p3, p2, p1)];
#undef IX
        G[t] = s / 6.0;
    }
}

// GAM4_ICR / symm_gam4 / calc_GAM_4_integral (gam_4_OMP.f:712-1192)
__constant__ int c_perm4[24][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {2, 0, 1, 3}, {0, 2, 1, 3}, {1, 2, 0, 3}, {2, 1, 0, 3},
                                   {3, 1, 0, 2}, {1, 3, 0, 2}, {0, 3, 1, 2}, {3, 0, 1, 2}, {1, 0, 3, 2}, {0, 1, 3, 2},
                                   {0, 2, 3, 1}, {2, 0, 3, 1}, {3, 0, 2, 1}, {0, 3, 2, 1}, {2, 3, 0, 1}, {3, 2, 0, 1},
                                   {3, 2, 1, 0}, {2, 3, 1, 0}, {1, 3, 2, 0}, {3, 1, 2, 0}, {2, 1, 3, 0}, {1, 2, 3, 0}};
__constant__ double c_w4[24] = {16, -8, 4, -8, 4, -8, 4, -2, 4, -2, 4, -8, 4, -2, 4, -8, 4, -2, 4, -2, 4, -8, 4, -2};

// calc_GAM_4_integral (gam_4_OMP.f:964-1192) of entry (ip,jp, br[k] bra /
// kt[k] ket of electron k): (1/6) sum over the six electron pairs of
// 1/2 GAM_ee(pair) GAM_2s(complement), every factor through underflow()
NRX_DI double gam4_term(i64 ne, i64 np, const double *__restrict__ G2s, const double *__restrict__ Gee, i64 ip,
                        i64 jp, const i64 *br, const i64 *kt)
{
    i64 ne2 = ne * ne;
    i64 gb = (ip * np + jp) * ne2 * ne2;
    i64 pr[4] = {br[0] * ne + kt[0], br[1] * ne + kt[1], br[2] * ne + kt[2], br[3] * ne + kt[3]};
    double ee34 = Gee[pr[2] * ne2 + pr[3]], ee24 = Gee[pr[1] * ne2 + pr[3]], ee23 = Gee[pr[1] * ne2 + pr[2]];
    double ee14 = Gee[pr[0] * ne2 + pr[3]], ee13 = Gee[pr[0] * ne2 + pr[2]], ee12 = Gee[pr[0] * ne2 + pr[1]];
    double gs12 = G2s[gb + pr[0] * ne2 + pr[1]], gs13 = G2s[gb + pr[0] * ne2 + pr[2]];
    double gs14 = G2s[gb + pr[0] * ne2 + pr[3]], gs23 = G2s[gb + pr[1] * ne2 + pr[2]];
    double gs24 = G2s[gb + pr[1] * ne2 + pr[3]], gs34 = G2s[gb + pr[2] * ne2 + pr[3]];
    underflow(ee34); underflow(ee24); underflow(ee23); underflow(ee14); underflow(ee13); underflow(ee12);
    underflow(gs12); underflow(gs13); underflow(gs14); underflow(gs23); underflow(gs24); underflow(gs34);
    return (0.5 * ee34 * gs12 + 0.5 * ee24 * gs13 + 0.5 * ee23 * gs14 + 0.5 * ee14 * gs23 + 0.5 * ee13 * gs24 +
            0.5 * ee12 * gs34) /
           6.0;
}

// symm_gam4/16 of entry (ip,jp, br[k] bra / kt[k] ket of electron k)
NRX_DI double gam4_value(i64 ne, i64 np, const double *__restrict__ G2s, const double *__restrict__ Gee, i64 ip,
                         i64 jp, const i64 *br, const i64 *kt)
{
    double ans = 0.0;
    for (int p = 0; p < 24; ++p) {
        i64 kp[4] = {kt[c_perm4[p][0]], kt[c_perm4[p][1]], kt[c_perm4[p][2]], kt[c_perm4[p][3]]};
        ans += c_w4[p] * gam4_term(ne, np, G2s, Gee, ip, jp, br, kp);
    }
    return ans / 16.0;
}

__global__ void __launch_bounds__(256) k_gam4(i64 ne, i64 np, i64 off, i64 n, const double *__restrict__ G2s,
                                              const double *__restrict__ Gee, double *__restrict__ G4)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    i64 r = off + t;
    i64 br[4], kt[4];
    for (int k = 3; k >= 0; --k) {
        kt[k] = r % ne; r /= ne;
        br[k] = r % ne; r /= ne;
    }
    i64 jp = r % np, ip = r / np;
    G4[off + t] = gam4_value(ne, np, G2s, Gee, ip, jp, br, kt);
}

// ------------------------------------------------------------------
// host helpers for builders
// ------------------------------------------------------------------

// ------------------------------------------------------------------
// Resident tensors.  The host-pointer builders (GAM2_IC1, GAM3_IC1,
// GAM4_ICR) fill the caller's array and keep their device copy, keyed by the
// host array's address and size; a later host-pointer contraction given the
// same array reads that copy instead of copying the tensor over PCIe again
// (C1's GM3 is 97.8 GB).  The caller's array may have changed in between:
// 4096 entries sampled over it are compared bit for bit with the values kept
// at registration, and any difference drops the copy.  Resident copies use
// at most 60 % of the device memory (the oldest are released first);
// NEORX_RESIDENT=0 turns the cache off; neorx_release_resident frees it.
// ------------------------------------------------------------------
namespace {
struct Resident {
    const void *host;
    i64 n; // doubles
    void *dev;
    std::vector<double> probe;
};
std::mutex g_res_mu;
std::vector<Resident> g_res; // oldest first
std::atomic<long long> g_h2d_bytes{0}, g_d2h_bytes{0};
constexpr i64 RES_PROBES = 4096;

i64 probe_pos(i64 n, i64 k) { return n <= RES_PROBES ? k : (k * (n - 1)) / (RES_PROBES - 1); }

std::vector<double> take_probe(const double *h, i64 n)
{
    const i64 m = std::min(n, RES_PROBES);
    std::vector<double> v(m);
    for (i64 k = 0; k < m; ++k) v[k] = h[probe_pos(n, k)];
    return v;
}

bool resident_on()
{
    const char *e = getenv("NEORX_RESIDENT");
    return !(e && atoi(e) == 0);
}

void res_drop_locked(size_t i)
{
    (void)hipFree(g_res[i].dev);
    g_res.erase(g_res.begin() + (long)i);
}

// device copy kept for host array h of n doubles, or null
const double *res_find(const double *h, i64 n)
{
    std::lock_guard<std::mutex> lk(g_res_mu);
    for (size_t i = 0; i < g_res.size(); ++i) {
        Resident &r = g_res[i];
        if (r.host != (const void *)h) continue;
        if (r.n != n) return nullptr;
        const i64 m = (i64)r.probe.size();
        for (i64 k = 0; k < m; ++k)
            if (__builtin_memcmp(&r.probe[k], &h[probe_pos(n, k)], 8) != 0) {
                res_drop_locked(i); // the caller changed its array
                return nullptr;
            }
        return (const double *)r.dev;
    }
    return nullptr;
}

// keep the device buffer b (n doubles, equal to host array h) as h's resident
// copy; b gives up its memory either way
void res_keep(DevBuf &b, const double *h, i64 n)
{
    if (!b.p || !b.own) return;
    if (!resident_on()) return;
    std::lock_guard<std::mutex> lk(g_res_mu);
    for (size_t i = 0; i < g_res.size(); ++i)
        if (g_res[i].host == (const void *)h) {
            res_drop_locked(i);
            break;
        }
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return;
    const long double cap = 0.6L * (long double)total;
    long double held = 0.0L;
    for (auto &r : g_res) held += (long double)r.n * 8;
    while (!g_res.empty() && held + (long double)n * 8 > cap) {
        held -= (long double)g_res.front().n * 8;
        res_drop_locked(0);
    }
    if (held + (long double)n * 8 > cap) return; // larger than the budget: not kept
    g_res.push_back(Resident{h, n, b.p, take_probe(h, n)});
    b.p = nullptr;
}
} // namespace

extern "C" int neorx_release_resident(void)
{
    std::lock_guard<std::mutex> lk(g_res_mu);
    while (!g_res.empty()) res_drop_locked(g_res.size() - 1);
    return NEORX_OK;
}

extern "C" int neorx_transfer_bytes(int64_t *h2d_bytes, int64_t *d2h_bytes)
{
    if (h2d_bytes) *h2d_bytes = g_h2d_bytes.load();
    if (d2h_bytes) *d2h_bytes = g_d2h_bytes.load();
    return NEORX_OK;
}

// input tensor h (n doubles) on the device: its resident copy (borrowed by b)
// or a fresh copy over PCIe
static int h2d(DevBuf &b, const double *h, i64 n)
{
    if (const double *r = res_find(h, n)) {
        b.p = (void *)r;
        b.own = false;
        return NEORX_OK;
    }
    int rc = b.alloc(n * 8);
    if (rc) return rc;
    HIPCHK(hipMemcpy(b.p, h, n * 8, hipMemcpyHostToDevice));
    g_h2d_bytes += n * 8;
    return NEORX_OK;
}

static const i64 CHUNK = (i64)256 * 65536; // lanes per launch of the contracted builders

// primitive values held per slab (each of E and S); a slab holds at least
// one leading index, whatever its size
static const i64 SLAB_CAP = (i64)1 << 26;
// contracted values held per block of (ip,jp) pairs by the direct-SCF routines
static const i64 BLOCK_CAP = (i64)1 << 27;

// first primitive tuple (original order) of leading index L of class cls
// (e-e-p: L = ip*npbf + jp; e-e-e-p: L = (ip*npbf + jp)*nebf + i1); the
// number of leading indices gives the end of the enumeration.  The
// RXCHFmult classes 5 and 6 count e1 over the first basis (pa primitives,
// nebf1 functions) and the other electrons over the special basis (pb).
static i64 slab_start(const System *s, int cls, i64 L)
{
    bool two = cls == 5 || cls == 6;
    i64 pa = two ? s->d.npebf1 : s->d.npebf, pb = two ? s->d.npebf - pa : pa, ne = two ? s->d.nebf1 : s->d.nebf;
    if (cls == 2 || cls == 5) return L * pa * pa * pb * pb;
    i64 pp = L / ne, i1 = L % ne;
    return (pp * pa + (s->kpestr[i1] - 1)) * pa * pb * pb * pb * pb; // i1 == 0 past the end: pp * pa^2 pb^4
}

// consecutive slabs [L0, L1) covering [La, Lb); cap = largest primitive span
static std::vector<std::pair<i64, i64>> make_slabs(const System *s, int cls, i64 La, i64 Lb, i64 &cap)
{
    std::vector<std::pair<i64, i64>> out;
    cap = 0;
    for (i64 L0 = La; L0 < Lb;) {
        i64 L1 = L0 + 1;
        while (L1 < Lb && slab_start(s, cls, L1 + 1) - slab_start(s, cls, L0) <= SLAB_CAP) ++L1;
        cap = std::max(cap, slab_start(s, cls, L1) - slab_start(s, cls, L0));
        out.push_back({L0, L1});
        L0 = L1;
    }
    return out;
}

// largest primitive span of any slab over all leading indices of cls
static i64 slab_cap(const System *s, int cls)
{
    i64 cap = 0;
    i64 nL = s->d.npbf * s->d.npbf * ((cls == 3 || cls == 6) ? s->d.nebf1 : 1);
    make_slabs(s, cls, 0, nL, cap);
    return cap;
}

// Unsymmetrised contracted e-e-p values of (ip,jp) pairs [p0, p1) into X, Xs
// (block-local, ne^4 per pair), through primitive slabs in Y, Ys.
template <int CLS = 2>
static int build2_block(System *s, i64 p0, i64 p1, double *X, double *Xs, double *Y, double *Ys, int flush,
                        hipStream_t st)
{
    i64 ne = s->d.nebf, ne4 = ne * ne * ne * ne, cap;
    for (auto &sl : make_slabs(s, 2, p0, p1, cap)) {
        i64 R0 = slab_start(s, 2, sl.first), R1 = slab_start(s, 2, sl.second);
        int rc = launch_prims<CLS, SRC_ORIG>(s->d, R0, R1 - R0, nullptr, Y, Ys, st);
        if (rc) return rc;
        i64 n = (sl.second - sl.first) * ne4;
        hipLaunchKernelGGL(k_contract2, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y, Ys, X, Xs,
                           p0, flush);
        HIPCHK(hipGetLastError());
    }
    return NEORX_OK;
}

// Unsymmetrised contracted e-e-e-p values of (ip,jp) pairs [p0, p1) into X
// (block-local, ne^6 per pair).
template <int CLS = 3>
static int build3_block(System *s, i64 p0, i64 p1, double *X, double *Y, hipStream_t st)
{
    i64 ne = s->d.nebf, ne5 = ne * ne * ne * ne * ne, cap;
    for (auto &sl : make_slabs(s, 3, p0 * ne, p1 * ne, cap)) {
        i64 R0 = slab_start(s, 3, sl.first), R1 = slab_start(s, 3, sl.second);
        int rc = launch_prims<CLS, SRC_ORIG>(s->d, R0, R1 - R0, nullptr, Y, nullptr, st);
        if (rc) return rc;
        i64 n = (sl.second - sl.first) * ne5;
        hipLaunchKernelGGL(k_contract3, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y, X, p0 * ne);
        HIPCHK(hipGetLastError());
    }
    return NEORX_OK;
}

extern "C" int neorx_gam2_ic1_dev(int64_t handle, double *d_gm2, double *d_gm2s, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (!d_gm2 || !d_gm2s) return fail(NEORX_E_ARG, "null output");
    i64 ne = s->d.nebf, np = s->d.npbf, ng2 = ne * ne * ne * ne * np * np;
    hipStream_t st = (hipStream_t)stream;
    i64 cap = slab_cap(s, 2);
    DevBuf X, Xs, Y, Ys;
    int rc = X.alloc(ng2 * 8);
    if (!rc) rc = Xs.alloc(ng2 * 8);
    if (!rc) rc = Y.alloc(cap * 8);
    if (!rc) rc = Ys.alloc(cap * 8);
    if (!rc) rc = build2_block(s, 0, np * np, X.d(), Xs.d(), Y.d(), Ys.d(), 0, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gam2_symm, dim3(grid_blocks(ng2, 256)), dim3(256), 0, st, ne, ng2, X.d(), Xs.d(), d_gm2, d_gm2s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st)); // scratch buffers are released on return
    return NEORX_OK;
}

// GAM3_IC1 (class 3) and RXCHFmult_GAM3_XCHF (class 7): contracted e-e-e-p
// values per (ip,jp) block, then the six-permutation average into d_gm3
template <int CLS> static int gam3_build_dev(System *s, double *d_gm3, void *stream)
{
    i64 ne = s->d.nebf, np = s->d.npbf, blk = ne * ne * ne * ne * ne * ne;
    hipStream_t st = (hipStream_t)stream;
    i64 cap = slab_cap(s, 3);
    // one (ip,jp) block of unsymmetrised values at a time: the scratch is
    // nebf^6 doubles, not the whole tensor again
    DevBuf X, Y;
    int rc = X.alloc(blk * 8);
    if (!rc) rc = Y.alloc(cap * 8);
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    for (i64 b = 0; b < np * np; ++b) {
        rc = build3_block<CLS>(s, b, b + 1, X.d(), Y.d(), st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gam3_symm, dim3(grid_blocks(blk, 256)), dim3(256), 0, st, ne, blk, X.d(), d_gm3 + b * blk);
        HIPCHK(hipGetLastError());
        if (g_verbose && ((b + 1) % 8 == 0 || b + 1 == np * np)) {
            HIPCHK(hipStreamSynchronize(st));
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            fprintf(stderr, "neorx: e-e-e-p build, %lld of %lld (ip,jp) blocks in %.1f s\n", (long long)(b + 1),
                    (long long)(np * np), el);
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_gam3_ic1_dev(int64_t handle, double *d_gm3, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (!d_gm3) return fail(NEORX_E_ARG, "null output");
    return gam3_build_dev<3>(s, d_gm3, stream);
}

static int copy_out(double *h, const double *d, i64 n)
{
    HIPCHK(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
    g_d2h_bytes += n * 8;
    return NEORX_OK;
}

template <int CLS> static int gam1_host(int64_t handle, double *gam1, double *gam1s)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, np = s->d.npbf, n = ne * ne * np * np;
    DevBuf G, Gs;
    int rc = G.alloc(n * 8);
    if (!rc) rc = Gs.alloc(n * 8);
    if (rc) return rc;
    if (s->maxl > 2) {
        rc = nrx_wide_gam1(CLS, s->d, G.d(), Gs.d(), 0);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL((k_gam1<CLS, 2>), dim3(nblocks(n, 256)), dim3(256), 0, 0, s->d, G.d(), Gs.d());
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    rc = copy_out(gam1, G.d(), n);
    if (!rc) rc = copy_out(gam1s, Gs.d(), n);
    return rc;
}

extern "C" int neorx_gam1(int64_t handle, double *gam1, double *gam1s) { return gam1_host<1>(handle, gam1, gam1s); }

extern "C" int neorx_gam2_ic1(int64_t handle, double *gm2icr, double *gm2sicr)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, np = s->d.npbf, n = ne * ne * ne * ne * np * np;
    DevBuf G, Gs;
    int rc = G.alloc(n * 8);
    if (!rc) rc = Gs.alloc(n * 8);
    if (!rc) rc = neorx_gam2_ic1_dev(handle, G.d(), Gs.d(), 0);
    if (!rc) rc = copy_out(gm2icr, G.d(), n);
    if (!rc) rc = copy_out(gm2sicr, Gs.d(), n);
    if (!rc) {
        res_keep(G, gm2icr, n);
        res_keep(Gs, gm2sicr, n);
    }
    return rc;
}

extern "C" int neorx_gam3_ic1(int64_t handle, double *gm3icr)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, np = s->d.npbf, n = ne * ne * ne * ne * ne * ne * np * np;
    DevBuf G;
    int rc = G.alloc(n * 8);
    if (!rc) rc = neorx_gam3_ic1_dev(handle, G.d(), 0);
    if (!rc) rc = copy_out(gm3icr, G.d(), n);
    if (!rc) res_keep(G, gm3icr, n);
    return rc;
}

extern "C" int neorx_gam_ee(int64_t handle, double *gam_ee)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, n = ne * ne * ne * ne;
    DevBuf G;
    int rc = G.alloc(n * 8);
    if (rc) return rc;
    rc = nrx_gam_ee_launch(s, G.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    return copy_out(gam_ee, G.d(), n);
}

int nrx_gam_ee_launch(System *s, double *d_gee, hipStream_t st)
{
    i64 ne = s->d.nebf, n = ne * ne * ne * ne;
    if (s->maxl > 2) return nrx_wide_gam_ee(s->d, d_gee, st);
    hipLaunchKernelGGL(k_gam_ee<2>, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, d_gee);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

int nrx_gam1_launch(System *s, double *d_g1, double *d_g1s, hipStream_t st)
{
    i64 ne = s->d.nebf, np = s->d.npbf, n = ne * ne * np * np;
    if (s->maxl > 2) return nrx_wide_gam1(1, s->d, d_g1, d_g1s, st);
    hipLaunchKernelGGL((k_gam1<1, 2>), dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, d_g1, d_g1s);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

extern "C" int neorx_gam4_icr(int64_t nebf, int64_t npbf, const double *gam_2s, const double *gam_ee, double *gm4icr)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    i64 ng2 = ne * ne * ne * ne * np * np, ngee = ne * ne * ne * ne, ng4 = ngee * ngee * np * np;
    DevBuf A, B, C;
    rc = h2d(A, gam_2s, ng2); // GM2sICR of GAM2_IC1: resident when it built it
    if (!rc) rc = h2d(B, gam_ee, ngee);
    if (!rc) rc = C.alloc(ng4 * 8);
    if (rc) return rc;
    for (i64 off = 0; off < ng4; off += CHUNK) {
        i64 n = std::min(CHUNK, ng4 - off);
        hipLaunchKernelGGL(k_gam4, dim3(nblocks(n, 256)), dim3(256), 0, 0, ne, np, off, n, A.d(), B.d(), C.d());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipDeviceSynchronize());
    rc = copy_out(gm4icr, C.d(), ng4);
    if (!rc) res_keep(C, gm4icr, ng4);
    return rc;
}

// ------------------------------------------------------------------
// Fock contractions
// ------------------------------------------------------------------
// sum over the 64 lanes of a wavefront (every lane gets the total)
NRX_DI double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// EG2IC1.  Block (ip,jp,a) reads the ne^3 slab G[ip][jp][a][b][c][d] once and
// forms dir[b] = sum_cd DE(c,d) G[b][c][d] and ex[d] = sum_bc DE(c,b) G[b][c][d]
// (the exchange read G(a,j2,i2,j1) of thread_EG2ICR re-indexed onto the same
// slab), for G and Gs.  Lane t owns the (c,d) columns cd = t + k NT (k < K)
// for every b: its ex partials stay in registers and meet in LDS once, dir[b]
// is summed over the wave by shuffles and over the waves in a fixed order, so
// no two runs differ in rounding.  Partial vectors [dir, ex, dirS, exS] go to
// a scratch array that k_eg2_reduce sums in a fixed order.
template <int NT, int K, class T>
__global__ void __launch_bounds__(NT) k_eg2_slab(i64 ne, const double *__restrict__ DE, const T *__restrict__ G,
                                                 const T *__restrict__ Gs, double *__restrict__ part)
{
    constexpr int NW = NT / 64;
    extern __shared__ double sm[]; // wd[NW][2 ne] (dir, dirS per wave), xc[2][ne^2] (ex column partials)
    const i64 blk = blockIdx.x, ne2 = ne * ne, ne3 = ne2 * ne;
    const int w = threadIdx.x >> 6;
    const bool lead = (threadIdx.x & 63) == 0;
    const T *g = G + blk * ne3;
    const T *gs = Gs ? Gs + blk * ne3 : nullptr;
    double *wd = sm, *xc = sm + 2 * NW * ne;
    double dcd[K], xa[K], xs[K];
    int cc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const i64 cd = threadIdx.x + (i64)k * NT;
        const bool on = cd < ne2;
        cc[k] = on ? (int)(cd / ne) : 0;
        dcd[k] = on ? DE[(cd % ne) * ne + cc[k]] : 0.0; // DE(c,d), column-major
        xa[k] = xs[k] = 0.0;
    }
    for (i64 b = 0; b < ne; ++b) {
        const T *row = g + b * ne2;
        const T *rows = gs ? gs + b * ne2 : nullptr;
        const double *deb = DE + b * ne; // DE(c,b) = deb[c]
        double ds = 0.0, dss = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const i64 cd = threadIdx.x + (i64)k * NT;
            if (cd < ne2) {
                const double gv = (double)row[cd], de_cb = deb[cc[k]];
                ds += dcd[k] * gv;
                xa[k] += de_cb * gv;
                if (rows) {
                    const double gsv = (double)rows[cd];
                    dss += dcd[k] * gsv;
                    xs[k] += de_cb * gsv;
                }
            }
        }
        ds = wave_sum(ds);
        dss = wave_sum(dss);
        if (lead) {
            wd[w * 2 * ne + b] = ds;
            wd[w * 2 * ne + ne + b] = dss;
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const i64 cd = threadIdx.x + (i64)k * NT;
        if (cd < ne2) {
            xc[cd] = xa[k];
            xc[ne2 + cd] = xs[k];
        }
    }
    __syncthreads();
    for (i64 t = threadIdx.x; t < 4 * ne; t += NT) {
        const i64 q = t / ne, x = t % ne;
        double v = 0.0;
        if (q == 0 || q == 2) {
            for (int u = 0; u < NW; ++u) v += wd[u * 2 * ne + (q == 2 ? ne : 0) + x];
        } else {
            const double *col = xc + (q == 3 ? ne2 : 0);
            for (i64 c = 0; c < ne; ++c) v += col[c * ne + x];
        }
        part[blk * 4 * ne + t] = v;
    }
}

// largest nebf of the EG2 slab kernel: ne^2 columns over 1024 lanes x 8
static const i64 EG2_MAX_NE = 90;

// k_eg2_slab over nblk slabs: the smallest (lanes, columns per lane) that
// holds ne^2 columns
template <class T>
static int eg2_slab_launch(i64 ne, i64 nblk, const double *d_de, const T *g, const T *gs, double *part,
                           hipStream_t st)
{
    if (ne <= 0 || ne > EG2_MAX_NE)
        return fail(NEORX_E_UNSUPPORTED, "EG2 slabs: nebf=%lld not in 1..%lld", (long long)ne, (long long)EG2_MAX_NE);
    const i64 ne2 = ne * ne;
    const dim3 grid((unsigned)nblk);
    auto lds = [&](int nt) { return (size_t)(2 * (nt / 64) * ne + 2 * ne2) * 8; };
    if (ne2 <= 256 * 4)
        hipLaunchKernelGGL((k_eg2_slab<256, 4, T>), grid, dim3(256), lds(256), st, ne, d_de, g, gs, part);
    else if (ne2 <= 256 * 8)
        hipLaunchKernelGGL((k_eg2_slab<256, 8, T>), grid, dim3(256), lds(256), st, ne, d_de, g, gs, part);
    else
        hipLaunchKernelGGL((k_eg2_slab<1024, 8, T>), grid, dim3(1024), lds(1024), st, ne, d_de, g, gs, part);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

int nrx_eg2_slabs(i64 ne, i64 nblk, const double *d_de, const double *g, const double *gs, double *part,
                  hipStream_t st)
{
    return eg2_slab_launch<double>(ne, nblk, d_de, g, gs, part, st);
}

// reduce partials: FE(a,b) += 2 sum_{ip,jp} DP(ip,jp) [dir_a[b] - ex_a[b]/2] etc.
__global__ void k_eg2_reduce(i64 ne, i64 np, const double *__restrict__ DE, const double *__restrict__ DP,
                             const double *__restrict__ part, int withS, int sacc, double *__restrict__ FE,
                             double *__restrict__ SE, double *__restrict__ FP, double *__restrict__ SP,
                             double *__restrict__ E, double *__restrict__ S)
{
    // SE is overwritten (EG2IC1 zeroes it first) or, with sacc, accumulated
    // (GAM2_DSCF's add2fock)
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    i64 ne2 = ne * ne, np2 = np * np;
    if (t < ne2) {
        // FE(a,b): Fortran element (a,b) at b*ne + a
        i64 a = t % ne, b = t / ne;
        double f = 0.0, fs = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) {
            i64 ip = pp / np, jp = pp % np;
            double dp = DP[jp * np + ip];
            const double *pr = part + (pp * ne + a) * 4 * ne;
            f += dp * 2.0 * (pr[b] - 0.5 * pr[ne + b]);
            fs += dp * (pr[2 * ne + b] - 0.5 * pr[3 * ne + b]);
        }
        FE[b * ne + a] += f;
        if (withS) SE[b * ne + a] = sacc ? SE[b * ne + a] + fs : fs;
    } else if (t < ne2 + np2) {
        i64 pp = t - ne2, ip = pp / np, jp = pp % np;
        double f = 0.0, fs = 0.0;
        for (i64 a = 0; a < ne; ++a) {
            const double *pr = part + (pp * ne + a) * 4 * ne;
            for (i64 b = 0; b < ne; ++b) {
                double dab = DE[b * ne + a];
                f += dab * (pr[b] - 0.5 * pr[ne + b]);
                fs += dab * (pr[2 * ne + b] - 0.5 * pr[3 * ne + b]);
            }
        }
        FP[jp * np + ip] += f;
        if (withS) SP[jp * np + ip] = fs;
    }
}

// y += x
__global__ void k_axpy(i64 n, const double *__restrict__ x, double *__restrict__ y)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] += x[t];
}

__global__ void k_energy(i64 np, const double *__restrict__ DP, const double *__restrict__ FPd,
                         const double *__restrict__ FPsd, double *__restrict__ E, double *__restrict__ S)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double e = 0.0, s = 0.0;
        for (i64 k = 0; k < np * np; ++k) {
            e += DP[k] * FPd[k];
            if (FPsd) s += DP[k] * FPsd[k];
        }
        *E = e;
        if (S) *S = s;
    }
}

template <class T>
static int eg2_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const T *d_gm2,
                   const T *d_gm2s, double *d_fe, double *d_se, double *d_fp, double *d_sp, double *d_e, double *d_s,
                   void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    hipStream_t st = (hipStream_t)stream;
    i64 nblk = np * np * ne;
    DevBuf part, fpd, fpsd;
    rc = part.alloc(nblk * 4 * ne * 8);
    if (!rc) rc = fpd.alloc(np * np * 8);
    if (!rc) rc = fpsd.alloc(np * np * 8);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(fpd.p, 0, np * np * 8, st));
    HIPCHK(hipMemsetAsync(fpsd.p, 0, np * np * 8, st));
    rc = eg2_slab_launch<T>(ne, nblk, d_de, d_gm2, d_gm2s, part.d(), st);
    if (rc) return rc;
    int withS = d_gm2s != nullptr;
    // FE/SE and the per-(ip,jp) proton terms into fpd/fpsd; then add fpd to FP
    hipLaunchKernelGGL(k_eg2_reduce, dim3(nblocks(ne * ne + np * np, 256)), dim3(256), 0, st, ne, np, d_de, d_dp,
                       part.d(), withS, 0, d_fe, d_se, fpd.d(), fpsd.d(), d_e, d_s);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st, np, d_dp, fpd.d(), withS ? fpsd.d() : nullptr, d_e,
                       withS ? d_s : nullptr);
    HIPCHK(hipGetLastError());
    // FP accumulates (add2fock), SP is overwritten
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(np * np, 256)), dim3(256), 0, st, np * np, fpd.d(), d_fp);
    HIPCHK(hipGetLastError());
    if (withS && d_sp) HIPCHK(hipMemcpyAsync(d_sp, fpsd.p, np * np * 8, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st)); // scratch buffers are released on return
    return NEORX_OK;
}

extern "C" int neorx_eg2ic1_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                                const double *d_gm2, const double *d_gm2s, double *d_fe, double *d_se, double *d_fp,
                                double *d_sp, double *d_e, double *d_s, void *stream)
{
    return eg2_dev<double>(nebf, npbf, d_de, d_dp, d_gm2, d_gm2s, d_fe, d_se, d_fp, d_sp, d_e, d_s, stream);
}

extern "C" int neorx_eg2ic1_f32_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                                    const float *d_gm2, const float *d_gm2s, double *d_fe, double *d_se,
                                    double *d_fp, double *d_sp, double *d_e, double *d_s, void *stream)
{
    return eg2_dev<float>(nebf, npbf, d_de, d_dp, d_gm2, d_gm2s, d_fe, d_se, d_fp, d_sp, d_e, d_s, stream);
}

extern "C" int neorx_eg2ic1(int64_t nebf, int64_t npbf, const double *de, const double *dp, const double *gm2icr,
                            const double *gm2sicr, double *fe, double *se, double *fp, double *sp, double *e_gam2,
                            double *s_gam2)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf, ng2 = ne * ne * ne * ne * np * np;
    DevBuf dDE, dDP, dG, dGs, dFE, dSE, dFP, dSP, dE;
    rc = dDE.alloc(ne * ne * 8);
    if (!rc) rc = dDP.alloc(np * np * 8);
    if (!rc) rc = dFE.alloc(ne * ne * 8);
    if (!rc) rc = dSE.alloc(ne * ne * 8);
    if (!rc) rc = dFP.alloc(np * np * 8);
    if (!rc) rc = dSP.alloc(np * np * 8);
    if (!rc) rc = dE.alloc(2 * 8);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dDE.p, de, ne * ne * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dDP.p, dp, np * np * 8, hipMemcpyHostToDevice));
    if (!rc) rc = h2d(dG, gm2icr, ng2);
    if (!rc) rc = h2d(dGs, gm2sicr, ng2);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dFE.p, fe, ne * ne * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dFP.p, fp, np * np * 8, hipMemcpyHostToDevice));
    rc = neorx_eg2ic1_dev(ne, np, dDE.d(), dDP.d(), dG.d(), dGs.d(), dFE.d(), dSE.d(), dFP.d(), dSP.d(), dE.d(),
                          dE.d() + 1, 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    double es[2];
    HIPCHK(hipMemcpy(fe, dFE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(se, dSE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fp, dFP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sp, dSP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(es, dE.p, 16, hipMemcpyDeviceToHost));
    *e_gam2 = es[0];
    *s_gam2 = es[1];
    return NEORX_OK;
}

// EG3IC1 (thread_EG3ICR, gam3_IC1_OMP.f:537-725).  Block = slab (ip,jp,i1=a)
// of GM3, entries g = G[a][b][c][d][e][f] (i1, j-slot 1, i2, j-slot 2, i3,
// j-slot 3), stored [a][b][c][d][e][f] with (e,f) fastest.  thread_EG3ICR's
// six ket-permuted reads, weighted (8,-4,2,-4,2,-4)/8 and times 3 for FE,
// become, per stored entry,
//   F[b] += 3 ( X(c,d)Y(e,f) - 1/2 X(c,f)Y(e,d) ) g
//   F[d] += 3 (-1/2 X(c,b)Y(e,f) + 1/4 X(c,f)Y(e,b) ) g
//   F[f] += 3 ( 1/4 X(c,b)Y(e,d) - 1/2 X(c,d)Y(e,b) ) g
// with X = Y = DE for EG3IC1 (F = FE row a before the DP(ip,jp) weight).
// Because GM3 is symmetric under exchanging its electron pairs (c,d) and
// (e,f), the rows are a symmetric bilinear form T(X,Y) of the two densities,
// so the difference of two Fock builds is one pass:
//   T(D,D) - T(D0,D0) = T(D - D0, D + D0)
// (the incremental build of the mixed-precision SCF).
// Each thread owns one (e,f) column pair, so every read of the slab is a
// contiguous row of ne^2 values whose base is the same for the whole
// wavefront (a scalar base plus the lane's 32-bit offset); the sums over c run
// in registers.  F[b] and F[d] are wave-reduced and added by lane 0 into its
// wavefront's own LDS row; F[f] is kept per (e,f) column; the rows and columns
// are summed in a fixed order at the end, so the result does not depend on
// wavefront timing.  T is the storage type of GM3 (double, float or BF16);
// all arithmetic is FP64.
// part[blk][0..ne-1] = F, part[blk][ne] = sum_x Z(a,x) F[x] / 3 (0 when Z is
// null).
// BF16 storage: the upper 16 bits of the FP32 value (round to nearest even)
struct bf16_t {
    uint16_t bits;
};

// nontemporal load of one stored tensor entry in its storage bits
// (ld_raw), and the stored value widened to FP64 (widen)
NRX_DI double ld_raw(const double *p) { return __builtin_nontemporal_load(p); }
NRX_DI float ld_raw(const float *p) { return __builtin_nontemporal_load(p); }
NRX_DI unsigned short ld_raw(const bf16_t *p) { return __builtin_nontemporal_load((const unsigned short *)p); }
NRX_DI double widen(double v) { return v; }
NRX_DI double widen(float v) { return (double)v; }
NRX_DI double widen(unsigned short b) { return (double)__uint_as_float((unsigned)b << 16); }
template <class T> NRX_DI double ld_entry(const T *p) { return widen(ld_raw(p)); }
template <class T> using raw_t = decltype(ld_raw((const T *)nullptr));

// Raw buffer loads of stored entries: a buffer resource over one run of the
// tensor (base and byte size, in scalar registers), a per-lane byte offset
// and a wave-uniform byte offset that the scalar unit computes, so a row
// address costs no vector instruction.  Aux bit 1 marks the load
// nontemporal.  Bytes past the resource's size read as zero.
constexpr int BUF_WORD3 = 0x00020000; // raw (untyped) buffer, gfx9 family
template <class T> NRX_DI __amdgpu_buffer_rsrc_t buf_rsrc(const T *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, BUF_WORD3);
}
NRX_DI double buf_ld(const double *, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}
NRX_DI float buf_ld(const float *, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));
}
NRX_DI unsigned short buf_ld(const bf16_t *, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 2);
}

// Compiler barrier between a batch of independent loads and their first
// use: memory operations cannot move across it, so every load of the batch
// is issued before the widening and FMAs that consume them (left to itself
// the scheduler issued narrow loads one at a time, each followed by a wait).
#define NRX_LOADS_ISSUED() asm volatile("" ::: "memory")

// loads in flight per lane in the EG3 slab loop when no tuned value is
// recorded: 16 for FP64 storage and 12 for FP32 (both measured best at
// nebf = 24), 16 for BF16
template <class T> constexpr int eg3_unroll() { return sizeof(T) == 4 ? 12 : 16; }

// most lanes of one EG3 slab block (its launch bound)
constexpr unsigned EG3_MAX_THREADS = 512;

// P adjacent (e,f) columns of one stored row in one buffer load: P = 1 for
// every storage type, P = 2 for BF16 (a 4-byte load, so a wavefront keeps
// twice the bytes in flight per load; the lower address is the low half)
template <class T, int P> struct Pack;
template <class T> struct Pack<T, 1> {
    using raw = raw_t<T>;
    static NRX_DI raw ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
    {
        return buf_ld((const T *)nullptr, r, voff, soff);
    }
    static NRX_DI double get(raw v, int) { return widen(v); }
};
template <> struct Pack<bf16_t, 1> {
    using raw = unsigned short;
    static NRX_DI raw ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
    {
        return buf_ld((const bf16_t *)nullptr, r, voff, soff);
    }
    static NRX_DI float get(raw v, int) { return __uint_as_float((unsigned)v << 16); }
};
template <> struct Pack<bf16_t, 2> {
    using raw = unsigned;
    static NRX_DI raw ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
    {
        return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2);
    }
    static NRX_DI float get(raw v, int p) { return __uint_as_float(p ? (v & 0xffff0000u) : (v << 16)); }
};

// 8-byte loads of two FP32 or four BF16 entries (the lower address in the
// low bits)
template <> struct Pack<float, 2> {
    using raw = unsigned long long;
    static NRX_DI raw ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
    {
        return __builtin_bit_cast(raw, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
    }
    static NRX_DI double get(raw v, int p) { return (double)__uint_as_float(p ? (unsigned)(v >> 32) : (unsigned)v); }
};
template <> struct Pack<bf16_t, 4> {
    using raw = unsigned long long;
    static NRX_DI raw ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
    {
        return __builtin_bit_cast(raw, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
    }
    static NRX_DI float get(raw v, int p)
    {
        const unsigned w = p < 2 ? (unsigned)v : (unsigned)(v >> 32);
        return __uint_as_float((p & 1) ? (w & 0xffff0000u) : (w << 16));
    }
};

// Arithmetic type of the sums over c: FP64, except for BF16 storage, whose
// sums run in FP32 (the storage rounding, 2^-9 relative, dwarfs FP32's; the
// pass is bound by instruction issue, and FP32 saves the widening to FP64).
// The sums are widened to FP64 before they are weighted and accumulated.
template <class T> struct SlabAcc {
    using type = double;
};
template <> struct SlabAcc<bf16_t> {
    using type = float;
};

// R rows of the slab, c = c0 .. c0+R-1, for one (b,d) and the lane's P
// columns: raw loads first (the buffer covers row block b, each row's byte
// offset is wave-uniform), then the widening and three FMAs per entry
template <class T, int R, int P, class A>
NRX_DI void eg3_rows(__amdgpu_buffer_rsrc_t rb, unsigned row_d, unsigned row_c, unsigned voff, unsigned c0,
                     const A *__restrict__ xd, const A *__restrict__ xb, const A *const *xf, A *s1, A *s2, A *s3)
{
    typename Pack<T, P>::raw r[R];
#pragma unroll
    for (int k = 0; k < R; ++k) r[k] = Pack<T, P>::ld(rb, voff, row_d + (c0 + k) * row_c);
    NRX_LOADS_ISSUED();
#pragma unroll
    for (int k = 0; k < R; ++k) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const A gv = (A)Pack<T, P>::get(r[k], p);
            s1[p] += xd[c0 + k] * gv;      // X(c,d)
            s2[p] += xb[c0 + k] * gv;      // X(c,b)
            s3[p] += xf[p][c0 + k] * gv;   // X(c,f)
        }
    }
}

// X, Y: the two densities (column-major ne x ne); Z: the density of the
// nuclear-row term, or null.  P columns per lane (P divides ne).
template <class T, int U, int P>
__global__ void __launch_bounds__(EG3_MAX_THREADS) k_eg3_slab(i64 ne_, const double *__restrict__ X,
                                                              const double *__restrict__ Y, const double *__restrict__ Z,
                                                              const T *__restrict__ G, double *__restrict__ part)
{
    // LDS: Fw[nw][ne] (per-wavefront rows), Ff[ne^2] (per-column partials),
    // Xf[ne][ne+1] with Xf[f*(ne+1) + c] = X(c,f) (a lane reads its own
    // column f at consecutive c; the odd stride spreads the lanes' columns
    // over the banks), Yt[ne^2] with Yt[e*ne + x] = Y(e,x), and Xs, Xf in
    // FP32 for the FP32 sums of BF16 storage
    using A = typename SlabAcc<T>::type;
    constexpr bool F32 = sizeof(A) == 4;
    extern __shared__ double sm[];
    const unsigned ne = (unsigned)ne_, xs = ne + 1;
    const i64 blk = blockIdx.x;
    const unsigned a = (unsigned)(blk % ne);
    const unsigned ne2 = ne * ne, ne3 = ne2 * ne, ne4 = ne3 * ne, ncol = ne2 / P;
    const unsigned nw = blockDim.x >> 6;
    double *Fw = sm + (threadIdx.x >> 6) * ne, *Ff = sm + nw * ne, *Xf = Ff + ne2, *Yt = Xf + ne * xs;
    float *Xs = (float *)(Yt + ne2);
    const T *g = G + blk * (i64)ne4 * ne;
    for (unsigned k = threadIdx.x; k < nw * ne + ne2; k += blockDim.x) sm[k] = 0.0;
    for (unsigned k = threadIdx.x; k < ne2; k += blockDim.x) {
        const unsigned r = k / ne, q = k % ne;
        Xf[r * xs + q] = X[k];     // X(q, r)
        Xs[r * xs + q] = (float)X[k];
        Yt[k] = Y[q * ne + r];     // Y(r, q)
    }
    __syncthreads();
    // the X columns of the sums: global FP64 (wave-uniform, scalar loads) or
    // LDS FP32
    const A *Xa = F32 ? (const A *)(const void *)Xs : (const A *)(const void *)X;
    const unsigned xstride = F32 ? xs : ne;
    const bool lead = (threadIdx.x & 63) == 0;
    for (unsigned q0 = 0; q0 < ncol; q0 += blockDim.x) {
        const unsigned qc = q0 + threadIdx.x;
        const bool on = qc < ncol;
        const unsigned ef = on ? qc * P : 0;
        const unsigned e = ef / ne, f = ef % ne; // the P columns share e
        const double *Ye = Yt + e * ne;          // Ye[x] = Y(e,x)
        const A *xf[P];
        double y_ef[P], accf[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            // xf[p][c] = X(c,f+p)
            xf[p] = F32 ? (const A *)(const void *)(Xs + (f + p) * xs) : (const A *)(const void *)(Xf + (f + p) * xs);
            y_ef[p] = on ? Ye[f + p] : 0.0;
            accf[p] = 0.0;
        }
        const unsigned voff = ef * (unsigned)sizeof(T);
        for (unsigned b = 0; b < ne; ++b) {
            const double y_eb = on ? Ye[b] : 0.0;
            // row block b of the slab: ne^4 entries, at most 2^31 bytes for
            // nebf <= EG3_MAX_NE
            const __amdgpu_buffer_rsrc_t rb = buf_rsrc(g + (size_t)b * ne4, ne4 * (unsigned)sizeof(T));
            const A *xb = Xa + b * xstride;
            double accb = 0.0;
            for (unsigned d = 0; d < ne; ++d) {
                const double y_ed = on ? Ye[d] : 0.0;
                A s1[P], s2[P], s3[P];
#pragma unroll
                for (int p = 0; p < P; ++p) s1[p] = s2[p] = s3[p] = (A)0;
                if (on) {
                    const unsigned row_d = d * ne2 * (unsigned)sizeof(T), row_c = ne3 * (unsigned)sizeof(T);
                    const A *xd = Xa + d * xstride;
                    unsigned c = 0;
                    // U (then eight) independent loads in flight per lane
                    // before the FMAs; the slab is read once, so the loads are
                    // nontemporal
                    for (; c + U <= ne; c += U)
                        eg3_rows<T, U, P, A>(rb, row_d, row_c, voff, c, xd, xb, xf, s1, s2, s3);
                    if (c + 8 <= ne) {
                        eg3_rows<T, 8, P, A>(rb, row_d, row_c, voff, c, xd, xb, xf, s1, s2, s3);
                        c += 8;
                    }
                    for (; c < ne; ++c) eg3_rows<T, 1, P, A>(rb, row_d, row_c, voff, c, xd, xb, xf, s1, s2, s3);
                }
                double vd = 0.0;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const double t1 = s1[p], t2 = s2[p], t3 = s3[p];
                    accb += y_ef[p] * t1 - 0.5 * y_ed * t3;
                    accf[p] += 0.25 * y_ed * t2 - 0.5 * y_eb * t1;
                    vd += -0.5 * y_ef[p] * t2 + 0.25 * y_eb * t3;
                }
                vd = wave_sum(vd);
                if (lead) Fw[d] += 3.0 * vd;
            }
            accb = wave_sum(accb);
            if (lead) Fw[b] += 3.0 * accb;
        }
        if (on) {
#pragma unroll
            for (int p = 0; p < P; ++p) Ff[ef + p] = 3.0 * accf[p]; // each column belongs to one lane
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        double p = 0.0;
        for (unsigned x = threadIdx.x; x < ne; x += 64) {
            double fx = 0.0;
            for (unsigned w = 0; w < nw; ++w) fx += sm[w * ne + x];
            for (unsigned e = 0; e < ne; ++e) fx += Ff[e * ne + x];
            part[blk * (ne + 1) + x] = fx;
            if (Z) p += Z[x * ne + a] * fx;
        }
        p = wave_sum(p);
        if (threadIdx.x == 0) part[blk * (ne + 1) + ne] = p / 3.0;
    }
}

// LDS bytes of k_eg3_slab: per-wavefront rows, per-column partials and the
// two density layouts
static inline size_t eg3_lds(i64 ne, unsigned threads)
{
    return (size_t)((threads / 64) * ne + 2 * ne * ne + ne * (ne + 1)) * 8 + (size_t)ne * (ne + 1) * 4;
}

// largest nebf whose e-e-e-p slab (nebf^5 entries) has 32-bit offsets
static const i64 EG3_MAX_NE = 73;

// lanes per slab block for n lane columns: at most EG3_MAX_THREADS (the block
// loops over its columns), the multiple of 64 that leaves the fewest idle
// lanes in the last pass, the larger one on a tie
static inline unsigned eg3_threads_n(i64 n)
{
    unsigned best = 64;
    i64 waste = ((n + 63) / 64) * 64 - n;
    for (unsigned t = 128; t <= EG3_MAX_THREADS; t += 64) {
        const i64 w = ((n + t - 1) / t) * t - n;
        if (w <= waste) {
            best = t;
            waste = w;
        }
    }
    return best;
}
static inline unsigned eg3_threads(i64 ne) { return eg3_threads_n(ne * ne); }

// columns per lane: two for BF16 storage when nebf is even (the pair shares
// e and its first entry lies on a 4-byte boundary), else one
static inline int eg3_pack(int dtype, i64 ne) { return dtype == NRX_BF16 && ne % 2 == 0 ? 2 : 1; }

// ------------------------------------------------------------------
// autotuned EG3 slab launches: the loads-in-flight depth U of k_eg3_slab per
// (storage type, nebf), chosen by neorx_autotune_eg3_dev from timed launches
// on the caller's tensor; 0 = not tuned (eg3_unroll's default)
// ------------------------------------------------------------------
static const int EG3_UNROLLS[4] = {8, 12, 16, 24};
static int g_eg3_tuned[3][EG3_MAX_NE + 1] = {};

// the densities of one launch: X on the (c,d) pair, Y on the (e,f) pair, Z
// for the nuclear-row term (or null)
struct Eg3Dens {
    const double *X, *Y, *Z;
};

template <class T, int P>
static int eg3_launch_u(int u, i64 ne, i64 nblk, Eg3Dens D, const T *g, double *part, hipStream_t st)
{
    const unsigned nt = eg3_threads_n(ne * ne / P);
    const size_t lds = eg3_lds(ne, nt);
    const dim3 grid((unsigned)nblk), blk(nt);
    switch (u) {
    case 8: hipLaunchKernelGGL((k_eg3_slab<T, 8, P>), grid, blk, lds, st, ne, D.X, D.Y, D.Z, g, part); break;
    case 12: hipLaunchKernelGGL((k_eg3_slab<T, 12, P>), grid, blk, lds, st, ne, D.X, D.Y, D.Z, g, part); break;
    case 16: hipLaunchKernelGGL((k_eg3_slab<T, 16, P>), grid, blk, lds, st, ne, D.X, D.Y, D.Z, g, part); break;
    case 24: hipLaunchKernelGGL((k_eg3_slab<T, 24, P>), grid, blk, lds, st, ne, D.X, D.Y, D.Z, g, part); break;
    default: return fail(NEORX_E_ARG, "EG3 slabs: no variant with %d loads in flight", u);
    }
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

static int eg3_launch(int dtype, int u, i64 ne, i64 nblk, Eg3Dens D, const void *g, double *part, hipStream_t st)
{
    switch (dtype) {
    case NRX_F64:
        return eg3_launch_u<double, 1>(u ? u : eg3_unroll<double>(), ne, nblk, D, (const double *)g, part, st);
    case NRX_F32: return eg3_launch_u<float, 1>(u ? u : eg3_unroll<float>(), ne, nblk, D, (const float *)g, part, st);
    case NRX_BF16:
        if (eg3_pack(dtype, ne) == 2)
            return eg3_launch_u<bf16_t, 2>(u ? u : eg3_unroll<bf16_t>(), ne, nblk, D, (const bf16_t *)g, part, st);
        return eg3_launch_u<bf16_t, 1>(u ? u : eg3_unroll<bf16_t>(), ne, nblk, D, (const bf16_t *)g, part, st);
    default: return fail(NEORX_E_ARG, "EG3 slabs: unknown storage type %d", dtype);
    }
}

int nrx_eg3_slabs2(i64 ne, i64 nblk, const double *d_x, const double *d_y, const double *d_z, const void *g, int dtype,
                   double *part, hipStream_t st)
{
    if (ne <= 0 || ne > EG3_MAX_NE) return fail(NEORX_E_UNSUPPORTED, "EG3 slabs: nebf=%lld not in 1..%lld",
                                                (long long)ne, (long long)EG3_MAX_NE);
    if (dtype < 0 || dtype > 2) return fail(NEORX_E_ARG, "EG3 slabs: unknown storage type %d", dtype);
    return eg3_launch(dtype, g_eg3_tuned[dtype][ne], ne, nblk, Eg3Dens{d_x, d_y, d_z}, g, part, st);
}

int nrx_eg3_slabs(i64 ne, i64 nblk, const double *d_de, const void *g, int dtype, double *part, hipStream_t st)
{
    return nrx_eg3_slabs2(ne, nblk, d_de, d_de, d_de, g, dtype, part, st);
}

extern "C" int neorx_autotune_eg3_dev(int64_t nebf, int64_t npbf, int64_t dtype, const double *d_de, const void *d_gm3,
                                      int64_t reps, int64_t *best_unroll, double *best_ms, double *all_ms)
{
    int rc = ensure_init();
    if (rc) return rc;
    const i64 ne = nebf, np = npbf;
    if (ne <= 0 || ne > EG3_MAX_NE || np <= 0) return fail(NEORX_E_ARG, "autotune EG3: bad sizes");
    if (dtype < 0 || dtype > 2) return fail(NEORX_E_ARG, "autotune EG3: unknown storage type %lld", (long long)dtype);
    if (!d_de || !d_gm3 || reps < 1) return fail(NEORX_E_ARG, "autotune EG3: bad arguments");
    const i64 nblk = np * np * ne;
    DevBuf part;
    rc = part.alloc(nblk * (ne + 1) * 8);
    if (rc) return rc;
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreate(&st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int best = 0;
    float bt = 0.0f;
    for (int k = 0; k < 4 && !rc; ++k) {
        const int u = EG3_UNROLLS[k];
        rc = eg3_launch((int)dtype, u, ne, nblk, Eg3Dens{d_de, d_de, d_de}, d_gm3, part.d(), st); // warm-up
        if (!rc) rc = hipEventRecord(e0, st) == hipSuccess ? 0 : fail(NEORX_E_HIP, "hipEventRecord failed");
        for (i64 r = 0; r < reps && !rc; ++r)
            rc = eg3_launch((int)dtype, u, ne, nblk, Eg3Dens{d_de, d_de, d_de}, d_gm3, part.d(), st);
        if (!rc) rc = hipEventRecord(e1, st) == hipSuccess ? 0 : fail(NEORX_E_HIP, "hipEventRecord failed");
        if (!rc) rc = hipEventSynchronize(e1) == hipSuccess ? 0 : fail(NEORX_E_HIP, "hipEventSynchronize failed");
        float ms = 0.0f;
        if (!rc) (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= (float)reps;
        if (all_ms) all_ms[k] = ms;
        if (!rc && (best == 0 || ms < bt)) {
            best = u;
            bt = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (rc) return rc;
    g_eg3_tuned[dtype][ne] = best;
    if (best_unroll) *best_unroll = best;
    if (best_ms) *best_ms = bt;
    return NEORX_OK;
}

__global__ void k_eg3_reduce(i64 ne, i64 np, const double *__restrict__ DP, const double *__restrict__ part,
                             double *__restrict__ FE, double *__restrict__ FPd)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    i64 ne2 = ne * ne, np2 = np * np;
    if (t < ne2) {
        i64 a = t % ne, b = t / ne;
        double f = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) {
            i64 ip = pp / np, jp = pp % np;
            f += DP[jp * np + ip] * part[(pp * ne + a) * (ne + 1) + b];
        }
        FE[b * ne + a] += f;
    } else if (t < ne2 + np2) {
        i64 pp = t - ne2, ip = pp / np, jp = pp % np;
        double f = 0.0;
        for (i64 a = 0; a < ne; ++a) f += part[(pp * ne + a) * (ne + 1) + ne];
        FPd[jp * np + ip] = f;
    }
}

template <class T>
static int eg3_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const T *d_gm3, double *d_fe,
                   double *d_fp, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    if (ne > EG3_MAX_NE) return fail(NEORX_E_UNSUPPORTED, "EG3IC1: nebf=%lld > %lld", (long long)ne, (long long)EG3_MAX_NE);
    hipStream_t st = (hipStream_t)stream;
    i64 nblk = np * np * ne;
    DevBuf part, fpd;
    rc = part.alloc(nblk * (ne + 1) * 8);
    if (!rc) rc = fpd.alloc(np * np * 8);
    if (rc) return rc;
    rc = nrx_eg3_slabs(ne, nblk, d_de, d_gm3, sizeof(T) == 8 ? NRX_F64 : NRX_F32, part.d(), st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_eg3_reduce, dim3(nblocks(ne * ne + np * np, 256)), dim3(256), 0, st, ne, np, d_dp, part.d(),
                       d_fe, fpd.d());
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st, np, d_dp, fpd.d(), nullptr, d_e, nullptr);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(np * np, 256)), dim3(256), 0, st, np * np, fpd.d(), d_fp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_eg3ic1_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                                const double *d_gm3, double *d_fe, double *d_fp, double *d_e, void *stream)
{
    return eg3_dev<double>(nebf, npbf, d_de, d_dp, d_gm3, d_fe, d_fp, d_e, stream);
}

// EG3IC1 without any allocation: the slab partials and the nuclear partial
// sums live in the current device's scratch array.  Every scratch element the
// kernels read is written earlier in the same call, so nothing is cleared and
// the previous call's contents never reach the results.  Enqueued on the
// stream and not synchronised (the existing entry point synchronises because
// it frees its buffers).  Calls that overlap on one device share the array
// and race.
extern "C" int neorx_eg3ic1_noalloc_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                                        const double *d_gm3, double *d_fe, double *d_fp, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    const i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    if (ne > EG3_MAX_NE) return fail(NEORX_E_UNSUPPORTED, "EG3IC1: nebf=%lld > %lld", (long long)ne, (long long)EG3_MAX_NE);
    if (!d_de || !d_dp || !d_gm3 || !d_fe || !d_fp || !d_e) return fail(NEORX_E_ARG, "null device pointer");
    const i64 nblk = np * np * ne;
    if (np > SCRATCH_DOUBLES || nblk > SCRATCH_DOUBLES / (ne + 1) ||
        nblk * (ne + 1) + np * np > SCRATCH_DOUBLES)
        return fail(NEORX_E_ARG,
                    "neorx_eg3ic1_noalloc_dev: nebf=%lld npbf=%lld needs npbf^2*nebf*(nebf+1)+npbf^2 scratch doubles, "
                    "more than the %lld the library holds",
                    (long long)ne, (long long)np, (long long)SCRATCH_DOUBLES);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES || !g_scratch[dev])
        return fail(NEORX_E_NODEVICE, "neorx_eg3ic1_noalloc_dev: device %d was not set up by neorx_init", dev);
    double *part = g_scratch[dev], *fpd = part + nblk * (ne + 1);
    hipStream_t st = (hipStream_t)stream;
    rc = nrx_eg3_slabs(ne, nblk, d_de, d_gm3, NRX_F64, part, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_eg3_reduce, dim3(nblocks(ne * ne + np * np, 256)), dim3(256), 0, st, ne, np, d_dp, part,
                       d_fe, fpd);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st, np, d_dp, fpd, nullptr, d_e, nullptr);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(np * np, 256)), dim3(256), 0, st, np * np, fpd, d_fp);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

extern "C" int neorx_eg3ic1_f32_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                                    const float *d_gm3, double *d_fe, double *d_fp, double *d_e, void *stream)
{
    return eg3_dev<float>(nebf, npbf, d_de, d_dp, d_gm3, d_fe, d_fp, d_e, stream);
}

extern "C" int neorx_eg3ic1(int64_t nebf, int64_t npbf, const double *de, const double *dp, const double *gm3icr,
                            double *focke, double *fockp, double *e_gam3)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf, ng3 = ne * ne * ne * ne * ne * ne * np * np;
    DevBuf dDE, dDP, dG, dFE, dFP, dE;
    rc = dDE.alloc(ne * ne * 8);
    if (!rc) rc = dDP.alloc(np * np * 8);
    if (!rc) rc = dFE.alloc(ne * ne * 8);
    if (!rc) rc = dFP.alloc(np * np * 8);
    if (!rc) rc = dE.alloc(8);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dDE.p, de, ne * ne * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dDP.p, dp, np * np * 8, hipMemcpyHostToDevice));
    rc = h2d(dG, gm3icr, ng3);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dFE.p, focke, ne * ne * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dFP.p, fockp, np * np * 8, hipMemcpyHostToDevice));
    rc = neorx_eg3ic1_dev(ne, np, dDE.d(), dDP.d(), dG.d(), dFE.d(), dFP.d(), dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(focke, dFE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fockp, dFP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e_gam3, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

// EG4ICR (thread_EG4ICR, gam4_ICR_OMP.f:390-490).  Block (ip,jp,i1=a) of
// GM4, entries G[a][j1][i2][j2][i3][j3][i4][j4]: for every j1 the block sums
// F(j1) = sum DE(i2,j2) DE(i3,j3) DE(i4,j4) G over the ne^6 entries behind it
// (lanes stride over them, the wave sums by shuffles, the waves meet in LDS
// slots summed in a fixed order), then writes the partial Fock row 4 F(j1)
// and the proton term sum_j1 DE(j1,a) F(j1).  ONFLY evaluates symm_gam4/16 per
// entry from GAM_2s and GAM_ee (thread_gam4_DSCF, gam4_DSCF_OMP.f:322-420)
// instead of reading the in-core tensor.
template <int NT, bool ONFLY, class T>
__global__ void __launch_bounds__(NT) k_eg4_slab(i64 ne, i64 np, const double *__restrict__ DE,
                                                 const T *__restrict__ G, const double *__restrict__ G2s,
                                                 const double *__restrict__ Gee, double *__restrict__ part)
{
    constexpr int NW = NT / 64;
    extern __shared__ double sm[]; // [NW][ne] per-wave sums
    const i64 blk = blockIdx.x, a = blk % ne, pp = blk / ne, ip = pp / np, jp = pp % np;
    const i64 ne2 = ne * ne, ne4 = ne2 * ne2, ne6 = ne4 * ne2;
    const int w = threadIdx.x >> 6;
    for (i64 j1 = 0; j1 < ne; ++j1) {
        double acc = 0.0;
        for (i64 r = threadIdx.x; r < ne6; r += NT) {
            const i64 j4 = r % ne, i4 = (r / ne) % ne, j3 = (r / ne2) % ne, i3 = (r / (ne2 * ne)) % ne;
            const i64 j2 = (r / ne4) % ne, i2 = r / (ne4 * ne);
            double gv;
            if constexpr (ONFLY) {
                i64 br[4] = {a, i2, i3, i4}, kt[4] = {j1, j2, j3, j4};
                gv = gam4_value(ne, np, G2s, Gee, ip, jp, br, kt);
            } else {
                gv = (double)G[(blk * ne + j1) * ne6 + r];
            }
            acc += DE[j2 * ne + i2] * DE[j3 * ne + i3] * DE[j4 * ne + i4] * gv;
        }
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0) sm[w * ne + j1] = acc;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        double p = 0.0;
        for (i64 j1 = threadIdx.x; j1 < ne; j1 += 64) {
            double f = 0.0;
            for (int u = 0; u < NW; ++u) f += sm[u * ne + j1];
            part[blk * (ne + 1) + j1] = 4.0 * f;
            p += DE[j1 * ne + a] * f;
        }
        p = wave_sum(p);
        if (threadIdx.x == 0) part[blk * (ne + 1) + ne] = p;
    }
}

template <class T>
static int eg4_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const T *d_gm4, double *d_fe,
                   double *d_fp, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    hipStream_t st = (hipStream_t)stream;
    i64 nblk = np * np * ne;
    DevBuf part, fpd;
    rc = part.alloc(nblk * (ne + 1) * 8);
    if (!rc) rc = fpd.alloc(np * np * 8);
    if (rc) return rc;
    hipLaunchKernelGGL((k_eg4_slab<256, false, T>), dim3((unsigned)nblk), dim3(256), 4 * ne * 8, st, ne, np, d_de,
                       d_gm4, nullptr, nullptr, part.d());
    hipLaunchKernelGGL(k_eg3_reduce, dim3(nblocks(ne * ne + np * np, 256)), dim3(256), 0, st, ne, np, d_dp, part.d(),
                       d_fe, fpd.d());
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st, np, d_dp, fpd.d(), nullptr, d_e, nullptr);
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(np * np, 256)), dim3(256), 0, st, np * np, fpd.d(), d_fp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_eg4icr_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                                const double *d_gm4, double *d_fe, double *d_fp, double *d_e, void *stream)
{
    return eg4_dev<double>(nebf, npbf, d_de, d_dp, d_gm4, d_fe, d_fp, d_e, stream);
}

extern "C" int neorx_eg4icr_f32_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                                    const float *d_gm4, double *d_fe, double *d_fp, double *d_e, void *stream)
{
    return eg4_dev<float>(nebf, npbf, d_de, d_dp, d_gm4, d_fe, d_fp, d_e, stream);
}

extern "C" int neorx_eg4icr(int64_t nebf, int64_t npbf, const double *de, const double *dp, const double *gm4icr,
                            double *focke, double *fockp, double *e_gam4)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf, ng4 = ne * ne * ne * ne * ne * ne * ne * ne * np * np;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    DevBuf dDE, dDP, dG, dFE, dFP, dE;
    rc = h2d(dDE, de, ne * ne);
    if (!rc) rc = h2d(dDP, dp, np * np);
    if (!rc) rc = h2d(dG, gm4icr, ng4);
    if (!rc) rc = h2d(dFE, focke, ne * ne);
    if (!rc) rc = h2d(dFP, fockp, np * np);
    if (!rc) rc = dE.alloc(8);
    if (!rc) rc = eg4_dev<double>(ne, np, dDE.d(), dDP.d(), dG.d(), dFE.d(), dFP.d(), dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipMemcpy(focke, dFE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fockp, dFP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e_gam4, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

// ------------------------------------------------------------------
// mixed precision: FP32 copies of the in-core tensors and the switch back
// to FP64
// ------------------------------------------------------------------
__global__ void k_to_f32(i64 n, const double *__restrict__ x, float *__restrict__ y)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x)
        y[t] = (float)x[t];
}

extern "C" int neorx_convert_f32_dev(int64_t n, const double *d_src, float *d_dst, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!d_src || !d_dst))) return fail(NEORX_E_ARG, "bad conversion arguments");
    if (n == 0) return NEORX_OK;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)std::min<i64>(nblocks(n, 256), 65536)), dim3(256), 0,
                       (hipStream_t)stream, n, d_src, d_dst);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

// state[0] = 1 once FP64 is selected (sticky), state[1] = 1 once a previous
// density is held, state[2..] = previous DE; out = max|DE - previous|
__global__ void k_switch(i64 n, const double *__restrict__ de, double *__restrict__ state, double threshold,
                         double *__restrict__ out)
{
    __shared__ double red[256];
    double m = 0.0;
    bool have = state[1] != 0.0;
    for (i64 k = threadIdx.x; k < n; k += blockDim.x) {
        double d = have ? fabs(de[k] - state[2 + k]) : INFINITY;
        m = d > m ? d : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    for (i64 k = threadIdx.x; k < n; k += blockDim.x) state[2 + k] = de[k];
    if (threadIdx.x == 0) {
        double delta = red[0];
        if (have && delta < threshold) state[0] = 1.0;
        state[1] = 1.0;
        out[0] = delta;
        out[1] = state[0];
    }
}

extern "C" int neorx_precision_switch_dev(int64_t nebf, const double *d_de, double *d_state, double threshold,
                                          int64_t *use_fp64, double *delta, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (nebf <= 0 || !d_de || !d_state || !use_fp64) return fail(NEORX_E_ARG, "bad switch arguments");
    hipStream_t st = (hipStream_t)stream;
    DevBuf out;
    rc = out.alloc(16);
    if (rc) return rc;
    hipLaunchKernelGGL(k_switch, dim3(1), dim3(256), 0, st, nebf * nebf, d_de, d_state, threshold, out.d());
    HIPCHK(hipGetLastError());
    double h[2];
    HIPCHK(hipMemcpyAsync(h, out.p, 16, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *use_fp64 = h[1] != 0.0;
    if (delta) *delta = h[0];
    return NEORX_OK;
}

// ------------------------------------------------------------------
// direct-SCF routines: contracted values are rebuilt per call, one block of
// (ip,jp) pairs at a time, symmetrised and contracted into the Fock partials
// of that block; no ng2/ng3/ng4 tensor is stored
// ------------------------------------------------------------------
extern "C" int neorx_gam2_dscf_dev(int64_t handle, const double *d_de, const double *d_dp, double *d_fe,
                                   double *d_fp, double *d_se2, double *d_sp2, double *d_e, double *d_s, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (!d_de || !d_dp || !d_fe || !d_fp || !d_se2 || !d_sp2 || !d_e || !d_s)
        return fail(NEORX_E_ARG, "null argument");
    i64 ne = s->d.nebf, np = s->d.npbf, ne4 = ne * ne * ne * ne, npp = np * np;
    hipStream_t st = (hipStream_t)stream;
    i64 nb = std::max<i64>(1, std::min<i64>(npp, BLOCK_CAP / ne4));
    DevBuf X, Xs, G, Gs, Y, Ys, part, fpd, fpsd;
    i64 cap = slab_cap(s, 2);
    int rc = X.alloc(nb * ne4 * 8);
    if (!rc) rc = Xs.alloc(nb * ne4 * 8);
    if (!rc) rc = G.alloc(nb * ne4 * 8);
    if (!rc) rc = Gs.alloc(nb * ne4 * 8);
    if (!rc) rc = Y.alloc(cap * 8);
    if (!rc) rc = Ys.alloc(cap * 8);
    if (!rc) rc = part.alloc(npp * ne * 4 * ne * 8);
    if (!rc) rc = fpd.alloc(npp * 8);
    if (!rc) rc = fpsd.alloc(npp * 8);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(fpd.p, 0, npp * 8, st));
    HIPCHK(hipMemsetAsync(fpsd.p, 0, npp * 8, st));
    for (i64 p0 = 0; p0 < npp; p0 += nb) {
        i64 p1 = std::min(npp, p0 + nb), n = (p1 - p0) * ne4;
        rc = build2_block(s, p0, p1, X.d(), Xs.d(), Y.d(), Ys.d(), 1, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gam2_symm, dim3(grid_blocks(n, 256)), dim3(256), 0, st, ne, n, X.d(), Xs.d(), G.d(), Gs.d());
        rc = eg2_slab_launch<double>(ne, (p1 - p0) * ne, d_de, G.d(), Gs.d(), part.d() + p0 * ne * 4 * ne, st);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_eg2_reduce, dim3(nblocks(ne * ne + npp, 256)), dim3(256), 0, st, ne, np, d_de, d_dp, part.d(),
                       1, 1, d_fe, d_se2, fpd.d(), fpsd.d(), d_e, d_s);
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st, np, d_dp, fpd.d(), fpsd.d(), d_e, d_s);
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(npp, 256)), dim3(256), 0, st, npp, fpd.d(), d_fp);
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(npp, 256)), dim3(256), 0, st, npp, fpsd.d(), d_sp2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_gam3_dscf_dev(int64_t handle, const double *d_de, const double *d_dp, double *d_fe,
                                   double *d_fp, double *d_e, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (s->d.nebf > EG3_MAX_NE)
        return fail(NEORX_E_UNSUPPORTED, "GAM3_DSCF: nebf=%lld > %lld", (long long)s->d.nebf, (long long)EG3_MAX_NE);
    if (!d_de || !d_dp || !d_fe || !d_fp || !d_e) return fail(NEORX_E_ARG, "null argument");
    i64 ne = s->d.nebf, np = s->d.npbf, ne6 = ne * ne * ne * ne * ne * ne, npp = np * np;
    hipStream_t st = (hipStream_t)stream;
    i64 nb = std::max<i64>(1, std::min<i64>(npp, BLOCK_CAP / ne6));
    DevBuf X, G, Y, part, fpd;
    int rc = X.alloc(nb * ne6 * 8);
    if (!rc) rc = G.alloc(nb * ne6 * 8);
    if (!rc) rc = Y.alloc(slab_cap(s, 3) * 8);
    if (!rc) rc = part.alloc(npp * ne * (ne + 1) * 8);
    if (!rc) rc = fpd.alloc(npp * 8);
    if (rc) return rc;
    for (i64 p0 = 0; p0 < npp; p0 += nb) {
        i64 p1 = std::min(npp, p0 + nb), n = (p1 - p0) * ne6;
        rc = build3_block(s, p0, p1, X.d(), Y.d(), st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_gam3_symm, dim3(grid_blocks(n, 256)), dim3(256), 0, st, ne, n, X.d(), G.d());
        hipLaunchKernelGGL((k_eg3_slab<double, 16, 1>), dim3((unsigned)((p1 - p0) * ne)), dim3(eg3_threads(ne)),
                           eg3_lds(ne, eg3_threads(ne)), st, ne, d_de, d_de, d_de, G.d(), part.d() + p0 * ne * (ne + 1));
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_eg3_reduce, dim3(nblocks(ne * ne + npp, 256)), dim3(256), 0, st, ne, np, d_dp, part.d(), d_fe,
                       fpd.d());
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st, np, d_dp, fpd.d(), nullptr, d_e, nullptr);
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(npp, 256)), dim3(256), 0, st, npp, fpd.d(), d_fp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_gam4_dscf_dev(int64_t nebf, int64_t npbf, const double *d_gam2s, const double *d_gamee,
                                   const double *d_de, const double *d_dp, double *d_fe, double *d_fp, double *d_e,
                                   void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf, npp = np * np;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    if (!d_gam2s || !d_gamee || !d_de || !d_dp || !d_fe || !d_fp || !d_e) return fail(NEORX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    DevBuf part, fpd;
    rc = part.alloc(npp * ne * (ne + 1) * 8);
    if (!rc) rc = fpd.alloc(npp * 8);
    if (rc) return rc;
    hipLaunchKernelGGL((k_eg4_slab<256, true, double>), dim3((unsigned)(npp * ne)), dim3(256), 4 * ne * 8, st, ne, np, d_de,
                       nullptr, d_gam2s, d_gamee, part.d());
    hipLaunchKernelGGL(k_eg3_reduce, dim3(nblocks(ne * ne + npp, 256)), dim3(256), 0, st, ne, np, d_dp, part.d(), d_fe,
                       fpd.d());
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, st, np, d_dp, fpd.d(), nullptr, d_e, nullptr);
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(npp, 256)), dim3(256), 0, st, npp, fpd.d(), d_fp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

// host-array wrappers: inputs copied to HBM, outputs copied back

extern "C" int neorx_gam2_dscf(int64_t handle, const double *de, const double *dp, double *focke, double *fockp,
                               double *se2, double *sp2, double *e_gam2, double *s_gam2)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, np = s->d.npbf;
    DevBuf dDE, dDP, dFE, dFP, dSE, dSP, dE;
    int rc = h2d(dDE, de, ne * ne);
    if (!rc) rc = h2d(dDP, dp, np * np);
    if (!rc) rc = h2d(dFE, focke, ne * ne);
    if (!rc) rc = h2d(dFP, fockp, np * np);
    if (!rc) rc = h2d(dSE, se2, ne * ne);
    if (!rc) rc = h2d(dSP, sp2, np * np);
    if (!rc) rc = dE.alloc(16);
    if (!rc) rc = neorx_gam2_dscf_dev(handle, dDE.d(), dDP.d(), dFE.d(), dFP.d(), dSE.d(), dSP.d(), dE.d(), dE.d() + 1, 0);
    if (rc) return rc;
    double es[2];
    HIPCHK(hipMemcpy(focke, dFE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fockp, dFP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(se2, dSE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sp2, dSP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(es, dE.p, 16, hipMemcpyDeviceToHost));
    *e_gam2 = es[0];
    *s_gam2 = es[1];
    return NEORX_OK;
}

extern "C" int neorx_gam3_dscf(int64_t handle, const double *de, const double *dp, double *focke, double *fockp,
                               double *e_gam3)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, np = s->d.npbf;
    DevBuf dDE, dDP, dFE, dFP, dE;
    int rc = h2d(dDE, de, ne * ne);
    if (!rc) rc = h2d(dDP, dp, np * np);
    if (!rc) rc = h2d(dFE, focke, ne * ne);
    if (!rc) rc = h2d(dFP, fockp, np * np);
    if (!rc) rc = dE.alloc(8);
    if (!rc) rc = neorx_gam3_dscf_dev(handle, dDE.d(), dDP.d(), dFE.d(), dFP.d(), dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipMemcpy(focke, dFE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fockp, dFP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e_gam3, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

extern "C" int neorx_gam4_dscf(int64_t nebf, int64_t npbf, const double *gam_2s, const double *gam_ee,
                               const double *de, const double *dp, double *focke, double *fockp, double *e_gam4)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    i64 ng2 = ne * ne * ne * ne * np * np, ngee = ne * ne * ne * ne;
    DevBuf dG2, dGee, dDE, dDP, dFE, dFP, dE;
    rc = h2d(dG2, gam_2s, ng2);
    if (!rc) rc = h2d(dGee, gam_ee, ngee);
    if (!rc) rc = h2d(dDE, de, ne * ne);
    if (!rc) rc = h2d(dDP, dp, np * np);
    if (!rc) rc = h2d(dFE, focke, ne * ne);
    if (!rc) rc = h2d(dFP, fockp, np * np);
    if (!rc) rc = dE.alloc(8);
    if (!rc) rc = neorx_gam4_dscf_dev(ne, np, dG2.d(), dGee.d(), dDE.d(), dDP.d(), dFE.d(), dFP.d(), dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipMemcpy(focke, dFE.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fockp, dFP.p, np * np * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e_gam4, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

// ------------------------------------------------------------------
// XCUHF contractions (source/xcuhf/xcuhf_omg{2,3,4}.f).  One block per slab
// (ip,jp,i1); each entry reads its ket-permuted partners inside the same slab
// and forms the spin-resolved value of the reference routine.  FAE variants
// add EFAC*t to row j1 of the slab's Fock row and DAE(i1,j1)*t to its energy
// partial; UFP variants add t to the slab's scalar.  part[blk] = [F row (ne),
// S row (ne), energy, overlap energy]; k_uhf_reduce applies DP(ip,jp).
// ------------------------------------------------------------------
// ket permutations of the 24 reads of calc_FAEOMG4 / calc_UFPOMG4, in the
// reference's order (OMG, OMG_23, OMG_12, OMG_12_23, OMG_13_23, OMG_13,
// OMG_24, OMG_12_24, OMG_12_14, OMG_14, OMG_34, OMG_12_34, OMG_13_34,
// OMG_14_34, OMG_13_24, OMG_14_23, OMG_23_34, OMG_24_34, OMG_12_23_34,
// OMG_12_23_24, OMG_12_13_34, OMG_13_23_24, OMG_14_24_34, OMG_14_23_34)
__constant__ int c_omg4[24][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {1, 0, 2, 3}, {1, 2, 0, 3}, {2, 0, 1, 3}, {2, 1, 0, 3},
                                  {0, 3, 2, 1}, {1, 3, 2, 0}, {3, 0, 2, 1}, {3, 1, 2, 0}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                  {2, 1, 3, 0}, {3, 1, 0, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}, {0, 2, 3, 1}, {0, 3, 1, 2},
                                  {1, 2, 3, 0}, {1, 3, 0, 2}, {2, 0, 3, 1}, {2, 3, 1, 0}, {3, 0, 1, 2}, {3, 2, 0, 1}};
// signs of the antisymmetriser over those reads, and the 6-term and 4-term
// combinations of the mixed-spin terms (indices into the 24 reads)
__constant__ double c_anti[24] = {1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1};

enum UhfFlag {
    U_AA = 1, U_AB = 2, U_BB = 4,              // OMG2 (FAE: AA, AB; UFP: AA, AB, BB)
    U_AAA = 8, U_BBB = 16, U_AAB = 32, U_BBA = 64, // OMG3/OMG4 (FAE4: AAA, BBB, AAB, BBA; UFP3 likewise)
    U_4A = 128, U_4B = 256, U_3A1B = 512, U_1A3B = 1024, U_2A2B = 2048, // UFP4
};

template <int N, bool FAE, class T>
__global__ void __launch_bounds__(256) k_uhf_slab(i64 ne, i64 np, int flags, const double *__restrict__ DA,
                                                  const double *__restrict__ DB, const T *__restrict__ G,
                                                  const T *__restrict__ Gs, double *__restrict__ part)
{
    extern __shared__ double sm[]; // [2 ne + 2]
    i64 blk = blockIdx.x, a = blk % ne, pp = blk / ne;
    i64 ne2 = ne * ne;
    i64 nrem = 1, nblock = 1;
    for (int k = 0; k < 2 * N - 1; ++k) nrem *= ne;
    nblock = nrem * ne;
    const T *Gb = G + pp * nblock;
    const T *Gsb = Gs ? Gs + pp * nblock : nullptr;
    // Fixed-order sums throughout (no floating-point atomics): the slab is
    // walked one ket index j0 of the first electron at a time; the row
    // entries sm[j0], sm[ne + j0] are block sums over the other indices
    // (wave_sum, then the four waves' slots in order).
    __shared__ double wslot[4][2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const i64 ninner = nrem / ne;
    double pe = 0.0, ps = 0.0;
#define DA_(i, j) DA[(j) * ne + (i)]
#define DB_(i, j) DB[(j) * ne + (i)]
    for (i64 j0 = 0; j0 < ne; ++j0) {
        double fe = 0.0, fs = 0.0;
        for (i64 r0 = threadIdx.x; r0 < ninner; r0 += blockDim.x) {
            i64 ii[N], jj[N], r = r0;
            for (int k = N - 1; k >= 1; --k) {
                jj[k] = r % ne;
                r /= ne;
                ii[k] = r % ne;
                r /= ne;
            }
            jj[0] = j0;
            ii[0] = a;
            // read with ket of electron k = jj[perm[k]]
            auto rd = [&](const T *B, const int *perm) -> double {
                i64 off = 0;
                for (int k = 0; k < N; ++k) off = off * ne2 + ii[k] * ne + jj[perm[k]];
                return (double)B[off];
            };
            double t = 0.0, ts = 0.0;
            if constexpr (N == 2) {
                const int p12[2] = {0, 1}, p21[2] = {1, 0};
                double o12 = rd(Gb, p12), o21 = rd(Gb, p21), so12 = rd(Gsb, p12), so21 = rd(Gsb, p21);
                if constexpr (FAE) {
                    double c1 = (flags & U_AA) ? DA_(ii[1], jj[1]) : 0.0, c2 = (flags & U_AB) ? DB_(ii[1], jj[1]) : 0.0;
                    t = c1 * (o12 - o21) + c2 * o12;
                    ts = c1 * (so12 - so21) + c2 * so12;
                } else {
                    double a1 = DA_(ii[0], jj[0]), b1 = DB_(ii[0], jj[0]), a2 = DA_(ii[1], jj[1]), b2 = DB_(ii[1], jj[1]);
                    double c1 = (flags & U_AB) ? a1 * b2 + b1 * a2 : 0.0;
                    double c2 = (flags & U_AA) ? a1 * a2 : 0.0, c3 = (flags & U_BB) ? b1 * b2 : 0.0;
                    t = c1 * o12 + (c2 + c3) * (o12 - o21);
                    ts = c1 * so12 + (c2 + c3) * (so12 - so21);
                }
            } else if constexpr (N == 3) {
                const int p123[3] = {0, 1, 2}, p213[3] = {1, 0, 2}, p312[3] = {2, 0, 1}, p132[3] = {0, 2, 1},
                          p231[3] = {1, 2, 0}, p321[3] = {2, 1, 0};
                double g123 = rd(Gb, p123), g213 = rd(Gb, p213), g312 = rd(Gb, p312), g132 = rd(Gb, p132),
                       g231 = rd(Gb, p231), g321 = rd(Gb, p321);
                double o6 = g123 - g213 + g312 - g132 + g231 - g321;
                double o23 = g123 - g132, o12 = g123 - g213, o13 = g123 - g321;
                double a1 = DA_(ii[0], jj[0]), b1 = DB_(ii[0], jj[0]), a2 = DA_(ii[1], jj[1]), b2 = DB_(ii[1], jj[1]);
                double a3 = DA_(ii[2], jj[2]), b3 = DB_(ii[2], jj[2]);
                if constexpr (FAE) {
                    // FAE_OMG3 flags: U_AAA = LAA (NAE>=3), U_BBB = LBB (NBE>=2), U_AAB = LAB
                    double xaa = (flags & U_AAA) ? a2 * a3 * o6 : 0.0;
                    double xbb = (flags & U_BBB) ? b2 * b3 * o23 : 0.0;
                    double xab = (flags & U_AAB) ? a2 * b3 * o12 : 0.0;
                    double xba = (flags & U_AAB) ? b2 * a3 * o13 : 0.0;
                    t = xaa + xbb + xab + xba;
                } else {
                    double caaa = (flags & U_AAA) ? a1 * a2 * a3 : 0.0, cbbb = (flags & U_BBB) ? b1 * b2 * b3 : 0.0;
                    double caab = 0, caba = 0, cbaa = 0, cbba = 0, cbab = 0, cabb = 0;
                    if (flags & U_AAB) {
                        caab = a1 * a2 * b3;
                        caba = a1 * b2 * a3;
                        cbaa = b1 * a2 * a3;
                    }
                    if (flags & U_BBA) {
                        cbba = b1 * b2 * a3;
                        cbab = b1 * a2 * b3;
                        cabb = a1 * b2 * b3;
                    }
                    t = (caaa + cbbb) * o6 + (caab + cbba) * o12 + (caba + cbab) * o13 + (cabb + cbaa) * o23;
                }
            } else {
                double o[24];
                for (int p = 0; p < 24; ++p) o[p] = rd(Gb, c_omg4[p]);
                double anti = 0.0;
                for (int p = 0; p < 24; ++p) anti += c_anti[p] * o[p];
                // 0 OMG 1 23 2 12 3 12_23 4 13_23 5 13 6 24 7 12_24 8 12_14 9 14 10 34 11 12_34 12 13_34 13 14_34
                // 14 13_24 15 14_23 16 23_34 17 24_34
                double aab = o[0] - o[1] - o[2] + o[3] + o[4] - o[5];
                double aba = o[0] - o[6] - o[2] + o[7] + o[8] - o[9];
                double baa = o[0] - o[10] - o[5] + o[12] + o[13] - o[9];
                double abb = o[0] - o[10] - o[2] + o[11];
                double bab = o[0] - o[6] - o[5] + o[14];
                double bba = o[0] - o[1] - o[9] + o[15];
                double bbb = o[0] - o[10] - o[1] + o[16] + o[17] - o[6];
                double A[4], B[4];
                for (int k = 0; k < 4; ++k) {
                    A[k] = DA_(ii[k], jj[k]);
                    B[k] = DB_(ii[k], jj[k]);
                }
                if constexpr (FAE) {
                    double x = 0.0;
                    x += (flags & U_AAA) ? A[1] * A[2] * A[3] * anti : 0.0;
                    x += (flags & U_AAB) ? A[1] * A[2] * B[3] * aab : 0.0;
                    x += (flags & U_AAB) ? A[1] * B[2] * A[3] * aba : 0.0;
                    x += (flags & U_AAB) ? B[1] * A[2] * A[3] * baa : 0.0;
                    x += (flags & U_BBA) ? A[1] * B[2] * B[3] * abb : 0.0;
                    x += (flags & U_BBA) ? B[1] * A[2] * B[3] * bab : 0.0;
                    x += (flags & U_BBA) ? B[1] * B[2] * A[3] * bba : 0.0;
                    x += (flags & U_BBB) ? B[1] * B[2] * B[3] * bbb : 0.0;
                    t = x;
                } else {
                    double x = 0.0;
                    x += (flags & U_4A) ? A[0] * A[1] * A[2] * A[3] * anti : 0.0;
                    x += (flags & U_3A1B) ? A[0] * A[1] * A[2] * B[3] * aab : 0.0;
                    x += (flags & U_3A1B) ? A[0] * A[1] * B[2] * A[3] * aba : 0.0;
                    x += (flags & U_3A1B) ? A[0] * B[1] * A[2] * A[3] * baa : 0.0;
                    x += (flags & U_2A2B) ? A[0] * A[1] * B[2] * B[3] * abb : 0.0;
                    x += (flags & U_2A2B) ? A[0] * B[1] * A[2] * B[3] * bab : 0.0;
                    x += (flags & U_2A2B) ? A[0] * B[1] * B[2] * A[3] * bba : 0.0;
                    x += (flags & U_1A3B) ? A[0] * B[1] * B[2] * B[3] * bbb : 0.0;
                    x += (flags & U_4B) ? B[0] * B[1] * B[2] * B[3] * anti : 0.0;
                    x += (flags & U_1A3B) ? B[0] * B[1] * B[2] * A[3] * aab : 0.0;
                    x += (flags & U_1A3B) ? B[0] * B[1] * A[2] * B[3] * aba : 0.0;
                    x += (flags & U_1A3B) ? B[0] * A[1] * B[2] * B[3] * baa : 0.0;
                    x += (flags & U_2A2B) ? B[0] * B[1] * A[2] * A[3] * abb : 0.0;
                    x += (flags & U_2A2B) ? B[0] * A[1] * B[2] * A[3] * bab : 0.0;
                    x += (flags & U_2A2B) ? B[0] * A[1] * A[2] * B[3] * bba : 0.0;
                    x += (flags & U_3A1B) ? B[0] * A[1] * A[2] * A[3] * bbb : 0.0;
                    t = x;
                }
            }
            if constexpr (FAE) {
                fe += (double)N * t;
                fs += 2.0 * ts;
                pe += DA_(a, jj[0]) * t;
                ps += DA_(a, jj[0]) * ts;
            } else {
                pe += t;
                ps += ts;
            }
        }
        if constexpr (FAE) {
            fe = wave_sum(fe);
            fs = wave_sum(fs);
            if (lane == 0) {
                wslot[wv][0] = fe;
                wslot[wv][1] = fs;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                sm[j0] = ((wslot[0][0] + wslot[1][0]) + wslot[2][0]) + wslot[3][0];
                sm[ne + j0] = N == 2 ? ((wslot[0][1] + wslot[1][1]) + wslot[2][1]) + wslot[3][1] : 0.0;
            }
            __syncthreads();
        } else if (threadIdx.x < 2) {
            sm[threadIdx.x * ne + j0] = 0.0;
        }
    }
#undef DA_
#undef DB_
    pe = wave_sum(pe);
    ps = wave_sum(ps);
    if (lane == 0) {
        wslot[wv][0] = pe;
        wslot[wv][1] = ps;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sm[2 * ne] = ((wslot[0][0] + wslot[1][0]) + wslot[2][0]) + wslot[3][0];
        sm[2 * ne + 1] = ((wslot[0][1] + wslot[1][1]) + wslot[2][1]) + wslot[3][1];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * ne + 2; k += blockDim.x) part[blk * (2 * ne + 2) + k] = sm[k];
}

// FAE: FE(a,b) += sum_pp DP(pp) F_pp,a[b], SE(a,b) = sum_pp DP(pp) S_pp,a[b];
// UFP: FP(pp) += sum_a e_pp,a, SP(pp) = sum_a s_pp,a;
// both: E = sum_pp DP(pp) sum_a e_pp,a, S likewise
__global__ void k_uhf_reduce(i64 ne, i64 np, const double *__restrict__ DP, const double *__restrict__ part, int fae,
                             double *__restrict__ FE, double *__restrict__ SE, double *__restrict__ FP,
                             double *__restrict__ SP, double *__restrict__ E, double *__restrict__ S)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    i64 ne2 = ne * ne, np2 = np * np, w = 2 * ne + 2;
    if (fae && t < ne2) {
        i64 a = t % ne, b = t / ne;
        double f = 0.0, fs = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) {
            double dp = DP[(pp % np) * np + pp / np];
            f += dp * part[(pp * ne + a) * w + b];
            fs += dp * part[(pp * ne + a) * w + ne + b];
        }
        FE[b * ne + a] += f;
        if (SE) SE[b * ne + a] = fs;
    }
    if (!fae && t < np2) {
        i64 ip = t / np, jp = t % np;
        double f = 0.0, fs = 0.0;
        for (i64 a = 0; a < ne; ++a) {
            f += part[(t * ne + a) * w + 2 * ne];
            fs += part[(t * ne + a) * w + 2 * ne + 1];
        }
        FP[jp * np + ip] += f;
        if (SP) SP[jp * np + ip] = fs;
    }
    if (t == 0) {
        double e = 0.0, s = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) {
            double dp = DP[(pp % np) * np + pp / np], pe = 0.0, ps = 0.0;
            for (i64 a = 0; a < ne; ++a) {
                pe += part[(pp * ne + a) * w + 2 * ne];
                ps += part[(pp * ne + a) * w + 2 * ne + 1];
            }
            e += dp * pe;
            s += dp * ps;
        }
        *E = e;
        if (S) *S = s;
    }
}

template <int N, bool FAE>
static int uhf_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                   const double *d_dp, const double *d_g, const double *d_gs, double *d_f, double *d_s2, double *d_e,
                   double *d_s, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0 || nae < 0 || nbe < 0) return fail(NEORX_E_ARG, "bad sizes");
    if (!d_dae || !d_dbe || !d_dp || !d_g || !d_f || !d_e || (N == 2 && (!d_gs || !d_s2 || !d_s)))
        return fail(NEORX_E_ARG, "null argument");
    int fl = 0;
    if (N == 2 && FAE) fl = (nae >= 2 ? U_AA : 0) | (nae >= 1 && nbe >= 1 ? U_AB : 0);
    if (N == 2 && !FAE) fl = (nae >= 1 && nbe >= 1 ? U_AB : 0) | (nae >= 2 ? U_AA : 0) | (nbe >= 2 ? U_BB : 0);
    if (N == 3 && FAE) fl = (nbe >= 1 && nae >= 2 ? U_AAB : 0) | (nbe >= 2 ? U_BBB : 0) | (nae >= 3 ? U_AAA : 0);
    if (N == 3 && !FAE)
        fl = (nae >= 3 ? U_AAA : 0) | (nbe >= 3 ? U_BBB : 0) | (nae >= 2 && nbe >= 1 ? U_AAB : 0) |
             (nbe >= 2 && nae >= 1 ? U_BBA : 0);
    if (N == 4 && FAE)
        fl = (nae >= 4 ? U_AAA : 0) | (nbe >= 3 ? U_BBB : 0) | (nae >= 3 && nbe >= 1 ? U_AAB : 0) |
             (nbe >= 2 && nae >= 2 ? U_BBA : 0);
    if (N == 4 && !FAE)
        fl = (nae >= 4 ? U_4A : 0) | (nbe >= 4 ? U_4B : 0) | (nae >= 3 && nbe >= 1 ? U_3A1B : 0) |
             (nae >= 1 && nbe >= 3 ? U_1A3B : 0) | (nae >= 2 && nbe >= 2 ? U_2A2B : 0);
    hipStream_t st = (hipStream_t)stream;
    i64 nblk = np * np * ne, w = 2 * ne + 2;
    DevBuf part;
    rc = part.alloc(nblk * w * 8);
    if (rc) return rc;
    hipLaunchKernelGGL((k_uhf_slab<N, FAE, double>), dim3((unsigned)nblk), dim3(256), w * 8, st, ne, np, fl, d_dae,
                       d_dbe, d_g, N == 2 ? d_gs : nullptr, part.d());
    hipLaunchKernelGGL(k_uhf_reduce, dim3(nblocks(std::max(ne * ne, np * np), 256)), dim3(256), 0, st, ne, np, d_dp,
                       part.d(), FAE ? 1 : 0, FAE ? d_f : nullptr, FAE ? d_s2 : nullptr, FAE ? nullptr : d_f,
                       FAE ? nullptr : d_s2, d_e, N == 2 ? d_s : nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

#define UHF_DEV(NAME, N, FAE)                                                                                      \
    extern "C" int NAME(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae,                \
                        const double *d_dbe, const double *d_dp, const double *d_g, double *d_f, double *d_e,     \
                        void *stream)                                                                             \
    {                                                                                                             \
        return uhf_dev<N, FAE>(nae, nbe, nebf, npbf, d_dae, d_dbe, d_dp, d_g, nullptr, d_f, nullptr, d_e, nullptr, \
                               stream);                                                                           \
    }
UHF_DEV(neorx_fae_omg3_dev, 3, true)
UHF_DEV(neorx_ufp_omg3_dev, 3, false)
UHF_DEV(neorx_fae_omg4_dev, 4, true)
UHF_DEV(neorx_ufp_omg4_dev, 4, false)
#undef UHF_DEV

extern "C" int neorx_fae_omg2_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae,
                                  const double *d_dbe, const double *d_dp, const double *d_gm2, const double *d_gm2s,
                                  double *d_fae, double *d_sae2, double *d_e, double *d_s, void *stream)
{
    return uhf_dev<2, true>(nae, nbe, nebf, npbf, d_dae, d_dbe, d_dp, d_gm2, d_gm2s, d_fae, d_sae2, d_e, d_s, stream);
}

extern "C" int neorx_ufp_omg2_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae,
                                  const double *d_dbe, const double *d_dp, const double *d_gm2, const double *d_gm2s,
                                  double *d_fp, double *d_sp2, double *d_e, double *d_s, void *stream)
{
    return uhf_dev<2, false>(nae, nbe, nebf, npbf, d_dae, d_dbe, d_dp, d_gm2, d_gm2s, d_fp, d_sp2, d_e, d_s, stream);
}

// host-array forms, as the Fortran routines take them
template <int N, bool FAE>
static int uhf_host(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                    const double *dp, const double *g, const double *gs, double *f, double *s2, double *e, double *s)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf, ng = np * np;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    for (int k = 0; k < 2 * N; ++k) ng *= ne;
    i64 nf = FAE ? ne * ne : np * np;
    DevBuf dA, dB, dP, dG, dGs, dF, dS2, dE;
    rc = h2d(dA, dae, ne * ne);
    if (!rc) rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(dG, g, ng);
    if (!rc && N == 2) rc = h2d(dGs, gs, ng);
    if (!rc) rc = h2d(dF, f, nf);
    if (!rc && N == 2) rc = dS2.alloc(nf * 8);
    if (!rc) rc = dE.alloc(16);
    if (!rc)
        rc = uhf_dev<N, FAE>(nae, nbe, ne, np, dA.d(), dB.d(), dP.d(), dG.d(), N == 2 ? dGs.d() : nullptr, dF.d(),
                             N == 2 ? dS2.d() : nullptr, dE.d(), N == 2 ? dE.d() + 1 : nullptr, 0);
    if (rc) return rc;
    double es[2];
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    if (N == 2) HIPCHK(hipMemcpy(s2, dS2.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(es, dE.p, 16, hipMemcpyDeviceToHost));
    *e = es[0];
    if (N == 2) *s = es[1];
    return NEORX_OK;
}

extern "C" int neorx_fae_omg2(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae,
                              const double *dbe, const double *dp, const double *gm2, const double *gm2s, double *fae,
                              double *sae2, double *e, double *s)
{
    return uhf_host<2, true>(nae, nbe, nebf, npbf, dae, dbe, dp, gm2, gm2s, fae, sae2, e, s);
}
extern "C" int neorx_ufp_omg2(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae,
                              const double *dbe, const double *dp, const double *gm2, const double *gm2s, double *fp,
                              double *sp2, double *e, double *s)
{
    return uhf_host<2, false>(nae, nbe, nebf, npbf, dae, dbe, dp, gm2, gm2s, fp, sp2, e, s);
}
#define UHF_HOST(NAME, N, FAE)                                                                                     \
    extern "C" int NAME(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe, \
                        const double *dp, const double *g, double *f, double *e)                                  \
    {                                                                                                             \
        return uhf_host<N, FAE>(nae, nbe, nebf, npbf, dae, dbe, dp, g, nullptr, f, nullptr, e, nullptr);          \
    }
UHF_HOST(neorx_fae_omg3, 3, true)
UHF_HOST(neorx_ufp_omg3, 3, false)
UHF_HOST(neorx_fae_omg4, 4, true)
UHF_HOST(neorx_ufp_omg4, 4, false)
#undef UHF_HOST

// OMG4_ICR (xcuhf_omg4.f:1080-1320): the unsymmetrised 5-particle value
__global__ void __launch_bounds__(256) k_omg4(i64 ne, i64 np, i64 off, i64 n, const double *__restrict__ G2s,
                                              const double *__restrict__ Gee, double *__restrict__ G4)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    i64 r = off + t;
    i64 br[4], kt[4];
    for (int k = 3; k >= 0; --k) {
        kt[k] = r % ne;
        r /= ne;
        br[k] = r % ne;
        r /= ne;
    }
    i64 jp = r % np, ip = r / np;
    G4[off + t] = gam4_term(ne, np, G2s, Gee, ip, jp, br, kt);
}

extern "C" int neorx_omg4_icr(int64_t nebf, int64_t npbf, const double *gam_2s, const double *gam_ee, double *gm4)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    i64 ng2 = ne * ne * ne * ne * np * np, ngee = ne * ne * ne * ne, ng4 = ngee * ngee * np * np;
    DevBuf A, B, C;
    rc = h2d(A, gam_2s, ng2);
    if (!rc) rc = h2d(B, gam_ee, ngee);
    if (!rc) rc = C.alloc(ng4 * 8);
    if (rc) return rc;
    for (i64 off = 0; off < ng4; off += CHUNK) {
        i64 n = std::min(CHUNK, ng4 - off);
        hipLaunchKernelGGL(k_omg4, dim3(nblocks(n, 256)), dim3(256), 0, 0, ne, np, off, n, A.d(), B.d(), C.d());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(gm4, C.p, ng4 * 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

// ------------------------------------------------------------------
// RXCHF-ne (source/rxchfne/): RXCHFne_GAM2_IC1 and the contractions of the
// RXCHF-ne / RXCUHF-ne Fock builds
// ------------------------------------------------------------------
// GM2 = contracted ansE1 as is, GM2ex = (x12 + x21)/2 of contracted ansE2
// (rxchfne_gam2_IC1_OMP.f:169-205)
__global__ void __launch_bounds__(256) k_rxne_symm(i64 ne, i64 n, const double *__restrict__ X,
                                                   const double *__restrict__ Xs, double *__restrict__ G,
                                                   double *__restrict__ Gex)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 ne2 = ne * ne;
        i64 b = t / (ne2 * ne2), p1 = (t / ne2) % ne2, p2 = t % ne2;
        i64 t21 = (b * ne2 + p2) * ne2 + p1;
        G[t] = X[t];
        Gex[t] = (Xs[t] + Xs[t21]) / 2.0;
    }
}

extern "C" int neorx_rxchfne_gam2_ic1_dev(int64_t handle, double *d_gm2, double *d_gm2ex, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (!d_gm2 || !d_gm2ex) return fail(NEORX_E_ARG, "null output");
    i64 ne = s->d.nebf, np = s->d.npbf, ng2 = ne * ne * ne * ne * np * np;
    hipStream_t st = (hipStream_t)stream;
    i64 cap = slab_cap(s, 2);
    DevBuf X, Xs, Y, Ys;
    int rc = X.alloc(ng2 * 8);
    if (!rc) rc = Xs.alloc(ng2 * 8);
    if (!rc) rc = Y.alloc(cap * 8);
    if (!rc) rc = Ys.alloc(cap * 8);
    if (!rc) rc = build2_block<4>(s, 0, np * np, X.d(), Xs.d(), Y.d(), Ys.d(), 0, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rxne_symm, dim3(grid_blocks(ng2, 256)), dim3(256), 0, st, ne, ng2, X.d(), Xs.d(), d_gm2, d_gm2ex);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_rxchfne_gam2_ic1(int64_t handle, double *gm2icr, double *gm2exicr)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, np = s->d.npbf, n = ne * ne * ne * ne * np * np;
    DevBuf G, Gx;
    int rc = G.alloc(n * 8);
    if (!rc) rc = Gx.alloc(n * 8);
    if (!rc) rc = neorx_rxchfne_gam2_ic1_dev(handle, G.d(), Gx.d(), 0);
    if (!rc) rc = copy_out(gm2icr, G.d(), n);
    if (!rc) rc = copy_out(gm2exicr, Gx.d(), n);
    return rc;
}

// Block = (ip,jp); thread = (i2,j2) column of the block's ne^2 x ne^2 matrix
// G[(i1 j1),(i2 j2)].  Per block:
//   v12(i2,j2) = sum_{i1 j1} DBE(i1,j1) G(i1 j1 i2 j2)
//   v21(i2,j2) = sum_{i1 j1} DBE(i1,j1) Gx(i1 j2 i2 j1)
//   R(i1,j1)   = sum_{i2 j2} DT(i2,j2) G(i1 j1 i2 j2) - DX(i2,j2) Gx(i1 j2 i2 j1)
//                (k_rxne_slab_rows)
// (Gx, DX absent in the closed-shell routines).  part[pp] = [v12, v21, R].
__global__ void __launch_bounds__(1024) k_rxne_slab(i64 ne, const double *__restrict__ DBE,
                                                    const double *__restrict__ G, const double *__restrict__ Gx,
                                                    double *__restrict__ part)
{
    i64 pp = blockIdx.x, ne2 = ne * ne, ne4 = ne2 * ne2;
    const double *g = G + pp * ne4, *gx = Gx ? Gx + pp * ne4 : nullptr;
    double *out = part + pp * 3 * ne2;
    for (i64 c = threadIdx.x; c < ne2; c += blockDim.x) {
        i64 i2 = c / ne, j2 = c % ne;
        double v12 = 0.0, v21 = 0.0;
        for (i64 i1 = 0; i1 < ne; ++i1)
            for (i64 j1 = 0; j1 < ne; ++j1) {
                double d1 = DBE[j1 * ne + i1];
                v12 += d1 * g[(i1 * ne + j1) * ne2 + c];
                if (gx) v21 += d1 * gx[((i1 * ne + j2) * ne + i2) * ne + j1];
            }
        out[c] = v12;
        out[ne2 + c] = v21;
    }
}

// R of k_rxne_slab's block pp: one wave per row r = (i1,j1), lanes over the
// columns (i2,j2) in a fixed stride, then wave_sum.  Grid (np^2, ceil(ne^2/4)).
__global__ void __launch_bounds__(256) k_rxne_slab_rows(i64 ne, const double *__restrict__ DT,
                                                        const double *__restrict__ DX, const double *__restrict__ G,
                                                        const double *__restrict__ Gx, double *__restrict__ part)
{
    const i64 pp = blockIdx.x, ne2 = ne * ne, ne4 = ne2 * ne2;
    const i64 r = (i64)blockIdx.y * 4 + (threadIdx.x >> 6);
    if (r >= ne2) return; // whole waves leave together
    const i64 i1 = r / ne, j1 = r % ne;
    const double *g = G + pp * ne4 + r * ne2, *gx = Gx ? Gx + pp * ne4 : nullptr;
    double acc = 0.0;
    for (i64 c = threadIdx.x & 63; c < ne2; c += 64) {
        const i64 i2 = c / ne, j2 = c % ne;
        double v = DT[j2 * ne + i2] * g[c];
        if (gx) v -= DX[j2 * ne + i2] * gx[((i1 * ne + j2) * ne + i2) * ne + j1];
        acc += v;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[pp * 3 * ne2 + 2 * ne2 + r] = acc;
}

// mode 0: F1 += sum_pp DP (v12 - v21), F2 (may be null) += sum_pp DP v12;
// mode 1: F1 = FP, FP(pp) += sum DBE R_pp;  mode 2: F1(i1,j1) += sum_pp DP R_pp;
// every mode: E = sum_pp DP(pp) sum DBE R_pp
__global__ void k_rxne_reduce(i64 ne, i64 np, int mode, const double *__restrict__ DP, const double *__restrict__ DBE,
                              const double *__restrict__ part, double *__restrict__ F1, double *__restrict__ F2,
                              double *__restrict__ E)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    i64 ne2 = ne * ne, np2 = np * np;
    auto dpv = [&](i64 pp) { return DP[(pp % np) * np + pp / np]; }; // DP(ip,jp), pp = ip*np + jp
    auto fp_of = [&](i64 pp) {
        const double *R = part + pp * 3 * ne2 + 2 * ne2;
        double f = 0.0;
        for (i64 k = 0; k < ne2; ++k) f += DBE[(k % ne) * ne + k / ne] * R[k]; // R(i1,j1) at i1*ne + j1
        return f;
    };
    if (mode != 1 && t < ne2) {
        i64 i = t / ne, j = t % ne; // output element (i,j)
        double f = 0.0, f2 = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) {
            const double *o = part + pp * 3 * ne2;
            double dp = dpv(pp);
            if (mode == 0) {
                f += dp * (o[t] - o[ne2 + t]);
                f2 += dp * o[t];
            } else {
                f += dp * o[2 * ne2 + t];
            }
        }
        F1[j * ne + i] += f;
        if (mode == 0 && F2) F2[j * ne + i] += f2;
    }
    if (mode == 1 && t < np2) {
        i64 ip = t / np, jp = t % np;
        F1[jp * np + ip] += fp_of(t);
    }
    if (t == 0) {
        double e = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) e += dpv(pp) * fp_of(pp);
        *E = e;
    }
}

static int rxne_contract(int mode, i64 ne, i64 np, const double *dbe, const double *dt, const double *dx,
                         const double *dp, const double *gm2, const double *gm2ex, double *f1, double *f2, double *e)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    i64 ng2 = ne * ne * ne * ne * np * np, nf = mode == 1 ? np * np : ne * ne;
    DevBuf dB, dT, dX, dP, dG, dGx, dF1, dF2, dE, part;
    rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dT, dt, ne * ne);
    if (!rc && dx) rc = h2d(dX, dx, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(dG, gm2, ng2);
    if (!rc && gm2ex) rc = h2d(dGx, gm2ex, ng2);
    if (!rc) rc = h2d(dF1, f1, nf);
    if (!rc && f2) rc = h2d(dF2, f2, ne * ne);
    if (!rc) rc = dE.alloc(8);
    if (!rc) rc = part.alloc(np * np * 3 * ne * ne * 8);
    if (rc) return rc;
    unsigned nt = (unsigned)std::min<i64>(((ne * ne + 63) / 64) * 64, 1024);
    hipLaunchKernelGGL(k_rxne_slab, dim3((unsigned)(np * np)), dim3(nt), 0, 0, ne, dB.d(), dG.d(),
                       gm2ex ? dGx.d() : nullptr, part.d());
    hipLaunchKernelGGL(k_rxne_slab_rows, dim3((unsigned)(np * np), nblocks(ne * ne, 4)), dim3(256), 0, 0, ne, dT.d(),
                       dx ? dX.d() : nullptr, dG.d(), gm2ex ? dGx.d() : nullptr, part.d());
    hipLaunchKernelGGL(k_rxne_reduce, dim3(nblocks(std::max(ne * ne, np * np), 256)), dim3(256), 0, 0, ne, np, mode,
                       dP.d(), dB.d(), part.d(), dF1.d(), f2 ? dF2.d() : nullptr, dE.d());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(f1, dF1.p, nf * 8, hipMemcpyDeviceToHost));
    if (f2) HIPCHK(hipMemcpy(f2, dF2.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

// Closed-shell RXCHF-ne contractions (RXCHFne_FAE/FP/FBE_OMG2,
// rxchfne_omg2.f:2, :212, :409) in one pass over GM2 stored as T (FP64,
// FP32 or BF16).  Row r = (i1,j1) of the ne^2 x ne^2 block G[pp] is the
// special electron's pair, column c = (i2,j2) the regular electron's.
// Block (pp, s) takes split s of the rows; wave w of the block the rows
// r0 + w, r0 + w + RX_WAVES, ...  A lane loads V adjacent columns in one
// 8-byte buffer load (V = 1 for FP64, 2 for FP32, 4 for BF16, when a row's
// byte length is a multiple of 8; else V = 1), so every wave load moves 512
// bytes.  Per chunk of 64 KC V columns lane l owns the vectors at
// c = c0 + (l + 64 k) V (k < KC) and per row adds
//   vc[k][v] += DBE(r) G(r,c+v)          (column sums, in registers)
//   R(r)     += sum_c DAE(c) G(r,c)      (one wave reduction per row and chunk)
//   fp       += DBE(r) R(r)              (the FP sum of the wave's rows)
// Two rows per step keep 2 KC loads in flight per lane; each row's byte
// offset is wave-uniform (scalar).  Lanes past ne^2 read the row's first
// entries and weight them by zero.  Outputs:
//   vpart[pp][s][c]    = sum over the split's rows of DBE(r) G(r,c)
//   rpart[pp][ch][r]   = sum over chunk ch's columns of DAE(c) G(r,c)
//   fpart[pp][s][w]    = wave w's sum of DBE(r) R(r)
// Sums over a row run in SlabAcc<T> (FP32 for BF16 storage, else FP64); the
// column partials meet in LDS wave by wave, in a fixed order.
constexpr int RX_WAVES = 4;

template <class T, int KC, int V>
__global__ void __launch_bounds__(64 * RX_WAVES) k_rxne_pass(i64 ne, int nsplit, const double *__restrict__ DBE,
                                                             const double *__restrict__ DAE,
                                                             const T *__restrict__ G, double *__restrict__ vpart,
                                                             double *__restrict__ rpart, double *__restrict__ fpart)
{
    using A = typename SlabAcc<T>::type;
    using P = Pack<T, V>;
    constexpr int COLS = 64 * KC * V;
    __shared__ double cs[COLS];
    const i64 ne2 = ne * ne, pp = blockIdx.x;
    const int s = blockIdx.y, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 rpb = (ne2 + nsplit - 1) / nsplit, r0 = s * rpb, r1 = min(ne2, r0 + rpb);
    const int nch = (int)((ne2 + COLS - 1) / COLS);
    const auto rs = buf_rsrc(G + pp * ne2 * ne2, (unsigned)(ne2 * ne2 * (i64)sizeof(T)));
    double fp = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
        const i64 c0 = (i64)ch * COLS;
        A dt[KC][V], vc[KC][V];
        unsigned voff[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const i64 cb = c0 + (i64)(64 * k + lane) * V;
            const bool on = cb < ne2;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const i64 c = cb + v;
                dt[k][v] = on ? (A)DAE[(c % ne) * ne + c / ne] : (A)0;
                vc[k][v] = (A)0;
            }
            voff[k] = on ? (unsigned)(cb * (i64)sizeof(T)) : 0u;
        }
        i64 r = r0 + w;
        for (; r + RX_WAVES < r1; r += 2 * RX_WAVES) {
            typename P::raw va[KC], vb[KC];
            const i64 rb = r + RX_WAVES;
            const unsigned sa = (unsigned)(r * ne2 * (i64)sizeof(T)), sb = (unsigned)(rb * ne2 * (i64)sizeof(T));
#pragma unroll
            for (int k = 0; k < KC; ++k) va[k] = P::ld(rs, voff[k], sa);
#pragma unroll
            for (int k = 0; k < KC; ++k) vb[k] = P::ld(rs, voff[k], sb);
            NRX_LOADS_ISSUED();
            const double dad = DBE[(r % ne) * ne + r / ne], dbd = DBE[(rb % ne) * ne + rb / ne];
            const A da = (A)dad, db = (A)dbd;
            A acca = (A)0, accb = (A)0;
#pragma unroll
            for (int k = 0; k < KC; ++k)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const A a = (A)P::get(va[k], v), b = (A)P::get(vb[k], v);
                    vc[k][v] += da * a;
                    vc[k][v] += db * b;
                    acca += dt[k][v] * a;
                    accb += dt[k][v] * b;
                }
            const double ta = wave_sum((double)acca), tb = wave_sum((double)accb);
            fp += dad * ta;
            fp += dbd * tb;
            if (lane == 0) {
                rpart[(pp * nch + ch) * ne2 + r] = ta;
                rpart[(pp * nch + ch) * ne2 + rb] = tb;
            }
        }
        if (r < r1) {
            typename P::raw va[KC];
            const unsigned sa = (unsigned)(r * ne2 * (i64)sizeof(T));
#pragma unroll
            for (int k = 0; k < KC; ++k) va[k] = P::ld(rs, voff[k], sa);
            NRX_LOADS_ISSUED();
            const double dad = DBE[(r % ne) * ne + r / ne];
            const A da = (A)dad;
            A acca = (A)0;
#pragma unroll
            for (int k = 0; k < KC; ++k)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const A a = (A)P::get(va[k], v);
                    vc[k][v] += da * a;
                    acca += dt[k][v] * a;
                }
            const double ta = wave_sum((double)acca);
            fp += dad * ta;
            if (lane == 0) rpart[(pp * nch + ch) * ne2 + r] = ta;
        }
        // column partials: the waves add into one LDS row in turn
        for (int q = 0; q < RX_WAVES; ++q) {
            if (w == q)
#pragma unroll
                for (int k = 0; k < KC; ++k)
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const int j = (64 * k + lane) * V + v;
                        cs[j] = q ? cs[j] + (double)vc[k][v] : (double)vc[k][v];
                    }
            __syncthreads();
        }
        for (int j = threadIdx.x; j < COLS; j += 64 * RX_WAVES) {
            const i64 c = c0 + j;
            if (c < ne2) vpart[(pp * nsplit + s) * ne2 + c] = cs[j];
        }
        __syncthreads();
    }
    if (lane == 0) fpart[(pp * nsplit + s) * RX_WAVES + w] = fp;
}

// Weighted column sums of a row-major nrow x ncol matrix P, first stage:
// block (cb, g) sums the rows of group g (rows g*rpg .. +rpg) for the 64
// columns of block cb, each row q weighted by DP(pp) with pp = q / per
// (DP column-major npbf x npbf, pp = ip*npbf + jp; weight 1 when DP is
// null); 4 waves split the rows and meet in LDS in a fixed order.
// out1[g][c].
constexpr int WCS_GROUPS = 32;
__global__ void __launch_bounds__(256) k_wcolsum1(i64 nrow, i64 ncol, i64 per, i64 np, const double *__restrict__ DP,
                                                  const double *__restrict__ P, double *__restrict__ out1)
{
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 c = (i64)blockIdx.x * 64 + lane, g = blockIdx.y;
    const i64 rpg = (nrow + WCS_GROUPS - 1) / WCS_GROUPS, q0 = g * rpg, q1 = min(nrow, q0 + rpg);
    double acc = 0.0;
    if (c < ncol)
        for (i64 q = q0 + w; q < q1; q += 4) {
            const i64 pp = q / per;
            acc += (DP ? DP[(pp % np) * np + pp / np] : 1.0) * P[q * ncol + c];
        }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0 && c < ncol) out1[g * ncol + c] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// second stage: F(i,j) += sum_g out1[g][c], c = i*ne + j, F column-major
__global__ void k_wcolsum2(i64 ne, const double *__restrict__ out1, double *__restrict__ F)
{
    const i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x, ne2 = ne * ne;
    if (c >= ne2) return;
    double t = 0.0;
    for (int g = 0; g < WCS_GROUPS; ++g) t += out1[g * ne2 + c];
    F[(c % ne) * ne + c / ne] += t;
}

// FP(ip,jp) += sum_{s,w} fpart[pp][s][w] when FP is given;
// E = sum_pp DP(pp) FP-increment(pp) (overwritten)
__global__ void __launch_bounds__(64) k_rxne_e(i64 np, int nq, const double *__restrict__ DP,
                                               const double *__restrict__ fpart, double *__restrict__ FP,
                                               double *__restrict__ E)
{
    double acc = 0.0;
    for (i64 pp = threadIdx.x; pp < np * np; pp += 64) {
        double f = 0.0;
        for (int q = 0; q < nq; ++q) f += fpart[pp * nq + q];
        const i64 at = (pp % np) * np + pp / np;
        acc += DP[at] * f;
        if (FP) FP[at] += f;
    }
    acc = wave_sum(acc);
    if (threadIdx.x == 0) *E = acc;
}

// vector slots per lane and chunk (KC): of the candidates for the vector
// width V (at most 16 entries per lane and row), the one whose chunks of
// 64 KC V columns pad the ne^2 columns least (the larger on a tie)
static int rxne_kc(i64 ne2, int V)
{
    static const int c1[] = {16, 12, 9, 8}, c2[] = {8, 6, 4}, c4[] = {4, 3, 2};
    const int *c = V == 1 ? c1 : V == 2 ? c2 : c4;
    const int n = V == 1 ? 4 : 3;
    int best = c[0];
    i64 pad = -1;
    for (int i = 0; i < n; ++i) {
        const i64 cols = 64LL * c[i] * V, p = (ne2 + cols - 1) / cols * cols - ne2;
        if (pad < 0 || p < pad) best = c[i], pad = p;
    }
    return best;
}

template <class T, int V>
static void rxne_pass_kernel(int kc, i64 ne, i64 np2, int nsplit, const double *dbe, const double *dae, const T *g,
                             double *vpart, double *rpart, double *fpart, hipStream_t st)
{
    const dim3 grid((unsigned)np2, (unsigned)nsplit), block(64 * RX_WAVES);
#define NRX_RXNE_KC(K)                                                                                                \
    case K:                                                                                                           \
        hipLaunchKernelGGL((k_rxne_pass<T, K, V>), grid, block, 0, st, ne, nsplit, dbe, dae, g, vpart, rpart, fpart); \
        break
    if constexpr (V == 1) {
        switch (kc) {
            NRX_RXNE_KC(8);
            NRX_RXNE_KC(9);
            NRX_RXNE_KC(12);
            NRX_RXNE_KC(16);
        }
    } else if constexpr (V == 2) {
        switch (kc) {
            NRX_RXNE_KC(4);
            NRX_RXNE_KC(6);
            NRX_RXNE_KC(8);
        }
    } else {
        switch (kc) {
            NRX_RXNE_KC(2);
            NRX_RXNE_KC(3);
            NRX_RXNE_KC(4);
        }
    }
#undef NRX_RXNE_KC
}

template <class T>
static int rxne_pass_launch(i64 ne, i64 np, const double *dae, const double *dbe, const double *dp, const T *g,
                            double *fae, double *fp, double *fbe, double *e, hipStream_t st)
{
    const i64 ne2 = ne * ne, np2 = np * np;
    if (ne2 * ne2 * (i64)sizeof(T) > 0xffffffffLL)
        return fail(NEORX_E_ARG, "RXCHF-ne contraction: a GM2 block of nebf %lld exceeds 4 GiB", (long long)ne);
    // entries per 8-byte load when every row starts on 8 bytes
    const int V = (ne2 * (i64)sizeof(T)) % 8 == 0 ? (int)(8 / sizeof(T)) : 1;
    const int kc = rxne_kc(ne2, V);
    const int nch = (int)((ne2 + 64LL * kc * V - 1) / (64LL * kc * V));
    // about 2048 blocks, at least 2 RX_WAVES rows each
    const int nsplit = (int)std::max<i64>(1, std::min<i64>(ne2 / (2 * RX_WAVES), (2048 + np2 - 1) / np2));
    StreamBuf w;
    const i64 nv = np2 * nsplit * ne2, nr = np2 * nch * ne2, n1 = (i64)WCS_GROUPS * ne2,
              nf = np2 * nsplit * RX_WAVES;
    int rc = w.alloc((nv + nr + n1 + nf) * 8, st);
    if (rc) return rc;
    double *vpart = (double *)w.p, *rpart = vpart + nv, *o1 = rpart + nr, *fpart = o1 + n1;
    if constexpr (sizeof(T) == 8) {
        rxne_pass_kernel<T, 1>(kc, ne, np2, nsplit, dbe, dae, g, vpart, rpart, fpart, st);
    } else {
        if (V == 1)
            rxne_pass_kernel<T, 1>(kc, ne, np2, nsplit, dbe, dae, g, vpart, rpart, fpart, st);
        else
            rxne_pass_kernel<T, 8 / sizeof(T)>(kc, ne, np2, nsplit, dbe, dae, g, vpart, rpart, fpart, st);
    }
    const dim3 g1((unsigned)((ne2 + 63) / 64), WCS_GROUPS);
    if (fae) {
        hipLaunchKernelGGL(k_wcolsum1, g1, dim3(256), 0, st, np2 * nsplit, ne2, (i64)nsplit, np, dp, vpart, o1);
        hipLaunchKernelGGL(k_wcolsum2, dim3(nblocks(ne2, 256)), dim3(256), 0, st, ne, o1, fae);
    }
    if (fbe) {
        hipLaunchKernelGGL(k_wcolsum1, g1, dim3(256), 0, st, np2 * nch, ne2, (i64)nch, np, dp, rpart, o1);
        hipLaunchKernelGGL(k_wcolsum2, dim3(nblocks(ne2, 256)), dim3(256), 0, st, ne, o1, fbe);
    }
    hipLaunchKernelGGL(k_rxne_e, dim3(1), dim3(64), 0, st, np, nsplit * RX_WAVES, dp, fpart, fp, e);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

extern "C" int neorx_rxchfne_omg2_dev(int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                                      const double *d_dp, const void *d_gm2, int dtype, double *d_fae,
                                      double *d_fp, double *d_fbe, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (nebf <= 0 || npbf <= 0) return fail(NEORX_E_ARG, "bad sizes");
    if (!d_dae || !d_dbe || !d_dp || !d_gm2 || !d_e) return fail(NEORX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
    case 0: return rxne_pass_launch<double>(nebf, npbf, d_dae, d_dbe, d_dp, (const double *)d_gm2, d_fae, d_fp, d_fbe, d_e, st);
    case 1: return rxne_pass_launch<float>(nebf, npbf, d_dae, d_dbe, d_dp, (const float *)d_gm2, d_fae, d_fp, d_fbe, d_e, st);
    case 2: return rxne_pass_launch<bf16_t>(nebf, npbf, d_dae, d_dbe, d_dp, (const bf16_t *)d_gm2, d_fae, d_fp, d_fbe, d_e, st);
    default: return fail(NEORX_E_ARG, "storage type %d (0 FP64, 1 FP32, 2 BF16)", dtype);
    }
}

// host-pointer closed-shell routines: one matrix of the one-pass contraction
static int rxne_closed(int mode, i64 ne, i64 np, const double *dae, const double *dbe, const double *dp,
                       const double *gm2, double *f, double *e)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    const i64 ng2 = ne * ne * ne * ne * np * np, nf = mode == 1 ? np * np : ne * ne;
    DevBuf dA, dB, dP, dG, dF, dE;
    rc = h2d(dA, dae, ne * ne);
    if (!rc) rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(dG, gm2, ng2);
    if (!rc) rc = h2d(dF, f, nf);
    if (!rc) rc = dE.alloc(8);
    if (!rc)
        rc = rxne_pass_launch<double>(ne, np, dA.d(), dB.d(), dP.d(), dG.d(), mode == 0 ? dF.d() : nullptr,
                                      mode == 1 ? dF.d() : nullptr, mode == 2 ? dF.d() : nullptr, dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

extern "C" int neorx_rxchfne_fae_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                                      const double *dp, const double *gm2, double *fae, double *e)
{
    return rxne_closed(0, nebf, npbf, dae, dbe, dp, gm2, fae, e);
}
extern "C" int neorx_rxchfne_fp_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                                     const double *dp, const double *gm2, double *fp, double *e)
{
    return rxne_closed(1, nebf, npbf, dae, dbe, dp, gm2, fp, e);
}
extern "C" int neorx_rxchfne_fbe_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                                      const double *dp, const double *gm2, double *fbe, double *e)
{
    return rxne_closed(2, nebf, npbf, dae, dbe, dp, gm2, fbe, e);
}
extern "C" int neorx_rxcuhfne_fe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot,
                                      const double *dbe, const double *dp, const double *gm2, const double *gm2ex,
                                      double *faalp, double *fabet, double *e)
{
    if (!gm2ex || !fabet) return fail(NEORX_E_ARG, "null argument");
    return rxne_contract(0, nebf, npbf, dbe, datot, daalp, dp, gm2, gm2ex, faalp, fabet, e);
}
extern "C" int neorx_rxcuhfne_fp_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot,
                                      const double *dbe, const double *dp, const double *gm2, const double *gm2ex,
                                      double *fp, double *e)
{
    if (!gm2ex) return fail(NEORX_E_ARG, "null argument");
    return rxne_contract(1, nebf, npbf, dbe, datot, daalp, dp, gm2, gm2ex, fp, nullptr, e);
}
extern "C" int neorx_rxcuhfne_fbe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot,
                                       const double *dbe, const double *dp, const double *gm2, const double *gm2ex,
                                       double *fbe, double *e)
{
    if (!gm2ex) return fail(NEORX_E_ARG, "null argument");
    return rxne_contract(2, nebf, npbf, dbe, datot, daalp, dp, gm2, gm2ex, fbe, nullptr, e);
}

// RXCHFne_FAE_GAMee (rxchfne_hfterms.f:2-60): lane per FAE(i1,j1);
// part[t] = FAE(i1,j1) increment, energy from a second pass
__global__ void k_rxne_gamee(i64 ne, const double *__restrict__ Gee, const double *__restrict__ DAE,
                             double *__restrict__ F, double *__restrict__ xf)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ne * ne) return;
    i64 i1 = t / ne, j1 = t % ne, ne2 = ne * ne;
    double f = 0.0;
    for (i64 i2 = 0; i2 < ne; ++i2)
        for (i64 j2 = 0; j2 < ne; ++j2) {
            double v1 = Gee[(i1 * ne + j1) * ne2 + i2 * ne + j2], v2 = Gee[(i1 * ne + j2) * ne2 + i2 * ne + j1];
            f += DAE[j2 * ne + i2] * (v1 - 0.5 * v2);
        }
    xf[j1 * ne + i1] = f;
    F[j1 * ne + i1] += f;
}

__global__ void k_rxne_gamee_e(i64 ne, const double *__restrict__ DAE, const double *__restrict__ xf,
                               double *__restrict__ E)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double e = 0.0;
        for (i64 k = 0; k < ne * ne; ++k) e += 0.5 * DAE[k] * xf[k];
        *E = e;
    }
}

extern "C" int neorx_rxchfne_fae_gamee(int64_t nebf, const double *gam_ee, const double *dae, double *fae, double *e)
{
    int rc = ensure_init();
    if (rc) return rc;
    i64 ne = nebf;
    if (ne <= 0) return fail(NEORX_E_ARG, "bad sizes");
    DevBuf dG, dD, dF, dX, dE;
    rc = h2d(dG, gam_ee, ne * ne * ne * ne);
    if (!rc) rc = h2d(dD, dae, ne * ne);
    if (!rc) rc = h2d(dF, fae, ne * ne);
    if (!rc) rc = dX.alloc(ne * ne * 8);
    if (!rc) rc = dE.alloc(8);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rxne_gamee, dim3(nblocks(ne * ne, 256)), dim3(256), 0, 0, ne, dG.d(), dD.d(), dF.d(), dX.d());
    hipLaunchKernelGGL(k_rxne_gamee_e, dim3(1), dim3(64), 0, 0, ne, dD.d(), dX.d(), dE.d());
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(fae, dF.p, ne * ne * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

// ------------------------------------------------------------------
// RXCHF (rxchf/): one special electron.  Tuple class 9 is
// RXCHF_xcalc_GAM2_MD (ansE1, ansS, ansE2); RXCHF_GAM2_IC1 contracts the
// three outputs (RXCHF_contract_omega2_conv, rxchf_gam2_CONV_OMP.f:388-580)
// and keeps GM2_1 as is and GM2_2, GM2s as pair-swap averages
// (rxchf_gam2_IC1_OMP.f:158-190).
// ------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_rxchf_symm(i64 ne, i64 n, const double *__restrict__ X1,
                                                    const double *__restrict__ X2, const double *__restrict__ XS,
                                                    double *__restrict__ G1, double *__restrict__ G2,
                                                    double *__restrict__ GS)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        const i64 ne2 = ne * ne;
        const i64 b = t / (ne2 * ne2), p1 = (t / ne2) % ne2, p2 = t % ne2;
        const i64 t21 = (b * ne2 + p2) * ne2 + p1;
        G1[t] = X1[t];
        G2[t] = (X2[t] + X2[t21]) / 2.0;
        GS[t] = (XS[t] + XS[t21]) / 2.0;
    }
}

extern "C" int neorx_rxchf_gam2_ic1_dev(int64_t handle, double *d_gm2_1, double *d_gm2_2, double *d_gm2s, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (!d_gm2_1 || !d_gm2_2 || !d_gm2s) return fail(NEORX_E_ARG, "null output");
    if (s->d.nebf1 != s->d.nebf) return fail(NEORX_E_ARG, "RXCHF_GAM2_IC1 takes a one-basis handle");
    const i64 ne = s->d.nebf, np = s->d.npbf, ne4 = ne * ne * ne * ne, ng2 = ne4 * np * np;
    hipStream_t st = (hipStream_t)stream;
    i64 cap = slab_cap(s, 2);
    DevBuf X1, X2, XS, Y1, YS, Y2;
    int rc = X1.alloc(ng2 * 8);
    if (!rc) rc = X2.alloc(ng2 * 8);
    if (!rc) rc = XS.alloc(ng2 * 8);
    if (!rc) rc = Y1.alloc(cap * 8);
    if (!rc) rc = YS.alloc(cap * 8);
    if (!rc) rc = Y2.alloc(cap * 8);
    if (rc) return rc;
    for (auto &sl : make_slabs(s, 2, 0, np * np, cap)) {
        const i64 R0 = slab_start(s, 2, sl.first), R1 = slab_start(s, 2, sl.second);
        // primitive values of the slab: ansE1, ansS, ansE2
        rc = launch_prims<9, SRC_ORIG>(s->d, R0, R1 - R0, nullptr, Y1.d(), YS.d(), st, Y2.d());
        if (rc) return rc;
        const i64 n = (sl.second - sl.first) * ne4;
        hipLaunchKernelGGL(k_contract2, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y1.d(), YS.d(),
                           X1.d(), XS.d(), (i64)0, 0);
        // the third output through the same loops (both of its outputs are X2)
        hipLaunchKernelGGL(k_contract2, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y2.d(), Y2.d(),
                           X2.d(), X2.d(), (i64)0, 0);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rxchf_symm, dim3(grid_blocks(ng2, 256)), dim3(256), 0, st, ne, ng2, X1.d(), X2.d(), XS.d(),
                       d_gm2_1, d_gm2_2, d_gm2s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_rxchf_gam2_ic1(int64_t handle, double *gm2_1, double *gm2_2, double *gm2s)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    const i64 ne = s->d.nebf, np = s->d.npbf, n = ne * ne * ne * ne * np * np;
    DevBuf G1, G2, GS;
    int rc = G1.alloc(n * 8);
    if (!rc) rc = G2.alloc(n * 8);
    if (!rc) rc = GS.alloc(n * 8);
    if (!rc) rc = neorx_rxchf_gam2_ic1_dev(handle, G1.d(), G2.d(), GS.d(), 0);
    if (!rc) rc = copy_out(gm2_1, G1.d(), n);
    if (!rc) rc = copy_out(gm2_2, G2.d(), n);
    if (!rc) rc = copy_out(gm2s, GS.d(), n);
    if (!rc) {
        res_keep(G1, gm2_1, n);
        res_keep(G2, gm2_2, n);
        res_keep(GS, gm2s, n);
    }
    return rc;
}

// RXCHF_FAE_OMG2 / _FP_OMG2 / _FBE_OMG2 (rxchf_omg2.f:2-660).  Per entry of
// block (ip,jp): w1 = val1 - val2/2 and ws = -Sval/2 with val1 =
// GM2_1(i1 j1 i2 j2), val2 = GM2_2(i1 j2 i2 j1), Sval = GM2s(i1 j2 i2 j1).
// Row sums over (i2,j2) weighted by DAE and column sums over (i1,j1) weighted
// by DBE, per block, in a fixed order; then
//   FAE(i2,j2) += 2 sum_pq DP C1,   SAE += -2 sum_pq DP Cs   (FAE's 2 val1 - val2, SAE's Sval)
//   FP(pq)     += sum_r DBE R1,     SP  += sum_r DBE Rs
//   FBE(i1,j1) += sum_pq DP R1,     SBE += sum_pq DP Rs
//   E = sum_pq DP sum_r DBE R1,     S   = sum_pq DP sum_r DBE Rs.
// rows[pq][r][2] (r = i1*ne + j1), cols[pq][c][2] (c = i2*ne + j2)
__global__ void __launch_bounds__(64) k_rxchf_rows(i64 ne, i64 np, const double *__restrict__ DAE,
                                                   const double *__restrict__ G1, const double *__restrict__ G2,
                                                   const double *__restrict__ GS, double *__restrict__ rows)
{
    const i64 ne2 = ne * ne, row = blockIdx.x; // (pq, i1, j1)
    if (row >= np * np * ne2) return;
    const i64 pq = row / ne2, i1 = (row / ne) % ne, j1 = row % ne;
    const double *g1 = G1 + row * ne2, *g2 = G2 + pq * ne2 * ne2, *gs = GS + pq * ne2 * ne2;
    double r1 = 0.0, rs = 0.0;
    for (i64 c = threadIdx.x; c < ne2; c += 64) {
        const i64 i2 = c / ne, j2 = c % ne;
        const i64 x21 = ((i1 * ne + j2) * ne + i2) * ne + j1;
        const double dae = DAE[j2 * ne + i2];
        r1 += dae * (g1[c] - 0.5 * g2[x21]);
        rs += dae * (-0.5 * gs[x21]);
    }
    r1 = wave_sum(r1);
    rs = wave_sum(rs);
    if (threadIdx.x == 0) {
        rows[2 * row] = r1;
        rows[2 * row + 1] = rs;
    }
}

__global__ void __launch_bounds__(256) k_rxchf_cols(i64 ne, i64 np, const double *__restrict__ DBE,
                                                    const double *__restrict__ G1, const double *__restrict__ G2,
                                                    const double *__restrict__ GS, double *__restrict__ cols)
{
    const i64 ne2 = ne * ne;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; // (pq, i2, j2)
    if (t >= np * np * ne2) return;
    const i64 pq = t / ne2, c = t % ne2, i2 = c / ne, j2 = c % ne;
    const double *g1 = G1 + pq * ne2 * ne2, *g2 = G2 + pq * ne2 * ne2, *gs = GS + pq * ne2 * ne2;
    double c1 = 0.0, cs = 0.0;
    for (i64 i1 = 0; i1 < ne; ++i1)
        for (i64 j1 = 0; j1 < ne; ++j1) {
            const i64 x12 = (i1 * ne + j1) * ne2 + c, x21 = ((i1 * ne + j2) * ne + i2) * ne + j1;
            const double dbe = DBE[j1 * ne + i1];
            c1 += dbe * (g1[x12] - 0.5 * g2[x21]);
            cs += dbe * (-0.5 * gs[x21]);
        }
    cols[2 * t] = c1;
    cols[2 * t + 1] = cs;
}

// which: 0 FAE/SAE, 1 FP/SP, 2 FBE/SBE, 3 the energies only; F, SF
// accumulate; es[0] = E, es[1] = S
__global__ void __launch_bounds__(64) k_rxchf_finish(int which, i64 ne, i64 np, const double *__restrict__ DBE,
                                                     const double *__restrict__ DP, const double *__restrict__ rows,
                                                     const double *__restrict__ cols, double *__restrict__ F,
                                                     double *__restrict__ SF, double *__restrict__ es)
{
    const i64 ne2 = ne * ne, np2 = np * np;
    if (which == 0) {
        for (i64 c = threadIdx.x; c < ne2; c += 64) {
            double f = 0.0, sf = 0.0;
            for (i64 pq = 0; pq < np2; ++pq) {
                const double dp = DP[(pq % np) * np + pq / np];
                f += dp * cols[2 * (pq * ne2 + c)];
                sf += dp * cols[2 * (pq * ne2 + c) + 1];
            }
            const i64 i2 = c / ne, j2 = c % ne;
            F[j2 * ne + i2] += 2.0 * f;
            SF[j2 * ne + i2] += -2.0 * sf;
        }
    } else if (which == 1) {
        for (i64 pq = threadIdx.x; pq < np2; pq += 64) {
            double f = 0.0, sf = 0.0;
            for (i64 r = 0; r < ne2; ++r) {
                const double dbe = DBE[(r % ne) * ne + r / ne];
                f += dbe * rows[2 * (pq * ne2 + r)];
                sf += dbe * rows[2 * (pq * ne2 + r) + 1];
            }
            const i64 ip = pq / np, jp = pq % np;
            F[jp * np + ip] += f;
            SF[jp * np + ip] += sf;
        }
    } else if (which == 2) {
        for (i64 r = threadIdx.x; r < ne2; r += 64) {
            double f = 0.0, sf = 0.0;
            for (i64 pq = 0; pq < np2; ++pq) {
                const double dp = DP[(pq % np) * np + pq / np];
                f += dp * rows[2 * (pq * ne2 + r)];
                sf += dp * rows[2 * (pq * ne2 + r) + 1];
            }
            const i64 i1 = r / ne, j1 = r % ne;
            F[j1 * ne + i1] += f;
            SF[j1 * ne + i1] += sf;
        }
    }
    // energies: lane-strided partial sums over pq, then a fixed-order wave sum
    double e = 0.0, sv = 0.0;
    for (i64 pq = threadIdx.x; pq < np2; pq += 64) {
        const double dp = DP[(pq % np) * np + pq / np];
        double a = 0.0, b = 0.0;
        for (i64 r = 0; r < ne2; ++r) {
            const double dbe = DBE[(r % ne) * ne + r / ne];
            a += dbe * rows[2 * (pq * ne2 + r)];
            b += dbe * rows[2 * (pq * ne2 + r) + 1];
        }
        e += dp * a;
        sv += dp * b;
    }
    e = wave_sum(e);
    sv = wave_sum(sv);
    if (threadIdx.x == 0) {
        es[0] = e;
        es[1] = sv;
    }
}

extern "C" int neorx_rxchf_omg2_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_dae,
                                    const double *d_dbe, const double *d_dp, const double *d_gm2_1,
                                    const double *d_gm2_2, const double *d_gm2s, double *d_f, double *d_sf,
                                    double *d_es, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    const i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0 || which < 0 || which > 2) return fail(NEORX_E_ARG, "RXCHF OMG2: bad arguments");
    if (!d_dae || !d_dbe || !d_dp || !d_gm2_1 || !d_gm2_2 || !d_gm2s || !d_f || !d_sf || !d_es)
        return fail(NEORX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const i64 nrow = np * np * ne * ne;
    StreamBuf w;
    rc = w.alloc((size_t)(4 * nrow) * 8, st);
    if (rc) return rc;
    double *rows = (double *)w.p, *cols = rows + 2 * nrow;
    hipLaunchKernelGGL(k_rxchf_rows, dim3((unsigned)nrow), dim3(64), 0, st, ne, np, d_dae, d_gm2_1, d_gm2_2, d_gm2s,
                       rows);
    hipLaunchKernelGGL(k_rxchf_cols, dim3(nblocks(nrow, 256)), dim3(256), 0, st, ne, np, d_dbe, d_gm2_1, d_gm2_2,
                       d_gm2s, cols);
    hipLaunchKernelGGL(k_rxchf_finish, dim3(1), dim3(64), 0, st, (int)which, ne, np, d_dbe, d_dp, rows, cols, d_f,
                       d_sf, d_es);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

static int rxchf_omg2_host(int which, i64 ne, i64 np, const double *dae, const double *dbe, const double *dp,
                           const double *g1, const double *g2, const double *gs, double *f, double *sf, double *e,
                           double *sv)
{
    int rc = ensure_init();
    if (rc) return rc;
    const i64 ng2 = ne * ne * ne * ne * np * np, nf = which == 1 ? np * np : ne * ne;
    DevBuf dA, dB, dP, d1, d2, dS, dF, dSF, dE;
    rc = h2d(dA, dae, ne * ne);
    if (!rc) rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(d1, g1, ng2);
    if (!rc) rc = h2d(d2, g2, ng2);
    if (!rc) rc = h2d(dS, gs, ng2);
    if (!rc) rc = dF.alloc(nf * 8);
    if (!rc) rc = dSF.alloc(nf * 8);
    if (!rc) rc = dE.alloc(16);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dF.p, f, nf * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dSF.p, sf, nf * 8, hipMemcpyHostToDevice));
    rc = neorx_rxchf_omg2_dev(which, ne, np, dA.d(), dB.d(), dP.d(), d1.d(), d2.d(), dS.d(), dF.d(), dSF.d(), dE.d(),
                              0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    double es[2];
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sf, dSF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(es, dE.p, 16, hipMemcpyDeviceToHost));
    *e = es[0];
    *sv = es[1];
    return NEORX_OK;
}

extern "C" int neorx_rxchf_fae_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                                    const double *gm2_1, const double *gm2_2, const double *gm2s, double *fae,
                                    double *sae, double *e, double *s)
{
    return rxchf_omg2_host(0, nebf, npbf, dae, dbe, dp, gm2_1, gm2_2, gm2s, fae, sae, e, s);
}
extern "C" int neorx_rxchf_fp_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                                   const double *gm2_1, const double *gm2_2, const double *gm2s, double *fp, double *sp,
                                   double *e, double *s)
{
    return rxchf_omg2_host(1, nebf, npbf, dae, dbe, dp, gm2_1, gm2_2, gm2s, fp, sp, e, s);
}
extern "C" int neorx_rxchf_fbe_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                                    const double *gm2_1, const double *gm2_2, const double *gm2s, double *fbe,
                                    double *sbe, double *e, double *s)
{
    return rxchf_omg2_host(2, nebf, npbf, dae, dbe, dp, gm2_1, gm2_2, gm2s, fbe, sbe, e, s);
}

// ------------------------------------------------------------------
// RXCHF e-e-e-p tensors (rxchf/rxchf_gam3_IC1_OMP.f) and their contractions
// (rxchf_omg3.f), and the GAM_1 contractions (rxchf_omg1.f)
// ------------------------------------------------------------------
// RXCHF_GAM3_IC1's symmetrisation (rxchf_gam3_IC1_OMP.f:211-238) of one
// (ip,jp) block: GM3_1 = (x123 + x132)/2 of the contracted ans1 (pairs 2 and
// 3 exchanged), GM3_2 = the contracted ans2 as is;
// t = ((((i1 ne + j1) ne + i2) ne + j2) ne + i3) ne + j3
__global__ void __launch_bounds__(256) k_rxchf3_symm(i64 ne, i64 blk, const double *__restrict__ X1,
                                                     const double *__restrict__ X2, double *__restrict__ G1,
                                                     double *__restrict__ G2)
{
    const i64 ne2 = ne * ne;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < blk; t += (i64)gridDim.x * blockDim.x) {
        const i64 p3 = t % ne2, p2 = (t / ne2) % ne2, p1 = t / (ne2 * ne2);
        G1[t] = (X1[t] + X1[(p1 * ne2 + p3) * ne2 + p2]) * 0.5;
        G2[t] = X2[t];
    }
}

extern "C" int neorx_rxchf_gam3_ic1_dev(int64_t handle, double *d_gm3_1, double *d_gm3_2, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (!d_gm3_1 || !d_gm3_2) return fail(NEORX_E_ARG, "null output");
    if (s->d.nebf1 != s->d.nebf) return fail(NEORX_E_ARG, "RXCHF_GAM3_IC1 takes a one-basis handle");
    const i64 ne = s->d.nebf, np = s->d.npbf, ne5 = ne * ne * ne * ne * ne, blk = ne5 * ne;
    hipStream_t st = (hipStream_t)stream;
    const i64 cap = slab_cap(s, 3);
    // one (ip,jp) block of unsymmetrised values at a time
    DevBuf X1, X2, Y1, Y2;
    int rc = X1.alloc(blk * 8);
    if (!rc) rc = X2.alloc(blk * 8);
    if (!rc) rc = Y1.alloc(cap * 8);
    if (!rc) rc = Y2.alloc(cap * 8);
    if (rc) return rc;
    for (i64 b = 0; b < np * np; ++b) {
        i64 c2;
        for (auto &sl : make_slabs(s, 3, b * ne, (b + 1) * ne, c2)) {
            const i64 R0 = slab_start(s, 3, sl.first), R1 = slab_start(s, 3, sl.second);
            // primitive values of the slab: ans1, ans2
            rc = launch_prims<10, SRC_ORIG>(s->d, R0, R1 - R0, nullptr, Y1.d(), Y2.d(), st);
            if (rc) return rc;
            const i64 n = (sl.second - sl.first) * ne5;
            hipLaunchKernelGGL(k_contract3, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y1.d(),
                               X1.d(), b * ne);
            hipLaunchKernelGGL(k_contract3, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y2.d(),
                               X2.d(), b * ne);
            HIPCHK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_rxchf3_symm, dim3(grid_blocks(blk, 256)), dim3(256), 0, st, ne, blk, X1.d(), X2.d(),
                           d_gm3_1 + b * blk, d_gm3_2 + b * blk);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_rxchf_gam3_ic1(int64_t handle, double *gm3_1, double *gm3_2)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    const i64 ne = s->d.nebf, np = s->d.npbf, n = ne * ne * ne * ne * ne * ne * np * np;
    DevBuf G1, G2;
    int rc = G1.alloc(n * 8);
    if (!rc) rc = G2.alloc(n * 8);
    if (!rc) rc = neorx_rxchf_gam3_ic1_dev(handle, G1.d(), G2.d(), 0);
    if (!rc) rc = copy_out(gm3_1, G1.d(), n);
    if (!rc) rc = copy_out(gm3_2, G2.d(), n);
    if (!rc) {
        res_keep(G1, gm3_1, n);
        res_keep(G2, gm3_2, n);
    }
    return rc;
}

// RXCHF_FAE_OMG3 / _FP_OMG3 / _FBE_OMG3 (rxchf_omg3.f:2-708).  Thread per
// (pq, i1, j1, i2, j2); over (i3, j3), with
//   val1 = GM3_1(i1 j1 i2 j2 i3 j3) - GM3_1(i1 j1 i2 j3 i3 j2)/2          (:215)
//   val2 = (GM3_2(i1 j2 i2 j1 i3 j3) + GM3_2(i1 i2 j2 j1 i3 j3))
//          - (GM3_2(i1 j3 i2 j1 i3 j2) + GM3_2(i1 j3 j2 j1 i3 i2))/2       (:239-240)
//   val3 = (GM3_2(i1 j3 i3 j1 i2 j2) + GM3_2(i1 j3 i3 j1 j2 i2))
//          - (GM3_2(i1 j2 i3 j1 i2 j3) + GM3_2(i1 i2 i3 j1 j2 j3))/2       (:241-242)
//   v2p  = GM3_2(i1 j2 i2 j1 i3 j3) - GM3_2(i1 j3 i2 j1 i3 j2)/2           (:466)
// it sums a = DAE(i3,j3) (2 val1 - (val2 + val3)/2) (FAE, :245-247),
// w = DAE(i3,j3) (val1 - v2p)/2 (FP and FBE, :469-475, :692-698) and
// v = DAE(i3,j3) (val1 - val2/2)/2 (FAE's energy, :249-251), and stores, per
// row q = (pq, i1, j1) and column c = (i2, j2):
//   P[q][c] = DP(pq) DBE(i1,j1) a,  W[q][c] = DAE(i2,j2) w,
//   V[q][c] = DP(pq) DBE(i1,j1) DAE(i2,j2) v.
__global__ void __launch_bounds__(256) k_rxchf3_terms(i64 ne, i64 np, const double *__restrict__ DAE,
                                                      const double *__restrict__ DBE, const double *__restrict__ DP,
                                                      const double *__restrict__ G1, const double *__restrict__ G2,
                                                      double *__restrict__ P, double *__restrict__ W,
                                                      double *__restrict__ V)
{
    const i64 ne2 = ne * ne, ne4 = ne2 * ne2;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np * np * ne4) return;
    const i64 pq = t / ne4, r = t % ne4, i1 = r / (ne2 * ne), j1 = (r / ne2) % ne, i2 = (r / ne) % ne, j2 = r % ne;
    const double *g1 = G1 + pq * ne4 * ne2, *g2 = G2 + pq * ne4 * ne2;
    auto X = [ne](i64 a1, i64 b1, i64 a2, i64 b2, i64 a3, i64 b3) {
        return ((((a1 * ne + b1) * ne + a2) * ne + b2) * ne + a3) * ne + b3;
    };
    double a = 0.0, w = 0.0, v = 0.0;
    for (i64 i3 = 0; i3 < ne; ++i3)
        for (i64 j3 = 0; j3 < ne; ++j3) {
            const double d3 = DAE[j3 * ne + i3];
            const double val1 = g1[X(i1, j1, i2, j2, i3, j3)] - 0.5 * g1[X(i1, j1, i2, j3, i3, j2)];
            const double g213 = g2[X(i1, j2, i2, j1, i3, j3)], g312 = g2[X(i1, j3, i2, j1, i3, j2)];
            const double val2 = (g213 + g2[X(i1, i2, j2, j1, i3, j3)]) - 0.5 * (g312 + g2[X(i1, j3, j2, j1, i3, i2)]);
            const double val3 = (g2[X(i1, j3, i3, j1, i2, j2)] + g2[X(i1, j3, i3, j1, j2, i2)]) -
                                0.5 * (g2[X(i1, j2, i3, j1, i2, j3)] + g2[X(i1, i2, i3, j1, j2, j3)]);
            const double v2p = g213 - 0.5 * g312;
            a += d3 * (2.0 * val1 - 0.5 * (val2 + val3));
            w += d3 * (0.5 * (val1 - v2p));
            v += d3 * (0.5 * (val1 - 0.5 * val2));
        }
    const double dpb = DP[(pq % np) * np + pq / np] * DBE[j1 * ne + i1], d2 = DAE[j2 * ne + i2];
    P[t] = dpb * a;
    W[t] = d2 * w;
    V[t] = dpb * d2 * v;
}

// wave per row q: R[q] = sum_c W[q][c], Ev[q] = sum_c V[q][c]
__global__ void __launch_bounds__(64) k_rxchf3_rows(i64 ncol, const double *__restrict__ W,
                                                    const double *__restrict__ V, double *__restrict__ R,
                                                    double *__restrict__ Ev)
{
    const i64 q = blockIdx.x;
    double x = 0.0, y = 0.0;
    for (i64 c = threadIdx.x; c < ncol; c += 64) {
        x += W[q * ncol + c];
        y += V[q * ncol + c];
    }
    x = wave_sum(x);
    y = wave_sum(y);
    if (threadIdx.x == 0) {
        R[q] = x;
        Ev[q] = y;
    }
}

// which 1: FP(pq) += sum_r DBE(r) R[pq][r]; which 2: FBE(r) += sum_pq DP(pq)
// R[pq][r]; E = sum_q Ev[q] for which 0, else (1, 2, and 3: the energy
// only) sum_pq DP sum_r DBE R
__global__ void __launch_bounds__(64) k_rxchf3_fin(int which, i64 ne, i64 np, const double *__restrict__ DBE,
                                                   const double *__restrict__ DP, const double *__restrict__ R,
                                                   const double *__restrict__ Ev, double *__restrict__ F,
                                                   double *__restrict__ E)
{
    const i64 ne2 = ne * ne, np2 = np * np;
    if (which == 1)
        for (i64 pq = threadIdx.x; pq < np2; pq += 64) {
            double f = 0.0;
            for (i64 r = 0; r < ne2; ++r) f += DBE[(r % ne) * ne + r / ne] * R[pq * ne2 + r];
            F[(pq % np) * np + pq / np] += f;
        }
    else if (which == 2)
        for (i64 r = threadIdx.x; r < ne2; r += 64) {
            double f = 0.0;
            for (i64 pq = 0; pq < np2; ++pq) f += DP[(pq % np) * np + pq / np] * R[pq * ne2 + r];
            F[(r % ne) * ne + r / ne] += f;
        }
    double e = 0.0;
    for (i64 q = threadIdx.x; q < np2 * ne2; q += 64) {
        const i64 pq = q / ne2, r = q % ne2;
        e += which == 0 ? Ev[q] : DP[(pq % np) * np + pq / np] * DBE[(r % ne) * ne + r / ne] * R[q];
    }
    e = wave_sum(e);
    if (threadIdx.x == 0) *E = e;
}

extern "C" int neorx_rxchf_omg3_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_dae,
                                    const double *d_dbe, const double *d_dp, const double *d_gm3_1,
                                    const double *d_gm3_2, double *d_f, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    const i64 ne = nebf, np = npbf, ne2 = ne * ne;
    if (ne <= 0 || np <= 0 || which < 0 || which > 2) return fail(NEORX_E_ARG, "RXCHF OMG3: bad arguments");
    if (!d_dae || !d_dbe || !d_dp || !d_gm3_1 || !d_gm3_2 || !d_f || !d_e) return fail(NEORX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const i64 nq = np * np * ne2, n5 = nq * ne2;
    StreamBuf w;
    rc = w.alloc((size_t)(3 * n5 + 2 * nq + WCS_GROUPS * ne2) * 8, st);
    if (rc) return rc;
    double *P = (double *)w.p, *W = P + n5, *V = W + n5, *R = V + n5, *Ev = R + nq, *o1 = Ev + nq;
    hipLaunchKernelGGL(k_rxchf3_terms, dim3(nblocks(n5, 256)), dim3(256), 0, st, ne, np, d_dae, d_dbe, d_dp, d_gm3_1,
                       d_gm3_2, P, W, V);
    hipLaunchKernelGGL(k_rxchf3_rows, dim3((unsigned)nq), dim3(64), 0, st, ne2, W, V, R, Ev);
    if (which == 0) {
        hipLaunchKernelGGL(k_wcolsum1, dim3((unsigned)((ne2 + 63) / 64), WCS_GROUPS), dim3(256), 0, st, nq, ne2, (i64)1,
                           np, (const double *)nullptr, P, o1);
        hipLaunchKernelGGL(k_wcolsum2, dim3(nblocks(ne2, 256)), dim3(256), 0, st, ne, o1, d_f);
    }
    hipLaunchKernelGGL(k_rxchf3_fin, dim3(1), dim3(64), 0, st, (int)which, ne, np, d_dbe, d_dp, R, Ev, d_f, d_e);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

static int rxchf_omg3_host(int which, i64 ne, i64 np, const double *dae, const double *dbe, const double *dp,
                           const double *g1, const double *g2, double *f, double *e)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    const i64 ng3 = ne * ne * ne * ne * ne * ne * np * np, nf = which == 1 ? np * np : ne * ne;
    DevBuf dA, dB, dP, d1, d2, dF, dE;
    rc = h2d(dA, dae, ne * ne);
    if (!rc) rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(d1, g1, ng3);
    if (!rc) rc = h2d(d2, g2, ng3);
    if (!rc) rc = dF.alloc(nf * 8);
    if (!rc) rc = dE.alloc(8);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dF.p, f, nf * 8, hipMemcpyHostToDevice));
    rc = neorx_rxchf_omg3_dev(which, ne, np, dA.d(), dB.d(), dP.d(), d1.d(), d2.d(), dF.d(), dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

extern "C" int neorx_rxchf_fae_omg3(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                                    const double *gm3_1, const double *gm3_2, double *fae, double *e)
{
    return rxchf_omg3_host(0, nebf, npbf, dae, dbe, dp, gm3_1, gm3_2, fae, e);
}
extern "C" int neorx_rxchf_fp_omg3(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                                   const double *gm3_1, const double *gm3_2, double *fp, double *e)
{
    return rxchf_omg3_host(1, nebf, npbf, dae, dbe, dp, gm3_1, gm3_2, fp, e);
}
extern "C" int neorx_rxchf_fbe_omg3(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                                    const double *gm3_1, const double *gm3_2, double *fbe, double *e)
{
    return rxchf_omg3_host(2, nebf, npbf, dae, dbe, dp, gm3_1, gm3_2, fbe, e);
}

// RXCHF_FP_OMG1 / RXCHF_FBE_OMG1 (rxchf_omg1.f:2-165, per entry :60-66,
// :142-148) over GAM_1 / GAM_1s (index ((ip np + jp) ne + i1) ne + j1):
// which 1: FP(ip,jp) += sum DBE(i1,j1) GAM_1, SP = sum DBE GAM_1s;
// which 2: FBE(i1,j1) += sum DP(ip,jp) GAM_1, SBE = sum DP GAM_1s;
// es[0] = sum DP DBE GAM_1, es[1] = sum DP DBE GAM_1s.  One block.
__global__ void __launch_bounds__(64) k_rxchf_omg1(int which, i64 ne, i64 np, const double *__restrict__ DBE,
                                                   const double *__restrict__ DP, const double *__restrict__ G,
                                                   const double *__restrict__ Gs, double *__restrict__ F,
                                                   double *__restrict__ SF, double *__restrict__ es)
{
    const i64 ne2 = ne * ne, np2 = np * np;
    if (which == 1)
        for (i64 pq = threadIdx.x; pq < np2; pq += 64) {
            double f = 0.0, sf = 0.0;
            for (i64 r = 0; r < ne2; ++r) {
                const double d = DBE[(r % ne) * ne + r / ne];
                f += d * G[pq * ne2 + r];
                sf += d * Gs[pq * ne2 + r];
            }
            const i64 at = (pq % np) * np + pq / np;
            F[at] += f;
            SF[at] = sf;
        }
    else
        for (i64 r = threadIdx.x; r < ne2; r += 64) {
            double f = 0.0, sf = 0.0;
            for (i64 pq = 0; pq < np2; ++pq) {
                const double d = DP[(pq % np) * np + pq / np];
                f += d * G[pq * ne2 + r];
                sf += d * Gs[pq * ne2 + r];
            }
            const i64 at = (r % ne) * ne + r / ne;
            F[at] += f;
            SF[at] = sf;
        }
    double e = 0.0, sv = 0.0;
    for (i64 q = threadIdx.x; q < np2 * ne2; q += 64) {
        const i64 pq = q / ne2, r = q % ne2;
        const double d = DP[(pq % np) * np + pq / np] * DBE[(r % ne) * ne + r / ne];
        e += d * G[q];
        sv += d * Gs[q];
    }
    e = wave_sum(e);
    sv = wave_sum(sv);
    if (threadIdx.x == 0) {
        es[0] = e;
        es[1] = sv;
    }
}

extern "C" int neorx_rxchf_omg1_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_dbe,
                                    const double *d_dp, const double *d_gam1, const double *d_gam1s, double *d_f,
                                    double *d_sf, double *d_es, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (nebf <= 0 || npbf <= 0 || (which != 1 && which != 2)) return fail(NEORX_E_ARG, "RXCHF OMG1: bad arguments");
    if (!d_dbe || !d_dp || !d_gam1 || !d_gam1s || !d_f || !d_sf || !d_es) return fail(NEORX_E_ARG, "null argument");
    hipLaunchKernelGGL(k_rxchf_omg1, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)which, nebf, npbf, d_dbe, d_dp,
                       d_gam1, d_gam1s, d_f, d_sf, d_es);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

static int rxchf_omg1_host(int which, i64 ne, i64 np, const double *dbe, const double *dp, const double *g1,
                           const double *g1s, double *f, double *sf, double *e, double *sv)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    const i64 ng1 = ne * ne * np * np, nf = which == 1 ? np * np : ne * ne;
    DevBuf dB, dP, d1, d1s, dF, dSF, dE;
    rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(d1, g1, ng1);
    if (!rc) rc = h2d(d1s, g1s, ng1);
    if (!rc) rc = dF.alloc(nf * 8);
    if (!rc) rc = dSF.alloc(nf * 8);
    if (!rc) rc = dE.alloc(16);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dF.p, f, nf * 8, hipMemcpyHostToDevice));
    rc = neorx_rxchf_omg1_dev(which, ne, np, dB.d(), dP.d(), d1.d(), d1s.d(), dF.d(), dSF.d(), dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    double es[2];
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sf, dSF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(es, dE.p, 16, hipMemcpyDeviceToHost));
    *e = es[0];
    *sv = es[1];
    return NEORX_OK;
}

extern "C" int neorx_rxchf_fp_omg1(int64_t nebf, int64_t npbf, const double *dbe, const double *dp,
                                   const double *gam1, const double *gam1s, double *fp, double *sp, double *e,
                                   double *s)
{
    return rxchf_omg1_host(1, nebf, npbf, dbe, dp, gam1, gam1s, fp, sp, e, s);
}
extern "C" int neorx_rxchf_fbe_omg1(int64_t nebf, int64_t npbf, const double *dbe, const double *dp,
                                    const double *gam1, const double *gam1s, double *fbe, double *sbe, double *e,
                                    double *s)
{
    return rxchf_omg1_host(2, nebf, npbf, dbe, dp, gam1, gam1s, fbe, sbe, e, s);
}

// ------------------------------------------------------------------
// RXCUHF: open-shell regular electrons (rxchf/rxcuhf_omg2.f, rxcuhf_omg3.f)
// ------------------------------------------------------------------
// RXCUHF_FE/FP/FBE_OMG2 (rxcuhf_omg2.f:2-694).  Per entry of block pq:
// val1 = GM2_1(i1 j1 i2 j2), val2 = GM2_2(i1 j2 i2 j1), Sval = GM2s(i1 j2 i2 j1)
// (:219-243, :438-454, :664-680).  Rows (pq, i1, j1), in k_rxchf_finish's
// layout: r = sum_c DAtot(c) val1 - DAalp(c) val2, rs = -sum_c DAalp(c) Sval;
// columns (pq, i2, j2): DBE-weighted sums of val1 - val2, Sval and val1.
__global__ void __launch_bounds__(64) k_rxcuhf2_rows(i64 ne, i64 np, const double *__restrict__ Dal,
                                                     const double *__restrict__ Dto, const double *__restrict__ G1,
                                                     const double *__restrict__ G2, const double *__restrict__ GS,
                                                     double *__restrict__ rows)
{
    const i64 ne2 = ne * ne, row = blockIdx.x; // (pq, i1, j1)
    if (row >= np * np * ne2) return;
    const i64 pq = row / ne2, i1 = (row / ne) % ne, j1 = row % ne;
    const double *g1 = G1 + row * ne2, *g2 = G2 + pq * ne2 * ne2, *gs = GS + pq * ne2 * ne2;
    double r1 = 0.0, rs = 0.0;
    for (i64 c = threadIdx.x; c < ne2; c += 64) {
        const i64 i2 = c / ne, j2 = c % ne;
        const i64 x21 = ((i1 * ne + j2) * ne + i2) * ne + j1;
        const double dal = Dal[j2 * ne + i2];
        r1 += Dto[j2 * ne + i2] * g1[c] - dal * g2[x21];
        rs += -dal * gs[x21];
    }
    r1 = wave_sum(r1);
    rs = wave_sum(rs);
    if (threadIdx.x == 0) {
        rows[2 * row] = r1;
        rows[2 * row + 1] = rs;
    }
}

__global__ void __launch_bounds__(256) k_rxcuhf2_cols(i64 ne, i64 np, const double *__restrict__ DBE,
                                                      const double *__restrict__ G1, const double *__restrict__ G2,
                                                      const double *__restrict__ GS, double *__restrict__ cols)
{
    const i64 ne2 = ne * ne;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; // (pq, i2, j2)
    if (t >= np * np * ne2) return;
    const i64 pq = t / ne2, c = t % ne2, i2 = c / ne, j2 = c % ne;
    const double *g1 = G1 + pq * ne2 * ne2, *g2 = G2 + pq * ne2 * ne2, *gs = GS + pq * ne2 * ne2;
    double ca = 0.0, cs = 0.0, cb = 0.0;
    for (i64 i1 = 0; i1 < ne; ++i1)
        for (i64 j1 = 0; j1 < ne; ++j1) {
            const i64 x12 = (i1 * ne + j1) * ne2 + c, x21 = ((i1 * ne + j2) * ne + i2) * ne + j1;
            const double dbe = DBE[j1 * ne + i1];
            ca += dbe * (g1[x12] - g2[x21]);
            cs += dbe * gs[x21];
            cb += dbe * g1[x12];
        }
    cols[3 * t] = ca;
    cols[3 * t + 1] = cs;
    cols[3 * t + 2] = cb;
}

// FE: FAalp(i2,j2) += sum_pq DP ca, SAalp += sum_pq DP cs, FAbet += sum_pq DP cb
__global__ void __launch_bounds__(64) k_rxcuhf2_fe(i64 ne, i64 np, const double *__restrict__ DP,
                                                   const double *__restrict__ cols, double *__restrict__ Fa,
                                                   double *__restrict__ Fb, double *__restrict__ Sa)
{
    const i64 ne2 = ne * ne, np2 = np * np;
    for (i64 c = threadIdx.x; c < ne2; c += 64) {
        double a = 0.0, sa = 0.0, b = 0.0;
        for (i64 pq = 0; pq < np2; ++pq) {
            const double dp = DP[(pq % np) * np + pq / np];
            a += dp * cols[3 * (pq * ne2 + c)];
            sa += dp * cols[3 * (pq * ne2 + c) + 1];
            b += dp * cols[3 * (pq * ne2 + c) + 2];
        }
        const i64 at = (c % ne) * ne + c / ne;
        Fa[at] += a;
        Sa[at] += sa;
        Fb[at] += b;
    }
}

extern "C" int neorx_rxcuhf_omg2_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_daalp,
                                     const double *d_datot, const double *d_dbe, const double *d_dp,
                                     const double *d_gm2_1, const double *d_gm2_2, const double *d_gm2s, double *d_f,
                                     double *d_f2, double *d_sf, double *d_es, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    const i64 ne = nebf, np = npbf;
    if (ne <= 0 || np <= 0 || which < 0 || which > 2) return fail(NEORX_E_ARG, "RXCUHF OMG2: bad arguments");
    if (!d_daalp || !d_datot || !d_dbe || !d_dp || !d_gm2_1 || !d_gm2_2 || !d_gm2s || !d_f || !d_sf || !d_es ||
        (which == 0 && !d_f2))
        return fail(NEORX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const i64 nrow = np * np * ne * ne;
    StreamBuf w;
    rc = w.alloc((size_t)(5 * nrow) * 8, st);
    if (rc) return rc;
    double *rows = (double *)w.p, *cols = rows + 2 * nrow;
    hipLaunchKernelGGL(k_rxcuhf2_rows, dim3((unsigned)nrow), dim3(64), 0, st, ne, np, d_daalp, d_datot, d_gm2_1,
                       d_gm2_2, d_gm2s, rows);
    if (which == 0) {
        hipLaunchKernelGGL(k_rxcuhf2_cols, dim3(nblocks(nrow, 256)), dim3(256), 0, st, ne, np, d_dbe, d_gm2_1, d_gm2_2,
                           d_gm2s, cols);
        hipLaunchKernelGGL(k_rxcuhf2_fe, dim3(1), dim3(64), 0, st, ne, np, d_dp, cols, d_f, d_f2, d_sf);
    }
    hipLaunchKernelGGL(k_rxchf_finish, dim3(1), dim3(64), 0, st, which == 0 ? 3 : (int)which, ne, np, d_dbe, d_dp,
                       rows, cols, d_f, d_sf, d_es);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

static int rxcuhf_omg2_host(int which, i64 ne, i64 np, const double *dal, const double *dto, const double *dbe,
                            const double *dp, const double *g1, const double *g2, const double *gs, double *f,
                            double *f2, double *sf, double *e, double *sv)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    const i64 ng2 = ne * ne * ne * ne * np * np, nf = which == 1 ? np * np : ne * ne;
    DevBuf dL, dT, dB, dP, d1, d2, dS, dF, dF2, dSF, dE;
    rc = h2d(dL, dal, ne * ne);
    if (!rc) rc = h2d(dT, dto, ne * ne);
    if (!rc) rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(d1, g1, ng2);
    if (!rc) rc = h2d(d2, g2, ng2);
    if (!rc) rc = h2d(dS, gs, ng2);
    if (!rc) rc = dF.alloc(nf * 8);
    if (!rc) rc = dF2.alloc(nf * 8);
    if (!rc) rc = dSF.alloc(nf * 8);
    if (!rc) rc = dE.alloc(16);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dF.p, f, nf * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dSF.p, sf, nf * 8, hipMemcpyHostToDevice));
    if (which == 0) HIPCHK(hipMemcpy(dF2.p, f2, nf * 8, hipMemcpyHostToDevice));
    rc = neorx_rxcuhf_omg2_dev(which, ne, np, dL.d(), dT.d(), dB.d(), dP.d(), d1.d(), d2.d(), dS.d(), dF.d(), dF2.d(),
                               dSF.d(), dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    double es[2];
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sf, dSF.p, nf * 8, hipMemcpyDeviceToHost));
    if (which == 0) HIPCHK(hipMemcpy(f2, dF2.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(es, dE.p, 16, hipMemcpyDeviceToHost));
    *e = es[0];
    *sv = es[1];
    return NEORX_OK;
}

extern "C" int neorx_rxcuhf_fe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet,
                                    const double *datot, const double *dbe, const double *dp, const double *gm2_1,
                                    const double *gm2_2, const double *gm2s, double *faalp, double *fabet,
                                    double *saalp, double *e, double *s)
{
    (void)dabet; // RXCUHF_FE_OMG2 takes DAbetE and does not read it
    if (!fabet) return fail(NEORX_E_ARG, "null argument");
    return rxcuhf_omg2_host(0, nebf, npbf, daalp, datot, dbe, dp, gm2_1, gm2_2, gm2s, faalp, fabet, saalp, e, s);
}
extern "C" int neorx_rxcuhf_fp_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot,
                                    const double *dbe, const double *dp, const double *gm2_1, const double *gm2_2,
                                    const double *gm2s, double *fp, double *sp, double *e, double *s)
{
    return rxcuhf_omg2_host(1, nebf, npbf, daalp, datot, dbe, dp, gm2_1, gm2_2, gm2s, fp, nullptr, sp, e, s);
}
extern "C" int neorx_rxcuhf_fbe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot,
                                     const double *dbe, const double *dp, const double *gm2_1, const double *gm2_2,
                                     const double *gm2s, double *fbe, double *sbe, double *e, double *s)
{
    return rxcuhf_omg2_host(2, nebf, npbf, daalp, datot, dbe, dp, gm2_1, gm2_2, gm2s, fbe, nullptr, sbe, e, s);
}

// RXCUHF_FE/FP/FBE_OMG3 (rxcuhf_omg3.f:2-728).  Thread per (pq, i1, j1, i2, j2);
// over (i3, j3), with val1a = GM3_1(i1 j1 i2 j2 i3 j3), val1b = GM3_1(i1 j1 i2 j3
// i3 j2), g213 = GM3_2(i1 j2 i2 j1 i3 j3), g312 = GM3_2(i1 j3 i2 j1 i3 j2) and the
// FE-only pair averages val2a..val2d (:233-251):
//   fa = DAtot3 val1a - DAalp3 val1b - DAtot3 val2a + DAalp3 val2b
//        - DAalp3 val2c + DAalp3 val2d                         (FAalpE, :240-246)
//   fb = DAtot3 val1a - DAbet3 val1b - DAalp3 val2c          (FAbetE, :247-250)
//   w  = DAtot2 DAtot3 val1a / 2 - (DAalp2 DAalp3 + DAbet2 DAbet3) val1b / 2
//        - DAalp2 DAtot3 g213 + DAalp2 DAalp3 g312            (FP, FBE, E: :476-489)
// stores Pa = DBE DP fa, Pb = DBE DP fb (sums over (i3,j3)) and W = w.
__global__ void __launch_bounds__(256) k_rxcuhf3_terms(int fe, i64 ne, i64 np, const double *__restrict__ Dal,
                                                       const double *__restrict__ Dbt, const double *__restrict__ Dto,
                                                       const double *__restrict__ DBE, const double *__restrict__ DP,
                                                       const double *__restrict__ G1, const double *__restrict__ G2,
                                                       double *__restrict__ Pa, double *__restrict__ Pb,
                                                       double *__restrict__ W)
{
    const i64 ne2 = ne * ne, ne4 = ne2 * ne2;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np * np * ne4) return;
    const i64 pq = t / ne4, r = t % ne4, i1 = r / (ne2 * ne), j1 = (r / ne2) % ne, i2 = (r / ne) % ne, j2 = r % ne;
    const double *g1 = G1 + pq * ne4 * ne2, *g2 = G2 + pq * ne4 * ne2;
    auto X = [ne](i64 a1, i64 b1, i64 a2, i64 b2, i64 a3, i64 b3) {
        return ((((a1 * ne + b1) * ne + a2) * ne + b2) * ne + a3) * ne + b3;
    };
    const double al2 = Dal[j2 * ne + i2], bt2 = Dbt[j2 * ne + i2], to2 = Dto[j2 * ne + i2];
    double fa = 0.0, fb = 0.0, w = 0.0;
    for (i64 i3 = 0; i3 < ne; ++i3)
        for (i64 j3 = 0; j3 < ne; ++j3) {
            const double al3 = Dal[j3 * ne + i3], bt3 = Dbt[j3 * ne + i3], to3 = Dto[j3 * ne + i3];
            const double val1a = g1[X(i1, j1, i2, j2, i3, j3)], val1b = g1[X(i1, j1, i2, j3, i3, j2)];
            const double g213 = g2[X(i1, j2, i2, j1, i3, j3)], g312 = g2[X(i1, j3, i2, j1, i3, j2)];
            w += 0.5 * to2 * to3 * val1a - 0.5 * (al2 * al3 + bt2 * bt3) * val1b - al2 * to3 * g213 +
                 al2 * al3 * g312;
            if (fe) {
                const double val2a = (g213 + g2[X(i1, i2, j2, j1, i3, j3)]) * 0.5;
                const double val2b = (g312 + g2[X(i1, j3, j2, j1, i3, i2)]) * 0.5;
                const double val2c = (g2[X(i1, j3, i3, j1, i2, j2)] + g2[X(i1, j3, i3, j1, j2, i2)]) * 0.5;
                const double val2d = (g2[X(i1, j2, i3, j1, i2, j3)] + g2[X(i1, i2, i3, j1, j2, j3)]) * 0.5;
                fa += to3 * val1a - al3 * val1b - to3 * val2a + al3 * val2b - al3 * val2c + al3 * val2d;
                fb += to3 * val1a - bt3 * val1b - al3 * val2c;
            }
        }
    const double f = DBE[j1 * ne + i1] * DP[(pq % np) * np + pq / np];
    Pa[t] = f * fa;
    Pb[t] = f * fb;
    W[t] = w;
}

extern "C" int neorx_rxcuhf_omg3_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_daalp,
                                     const double *d_dabet, const double *d_datot, const double *d_dbe,
                                     const double *d_dp, const double *d_gm3_1, const double *d_gm3_2, double *d_f,
                                     double *d_f2, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    const i64 ne = nebf, np = npbf, ne2 = ne * ne;
    if (ne <= 0 || np <= 0 || which < 0 || which > 2) return fail(NEORX_E_ARG, "RXCUHF OMG3: bad arguments");
    if (!d_daalp || !d_dabet || !d_datot || !d_dbe || !d_dp || !d_gm3_1 || !d_gm3_2 || !d_f || !d_e ||
        (which == 0 && !d_f2))
        return fail(NEORX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const i64 nq = np * np * ne2, n5 = nq * ne2;
    StreamBuf w;
    rc = w.alloc((size_t)(3 * n5 + 2 * nq + WCS_GROUPS * ne2) * 8, st);
    if (rc) return rc;
    double *Pa = (double *)w.p, *Pb = Pa + n5, *W = Pb + n5, *R = W + n5, *Ev = R + nq, *o1 = Ev + nq;
    hipLaunchKernelGGL(k_rxcuhf3_terms, dim3(nblocks(n5, 256)), dim3(256), 0, st, (int)(which == 0), ne, np, d_daalp,
                       d_dabet, d_datot, d_dbe, d_dp, d_gm3_1, d_gm3_2, Pa, Pb, W);
    hipLaunchKernelGGL(k_rxchf3_rows, dim3((unsigned)nq), dim3(64), 0, st, ne2, W, W, R, Ev);
    if (which == 0) {
        const dim3 g1((unsigned)((ne2 + 63) / 64), WCS_GROUPS);
        hipLaunchKernelGGL(k_wcolsum1, g1, dim3(256), 0, st, nq, ne2, (i64)1, np, (const double *)nullptr, Pa, o1);
        hipLaunchKernelGGL(k_wcolsum2, dim3(nblocks(ne2, 256)), dim3(256), 0, st, ne, o1, d_f);
        hipLaunchKernelGGL(k_wcolsum1, g1, dim3(256), 0, st, nq, ne2, (i64)1, np, (const double *)nullptr, Pb, o1);
        hipLaunchKernelGGL(k_wcolsum2, dim3(nblocks(ne2, 256)), dim3(256), 0, st, ne, o1, d_f2);
    }
    // which 3: the energy from the rows only (FE's energy is FP's)
    hipLaunchKernelGGL(k_rxchf3_fin, dim3(1), dim3(64), 0, st, which == 0 ? 3 : (int)which, ne, np, d_dbe, d_dp, R, Ev,
                       d_f, d_e);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

static int rxcuhf_omg3_host(int which, i64 ne, i64 np, const double *dal, const double *dbt, const double *dto,
                            const double *dbe, const double *dp, const double *g1, const double *g2, double *f,
                            double *f2, double *e)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    const i64 ng3 = ne * ne * ne * ne * ne * ne * np * np, nf = which == 1 ? np * np : ne * ne;
    DevBuf dL, dBt, dT, dB, dP, d1, d2, dF, dF2, dE;
    rc = h2d(dL, dal, ne * ne);
    if (!rc) rc = h2d(dBt, dbt, ne * ne);
    if (!rc) rc = h2d(dT, dto, ne * ne);
    if (!rc) rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(d1, g1, ng3);
    if (!rc) rc = h2d(d2, g2, ng3);
    if (!rc) rc = dF.alloc(nf * 8);
    if (!rc) rc = dF2.alloc(nf * 8);
    if (!rc) rc = dE.alloc(8);
    if (rc) return rc;
    HIPCHK(hipMemcpy(dF.p, f, nf * 8, hipMemcpyHostToDevice));
    if (which == 0) HIPCHK(hipMemcpy(dF2.p, f2, nf * 8, hipMemcpyHostToDevice));
    rc = neorx_rxcuhf_omg3_dev(which, ne, np, dL.d(), dBt.d(), dT.d(), dB.d(), dP.d(), d1.d(), d2.d(), dF.d(), dF2.d(),
                               dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    if (which == 0) HIPCHK(hipMemcpy(f2, dF2.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

extern "C" int neorx_rxcuhf_fe_omg3(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet,
                                    const double *datot, const double *dbe, const double *dp, const double *gm3_1,
                                    const double *gm3_2, double *faalp, double *fabet, double *e)
{
    if (!fabet) return fail(NEORX_E_ARG, "null argument");
    return rxcuhf_omg3_host(0, nebf, npbf, daalp, dabet, datot, dbe, dp, gm3_1, gm3_2, faalp, fabet, e);
}
extern "C" int neorx_rxcuhf_fp_omg3(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet,
                                    const double *datot, const double *dbe, const double *dp, const double *gm3_1,
                                    const double *gm3_2, double *fp, double *e)
{
    return rxcuhf_omg3_host(1, nebf, npbf, daalp, dabet, datot, dbe, dp, gm3_1, gm3_2, fp, nullptr, e);
}
extern "C" int neorx_rxcuhf_fbe_omg3(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet,
                                     const double *datot, const double *dbe, const double *dp, const double *gm3_1,
                                     const double *gm3_2, double *fbe, double *e)
{
    return rxcuhf_omg3_host(2, nebf, npbf, daalp, dabet, datot, dbe, dp, gm3_1, gm3_2, fbe, nullptr, e);
}

// ------------------------------------------------------------------
// RXCHFmult (rxchfmult/): more than one special electron.  One system handle
// holds both electron bases: primitives [0, npebf) and contracted functions
// [0, nebf) are the regular basis, the special-electron basis follows them.
// Tensors use the RXCHFmult packings (rxchfmult/rxchfmult_basis.f:2-134):
// C row-major [ip][jp][ie1][je1] over the regular basis, then [ie_k][je_k]
// over the special basis for every further electron.
// ------------------------------------------------------------------
extern "C" int neorx_mult_system_create(
    int64_t nebf, int64_t npebf, int64_t nebfbe, int64_t npebfbe, int64_t npbf, int64_t nat, int64_t ngtg, double pmass,
    const double *cat, const double *zan, const double *bcoef, const double *gamma, const int64_t *kpestr,
    const int64_t *kpeend, const int64_t *ampeb2c, const double *agebfcc, const double *elcex, const int64_t *elcam,
    const double *elcbfc, const int64_t *kpestr_be, const int64_t *kpeend_be, const int64_t *ampeb2c_be,
    const double *agebfcc_be, const double *elcex_be, const int64_t *elcam_be, const double *elcbfc_be,
    const double *agnbfcc, const double *nucex, const int64_t *nucam, const double *nucbfc, int64_t *handle)
{
    if (nebf <= 0 || npebf <= 0 || nebfbe <= 0 || npebfbe <= 0)
        return fail(NEORX_E_ARG, "bad basis sizes nebf=%lld npebf=%lld nebfBE=%lld npebfBE=%lld", (long long)nebf,
                    (long long)npebf, (long long)nebfbe, (long long)npebfbe);
    i64 n = npebf + npebfbe, nc = nebf + nebfbe;
    std::vector<int64_t> kps(nc), kpe(nc), amp(n), am(3 * n);
    std::vector<double> cc(n), ex(n), cen(3 * n);
    for (i64 c = 0; c < nebf; ++c) {
        kps[c] = kpestr[c];
        kpe[c] = kpeend[c];
    }
    for (i64 c = 0; c < nebfbe; ++c) {
        kps[nebf + c] = kpestr_be[c] + npebf;
        kpe[nebf + c] = kpeend_be[c] + npebf;
    }
    for (i64 i = 0; i < npebf; ++i) {
        amp[i] = ampeb2c[i];
        cc[i] = agebfcc[i];
        ex[i] = elcex[i];
        for (int d = 0; d < 3; ++d) {
            am[d * n + i] = elcam[d * npebf + i];
            cen[d * n + i] = elcbfc[d * npebf + i];
        }
    }
    for (i64 i = 0; i < npebfbe; ++i) {
        amp[npebf + i] = ampeb2c_be[i] + nebf;
        cc[npebf + i] = agebfcc_be[i];
        ex[npebf + i] = elcex_be[i];
        for (int d = 0; d < 3; ++d) {
            am[d * n + npebf + i] = elcam_be[d * npebfbe + i];
            cen[d * n + npebf + i] = elcbfc_be[d * npebfbe + i];
        }
    }
    int rc = neorx_system_create(nc, n, npbf, nat, ngtg, pmass, cat, zan, bcoef, gamma, kps.data(), kpe.data(),
                                 amp.data(), cc.data(), agnbfcc, ex.data(), nucex, am.data(), nucam, cen.data(),
                                 nucbfc, handle);
    if (rc) return rc;
    System *s = get_sys(*handle);
    s->d.nebf1 = nebf;
    s->d.npebf1 = npebf;
    return NEORX_OK;
}

static System *get_mult(int64_t handle, int &rc)
{
    System *s = get_sys(handle);
    rc = NEORX_OK;
    if (!s) rc = fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    else if (s->d.nebf1 == s->d.nebf) rc = fail(NEORX_E_ARG, "handle %lld has no special-electron basis "
                                                "(create it with neorx_mult_system_create)", (long long)handle);
    return rc ? nullptr : s;
}

// contract_omega2 (rxchfmult_gam2.f:413-626): lane per contracted
// (L = ip*npbf + jp, i1, j1, i2, j2); the primitive values ansE1, ansS, ansE2
// (and ansE3 of the exchange form, contract_omega2_convex,
// rxchfmult_gam2ex.f:354-558) of the slab in Y, Ys, Yx, Yy (class 5
// enumeration, from R0); Yy and X3 may be null
__global__ void __launch_bounds__(256) k_mcontract2(DevSys S, i64 L0, i64 n, i64 R0, const double *__restrict__ Y,
                                                    const double *__restrict__ Ys, const double *__restrict__ Yx,
                                                    const double *__restrict__ Yy, double *__restrict__ X1,
                                                    double *__restrict__ Xs, double *__restrict__ X2,
                                                    double *__restrict__ X3)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    i64 na = S.nebf1, nb = S.nebf - na, pa = S.npebf1, pb = S.npebf - pa, np = S.npbf;
    i64 blk = na * na * nb * nb, pblk = pa * pa * pb * pb;
    i64 L = L0 + t / blk, rem = t % blk;
    i64 j2 = rem % nb, i2 = (rem / nb) % nb, j1 = (rem / (nb * nb)) % na, i1 = rem / (nb * nb * na);
    i64 ip = L / np, jp = L % np;
    double cp = S.agnbfcc[ip] * S.agnbfcc[jp];
    double o1 = 0.0, o2 = 0.0, o3 = 0.0, os = 0.0;
    for (i64 a1 = S.kpestr[i1] - 1; a1 < S.kpeend[i1]; ++a1)
        for (i64 c1 = S.kpestr[j1] - 1; c1 < S.kpeend[j1]; ++c1)
            for (i64 a2 = S.kpestr[na + i2] - 1; a2 < S.kpeend[na + i2]; ++a2)
                for (i64 c2 = S.kpestr[na + j2] - 1; c2 < S.kpeend[na + j2]; ++c2) {
                    i64 r = L * pblk + ((a1 * pa + c1) * pb + (a2 - pa)) * pb + (c2 - pa) - R0;
                    double cf = cp * S.agebfcc[a1] * S.agebfcc[c1] * S.agebfcc[a2] * S.agebfcc[c2];
                    o1 += Y[r] * cf;
                    o2 += Yx[r] * cf;
                    os += Ys[r] * cf;
                    if (Yy) o3 += Yy[r] * cf;
                }
    i64 o = L * blk + rem;
    X1[o] = o1;
    X2[o] = o2;
    Xs[o] = os;
    if (X3) X3[o] = o3;
}

// XCHF_GAM2 = (x12 + x21)/2 over special-basis indices (rxchfmult_gam2.f:173-228):
// x12 at (ip,jp,i1,j1,i2,j2), x21 at (ip,jp,i2,j2,i1,j1) of the unsymmetrised
// two-basis tensor, the special indices taken as regular-basis ones.
__global__ void __launch_bounds__(256) k_mult_symm2(i64 na, i64 nb, i64 n, const double *__restrict__ X,
                                                    const double *__restrict__ Xs, double *__restrict__ G,
                                                    double *__restrict__ Gs)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 j2 = t % nb, i2 = (t / nb) % nb, j1 = (t / (nb * nb)) % nb, i1 = (t / (nb * nb * nb)) % nb,
            pp = t / (nb * nb * nb * nb);
        i64 a = ((pp * na + i1) * na + j1) * nb * nb + i2 * nb + j2;
        i64 b = ((pp * na + i2) * na + j2) * nb * nb + i1 * nb + j1;
        G[t] = (X[a] + X[b]) / 2.0;
        Gs[t] = (Xs[a] + Xs[b]) / 2.0;
    }
}

// The contracted e-e-p tensors of a two-basis handle: X1, Xs (unsymmetrised
// ansE1, ansS), X2 (ansE2) and, if X3 is given, X3 (ansE3).
static int mult_build2(System *s, double *X1, double *Xs, double *X2, double *X3, hipStream_t st)
{
    i64 na = s->d.nebf1, nb = s->d.nebf - na, np = s->d.npbf;
    i64 cap;
    auto slabs = make_slabs(s, 5, 0, np * np, cap);
    DevBuf Y, Ys, Yx, Yy;
    int rc = Y.alloc(cap * 8);
    if (!rc) rc = Ys.alloc(cap * 8);
    if (!rc) rc = Yx.alloc(cap * 8);
    if (!rc && X3) rc = Yy.alloc(cap * 8);
    if (rc) return rc;
    i64 blk = na * na * nb * nb;
    for (auto &sl : slabs) {
        i64 R0 = slab_start(s, 5, sl.first), R1 = slab_start(s, 5, sl.second);
        rc = launch_prims<5, SRC_ORIG>(s->d, R0, R1 - R0, nullptr, Y.d(), Ys.d(), st, Yx.d(), X3 ? Yy.d() : nullptr);
        if (rc) return rc;
        i64 n = (sl.second - sl.first) * blk;
        hipLaunchKernelGGL(k_mcontract2, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y.d(),
                           Ys.d(), Yx.d(), X3 ? Yy.d() : nullptr, X1, Xs, X2, X3);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st)); // slab buffers are released on return
    return NEORX_OK;
}

extern "C" int neorx_rxchfmult_gam2_dev(int64_t handle, double *d_int_gam2, double *d_xchf_gam2,
                                        double *d_xchf_gam2s, void *stream)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    if (!d_int_gam2 || !d_xchf_gam2 || !d_xchf_gam2s) return fail(NEORX_E_ARG, "null output");
    i64 na = s->d.nebf1, nb = s->d.nebf - na, np = s->d.npbf;
    if (nb > na) return fail(NEORX_E_ARG, "nebfBE=%lld > nebf=%lld: XCHF_GAM2 reads regular-basis slots with "
                             "special-basis indices", (long long)nb, (long long)na);
    i64 nint = np * np * na * na * nb * nb, nx = np * np * nb * nb * nb * nb;
    hipStream_t st = (hipStream_t)stream;
    DevBuf X1, Xs;
    rc = X1.alloc(nint * 8);
    if (!rc) rc = Xs.alloc(nint * 8);
    if (!rc) rc = mult_build2(s, X1.d(), Xs.d(), d_int_gam2, nullptr, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mult_symm2, dim3(grid_blocks(nx, 256)), dim3(256), 0, st, na, nb, nx, X1.d(), Xs.d(),
                       d_xchf_gam2, d_xchf_gam2s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

// RXCHFmult_GAM2_IC1ex (rxchfmult_gam2ex.f:2-229, symmetrised at :160-186),
// one basis (a two-basis handle whose two bases are the same, nebfBE = nebf):
// GM2_1ICR, GM2_3ICR, GM2sICR = (x12 + x21)/2 of ansE1, ansE3, ansS;
// GM2_2ICR = ansE2 as contracted
extern "C" int neorx_rxchfmult_gam2_ic1ex_dev(int64_t handle, double *d_gm2_1, double *d_gm2_2, double *d_gm2_3,
                                              double *d_gm2s, void *stream)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    if (!d_gm2_1 || !d_gm2_2 || !d_gm2_3 || !d_gm2s) return fail(NEORX_E_ARG, "null output");
    i64 na = s->d.nebf1, nb = s->d.nebf - na, np = s->d.npbf;
    if (nb != na) return fail(NEORX_E_ARG, "RXCHFmult_GAM2_IC1ex takes one basis: nebfBE=%lld != nebf=%lld",
                              (long long)nb, (long long)na);
    i64 ng2 = np * np * na * na * na * na;
    hipStream_t st = (hipStream_t)stream;
    DevBuf X1, Xs, X3;
    rc = X1.alloc(ng2 * 8);
    if (!rc) rc = Xs.alloc(ng2 * 8);
    if (!rc) rc = X3.alloc(ng2 * 8);
    if (!rc) rc = mult_build2(s, X1.d(), Xs.d(), d_gm2_2, X3.d(), st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gam2_symm, dim3(grid_blocks(ng2, 256)), dim3(256), 0, st, na, ng2, X1.d(), Xs.d(), d_gm2_1,
                       d_gm2s);
    hipLaunchKernelGGL(k_gam2_symm, dim3(grid_blocks(ng2, 256)), dim3(256), 0, st, na, ng2, X3.d(), X3.d(), d_gm2_3,
                       d_gm2_3);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_rxchfmult_gam2_ic1ex(int64_t handle, double *gm2_1, double *gm2_2, double *gm2_3, double *gm2s)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    i64 na = s->d.nebf1, np = s->d.npbf, ng2 = np * np * na * na * na * na;
    DevBuf A, B, C, D;
    rc = A.alloc(ng2 * 8);
    if (!rc) rc = B.alloc(ng2 * 8);
    if (!rc) rc = C.alloc(ng2 * 8);
    if (!rc) rc = D.alloc(ng2 * 8);
    if (!rc) rc = neorx_rxchfmult_gam2_ic1ex_dev(handle, A.d(), B.d(), C.d(), D.d(), 0);
    if (!rc) rc = copy_out(gm2_1, A.d(), ng2);
    if (!rc) rc = copy_out(gm2_2, B.d(), ng2);
    if (!rc) rc = copy_out(gm2_3, C.d(), ng2);
    if (!rc) rc = copy_out(gm2s, D.d(), ng2);
    return rc;
}

extern "C" int neorx_rxchfmult_gam2(int64_t handle, double *int_gam2, double *xchf_gam2, double *xchf_gam2s)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    i64 na = s->d.nebf1, nb = s->d.nebf - na, np = s->d.npbf;
    i64 nint = np * np * na * na * nb * nb, nx = np * np * nb * nb * nb * nb;
    DevBuf A, B, C;
    rc = A.alloc(nint * 8);
    if (!rc) rc = B.alloc(nx * 8);
    if (!rc) rc = C.alloc(nx * 8);
    if (!rc) rc = neorx_rxchfmult_gam2_dev(handle, A.d(), B.d(), C.d(), 0);
    if (!rc) rc = copy_out(int_gam2, A.d(), nint);
    if (!rc) rc = copy_out(xchf_gam2, B.d(), nx);
    if (!rc) rc = copy_out(xchf_gam2s, C.d(), nx);
    return rc;
}

// contract_omega3_INT (rxchfmult_gam3-int.f:417-660): lane per contracted
// (L = (ip*npbf + jp)*nebf + i1, j1, i2, j2, i3, j3); with Ys, Yx given also
// ans3, ans4 of the exchange form (contract_omega3_convex,
// rxchfmult_gam3ex.f:459-711) into X3, X4
__global__ void __launch_bounds__(256) k_mcontract3(DevSys S, i64 L0, i64 n, i64 R0, const double *__restrict__ Y,
                                                    const double *__restrict__ Ys, const double *__restrict__ Yx,
                                                    double *__restrict__ X, double *__restrict__ X3,
                                                    double *__restrict__ X4)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    i64 na = S.nebf1, nb = S.nebf - na, pa = S.npebf1, pb = S.npebf - pa, np = S.npbf;
    i64 nb2 = nb * nb, blk = na * nb2 * nb2, pb4 = pb * pb * pb * pb;
    i64 L = L0 + t / blk, rem = t % blk;
    i64 j3 = rem % nb, i3 = (rem / nb) % nb, j2 = (rem / nb2) % nb, i2 = (rem / (nb2 * nb)) % nb,
        j1 = rem / (nb2 * nb2);
    i64 i1 = L % na, pp = L / na, ip = pp / np, jp = pp % np;
    double cp = S.agnbfcc[ip] * S.agnbfcc[jp];
    double om = 0.0, o3 = 0.0, o4 = 0.0;
    for (i64 a1 = S.kpestr[i1] - 1; a1 < S.kpeend[i1]; ++a1)
        for (i64 c1 = S.kpestr[j1] - 1; c1 < S.kpeend[j1]; ++c1)
            for (i64 a2 = S.kpestr[na + i2] - 1; a2 < S.kpeend[na + i2]; ++a2)
                for (i64 c2 = S.kpestr[na + j2] - 1; c2 < S.kpeend[na + j2]; ++c2)
                    for (i64 a3 = S.kpestr[na + i3] - 1; a3 < S.kpeend[na + i3]; ++a3)
                        for (i64 c3 = S.kpestr[na + j3] - 1; c3 < S.kpeend[na + j3]; ++c3) {
                            i64 r = (pp * pa + a1) * pa * pb4 +
                                    (((c1 * pb + (a2 - pa)) * pb + (c2 - pa)) * pb + (a3 - pa)) * pb + (c3 - pa) - R0;
                            double cf = cp * S.agebfcc[a1] * S.agebfcc[c1] * S.agebfcc[a2] * S.agebfcc[c2] *
                                        S.agebfcc[a3] * S.agebfcc[c3];
                            om += Y[r] * cf;
                            if (Ys) {
                                o3 += Ys[r] * cf;
                                o4 += Yx[r] * cf;
                            }
                        }
    X[L * blk + rem] = om;
    if (X3) {
        X3[L * blk + rem] = o3;
        X4[L * blk + rem] = o4;
    }
}

// INT_GAM3 = (x123 + x132)/2, electrons 2 and 3 exchanged (rxchfmult_gam3-int.f:173-218)
__global__ void __launch_bounds__(256) k_mult_symm3(i64 nb, i64 n, const double *__restrict__ X,
                                                    double *__restrict__ G)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 nb2 = nb * nb;
        i64 p3 = t % nb2, p2 = (t / nb2) % nb2, head = t / (nb2 * nb2);
        G[t] = (X[t] + X[(head * nb2 + p3) * nb2 + p2]) * 0.5;
    }
}

// one basis (ne): (x123 + x213)/2 (swap = 1) or (x123 + x321)/2 (swap = 2),
// rxchfmult_gam3ex.f:248-269
__global__ void __launch_bounds__(256) k_mult_symm3x(i64 ne, i64 n, int swap, const double *__restrict__ X,
                                                     double *__restrict__ G)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 ne2 = ne * ne;
        i64 p3 = t % ne2, p2 = (t / ne2) % ne2, p1 = (t / (ne2 * ne2)) % ne2, pp = t / (ne2 * ne2 * ne2);
        i64 o = swap == 1 ? ((pp * ne2 + p2) * ne2 + p1) * ne2 + p3 : ((pp * ne2 + p3) * ne2 + p2) * ne2 + p1;
        G[t] = (X[t] + X[o]) * 0.5;
    }
}

extern "C" int neorx_rxchfmult_gam3_int_dev(int64_t handle, double *d_int_gam3, void *stream)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    if (!d_int_gam3) return fail(NEORX_E_ARG, "null output");
    i64 na = s->d.nebf1, nb = s->d.nebf - na, np = s->d.npbf;
    i64 nb2 = nb * nb, blk = na * nb2 * nb2, n3 = np * np * na * blk;
    hipStream_t st = (hipStream_t)stream;
    i64 cap;
    auto slabs = make_slabs(s, 6, 0, np * np * na, cap);
    DevBuf X, Y;
    rc = X.alloc(n3 * 8);
    if (!rc) rc = Y.alloc(cap * 8);
    if (rc) return rc;
    for (auto &sl : slabs) {
        i64 R0 = slab_start(s, 6, sl.first), R1 = slab_start(s, 6, sl.second);
        rc = launch_prims<6, SRC_ORIG>(s->d, R0, R1 - R0, nullptr, Y.d(), nullptr, st);
        if (rc) return rc;
        i64 n = (sl.second - sl.first) * blk;
        hipLaunchKernelGGL(k_mcontract3, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y.d(),
                           nullptr, nullptr, X.d(), nullptr, nullptr);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mult_symm3, dim3(grid_blocks(n3, 256)), dim3(256), 0, st, nb, n3, X.d(), d_int_gam3);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

// RXCHFmult_GAM3_IC1ex (rxchfmult_gam3ex.f:2-316), one basis (nebfBE = nebf):
// INT_GAM3 = (x123 + x132)/2 of ans, INT_GAM3ex1 = (x123 + x213)/2 of
// ans3 = 2 gVEEg2, INT_GAM3ex2 = (x123 + x321)/2 of ans4 = gVEEg1
extern "C" int neorx_rxchfmult_gam3_ic1ex_dev(int64_t handle, double *d_int_gam3, double *d_ex1, double *d_ex2,
                                              void *stream)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    if (!d_int_gam3 || !d_ex1 || !d_ex2) return fail(NEORX_E_ARG, "null output");
    i64 na = s->d.nebf1, nb = s->d.nebf - na, np = s->d.npbf;
    if (nb != na) return fail(NEORX_E_ARG, "RXCHFmult_GAM3_IC1ex takes one basis: nebfBE=%lld != nebf=%lld",
                              (long long)nb, (long long)na);
    i64 nb2 = nb * nb, blk = na * nb2 * nb2, n3 = np * np * na * blk;
    hipStream_t st = (hipStream_t)stream;
    i64 cap;
    auto slabs = make_slabs(s, 6, 0, np * np * na, cap);
    DevBuf X, X3, X4, Y, Ys, Yx;
    rc = X.alloc(n3 * 8);
    if (!rc) rc = X3.alloc(n3 * 8);
    if (!rc) rc = X4.alloc(n3 * 8);
    if (!rc) rc = Y.alloc(cap * 8);
    if (!rc) rc = Ys.alloc(cap * 8);
    if (!rc) rc = Yx.alloc(cap * 8);
    if (rc) return rc;
    for (auto &sl : slabs) {
        i64 R0 = slab_start(s, 6, sl.first), R1 = slab_start(s, 6, sl.second);
        rc = launch_prims<6, SRC_ORIG>(s->d, R0, R1 - R0, nullptr, Y.d(), Ys.d(), st, Yx.d());
        if (rc) return rc;
        i64 n = (sl.second - sl.first) * blk;
        hipLaunchKernelGGL(k_mcontract3, dim3(nblocks(n, 256)), dim3(256), 0, st, s->d, sl.first, n, R0, Y.d(),
                           Ys.d(), Yx.d(), X.d(), X3.d(), X4.d());
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mult_symm3, dim3(grid_blocks(n3, 256)), dim3(256), 0, st, nb, n3, X.d(), d_int_gam3);
    hipLaunchKernelGGL(k_mult_symm3x, dim3(grid_blocks(n3, 256)), dim3(256), 0, st, na, n3, 1, X3.d(), d_ex1);
    hipLaunchKernelGGL(k_mult_symm3x, dim3(grid_blocks(n3, 256)), dim3(256), 0, st, na, n3, 2, X4.d(), d_ex2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

extern "C" int neorx_rxchfmult_gam3_ic1ex(int64_t handle, double *int_gam3, double *ex1, double *ex2)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    i64 na = s->d.nebf1, np = s->d.npbf, n3 = np * np * na * na * na * na * na * na;
    DevBuf A, B, C;
    rc = A.alloc(n3 * 8);
    if (!rc) rc = B.alloc(n3 * 8);
    if (!rc) rc = C.alloc(n3 * 8);
    if (!rc) rc = neorx_rxchfmult_gam3_ic1ex_dev(handle, A.d(), B.d(), C.d(), 0);
    if (!rc) rc = copy_out(int_gam3, A.d(), n3);
    if (!rc) rc = copy_out(ex1, B.d(), n3);
    if (!rc) rc = copy_out(ex2, C.d(), n3);
    return rc;
}

extern "C" int neorx_rxchfmult_gam3_int(int64_t handle, double *int_gam3)
{
    int rc;
    System *s = get_mult(handle, rc);
    if (!s) return rc;
    i64 na = s->d.nebf1, nb = s->d.nebf - na, np = s->d.npbf;
    i64 n3 = np * np * na * na * nb * nb * nb * nb;
    DevBuf A;
    rc = A.alloc(n3 * 8);
    if (!rc) rc = neorx_rxchfmult_gam3_int_dev(handle, A.d(), 0);
    if (!rc) rc = copy_out(int_gam3, A.d(), n3);
    return rc;
}

// RXCHFmult_GAM1_MD (rxchfmult/rxchfmult_gam1.f:2-398) over the basis of a
// one-basis handle (the driver passes the special-electron basis,
// rxchfmult_driver.f:324-332)
extern "C" int neorx_rxchfmult_gam1(int64_t handle, double *gam1, double *gam1s)
{
    return gam1_host<8>(handle, gam1, gam1s);
}

// RXCHFmult_GAM3_XCHF (rxchfmult/rxchfmult_gam3-xchf.f:2-268) over the basis
// of a one-basis handle: contract_omega3_XCHF, then the six-permutation average
extern "C" int neorx_rxchfmult_gam3_xchf_dev(int64_t handle, double *d_gm3, void *stream)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (s->d.nebf1 != s->d.nebf) return fail(NEORX_E_ARG, "RXCHFmult_GAM3_XCHF takes a one-basis handle");
    if (!d_gm3) return fail(NEORX_E_ARG, "null output");
    return gam3_build_dev<7>(s, d_gm3, stream);
}

extern "C" int neorx_rxchfmult_gam3_xchf(int64_t handle, double *gm3)
{
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    i64 ne = s->d.nebf, np = s->d.npbf, ng3 = ne * ne * ne * ne * ne * ne * np * np;
    DevBuf A;
    int rc = A.alloc(ng3 * 8);
    if (!rc) rc = neorx_rxchfmult_gam3_xchf_dev(handle, A.d(), 0);
    if (!rc) rc = copy_out(gm3, A.d(), ng3);
    return rc;
}

// RXCHFmult_symm_gam4 / RXCHFmult_calc_GAM_4_integral (rxchfmult_gam4-int.f:290-486):
// lane per INT_GAM4 entry; electron 1 fixed, kets of electrons 2..4 permuted
// with weights 1, -1/2, -1/2, 1/4, 1/4, -1/2.  Each term is
// (ee14 gs23 + ee13 gs24 + ee12 gs34)/3 from GAM_ee over the regular basis
// (special indices taken as regular ones) and GAM_2s over the special basis.
__global__ void __launch_bounds__(256) k_mult_gam4(i64 na, i64 nb, i64 np, i64 off, i64 n,
                                                   const double *__restrict__ G2s, const double *__restrict__ Gee,
                                                   double *__restrict__ G4)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    i64 r = off + t;
    i64 br[4], kt[4];
    for (int k = 3; k >= 1; --k) {
        kt[k] = r % nb; r /= nb;
        br[k] = r % nb; r /= nb;
    }
    kt[0] = r % na; r /= na;
    br[0] = r % na; r /= na;
    i64 jp = r % np, ip = r / np;
    auto ee = [&](i64 x, const i64 *k) {
        double v = Gee[((br[0] * na + k[0]) * na + br[x]) * na + k[x]];
        underflow(v);
        return v;
    };
    auto gs = [&](i64 x, i64 y, const i64 *k) {
        double v = G2s[((((ip * np + jp) * nb + br[x]) * nb + k[x]) * nb + br[y]) * nb + k[y]];
        underflow(v);
        return v;
    };
    auto term = [&](i64 a, i64 b, i64 c) {
        i64 k[4] = {kt[0], kt[a], kt[b], kt[c]};
        double x1 = ee(3, k) * gs(1, 2, k), x2 = ee(2, k) * gs(1, 3, k), x3 = ee(1, k) * gs(2, 3, k);
        return (x1 + x2 + x3) / 3.0;
    };
    double ans = term(1, 2, 3) - term(1, 3, 2) / 2.0 - term(2, 1, 3) / 2.0 + term(2, 3, 1) / 4.0 +
                 term(3, 1, 2) / 4.0 - term(3, 2, 1) / 2.0;
    G4[off + t] = ans;
}

extern "C" int neorx_rxchfmult_gam4_int_dev(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *d_gam_2s,
                                            const double *d_gam_ee, double *d_gam4, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (nebf <= 0 || nebfbe <= 0 || npbf <= 0) return fail(NEORX_E_ARG, "bad sizes");
    if (nebfbe > nebf) return fail(NEORX_E_ARG, "nebfBE=%lld > nebf=%lld: GAM_ee is read with special-basis indices",
                                   (long long)nebfbe, (long long)nebf);
    if (!d_gam_2s || !d_gam_ee || !d_gam4) return fail(NEORX_E_ARG, "null argument");
    i64 na = nebf, nb = nebfbe, np = npbf, nb2 = nb * nb;
    i64 n4 = np * np * na * na * nb2 * nb2 * nb2;
    hipStream_t st = (hipStream_t)stream;
    for (i64 off = 0; off < n4; off += CHUNK) {
        i64 n = std::min(CHUNK, n4 - off);
        hipLaunchKernelGGL(k_mult_gam4, dim3(nblocks(n, 256)), dim3(256), 0, st, na, nb, np, off, n, d_gam_2s,
                           d_gam_ee, d_gam4);
        HIPCHK(hipGetLastError());
    }
    return NEORX_OK;
}

extern "C" int neorx_rxchfmult_gam4_int(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *gam_2s,
                                        const double *gam_ee, double *gam4)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (nebf <= 0 || nebfbe <= 0 || npbf <= 0) return fail(NEORX_E_ARG, "bad sizes");
    i64 na = nebf, nb = nebfbe, np = npbf, nb2 = nb * nb;
    i64 n4 = np * np * na * na * nb2 * nb2 * nb2;
    DevBuf G2, Ge, G4;
    rc = h2d(G2, gam_2s, np * np * nb2 * nb2);
    if (!rc) rc = h2d(Ge, gam_ee, na * na * na * na);
    if (!rc) rc = G4.alloc(n4 * 8);
    if (!rc) rc = neorx_rxchfmult_gam4_int_dev(nebf, nebfbe, npbf, G2.d(), Ge.d(), G4.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    return copy_out(gam4, G4.d(), n4);
}

// ---- RXCHFmult Fock contractions (rxchfmult_omg{2,3,4}.f) --------------------
// A tensor of order k (k = 1, 2, 3 special electrons besides electron 1) is a
// row-major matrix G[R][c]: rows R = (ip,jp,i1,j1), columns c = the special
// indices.  One pass over G gives
//   rowsum[R] = sum_c W(c) G[R][c]                    (FAE, FP and E)
//   HH[c]     = sum_R DP(ip,jp) DAE(i1,j1) G[R][c]    (FBE)
// with column weights W = DBE(i2,j2) (k=1); DBE(p,q)DBE(r,s) - 1/2 DBE(p,s)DBE(r,q)
// at c = (p,q,r,s) (k=2: the exchanged read val = G(..i2 j3 i3 j2)/2 of
// thread_F*_OMG3, rxchfmult_omg3.f:187, moved onto the weight); the product
// of three DBE (k=3).
__global__ void __launch_bounds__(256) k_mult_weights(int k, i64 nb, i64 nc, const double *__restrict__ DBE,
                                                      double *__restrict__ W)
{
    i64 c = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    auto D = [&](i64 i, i64 j) { return DBE[j * nb + i]; }; // DBE(i,j)
    double w;
    if (k == 1) {
        w = D(c / nb, c % nb);
    } else if (k == 2) {
        i64 s = c % nb, r = (c / nb) % nb, q = (c / (nb * nb)) % nb, p = c / (nb * nb * nb);
        w = D(p, q) * D(r, s) - 0.5 * D(p, s) * D(r, q);
    } else if (k == 4) { // DBE(i2,j2) DBE(i3,j3): the exchange forms, exchange already in the tensor
        i64 c2 = c % (nb * nb), c1 = c / (nb * nb);
        w = D(c1 / nb, c1 % nb) * D(c2 / nb, c2 % nb);
    } else {
        i64 nb2 = nb * nb, c3 = c % nb2, c2 = (c / nb2) % nb2, c1 = c / (nb2 * nb2);
        w = D(c1 / nb, c1 % nb) * D(c2 / nb, c2 % nb) * D(c3 / nb, c3 % nb);
    }
    W[c] = w;
}

static const i64 MULT_ROWS = 64; // rows of G per block of k_mult_rows

// block = (run of MULT_ROWS rows, column tile of ntile); lane per column.
// Fixed-order sums: a row's sum over the tile's columns goes to
// rowpart[tile][row] (wave_sum, then the waves' slots in order), a column's
// sum over the block's rows to hhpart[row run][column]; k_sum_parts adds the
// partials in order.
__global__ void __launch_bounds__(256) k_mult_rows(i64 na, i64 np, i64 nrows, i64 ncols, i64 ntile,
                                                   const double *__restrict__ DAE, const double *__restrict__ DP,
                                                   const double *__restrict__ W, const double *__restrict__ G,
                                                   double *__restrict__ rowpart, double *__restrict__ hhpart)
{
    __shared__ double rsw[4][MULT_ROWS];
    const i64 tile = (i64)blockIdx.x % ntile, grp = (i64)blockIdx.x / ntile;
    i64 c = tile * blockDim.x + threadIdx.x;
    i64 R0 = grp * MULT_ROWS, nr = std::min<i64>(MULT_ROWS, nrows - R0);
    bool on = c < ncols;
    double w = (on && W) ? W[c] : 0.0, h = 0.0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    i64 na2 = na * na;
    for (i64 k = 0; k < nr; ++k) {
        i64 R = R0 + k, pp = R / na2, a = R % na2;
        double g = on ? G[R * ncols + c] : 0.0;
        double wr = DP[(pp % np) * np + pp / np] * DAE[(a % na) * na + a / na];
        h += wr * g;
        if (rowpart) {
            double v = wave_sum(w * g);
            if (lane == 0) rsw[wv][k] = v;
        }
    }
    __syncthreads();
    if (rowpart && threadIdx.x < nr) {
        double t = 0.0;
        for (int q = 0; q < nw; ++q) t += rsw[q][threadIdx.x];
        rowpart[tile * nrows + R0 + threadIdx.x] = t;
    }
    if (hhpart && on) hhpart[grp * ncols + c] = h;
}

// out[i] = sum over p of part[p][i], p in order
__global__ void k_sum_parts(i64 n, i64 nparts, const double *__restrict__ part, double *__restrict__ out)
{
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = 0.0;
    for (i64 p = 0; p < nparts; ++p) t += part[p * n + i];
    out[i] = t;
}

// FAE(i1,j1) += sum_pp DP rowsum;  FP(ip,jp) += sum_a DAE rowsum;
// FBE(i2,j2) += k * (HH contracted with DBE over the other special electrons);
// E = sum_pp DP sum_a DAE rowsum
__global__ void __launch_bounds__(256) k_mult_finish(int k, i64 na, i64 nb, i64 np, const double *__restrict__ DAE,
                                                     const double *__restrict__ DBE, const double *__restrict__ DP,
                                                     const double *__restrict__ rowsum, const double *__restrict__ HH,
                                                     double *__restrict__ FAE, double *__restrict__ FP,
                                                     double *__restrict__ FBE, double *__restrict__ E)
{
    i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    i64 na2 = na * na, np2 = np * np, nb2 = nb * nb;
    auto dp = [&](i64 pp) { return DP[(pp % np) * np + pp / np]; };
    auto dae = [&](i64 a) { return DAE[(a % na) * na + a / na]; };
    auto dbe = [&](i64 i, i64 j) { return DBE[j * nb + i]; };
    auto fp_of = [&](i64 pp) {
        double f = 0.0;
        for (i64 a = 0; a < na2; ++a) f += dae(a) * rowsum[pp * na2 + a];
        return f;
    };
    if (FAE && t < na2) {
        double f = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) f += dp(pp) * rowsum[pp * na2 + t];
        FAE[(t % na) * na + t / na] += f;
    }
    if (FP && t < np2) FP[(t % np) * np + t / np] += fp_of(t);
    if (FBE && t < nb2) {
        i64 i2 = t / nb, j2 = t % nb;
        double f = 0.0;
        if (k == 1) {
            f = HH[t];
        } else if (k == 2) {
            for (i64 i3 = 0; i3 < nb; ++i3)
                for (i64 j3 = 0; j3 < nb; ++j3)
                    f += dbe(i3, j3) * (HH[t * nb2 + i3 * nb + j3] - 0.5 * HH[(i2 * nb + j3) * nb2 + i3 * nb + j2]);
            f *= 2.0;
        } else if (k == 3) {
            for (i64 c = 0; c < nb2 * nb2; ++c) {
                i64 c3 = c % nb2, c2 = c / nb2;
                f += dbe(c2 / nb, c2 % nb) * dbe(c3 / nb, c3 % nb) * HH[t * nb2 * nb2 + c];
            }
            f *= 3.0;
        } else { // k == 4: the FBE tensor of the exchange form carries its factors
            for (i64 c = 0; c < nb2; ++c) f += dbe(c / nb, c % nb) * HH[t * nb2 + c];
        }
        FBE[j2 * nb + i2] += f;
    }
    if (t == 0) {
        double e = 0.0;
        for (i64 pp = 0; pp < np2; ++pp) e += dp(pp) * fp_of(pp);
        *E = e;
    }
}

// One pass (two for the exchange form of order 2, where FBE reads its own
// tensor) over the device tensors: mode 1..3 = order k of the plain forms,
// mode 4 = the exchange form of order 2 (G: the FAE/FP/E tensor, K: the FBE
// tensor, both built by k_omg3ex_combine).
static int mult_omg_core(int mode, i64 na, i64 nb, i64 np, const double *d_dae, const double *d_dbe,
                         const double *d_dp, const double *G, const double *K, double *d_fae, double *d_fp,
                         double *d_fbe, double *d_e, hipStream_t st)
{
    int k = mode == 4 ? 2 : mode;
    i64 nrows = np * np * na * na, ncols = 1;
    for (int i = 0; i < k; ++i) ncols *= nb * nb;
    unsigned nt = (unsigned)std::min<i64>(256, ((ncols + 63) / 64) * 64);
    const i64 ntile = (ncols + nt - 1) / nt, ngrp = (nrows + MULT_ROWS - 1) / MULT_ROWS, nblk = ntile * ngrp;
    if (nblk > 0x7fffffff) return fail(NEORX_E_UNSUPPORTED, "tensor too large for one launch");
    // rowsum, W, HH, then the partials of the row and column sums
    StreamBuf ws;
    int rc = ws.alloc((nrows + 2 * ncols + ntile * nrows + (d_fbe ? ngrp * ncols : 0)) * 8, st);
    if (rc) return rc;
    double *rowsum = (double *)ws.p, *W = rowsum + nrows, *HH = d_fbe ? W + ncols : nullptr;
    double *rowpart = W + 2 * ncols, *hhpart = d_fbe ? rowpart + ntile * nrows : nullptr;
    hipLaunchKernelGGL(k_mult_weights, dim3(nblocks(ncols, 256)), dim3(256), 0, st, mode, nb, ncols, d_dbe, W);
    HIPCHK(hipGetLastError());
    bool sep = K != G;
    hipLaunchKernelGGL(k_mult_rows, dim3((unsigned)nblk), dim3(nt), 0, st, na, np, nrows, ncols, ntile, d_dae, d_dp, W,
                       G, rowpart, sep ? nullptr : hhpart);
    if (sep && HH)
        hipLaunchKernelGGL(k_mult_rows, dim3((unsigned)nblk), dim3(nt), 0, st, na, np, nrows, ncols, ntile, d_dae,
                           d_dp, nullptr, K, nullptr, hhpart);
    hipLaunchKernelGGL(k_sum_parts, dim3(nblocks(nrows, 256)), dim3(256), 0, st, nrows, ntile, rowpart, rowsum);
    if (HH) hipLaunchKernelGGL(k_sum_parts, dim3(nblocks(ncols, 256)), dim3(256), 0, st, ncols, ngrp, hhpart, HH);
    HIPCHK(hipGetLastError());
    i64 nf = std::max(std::max(na * na, np * np), nb * nb);
    hipLaunchKernelGGL(k_mult_finish, dim3(nblocks(nf, 256)), dim3(256), 0, st, mode, na, nb, np, d_dae, d_dbe, d_dp,
                       rowsum, HH, d_fae, d_fp, d_fbe, d_e);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

extern "C" int neorx_rxchfmult_omg_dev(int64_t k, int64_t nebf, int64_t nebfbe, int64_t npbf, const double *d_dae,
                                       const double *d_dbe, const double *d_dp, const double *d_gm, double *d_fae,
                                       double *d_fp, double *d_fbe, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (k < 1 || k > 3) return fail(NEORX_E_ARG, "order %lld not in 1..3", (long long)k);
    if (nebf <= 0 || nebfbe <= 0 || npbf <= 0) return fail(NEORX_E_ARG, "bad sizes");
    if (!d_dae || !d_dbe || !d_dp || !d_gm || !d_e) return fail(NEORX_E_ARG, "null argument");
    return mult_omg_core((int)k, nebf, nebfbe, npbf, d_dae, d_dbe, d_dp, d_gm, d_gm, d_fae, d_fp, d_fbe, d_e,
                         (hipStream_t)stream);
}

// thread_F*_OMG2ex (rxchfmult_omg2ex.f:94-180): val = G(i1 j1 i2 j2) - G2ex(i1 j2 i2 j1)/2,
// then the contractions of the plain order-1 form
__global__ void __launch_bounds__(256) k_omg2ex_combine(i64 ne, i64 n, const double *__restrict__ G,
                                                        const double *__restrict__ Gx, double *__restrict__ H)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 j2 = t % ne, i2 = (t / ne) % ne, j1 = (t / (ne * ne)) % ne, i1 = (t / (ne * ne * ne)) % ne,
            pp = t / (ne * ne * ne * ne);
        H[t] = G[t] - 0.5 * Gx[(((pp * ne + i1) * ne + j2) * ne + i2) * ne + j1];
    }
}

// thread_F*_OMG3ex (rxchfmult_omg3ex.f:101-216, :316-431, :531-663): per entry
//   val = G(123) - G(1,i2 j3,i3 j2)/2
//   v2  = X1(i1 j2,i2 j1,i3 j3) - X1(i1 j3,i2 j1,i3 j2)/2 + X2(i1 j3,i2 j2,i3 j1) - X2(i1 j2,i2 j3,i3 j1)/2
//   v3  = X1(i1 j3,i3 j1,i2 j2) - X1(i1 j2,i3 j1,i2 j3)/2 + X2(i1 j2,i3 j3,i2 j1) - X2(i1 j3,i3 j2,i2 j1)/2
//   H = val - v2/2 (FAE, FP and every energy), K = 2 val - (v2 + v3)/2 (FBE)
__global__ void __launch_bounds__(256) k_omg3ex_combine(i64 ne, i64 n, const double *__restrict__ G,
                                                        const double *__restrict__ X1, const double *__restrict__ X2,
                                                        double *__restrict__ H, double *__restrict__ K)
{
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
        i64 r = t;
        i64 j3 = r % ne; r /= ne;
        i64 i3 = r % ne; r /= ne;
        i64 j2 = r % ne; r /= ne;
        i64 i2 = r % ne; r /= ne;
        i64 j1 = r % ne; r /= ne;
        i64 i1 = r % ne, pp = r / ne;
        auto ix = [&](i64 a, i64 b, i64 c, i64 d, i64 e, i64 f) {
            return (((((pp * ne + a) * ne + b) * ne + c) * ne + d) * ne + e) * ne + f;
        };
        double val = G[t] - 0.5 * G[ix(i1, j1, i2, j3, i3, j2)];
        double v2 = X1[ix(i1, j2, i2, j1, i3, j3)] - 0.5 * X1[ix(i1, j3, i2, j1, i3, j2)] + X2[ix(i1, j3, i2, j2, i3, j1)] -
                    0.5 * X2[ix(i1, j2, i2, j3, i3, j1)];
        double v3 = X1[ix(i1, j3, i3, j1, i2, j2)] - 0.5 * X1[ix(i1, j2, i3, j1, i2, j3)] + X2[ix(i1, j2, i3, j3, i2, j1)] -
                    0.5 * X2[ix(i1, j3, i3, j2, i2, j1)];
        H[t] = val - 0.5 * v2;
        K[t] = 2.0 * val - 0.5 * (v2 + v3);
    }
}

// order 2 or 3 (OMG2ex, OMG3ex), one basis; d_gx2 only for order 3
extern "C" int neorx_rxchfmult_omgex_dev(int64_t order, int64_t nebf, int64_t npbf, const double *d_dae,
                                         const double *d_dbe, const double *d_dp, const double *d_gm,
                                         const double *d_gx1, const double *d_gx2, double *d_fae, double *d_fp,
                                         double *d_fbe, double *d_e, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (order != 2 && order != 3) return fail(NEORX_E_ARG, "order %lld not 2 or 3", (long long)order);
    if (nebf <= 0 || npbf <= 0) return fail(NEORX_E_ARG, "bad sizes");
    if (!d_dae || !d_dbe || !d_dp || !d_gm || !d_gx1 || (order == 3 && !d_gx2) || !d_e)
        return fail(NEORX_E_ARG, "null argument");
    i64 ne = nebf, np = npbf, n = np * np * ne * ne * ne * ne * (order == 3 ? ne * ne : 1);
    hipStream_t st = (hipStream_t)stream;
    StreamBuf tb;
    rc = tb.alloc(n * 8 * (order == 3 ? 2 : 1), st);
    if (rc) return rc;
    double *H = (double *)tb.p, *K = order == 3 ? H + n : H;
    if (order == 2)
        hipLaunchKernelGGL(k_omg2ex_combine, dim3(grid_blocks(n, 256)), dim3(256), 0, st, ne, n, d_gm, d_gx1, H);
    else
        hipLaunchKernelGGL(k_omg3ex_combine, dim3(grid_blocks(n, 256)), dim3(256), 0, st, ne, n, d_gm, d_gx1, d_gx2, H,
                           K);
    HIPCHK(hipGetLastError());
    return mult_omg_core(order == 2 ? 1 : 4, ne, ne, np, d_dae, d_dbe, d_dp, H, K, d_fae, d_fp, d_fbe, d_e, st);
}

// host form of one reference routine: which = 0 (FAE), 1 (FP), 2 (FBE)
static int mult_omg_host(int k, int which, i64 na, i64 nb, i64 np, const double *dae, const double *dbe,
                         const double *dp, const double *gm, double *f, double *e)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (na <= 0 || nb <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    i64 ng = np * np * na * na, nf = which == 0 ? na * na : which == 1 ? np * np : nb * nb;
    for (int i = 0; i < k; ++i) ng *= nb * nb;
    DevBuf dA, dB, dP, dG, dF, dE;
    rc = h2d(dA, dae, na * na);
    if (!rc) rc = h2d(dB, dbe, nb * nb);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(dG, gm, ng);
    if (!rc) rc = h2d(dF, f, nf);
    if (!rc) rc = dE.alloc(8);
    if (!rc)
        rc = neorx_rxchfmult_omg_dev(k, na, nb, np, dA.d(), dB.d(), dP.d(), dG.d(), which == 0 ? dF.d() : nullptr,
                                     which == 1 ? dF.d() : nullptr, which == 2 ? dF.d() : nullptr, dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

static int mult_omgex_host(int order, int which, i64 ne, i64 np, const double *dae, const double *dbe, const double *dp,
                           const double *gm, const double *gx1, const double *gx2, double *f, double *e)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (ne <= 0 || np <= 0) return fail(NEORX_E_ARG, "bad sizes");
    i64 ng = np * np * ne * ne * ne * ne * (order == 3 ? ne * ne : 1), nf = which == 1 ? np * np : ne * ne;
    DevBuf dA, dB, dP, dG, dX1, dX2, dF, dE;
    rc = h2d(dA, dae, ne * ne);
    if (!rc) rc = h2d(dB, dbe, ne * ne);
    if (!rc) rc = h2d(dP, dp, np * np);
    if (!rc) rc = h2d(dG, gm, ng);
    if (!rc) rc = h2d(dX1, gx1, ng);
    if (!rc && order == 3) rc = h2d(dX2, gx2, ng);
    if (!rc) rc = h2d(dF, f, nf);
    if (!rc) rc = dE.alloc(8);
    if (!rc)
        rc = neorx_rxchfmult_omgex_dev(order, ne, np, dA.d(), dB.d(), dP.d(), dG.d(), dX1.d(),
                                       order == 3 ? dX2.d() : nullptr, which == 0 ? dF.d() : nullptr,
                                       which == 1 ? dF.d() : nullptr, which == 2 ? dF.d() : nullptr, dE.d(), 0);
    if (rc) return rc;
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(f, dF.p, nf * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(e, dE.p, 8, hipMemcpyDeviceToHost));
    return NEORX_OK;
}

#define NRX_MULT_OMG2EX(name, which)                                                                              \
    extern "C" int name(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,       \
                        const double *gm2, const double *gm2ex, double *f, double *e)                            \
    {                                                                                                             \
        return mult_omgex_host(2, which, nebf, npbf, dae, dbe, dp, gm2, gm2ex, nullptr, f, e);                   \
    }
#define NRX_MULT_OMG3EX(name, which)                                                                              \
    extern "C" int name(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,       \
                        const double *gm3, const double *gm3ex1, const double *gm3ex2, double *f, double *e)     \
    {                                                                                                             \
        return mult_omgex_host(3, which, nebf, npbf, dae, dbe, dp, gm3, gm3ex1, gm3ex2, f, e);                   \
    }
NRX_MULT_OMG2EX(neorx_rxchfmult_fae_omg2ex, 0)
NRX_MULT_OMG2EX(neorx_rxchfmult_fp_omg2ex, 1)
NRX_MULT_OMG2EX(neorx_rxchfmult_fbe_omg2ex, 2)
NRX_MULT_OMG3EX(neorx_rxchfmult_fae_omg3ex, 0)
NRX_MULT_OMG3EX(neorx_rxchfmult_fp_omg3ex, 1)
NRX_MULT_OMG3EX(neorx_rxchfmult_fbe_omg3ex, 2)
#undef NRX_MULT_OMG2EX
#undef NRX_MULT_OMG3EX

#define NRX_MULT_OMG(name, k, which)                                                                              \
    extern "C" int name(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,        \
                        const double *dp, const double *gm, double *f, double *e)                                \
    {                                                                                                             \
        return mult_omg_host(k, which, nebf, nebfbe, npbf, dae, dbe, dp, gm, f, e);                              \
    }
NRX_MULT_OMG(neorx_rxchfmult_fae_omg2, 1, 0)
NRX_MULT_OMG(neorx_rxchfmult_fp_omg2, 1, 1)
NRX_MULT_OMG(neorx_rxchfmult_fbe_omg2, 1, 2)
NRX_MULT_OMG(neorx_rxchfmult_fae_omg3, 2, 0)
NRX_MULT_OMG(neorx_rxchfmult_fp_omg3, 2, 1)
NRX_MULT_OMG(neorx_rxchfmult_fbe_omg3, 2, 2)
NRX_MULT_OMG(neorx_rxchfmult_fae_omg4, 3, 0)
NRX_MULT_OMG(neorx_rxchfmult_fp_omg4, 3, 1)
NRX_MULT_OMG(neorx_rxchfmult_fbe_omg4, 3, 2)
#undef NRX_MULT_OMG
