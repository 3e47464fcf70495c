// nrx_scf.hip -- the device-resident NEO-XCHF RHF Fock build and SCF of
// libneorx_hip.so.
//
// What it restates (source/xchf1_independent/):
//   * the integrals the SCF reads, built on the device and kept in HBM:
//     class_nuc_rep (class_nuc_rep.f:1), elec_ovlap / nuc_ovlap
//     (pure_ovlap.f:2, :107), calc_GAM_epcore (gam_epcore.f:1), calc_GAM_ep
//     (gam_ep.f:1), and through nrx_lib.hip calc_GAM_ee, GAM1_OMP_MD,
//     GAM2_IC1 and GAM3_IC1 (the call sequence of xcneo_HYBRID.f:663-1002);
//   * the Fock build xchf1_fock (xchf1_fock.f:2-360): E_from_GAM_ecore /
//     _pcore / _ep / _ee / _1, S_from_GAM_1s, EG2IC1, EG3IC1, the 5-particle
//     term and XCfock_correction (xchf1_fock.f:728-784);
//   * ROOTHAN (xcscf.f:766-952): symmetric orthogonalisation, diagonalisation
//     and the density of the occupied orbitals, by a one-workgroup Jacobi
//     eigensolver;
//   * the SCF loop of xcscf (xcscf.f:2-560) without its second-order (SOSCF)
//     step, with optional DIIS extrapolation of the electronic Fock matrix.
//
// Two changes of method, both exact:
//   * The 5-particle energy is never formed from an in-core GAM_4 (nebf^8
//     npbf^2 entries).  symm_gam4's 24 weighted ket permutations of products
//     GAM_ee x GAM_2s (gam_4_OMP.f:833-856, :1179-1186), contracted with
//     DE^4 DP as EG4ICR does (gam4_ICR_OMP.f:473-481), equal
//         E4 = sum_sigma w_sigma/32 sum_b A(b_t0,b0,b_t1,b1) B(b_t2,b2,b_t3,b3)
//     (t = sigma^-1), with A the electronic density carried onto both bra
//     indices of GAM_ee and B the same for GAM_2s contracted with DP.  The
//     transforms are FP64 MFMA GEMMs, the permutation sums O(24 nebf^4), the
//     gradient dE4/dDE (= EG4ICR's FE, because GAM_4 is symmetric in its four
//     electron pairs) two more GEMMs and dE4/dDP two passes over GAM_2s.
//   * Every contraction is linear in DP, so xcscf's two Fock builds of one
//     iteration (the nuclear one with DP, the electronic one with the updated
//     DP, both with the same DE) share one pass over GM3 and GM2: "phase A"
//     forms the DE-dependent slab partials, "phase B" weights them by a DP.
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>

#include "nrx_internal.h"

namespace {

// ------------------------------------------------------------------
// small device helpers
// ------------------------------------------------------------------
NRX_DI double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

NRX_DI double wmax(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// sum over the block (every thread gets it); red holds >= 16 doubles
NRX_DI double block_sum(double v, double *red)
{
    v = wsum(v);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[i];
    __syncthreads();
    return t;
}

NRX_DI double block_max(double v, double *red)
{
    v = wmax(v);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t = fmax(t, red[i]);
    __syncthreads();
    return t;
}

NRX_DI SysParams sys_params(const DevSys &S)
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

#define GRID_LOOP(t, n) for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < (n); t += (i64)gridDim.x * blockDim.x)

// ------------------------------------------------------------------
// integrals of the SCF that nrx_lib.hip does not build
// ------------------------------------------------------------------
constexpr int NL1 = 9; // polynomial capacity of the one-particle terms (degree <= 6 with the kinetic +2)

// calc_GAM_epcore (gam_epcore.f:1-178) and elec_ovlap / nuc_ovlap
// (pure_ovlap.f:2-204): lane per contracted electronic pair [ic][jc], then
// per nuclear pair [ip][jp] (pack_2D order).
//   GAM_ecore = sum c c (T(mass 1) - sum_C Z_C V_C),  GAM_pcore = sum c c (T(pmass) + sum_C Z_C V_C)
__global__ void __launch_bounds__(256) k_one_body(DevSys S, double *__restrict__ He, double *__restrict__ Se,
                                                  double *__restrict__ Hp, double *__restrict__ Sp)
{
    const i64 ne = S.nebf, np = S.npbf, pe = S.npebf;
    const SysParams P = sys_params(S);
    GRID_LOOP(t, ne * ne + np * np)
    {
        if (t < ne * ne) {
            const i64 ic = t / ne, jc = t % ne;
            double h = 0.0, o = 0.0;
            for (i64 ie = S.kpestr[ic] - 1; ie < S.kpeend[ic]; ++ie)
                for (i64 je = S.kpestr[jc] - 1; je < S.kpeend[jc]; ++je) {
                    Sys s1 = mksys(P, 1, S.ep + ie * pe + je, 0, 0, 0, 0.0, 0, 0);
                    const double cf = S.agebfcc[ie] * S.agebfcc[je];
                    h += cf * (kin<NL1>(s1, 0, -1, 0.0, 1.0) - nuc_sum<NL1>(s1, 0, P.nat, P.cat, P.zan, false));
                    o += cf * ovl<NL1>(s1);
                }
            He[t] = h;
            Se[t] = o;
        } else {
            const i64 u = t - ne * ne, ip = u / np, jp = u % np;
            Sys s1 = mksys(P, 1, S.pp + ip * np + jp, 0, 0, 0, 0.0, 0, 0);
            const double cf = S.agnbfcc[ip] * S.agnbfcc[jp];
            Hp[u] = cf * (kin<NL1>(s1, 0, -1, 0.0, S.pmass) + nuc_sum<NL1>(s1, 0, P.nat, P.cat, P.zan, false));
            Sp[u] = cf * ovl<NL1>(s1);
        }
    }
}

// calc_GAM_ep (gam_ep.f:1-141): GAM_ep = -(ie je | ip jp), lane per
// contracted (ip,jp,ic,jc), pack_4D(nebf,nebf,npbf) = C row-major [ip][jp][ic][jc]
__global__ void __launch_bounds__(256) k_gam_ep(DevSys S, double *__restrict__ G)
{
    const i64 ne = S.nebf, np = S.npbf, pe = S.npebf;
    const SysParams P = sys_params(S);
    GRID_LOOP(t, ne * ne * np * np)
    {
        const i64 jc = t % ne, ic = (t / ne) % ne, jp = (t / (ne * ne)) % np, ip = t / (ne * ne * np);
        double acc = 0.0;
        for (i64 ie = S.kpestr[ic] - 1; ie < S.kpeend[ic]; ++ie)
            for (i64 je = S.kpestr[jc] - 1; je < S.kpeend[jc]; ++je) {
                Sys s = mksys(P, 2, S.ep + ie * pe + je, S.pp + ip * np + jp, 0, 0, 0.0, 0, 0);
                acc += coul<NL_EE>(s, 0, 1) * S.agebfcc[ie] * S.agebfcc[je];
            }
        G[t] = -acc * S.agnbfcc[ip] * S.agnbfcc[jp];
    }
}

// copy with underflow() applied (the factors of calc_GAM_4_integral,
// gam_4_OMP.f:1076-1125)
__global__ void k_flush_copy(i64 n, const double *__restrict__ x, double *__restrict__ y)
{
    GRID_LOOP(t, n)
    {
        double v = x[t];
        underflow(v);
        y[t] = v;
    }
}

// FP64 -> BF16 (round to nearest even through FP32)
__global__ void k_to_bf16(i64 n, const double *__restrict__ x, uint16_t *__restrict__ y)
{
    GRID_LOOP(t, n)
    {
        unsigned u = __float_as_uint((float)x[t]);
        u += 0x7fffu + ((u >> 16) & 1u);
        y[t] = (uint16_t)(u >> 16);
    }
}

__global__ void k_to_f32s(i64 n, const double *__restrict__ x, float *__restrict__ y)
{
    GRID_LOOP(t, n) y[t] = (float)x[t];
}

// ------------------------------------------------------------------
// FP64 MFMA GEMM: C[z] = sum_{k in chunk z} A(m,k) B(k,n) for the density
// transforms of the factorized 5-particle term.  A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], C row-major with leading dimension ldc, chunk z
// at C + z*zstride.  One wavefront per 16x16 tile of C with
// v_mfma_f64_16x16x4_f64: lane l supplies A(m0 + l%16, k + l/16) and
// B(k + l/16, n0 + l%16) and holds C(m0 + l/16 + 4i, n0 + l%16) in its i-th
// accumulator, i < 4.
// ------------------------------------------------------------------
typedef double d4_t __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_gemm_f64(i64 M, i64 N, i64 K, i64 kchunk, const double *__restrict__ A,
                                                  i64 sam, i64 sak, const double *__restrict__ B, i64 sbk, i64 sbn,
                                                  double *__restrict__ C, i64 ldc, i64 zstride)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const i64 n0 = ((i64)blockIdx.x * (blockDim.x >> 6) + w) * 16, m0 = (i64)blockIdx.y * 16;
    if (n0 >= N) return; // the whole wavefront leaves together
    const i64 k0 = (i64)blockIdx.z * kchunk, k1 = std::min<i64>(K, k0 + kchunk);
    const int r = lane & 15, kq = lane >> 4;
    const i64 am = m0 + r, bn = n0 + r;
    const bool aon = am < M, bon = bn < N;
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    for (i64 k = k0; k < k1; k += 4) {
        const i64 kk = k + kq;
        const double a = (aon && kk < k1) ? A[am * sam + kk * sak] : 0.0;
        const double b = (bon && kk < k1) ? B[kk * sbk + bn * sbn] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    double *Cz = C + (i64)blockIdx.z * zstride;
    for (int i = 0; i < 4; ++i) {
        const i64 m = m0 + kq + 4 * i;
        if (m < M && bon) Cz[m * ldc + bn] = acc[i];
    }
}

// C (M x N) = sum_z Cz (the split-K chunks, in chunk order)
__global__ void k_sum_chunks(i64 n, i64 nz, const double *__restrict__ Cz, double *__restrict__ C)
{
    GRID_LOOP(t, n)
    {
        double s = 0.0;
        for (i64 z = 0; z < nz; ++z) s += Cz[z * n + t];
        C[t] = s;
    }
}

// Y[(c,d)][(a,b)] = X[(a,b)][(c,d)] over pairs of an ne^4 tensor
__global__ void k_pair_transpose(i64 ne, const double *__restrict__ X, double *__restrict__ Y)
{
    const i64 n2 = ne * ne;
    GRID_LOOP(t, n2 * n2)
    {
        const i64 ab = t / n2, cd = t % n2;
        Y[cd * n2 + ab] = X[t];
    }
}

// ------------------------------------------------------------------
// adjoint of the 24-permutation sum of symm_gam4 (gam_4_OMP.f:788-856, ket
// permutations sigma with weights 16, -8, 4, -8, 4, -8, 4, -2, ...):
//   Y(a0,a1,a2,a3) = sum_s w_s/32 sum_{b : b_t0=a0, b0=a1, b_t1=a2, b1=a3} X(b_t2, b2, b_t3, b3)
// (t = sigma^-1); Y = dE4/dA for X = B and, by the symmetry of the sum under
// exchanging the electron pairs (0,1) and (2,3), Y = dE4/dB for X = A.
// Resolving the constraints permutation by permutation: four permutations
// leave no free index (Y reads one permuted entry of X), sixteen fix one of
// b2, b3 only up to a Kronecker delta between two a's and sum the other over
// a trace of X, and four leave two deltas and a double trace.  With the
// partial traces (k_g4_traces)
//   R0(p,q) = sum_y X(p,q,y,y)   R1(p,q) = sum_y X(y,p,q,y)
//   R2(p,q) = sum_x X(p,x,x,q)   R3(p,q) = sum_x X(x,x,p,q)
//   T1 = sum_x R0(x,x)           T2 = sum_x R2(x,x)
// the sum is the closed form evaluated by k_g4_adjoint; a lane per entry
// makes four reads of X and at most sixteen of the ne^2 traces.
// ------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_g4_traces(i64 ne, const double *__restrict__ X, double *__restrict__ R)
{
    const i64 ne2 = ne * ne;
    GRID_LOOP(t, 4 * ne2)
    {
        const i64 k = t / ne2, p = (t % ne2) / ne, q = t % ne;
        double acc = 0.0;
        for (i64 y = 0; y < ne; ++y) {
            const i64 i = k == 0 ? p : k == 1 ? y : k == 2 ? p : y;
            const i64 j = k == 0 ? q : k == 1 ? p : y;
            const i64 l = k == 0 ? y : k == 1 ? q : k == 2 ? y : p;
            const i64 m = k == 0 ? y : k == 1 ? y : q;
            acc += X[((i * ne + j) * ne + l) * ne + m];
        }
        R[t] = acc;
    }
}

__global__ void __launch_bounds__(256) k_g4_adjoint(i64 ne, const double *__restrict__ X,
                                                    const double *__restrict__ R, double *__restrict__ Y)
{
    const i64 ne2 = ne * ne;
    const double *R0 = R, *R1 = R + ne2, *R2 = R + 2 * ne2, *R3 = R + 3 * ne2;
    GRID_LOOP(t, ne2 * ne2)
    {
        const i64 a3 = t % ne, a2 = (t / ne) % ne, a1 = (t / ne2) % ne, a0 = t / (ne2 * ne);
        auto x = [&](i64 i, i64 j, i64 k, i64 l) { return X[((i * ne + j) * ne + k) * ne + l]; };
        auto r = [&](const double *Rk, i64 p, i64 q) { return Rk[p * ne + q]; };
        // no free index: sigma = (2,3,0,1), (3,2,0,1), (3,2,1,0), (2,3,1,0) as tau
        double s = 4.0 * x(a1, a0, a3, a2) - 2.0 * x(a3, a0, a1, a2) + 4.0 * x(a3, a2, a1, a0) - 2.0 * x(a1, a2, a3, a0);
        // one delta, one trace
        if (a0 == a1) s += -8.0 * r(R0, a3, a2) + 4.0 * r(R1, a2, a3) + 4.0 * r(R2, a3, a2) - 8.0 * r(R3, a3, a2);
        if (a0 == a3) s += 4.0 * r(R0, a1, a2) - 2.0 * r(R1, a2, a1) - 2.0 * r(R2, a1, a2) + 4.0 * r(R3, a1, a2);
        if (a2 == a1) s += 4.0 * r(R0, a3, a0) - 2.0 * r(R1, a0, a3) - 2.0 * r(R2, a3, a0) + 4.0 * r(R3, a3, a0);
        if (a2 == a3) s += -8.0 * r(R0, a1, a0) + 4.0 * r(R1, a0, a1) + 4.0 * r(R2, a1, a0) - 8.0 * r(R3, a1, a0);
        // two deltas, double traces
        const bool d01 = a0 == a1 && a2 == a3, d03 = a0 == a3 && a2 == a1;
        if (d01 || d03) {
            double t1 = 0.0, t2 = 0.0;
            for (i64 z = 0; z < ne; ++z) {
                t1 += r(R0, z, z);
                t2 += r(R2, z, z);
            }
            if (d01) s += 16.0 * t1 - 8.0 * t2;
            if (d03) s += -8.0 * t1 + 4.0 * t2;
        }
        Y[t] = s / 32.0;
    }
}

// GP[r] = sum_{pq} DP(ip,jp) G2s[pq][r] (r over nebf^4), DP column-major
__global__ void __launch_bounds__(256) k_gp(i64 ne4, i64 np, const double *__restrict__ DP,
                                            const double *__restrict__ G2s, double *__restrict__ GP)
{
    GRID_LOOP(r, ne4)
    {
        double acc = 0.0;
        for (i64 ip = 0; ip < np; ++ip)
            for (i64 jp = 0; jp < np; ++jp) acc += DP[jp * np + ip] * G2s[(ip * np + jp) * ne4 + r];
        GP[r] = acc;
    }
}

// part[pq][chunk] = sum_{r in chunk} G2s[pq][r] W[r]; grid (nchunk, np^2)
__global__ void __launch_bounds__(256) k_gemv_rows(i64 ne4, i64 chunk, const double *__restrict__ G2s,
                                                   const double *__restrict__ W, double *__restrict__ part)
{
    __shared__ double red[16];
    const i64 pq = blockIdx.y, r0 = (i64)blockIdx.x * chunk, r1 = std::min<i64>(ne4, r0 + chunk);
    const double *g = G2s + pq * ne4;
    double acc = 0.0;
    for (i64 r = r0 + threadIdx.x; r < r1; r += blockDim.x) acc += g[r] * W[r];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[pq * gridDim.x + blockIdx.x] = acc;
}

// sum of a[i] b[i] into part[block] (then k_sum_rows)
__global__ void __launch_bounds__(256) k_dot_part(i64 n, const double *__restrict__ a, const double *__restrict__ b,
                                                  double *__restrict__ part)
{
    __shared__ double red[16];
    double acc = 0.0;
    GRID_LOOP(t, n) acc += a[t] * b[t];
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// out[row] = sum_c part[row][c] (c < ncol), rows < nrow
__global__ void k_sum_rows(i64 nrow, i64 ncol, const double *__restrict__ part, double *__restrict__ out)
{
    GRID_LOOP(t, nrow)
    {
        double s = 0.0;
        for (i64 c = 0; c < ncol; ++c) s += part[t * ncol + c];
        out[t] = s;
    }
}

// ------------------------------------------------------------------
// phase A: quantities that depend on DE only
// ------------------------------------------------------------------
// FPx[0][pq] = sum_ij DE(i,j) GAM_ep[pq][i][j], FPx[1] from GAM_1, FPx[2]
// from GAM_1s (E_from_GAM_ep / E_from_GAM_1 / S_from_GAM_1s,
// xchf1_fock.f:913-1090): block per pq
__global__ void __launch_bounds__(256) k_ep_rows(i64 ne, i64 np, const double *__restrict__ DE,
                                                 const double *__restrict__ Gep, const double *__restrict__ G1,
                                                 const double *__restrict__ G1s, double *__restrict__ FPx)
{
    __shared__ double red[16];
    const i64 pq = blockIdx.x, ne2 = ne * ne, np2 = np * np;
    double a = 0.0, b = 0.0, c = 0.0;
    for (i64 u = threadIdx.x; u < ne2; u += blockDim.x) {
        const i64 i = u / ne, j = u % ne;
        const double d = DE[j * ne + i];
        a += d * Gep[pq * ne2 + u];
        b += d * G1[pq * ne2 + u];
        c += d * G1s[pq * ne2 + u];
    }
    a = block_sum(a, red);
    b = block_sum(b, red);
    c = block_sum(c, red);
    if (threadIdx.x == 0) {
        FPx[pq] = a;
        FPx[np2 + pq] = b;
        FPx[2 * np2 + pq] = c;
    }
}

// E_from_GAM_ee (xchf1_fock.f:852-911): FEee(a,b) = sum_cd DE(c,d)
// [(ab|cd) - (ad|cb)/2] (column-major), GAM_ee = [a][b][c][d].  A wavefront
// per (a,b): the lanes stride over (c,d) (the (ab|cd) row is contiguous) and
// meet in a fixed-order wave sum.
__global__ void __launch_bounds__(64) k_fock_ee(i64 ne, const double *__restrict__ DE, const double *__restrict__ Gee,
                                                double *__restrict__ FEee)
{
    const i64 ne2 = ne * ne;
    for (i64 t = blockIdx.x; t < ne2; t += gridDim.x) {
        const i64 a = t % ne, b = t / ne;
        const double *row = Gee + (a * ne + b) * ne2;
        double acc = 0.0;
        for (i64 cd = threadIdx.x; cd < ne2; cd += 64) {
            const i64 c = cd / ne, d = cd % ne;
            acc += DE[d * ne + c] * (row[cd] - 0.5 * Gee[((a * ne + d) * ne + c) * ne + b]);
        }
        acc = wsum(acc);
        if (threadIdx.x == 0) FEee[b * ne + a] = acc;
    }
}

// ------------------------------------------------------------------
// phase B: weighting by DP and the assembly of xchf1_fock
// ------------------------------------------------------------------
// a wavefront per (a,b) of the electronic matrices (lanes over (ip,jp)) and
// per (ip,jp) of the nuclear ones (lanes over (a,b)), fixed-order wave sums:
//   FEs(a,b) = sum_pq DP(pq) [GAM_ep + GAM_1][pq][a][b] + EG2 + EG3 rows,
//   SEs(a,b) = sum_pq DP(pq) GAM_1s[pq][a][b] + 2 SE2 (se1 + 2 se2),
//   FPs(pq) = EG2 + EG3 nuclear sums; SP2(pq)
__global__ void __launch_bounds__(64) k_phaseB_rows(i64 ne, i64 np, const double *__restrict__ DE,
                                                    const double *__restrict__ DP, const double *__restrict__ Gep,
                                                    const double *__restrict__ G1, const double *__restrict__ G1s,
                                                    const double *__restrict__ part2, const double *__restrict__ part3,
                                                    int has2, int has3, double *__restrict__ FEs, double *__restrict__ SEs,
                                                    double *__restrict__ FP23, double *__restrict__ SP2)
{
    const i64 ne2 = ne * ne, np2 = np * np;
    for (i64 t = blockIdx.x; t < ne2 + np2; t += gridDim.x) {
        if (t < ne2) {
            const i64 a = t % ne, b = t / ne; // FE(a,b) at b*ne + a
            double f = 0.0, s1 = 0.0, f2 = 0.0, s2 = 0.0, f3 = 0.0;
            for (i64 pq = threadIdx.x; pq < np2; pq += 64) {
                const double dp = DP[(pq % np) * np + pq / np];
                const i64 u = pq * ne2 + a * ne + b;
                f += dp * (Gep[u] + G1[u]);
                s1 += dp * G1s[u];
                if (has2) {
                    const double *pr = part2 + (pq * ne + a) * 4 * ne;
                    f2 += dp * 2.0 * (pr[b] - 0.5 * pr[ne + b]);
                    s2 += dp * (pr[2 * ne + b] - 0.5 * pr[3 * ne + b]);
                }
                if (has3) f3 += dp * part3[(pq * ne + a) * (ne + 1) + b];
            }
            f = wsum(f);
            s1 = wsum(s1);
            f2 = wsum(f2);
            s2 = wsum(s2);
            f3 = wsum(f3);
            if (threadIdx.x == 0) {
                FEs[t] = f + f2 + f3;
                SEs[t] = s1 + 2.0 * s2;
            }
        } else {
            const i64 pq = t - ne2;
            double f = 0.0, fs = 0.0, f3 = 0.0;
            for (i64 ab = threadIdx.x; ab < ne2; ab += 64) {
                const i64 a = ab / ne, b = ab % ne;
                if (has2) {
                    const double *pr = part2 + (pq * ne + a) * 4 * ne;
                    const double dab = DE[b * ne + a];
                    f += dab * (pr[b] - 0.5 * pr[ne + b]);
                    fs += dab * (pr[2 * ne + b] - 0.5 * pr[3 * ne + b]);
                }
                if (has3 && b == 0) f3 += part3[(pq * ne + a) * (ne + 1) + ne];
            }
            f = wsum(f);
            fs = wsum(fs);
            f3 = wsum(f3);
            if (threadIdx.x == 0) {
                FP23[pq] = f + f3;
                FP23[np2 + pq] = f;      // EG2 alone (E_gam2)
                FP23[2 * np2 + pq] = f3; // EG3 alone (E_gam3)
                SP2[pq] = fs;
            }
        }
    }
}

// Energy components (comps) and the corrected Fock matrices of
// xchf1_fock + XCfock_correction, one block:
//   comps[0] E_total  [1] E_nuc  [2] E_ecore [3] E_pcore [4] E_ep [5] E_ee
//   [6] E_gam1 [7] E_gam2 [8] E_gam3 [9] E_gam4 [10] S_total [11] S_gam1
//   [12] S_gam2 [13] H_expect
// FE = (He + FEs + FEee + FE4)/S - H (SEs); FP = (Hp + FPx + FP23 + FP4)/S - H (sp1 + sp2)
__global__ void __launch_bounds__(1024) k_assemble(i64 ne, i64 np, int nelec, double enuc,
                                                   const double *__restrict__ DE, const double *__restrict__ DP,
                                                   const double *__restrict__ He, const double *__restrict__ Hp,
                                                   const double *__restrict__ FEs, const double *__restrict__ SEs,
                                                   const double *__restrict__ FEee, const double *__restrict__ FE4,
                                                   const double *__restrict__ FPx, const double *__restrict__ FP23,
                                                   const double *__restrict__ SP2, const double *__restrict__ FP4,
                                                   const double *__restrict__ E4, double *__restrict__ FE,
                                                   double *__restrict__ FP, double *__restrict__ comps)
{
    __shared__ double red[16];
    __shared__ double sc[4];
    const i64 ne2 = ne * ne, np2 = np * np;
    double ecore = 0.0, eee = 0.0;
    for (i64 t = threadIdx.x; t < ne2; t += blockDim.x) {
        const i64 i = t % ne, j = t / ne; // DE(i,j) at j*ne + i; GAM_ecore [i][j]
        const double d = DE[t];
        ecore += d * He[i * ne + j];
        if (nelec > 1) eee += d * FEee[t];
    }
    double pcore = 0.0, eep = 0.0, eg1 = 0.0, sg1 = 0.0, eg2 = 0.0, sg2 = 0.0, eg3 = 0.0;
    for (i64 t = threadIdx.x; t < np2; t += blockDim.x) {
        const i64 ip = t / np, jp = t % np; // pq = ip*np + jp; DP(ip,jp) at jp*np + ip
        const double dp = DP[jp * np + ip];
        pcore += DP[t] * Hp[(t % np) * np + t / np];
        eep += dp * FPx[t];
        eg1 += dp * FPx[np2 + t];
        sg1 += dp * FPx[2 * np2 + t];
        if (nelec > 1) {
            eg2 += dp * FP23[np2 + t];
            sg2 += dp * SP2[t];
        }
        if (nelec > 2) eg3 += dp * FP23[2 * np2 + t];
    }
    ecore = block_sum(ecore, red);
    eee = 0.5 * block_sum(eee, red);
    pcore = block_sum(pcore, red);
    eep = block_sum(eep, red);
    eg1 = block_sum(eg1, red);
    sg1 = block_sum(sg1, red);
    eg2 = block_sum(eg2, red);
    sg2 = block_sum(sg2, red);
    eg3 = block_sum(eg3, red);
    const double eg4 = nelec > 2 ? E4[0] : 0.0;
    const double S = 1.0 + sg1 + sg2;
    const double psi = (ecore + pcore + eep + eee + eg1 + eg2 + eg3 + eg4) / S;
    const double H = psi / S;
    if (threadIdx.x == 0) {
        comps[0] = psi + enuc;
        comps[1] = enuc;
        comps[2] = ecore;
        comps[3] = pcore;
        comps[4] = eep;
        comps[5] = eee;
        comps[6] = eg1;
        comps[7] = eg2;
        comps[8] = eg3;
        comps[9] = eg4;
        comps[10] = S;
        comps[11] = sg1;
        comps[12] = sg2;
        comps[13] = H;
    }
    const double iS = 1.0 / S;
    for (i64 t = threadIdx.x; t < ne2; t += blockDim.x) {
        const i64 i = t % ne, j = t / ne;
        double f = He[i * ne + j] + FEs[t];
        if (nelec > 1) f += FEee[t];
        if (nelec > 2) f += 2.0 * (FE4[i * ne + j] + FE4[ne2 + i * ne + j]);
        FE[t] = iS * f - H * SEs[t];
    }
    for (i64 t = threadIdx.x; t < np2; t += blockDim.x) {
        const i64 ip = t % np, jp = t / np; // FP(ip,jp) at jp*np + ip; pq = ip*np + jp
        const i64 pq = ip * np + jp;
        double f = Hp[ip * np + jp] + FPx[pq] + FPx[np2 + pq];
        double s = FPx[2 * np2 + pq];
        if (nelec > 1) {
            f += FP23[np2 + pq];
            s += SP2[pq];
        }
        if (nelec > 2) f += FP23[2 * np2 + pq] + FP4[pq];
        FP[t] = iS * f - H * s;
    }
}

// ------------------------------------------------------------------
// ROOTHAN (xcscf.f:766-952) in one workgroup: symmetric orthogonalisation
// X = S^-1/2 (once per basis), F' = X^T F X, Jacobi diagonalisation of F',
// C = X V (eigenvalues ascending, as EISPACK's RS returns them), and the
// density D = prefac sum_{k=first..last} C(:,k) C(:,k)^T.  Matrices are
// column-major n x n.  The parallel cyclic Jacobi method pairs the indices
// round-robin (n/2 disjoint rotations per step) and sweeps until the
// off-diagonal norm is below 1e-15 of the diagonal's, starting from the
// identity or, within an SCF, from the previous call's eigenvectors.
// ------------------------------------------------------------------
constexpr int RH_MAXN = 128;
constexpr int RH_BS = 256;

// a (symmetric) is rotated to diagonal form; v accumulates the rotations,
// from the identity when `identity` is set, else from the v given
NRX_DI void jacobi(int n, double *a, double *v, double *red, bool identity = true)
{
    __shared__ double rc[RH_MAXN / 2], rs[RH_MAXN / 2];
    __shared__ int rp[RH_MAXN / 2], rq[RH_MAXN / 2];
    const int m = n + (n & 1), half = m / 2;
    if (identity)
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) v[t] = (t % n == t / n) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            const double x = a[t];
            if (t % n == t / n)
                dia += x * x;
            else
                off += x * x;
        }
        off = block_sum(off, red);
        dia = block_sum(dia, red);
        if (off <= 1e-30 * dia || off == 0.0) break;
        for (int r = 0; r < m - 1; ++r) {
            for (int k = threadIdx.x; k < half; k += blockDim.x) {
                int p = (r + k) % (m - 1), q = k == 0 ? m - 1 : (r - k + m - 1) % (m - 1);
                if (p > q) {
                    const int x = p;
                    p = q;
                    q = x;
                }
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double apq = a[q * n + p];
                    if (apq != 0.0) {
                        const double th = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
                        const double tt = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                    }
                }
                rc[k] = c;
                rs[k] = s;
                rp[k] = p;
                rq[k] = q < n ? q : -1;
            }
            __syncthreads();
            // rows: a <- J^T a
            for (int t = threadIdx.x; t < half * n; t += blockDim.x) {
                const int k = t / n, j = t % n, p = rp[k], q = rq[k];
                if (q < 0) continue;
                const double c = rc[k], s = rs[k], x = a[j * n + p], y = a[j * n + q];
                a[j * n + p] = c * x - s * y;
                a[j * n + q] = s * x + c * y;
            }
            __syncthreads();
            // columns: a <- a J, v <- v J
            for (int t = threadIdx.x; t < half * n; t += blockDim.x) {
                const int k = t / n, i = t % n, p = rp[k], q = rq[k];
                if (q < 0) continue;
                const double c = rc[k], s = rs[k];
                double x = a[p * n + i], y = a[q * n + i];
                a[p * n + i] = c * x - s * y;
                a[q * n + i] = s * x + c * y;
                x = v[p * n + i];
                y = v[q * n + i];
                v[p * n + i] = c * x - s * y;
                v[q * n + i] = s * x + c * y;
            }
            __syncthreads();
        }
    }
}

// ascending order of the diagonal of a into ord: a total order (NaN sorts as
// +inf, ties by index), so the ranks are a permutation of 0..n-1 whatever
// the values
NRX_DI double eig_key(double x) { return x != x ? __builtin_huge_val() : x; }

NRX_DI void eig_order(int n, const double *a, int *ord)
{
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double e = eig_key(a[k * n + k]);
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double f = eig_key(a[j * n + j]);
            rank += (f < e) || (f == e && j < k);
        }
        ord[rank] = k;
    }
    __syncthreads();
}

// X = S^-1/2 (mode 0) of the n x n S; work: 2 n^2 doubles (LDS when small)
__global__ void __launch_bounds__(RH_BS) k_orth(int n, const double *__restrict__ S, double *__restrict__ X,
                                                double *__restrict__ gwork)
{
    extern __shared__ double lds[];
    __shared__ double red[16];
    __shared__ int ord[RH_MAXN];
    double *a = gwork ? gwork : lds, *v = a + n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) a[t] = S[t];
    __syncthreads();
    jacobi(n, a, v, red);
    (void)ord;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = 0; k < n; ++k) x += v[k * n + i] * v[k * n + j] / sqrt(a[k * n + k]);
        X[t] = x;
    }
}

// ROOTHAN with X = S^-1/2: C, eigenvalues and the density of orbitals
// first..last (0-based, inclusive).  Vw (n^2, may be null) keeps the
// eigenvectors in the orthogonal basis between the calls of an SCF on one
// basis: when *vflag is set, F' is first rotated into that basis
// (Vw^T F' Vw, nearly diagonal once the SCF settles) and the rotations start
// from Vw, so a sweep or two meet the same convergence test as a start from
// the identity; the eigenvectors found go back to Vw and the flag is set.
// Work: 4 n^2 doubles in LDS (a, v and copies of F and X) when gwork is
// null, else a and v in gwork (2 n^2) and F, X read where they lie.
__global__ void __launch_bounds__(RH_BS) k_roothan(int n, const double *__restrict__ X, const double *__restrict__ F,
                                                   int first, int last, double prefac, double *__restrict__ C,
                                                   double *__restrict__ ev, double *__restrict__ D,
                                                   double *__restrict__ gwork, double *__restrict__ Vw,
                                                   double *__restrict__ vflag)
{
    extern __shared__ double lds[];
    __shared__ double red[16];
    __shared__ int ord[RH_MAXN];
    const int nn = n * n;
    double *a = gwork ? gwork : lds, *v = a + nn;
    const double *Fm = F, *Xm = X;
    double *fl = gwork ? nullptr : lds + 2 * nn;
    if (!gwork) {
        double *xl = lds + 3 * nn;
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            fl[t] = F[t];
            xl[t] = X[t];
        }
        __syncthreads();
        Fm = fl;
        Xm = xl;
    }
    // v = F X, a = X^T v
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = 0; k < n; ++k) x += Fm[k * n + i] * Xm[j * n + k];
        v[t] = x;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = 0; k < n; ++k) x += Xm[i * n + k] * v[j * n + k];
        a[t] = x;
    }
    __syncthreads();
    const bool warm = Vw && *vflag != 0.0;
    if (warm) {
        // a <- W^T a W with W = Vw (staged in F's place when F is in LDS)
        const double *W = Vw;
        if (fl) {
            for (int t = threadIdx.x; t < nn; t += blockDim.x) fl[t] = Vw[t];
            __syncthreads();
            W = fl;
        }
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            const int i = t % n, j = t / n;
            double x = 0.0;
            for (int k = 0; k < n; ++k) x += a[k * n + i] * W[j * n + k];
            v[t] = x;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            const int i = t % n, j = t / n;
            double x = 0.0;
            for (int k = 0; k < n; ++k) x += W[i * n + k] * v[j * n + k];
            a[t] = x;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nn; t += blockDim.x) v[t] = W[t];
        __syncthreads();
    }
    // symmetrise against rounding before the rotations
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        if (i < j) {
            const double m = 0.5 * (a[t] + a[i * n + j]);
            a[t] = m;
            a[i * n + j] = m;
        }
    }
    __syncthreads();
    jacobi(n, a, v, red, !warm);
    eig_order(n, a, ord);
    for (int k = threadIdx.x; k < n; k += blockDim.x) ev[k] = a[ord[k] * n + ord[k]];
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, k = t / n, kk = ord[k];
        double x = 0.0;
        for (int l = 0; l < n; ++l) x += Xm[l * n + i] * v[kk * n + l];
        C[t] = x;
    }
    if (Vw) {
        for (int t = threadIdx.x; t < nn; t += blockDim.x) Vw[t] = v[t];
        if (threadIdx.x == 0) *vflag = 1.0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = first; k <= last; ++k) x += C[k * n + i] * C[k * n + j];
        D[t] = prefac * x;
    }
}

// LDS of k_orth (2 n^2 doubles) and of k_roothan (4 n^2), or 0: work in
// global memory
inline size_t roothan_lds(int n) { return 2 * (size_t)n * n * 8 <= 60 * 1024 ? 2 * (size_t)n * n * 8 : 0; }
inline size_t roothan_lds4(int n) { return 4 * (size_t)n * n * 8 <= 60 * 1024 ? 4 * (size_t)n * n * 8 : 0; }

// DENDIF (xcscf.f:719-738) and COPYDEN: out = max |D - D0|, then D0 = D
__global__ void __launch_bounds__(256) k_dendif(i64 n, const double *__restrict__ D, double *__restrict__ D0,
                                                double *__restrict__ out)
{
    __shared__ double red[16];
    double m = 0.0;
    for (i64 t = threadIdx.x; t < n; t += blockDim.x) m = fmax(m, fabs(D[t] - D0[t]));
    m = block_max(m, red);
    for (i64 t = threadIdx.x; t < n; t += blockDim.x) D0[t] = D[t];
    if (threadIdx.x == 0) *out = m;
}

// ------------------------------------------------------------------
// DIIS (Pulay) on the electronic Fock matrix: error e = X^T (F D S - S D F) X
// for the F built from D; the extrapolated F = sum_i c_i F_i with
// sum_i c_i = 1 minimising |sum_i c_i e_i|.  One block; history of at most
// DIIS_MAX pairs in a ring; hist[0] = count, hist[1] = next slot.
// ------------------------------------------------------------------
constexpr int DIIS_MAX = 8;

// LDS of k_diis: F, D, S, X and two work matrices (6 n^2 doubles), or 0:
// the work matrices w1, w2 in global memory and the inputs read where they lie
inline size_t diis_lds(int n) { return 6 * (size_t)n * n * 8 <= 60 * 1024 ? 6 * (size_t)n * n * 8 : 0; }

__global__ void __launch_bounds__(256) k_diis(int n, const double *__restrict__ F, const double *__restrict__ D,
                                              const double *__restrict__ S, const double *__restrict__ X,
                                              double *__restrict__ Fh, double *__restrict__ Eh, double *__restrict__ B,
                                              double *__restrict__ hist, double *__restrict__ Fout,
                                              double *__restrict__ errout, double *__restrict__ w1g,
                                              double *__restrict__ w2g, int staged)
{
    extern __shared__ double lds[];
    __shared__ double red[16];
    __shared__ double cvec[DIIS_MAX + 1];
    __shared__ double dots[DIIS_MAX][4];
    __shared__ double M[DIIS_MAX + 1][DIIS_MAX + 2];
    const int nn = n * n, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int cnt = (int)hist[0], slot = (int)hist[1];
    double *Fs = Fh + (i64)slot * nn, *Es = Eh + (i64)slot * nn;
    const double *Fm = F, *Dm = D, *Sm = S, *Xm = X;
    double *w1 = w1g, *w2 = w2g;
    if (staged) {
        double *f = lds, *d = lds + nn, *sm = lds + 2 * nn, *xm = lds + 3 * nn;
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            f[t] = F[t];
            d[t] = D[t];
            sm[t] = S[t];
            xm[t] = X[t];
        }
        __syncthreads();
        Fm = f;
        Dm = d;
        Sm = sm;
        Xm = xm;
        w1 = lds + 4 * nn;
        w2 = lds + 5 * nn;
    }
    // w1 = F D, w2 = w1 S (= F D S)
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = 0; k < n; ++k) x += Fm[k * n + i] * Dm[j * n + k];
        w1[t] = x;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = 0; k < n; ++k) x += w1[k * n + i] * Sm[j * n + k];
        w2[t] = x;
    }
    __syncthreads();
    // commutator FDS - (FDS)^T (S D F = (F D S)^T for symmetric F, D, S)
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        w1[t] = w2[t] - w2[i * n + j];
    }
    __syncthreads();
    // Es = X^T w1 X
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = 0; k < n; ++k) x += w1[k * n + i] * Xm[j * n + k];
        w2[t] = x;
    }
    __syncthreads();
    double em = 0.0;
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const int i = t % n, j = t / n;
        double x = 0.0;
        for (int k = 0; k < n; ++k) x += Xm[i * n + k] * w2[j * n + k];
        Es[t] = x;
        w1[t] = x; // Es again, where this block reads it fastest
        Fs[t] = Fm[t];
        em = fmax(em, fabs(x));
    }
    em = block_max(em, red);
    cnt = cnt < DIIS_MAX ? cnt + 1 : DIIS_MAX;
    // row `slot` of the error overlaps: every dot in one pass, per-wave
    // partials summed in a fixed order (the slot's own error from w1)
    double x[DIIS_MAX];
#pragma unroll
    for (int k = 0; k < DIIS_MAX; ++k) x[k] = 0.0;
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const double e = w1[t];
#pragma unroll
        for (int k = 0; k < DIIS_MAX; ++k)
            if (k < cnt) x[k] += e * (k == slot ? e : Eh[(i64)k * nn + t]);
    }
#pragma unroll
    for (int k = 0; k < DIIS_MAX; ++k) {
        const double v = wsum(x[k]);
        if (lane == 0 && k < cnt) dots[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < cnt) {
        const int k = threadIdx.x, nw = (blockDim.x + 63) >> 6;
        double v = 0.0;
        for (int q = 0; q < nw; ++q) v += dots[k][q];
        B[slot * DIIS_MAX + k] = v;
        B[k * DIIS_MAX + slot] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // solve [B 1; 1 0][c; l] = [0; 1] by Gaussian elimination with
        // partial pivoting (cnt + 1 unknowns), scaled by the diagonal
        const int m = cnt + 1;
        double sc = 0.0;
        for (int i = 0; i < cnt; ++i) sc = fmax(sc, fabs(B[i * DIIS_MAX + i]));
        sc = sc > 0.0 ? 1.0 / sc : 1.0;
        for (int i = 0; i < m; ++i) {
            for (int j = 0; j < m; ++j)
                M[i][j] = (i < cnt && j < cnt) ? B[i * DIIS_MAX + j] * sc : ((i == cnt) != (j == cnt) ? 1.0 : 0.0);
            M[i][m] = i == cnt ? 1.0 : 0.0;
        }
        bool ok = true;
        for (int c = 0; c < m && ok; ++c) {
            int piv = c;
            for (int r = c + 1; r < m; ++r)
                if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
            if (fabs(M[piv][c]) < 1e-300) {
                ok = false;
                break;
            }
            for (int j = 0; j <= m; ++j) {
                const double t = M[c][j];
                M[c][j] = M[piv][j];
                M[piv][j] = t;
            }
            for (int r = 0; r < m; ++r)
                if (r != c) {
                    const double f = M[r][c] / M[c][c];
                    for (int j = c; j <= m; ++j) M[r][j] -= f * M[c][j];
                }
        }
        for (int i = 0; i < cnt; ++i) cvec[i] = ok ? M[i][m] / M[i][i] : (i == slot ? 1.0 : 0.0);
        hist[0] = cnt;
        hist[1] = (slot + 1) % DIIS_MAX;
        *errout = em;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        double v = 0.0;
        for (int k = 0; k < cnt; ++k) v += cvec[k] * (k == slot ? Fm[t] : Fh[(i64)k * nn + t]);
        Fout[t] = v;
    }
}

__global__ void k_fill(i64 n, double v, double *__restrict__ p)
{
    GRID_LOOP(t, n) p[t] = v;
}

__global__ void k_copy(i64 n, const double *__restrict__ x, double *__restrict__ y)
{
    GRID_LOOP(t, n) y[t] = x[t];
}

// the two densities of an incremental EG3 pass: X = DE - DEref, Y = DE + DEref
__global__ void k_inc_dens(i64 n, const double *__restrict__ D, const double *__restrict__ D0, double *__restrict__ X,
                           double *__restrict__ Y)
{
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        X[i] = D[i] - D0[i];
        Y[i] = D[i] + D0[i];
    }
}

// EG3 slab rows of DE from the increment: part[blk][x] += ref[blk][x] (the
// rows of the reference density), then part[blk][ne] = sum_x DE(a,x) F[x] / 3
// with a = blk % ne, as the slab kernel forms it; a wavefront per slab, fixed
// order
__global__ void __launch_bounds__(64) k_inc_rows(i64 ne, i64 nblk, const double *__restrict__ DE,
                                                 const double *__restrict__ ref, double *__restrict__ part)
{
    const i64 blk = blockIdx.x;
    if (blk >= nblk) return;
    const i64 a = blk % ne;
    double *r = part + blk * (ne + 1);
    const double *q = ref + blk * (ne + 1);
    double p = 0.0;
    for (i64 x = threadIdx.x; x < ne; x += 64) {
        const double v = r[x] + q[x];
        r[x] = v;
        p += DE[x * ne + a] * v;
    }
    p = wsum(p);
    if (threadIdx.x == 0) r[ne] = p / 3.0;
}

// ------------------------------------------------------------------
// host side
// ------------------------------------------------------------------
struct Xchf {
    System *s = nullptr;
    i64 ne = 0, np = 0;
    int nelec = 0;
    double enuc = 0.0;
    std::vector<void *> allocs;
    // integrals (HBM)
    double *Se = nullptr, *Sp = nullptr, *He = nullptr, *Hp = nullptr, *Gep = nullptr, *Gee = nullptr;
    double *G1 = nullptr, *G1s = nullptr, *G2 = nullptr, *G2s = nullptr, *G3 = nullptr;
    void *G3f = nullptr, *G3h = nullptr;
    double *GeeU = nullptr, *G2sU = nullptr, *Xe = nullptr, *Xp = nullptr;
    // Fock-build workspace
    double *part2 = nullptr, *part3 = nullptr, *FPx = nullptr, *FEee = nullptr;
    double *A4 = nullptr, *HA4 = nullptr, *Bb4 = nullptr, *GPb4 = nullptr, *FP4 = nullptr, *t4 = nullptr;
    double *GP4 = nullptr, *HB4 = nullptr, *B4 = nullptr, *Ab4 = nullptr, *FE4 = nullptr, *E4 = nullptr;
    double *R4 = nullptr; // partial traces of the adjoint (4 nebf^2)
    double *zpart = nullptr, *rows = nullptr;
    // incremental EG3 builds: rows of the reference density, the reference
    // density and the pass's two densities
    double *part3ref = nullptr, *DEref = nullptr, *DXi = nullptr, *DYi = nullptr;
    i64 nz_fe = 1, nch_fp = 1, nz_dot = 1;
    double *FEs = nullptr, *SEs = nullptr, *FP23 = nullptr, *SP2 = nullptr;
    // SCF state
    double *DE = nullptr, *DP = nullptr, *DE0 = nullptr, *DP0 = nullptr, *FE = nullptr, *FP = nullptr;
    double *FEx = nullptr, *CE = nullptr, *CP = nullptr, *evE = nullptr, *evP = nullptr, *comps = nullptr;
    double *stat = nullptr, *rwork = nullptr;
    double *Fh = nullptr, *Eh = nullptr, *Bd = nullptr, *hist = nullptr, *dw1 = nullptr, *dw2 = nullptr;
    // eigenvectors of the last ROOTHAN per basis (orthogonal basis) and
    // their valid flags [electronic, nuclear]
    double *VwE = nullptr, *VwP = nullptr, *vflag = nullptr;
    double *hstat = nullptr; // pinned
    static constexpr int NGRAPH = 6;
    hipGraphExec_t graph[NGRAPH] = {};
    int graph_key[NGRAPH] = {-1, -1, -1, -1, -1, -1};
    hipStream_t st = nullptr;

    template <class T> int alloc(T *&p, i64 n)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, (size_t)std::max<i64>(n, 1) * sizeof(T));
        if (e != hipSuccess) return fail(NEORX_E_NOMEM, "hipMalloc(%lld x %zu) failed: %s", (long long)n, sizeof(T),
                                         hipGetErrorString(e));
        allocs.push_back(q);
        p = (T *)q;
        return NEORX_OK;
    }
    ~Xchf()
    {
        for (auto &g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        for (void *p : allocs) (void)hipFree(p);
        if (hstat) (void)hipHostFree(hstat);
        if (st) (void)hipStreamDestroy(st);
    }
};

std::mutex g_xmu;
std::map<i64, Xchf *> g_xchf;
i64 g_xnext = 1;

Xchf *get_x(i64 h)
{
    std::lock_guard<std::mutex> lk(g_xmu);
    auto it = g_xchf.find(h);
    return it == g_xchf.end() ? nullptr : it->second;
}

// C (M x N) = A(M x K) B(K x N) through the MFMA GEMM, split over K when the
// output is small; zpart holds the chunks
int gemm(hipStream_t st, i64 M, i64 N, i64 K, const double *A, i64 sam, i64 sak, const double *B, i64 sbk, i64 sbn,
         double *C, double *zpart, i64 zcap)
{
    const i64 tm = (M + 15) / 16, tn = (N + 15) / 16;
    const i64 wpb = std::min<i64>(4, tn);
    i64 nz = 1;
    if (tm * tn < 64 && K > 256) nz = std::min<i64>(zcap, (K + 255) / 256);
    const i64 kchunk = ((K + nz - 1) / nz + 3) / 4 * 4;
    nz = (K + kchunk - 1) / kchunk;
    dim3 grid((unsigned)((tn + wpb - 1) / wpb), (unsigned)tm, (unsigned)nz);
    if (nz == 1) {
        hipLaunchKernelGGL(k_gemm_f64, grid, dim3((unsigned)(64 * wpb)), 0, st, M, N, K, kchunk, A, sam, sak, B, sbk, sbn,
                           C, N, (i64)0);
    } else {
        hipLaunchKernelGGL(k_gemm_f64, grid, dim3((unsigned)(64 * wpb)), 0, st, M, N, K, kchunk, A, sam, sak, B, sbk, sbn,
                           zpart, N, M * N);
        hipLaunchKernelGGL(k_sum_chunks, dim3(nblocks(M * N, 256)), dim3(256), 0, st, M * N, nz, zpart, C);
    }
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

// Y[c][r] = sum_i D(i,c) X[i][r] over the leading index of an ne x R tensor
// (D column-major: D(i,c) = D[c*ne + i], read row-major as [c][i])
int lead_transform(Xchf *x, const double *D, const double *X, double *Y, i64 R)
{
    return gemm(x->st, x->ne, R, x->ne, D, 1 * x->ne, 1, X, R, 1, Y, x->zpart, 1);
}

// the inverse direction: Y[i][r] = sum_c D(i,c) X[c][r]
int lead_transform_inv(Xchf *x, const double *D, const double *X, double *Y, i64 R)
{
    return gemm(x->st, x->ne, R, x->ne, D, 1, x->ne, X, R, 1, Y, x->zpart, 1);
}

// both bra indices of an ne^4 tensor [i0][b0][i1][b1] carried by D:
//   H = slot i1 only, T = both slots (t4: scratch)
int both_bras(Xchf *x, const double *D, const double *G, double *H, double *T)
{
    const i64 ne = x->ne, ne3 = ne * ne * ne, ne4 = ne3 * ne;
    hipLaunchKernelGGL(k_pair_transpose, dim3(grid_blocks(ne4, 256)), dim3(256), 0, x->st, ne, G, x->t4);
    int rc = lead_transform(x, D, x->t4, T, ne3); // T[(c1 b1)][(i0 b0)]
    if (rc) return rc;
    hipLaunchKernelGGL(k_pair_transpose, dim3(grid_blocks(ne4, 256)), dim3(256), 0, x->st, ne, T, H);
    return lead_transform(x, D, H, T, ne3);
}

// how one phase A reads GAM_3
enum G3Pass {
    G3_FULL = 0,    // the rows of DE
    G3_REFRESH = 1, // the rows of DE, kept as the reference of later increments
    G3_INC = 2,     // the reference rows plus the increment T(DE - DEref, DE + DEref)
};

// phase A (depends on DE): dtype = storage of GAM_3 read.  Everything but
// the GM3 pass (GAM_ep/GAM_1 rows, GAM_ee, the GM2 slabs and the DE side of
// the 5-particle term) goes first, then the GM3 pass, in one stream: run
// beside the pass, which takes every CU, the small kernels waited for CUs
// and lengthened the iteration.
int phase_a_rest(Xchf *x, const double *DE)
{
    hipStream_t st = x->st;
    const i64 ne = x->ne, np = x->np, np2 = np * np, ne4 = ne * ne * ne * ne;
    hipLaunchKernelGGL(k_ep_rows, dim3((unsigned)np2), dim3(256), 0, st, ne, np, DE, x->Gep, x->G1, x->G1s, x->FPx);
    HIPCHK(hipGetLastError());
    if (x->nelec > 1) {
        hipLaunchKernelGGL(k_fock_ee, dim3(grid_blocks(ne * ne, 1)), dim3(64), 0, st, ne, DE, x->Gee, x->FEee);
        int rc = nrx_eg2_slabs(ne, np2 * ne, DE, x->G2, x->G2s, x->part2, st);
        if (rc) return rc;
    }
    if (x->nelec > 2) {
        // 5-particle term, DE side: A (both bras), HA (one bra), Bbar = adj(A),
        // GPbar = Bbar carried back, FP4 = G2s . GPbar
        int rc = both_bras(x, DE, x->GeeU, x->HA4, x->A4);
        if (rc) return rc;
        hipLaunchKernelGGL(k_g4_traces, dim3(grid_blocks(4 * ne * ne, 256)), dim3(256), 0, st, ne, x->A4, x->R4);
        hipLaunchKernelGGL(k_g4_adjoint, dim3(grid_blocks(ne4, 256)), dim3(256), 0, st, ne, x->A4, x->R4, x->Bb4);
        // GPbar(i2,b2,i3,b3) = sum D(i2,c2) D(i3,c3) Bbar(c2,b2,c3,b3)
        hipLaunchKernelGGL(k_pair_transpose, dim3(grid_blocks(ne4, 256)), dim3(256), 0, st, ne, x->Bb4, x->t4);
        rc = lead_transform_inv(x, DE, x->t4, x->GPb4, ne * ne * ne);
        if (rc) return rc;
        hipLaunchKernelGGL(k_pair_transpose, dim3(grid_blocks(ne4, 256)), dim3(256), 0, st, ne, x->GPb4, x->t4);
        rc = lead_transform_inv(x, DE, x->t4, x->GPb4, ne * ne * ne);
        if (rc) return rc;
        const i64 chunk = (ne4 + x->nch_fp - 1) / x->nch_fp;
        hipLaunchKernelGGL(k_gemv_rows, dim3((unsigned)x->nch_fp, (unsigned)np2), dim3(256), 0, st, ne4, chunk, x->G2sU,
                           x->GPb4, x->rows);
        hipLaunchKernelGGL(k_sum_rows, dim3(nblocks(np2, 256)), dim3(256), 0, st, np2, x->nch_fp, x->rows, x->FP4);
        HIPCHK(hipGetLastError());
    }
    return NEORX_OK;
}

int phase_a(Xchf *x, const double *DE, int dtype, int pass = G3_FULL)
{
    hipStream_t st = x->st;
    const i64 ne = x->ne, np = x->np, np2 = np * np;
    if (x->nelec <= 2) return phase_a_rest(x, DE);
    const void *g = dtype == NRX_F64 ? (const void *)x->G3 : dtype == NRX_F32 ? x->G3f : x->G3h;
    if (!g) return fail(NEORX_E_ARG, "GAM_3 is not held in storage type %d (see neorx_xchf_create store)", dtype);
    int rc = phase_a_rest(x, DE);
    if (rc) return rc;
    // the GM3 pass
    const i64 nblk = np2 * ne;
    if (pass == G3_INC) {
        hipLaunchKernelGGL(k_inc_dens, dim3(nblocks(ne * ne, 256)), dim3(256), 0, st, ne * ne, DE, x->DEref, x->DXi,
                           x->DYi);
        rc = nrx_eg3_slabs2(ne, nblk, x->DXi, x->DYi, nullptr, g, dtype, x->part3, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_inc_rows, dim3((unsigned)nblk), dim3(64), 0, st, ne, nblk, DE, x->part3ref, x->part3);
        HIPCHK(hipGetLastError());
    } else {
        rc = nrx_eg3_slabs(ne, nblk, DE, g, dtype, x->part3, st);
        if (rc) return rc;
        if (pass == G3_REFRESH) {
            hipLaunchKernelGGL(k_copy, dim3(nblocks(nblk * (ne + 1), 256)), dim3(256), 0, st, nblk * (ne + 1),
                               x->part3, x->part3ref);
            hipLaunchKernelGGL(k_copy, dim3(nblocks(ne * ne, 256)), dim3(256), 0, st, ne * ne, DE, x->DEref);
            HIPCHK(hipGetLastError());
        }
    }
    return NEORX_OK;
}

// phase B (depends on DP, after phase A with the same DE): FE, FP, comps
int phase_b(Xchf *x, const double *DE, const double *DP, double *FE, double *FP, double *comps)
{
    hipStream_t st = x->st;
    const i64 ne = x->ne, np = x->np, ne2 = ne * ne, np2 = np * np, ne3 = ne2 * ne, ne4 = ne3 * ne;
    hipLaunchKernelGGL(k_phaseB_rows, dim3(grid_blocks(ne2 + np2, 1)), dim3(64), 0, st, ne, np, DE, DP, x->Gep, x->G1,
                       x->G1s, x->part2, x->part3, (int)(x->nelec > 1), (int)(x->nelec > 2), x->FEs, x->SEs, x->FP23,
                       x->SP2);
    HIPCHK(hipGetLastError());
    if (x->nelec > 2) {
        // 5-particle term, DP side: GP = DP . G2s, B (both bras), HB, Abar = adj(B),
        // E4 = <A, Abar>, FE4(x,y) = 2 sum_r Abar[y][r] HA[x][r] + 2 sum_r Bbar[y][r] HB[x][r]
        hipLaunchKernelGGL(k_gp, dim3(grid_blocks(ne4, 256)), dim3(256), 0, st, ne4, np, DP, x->G2sU, x->GP4);
        int rc = both_bras(x, DE, x->GP4, x->HB4, x->B4);
        if (rc) return rc;
        hipLaunchKernelGGL(k_g4_traces, dim3(grid_blocks(4 * ne2, 256)), dim3(256), 0, st, ne, x->B4, x->R4);
        hipLaunchKernelGGL(k_g4_adjoint, dim3(grid_blocks(ne4, 256)), dim3(256), 0, st, ne, x->B4, x->R4, x->Ab4);
        hipLaunchKernelGGL(k_dot_part, dim3((unsigned)x->nz_dot), dim3(256), 0, st, ne4, x->A4, x->Ab4, x->rows);
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, st, (i64)1, x->nz_dot, x->rows, x->E4);
        // FE4 = dE4/dDE: C1[x][y] = sum_r HA[x][r] Abar[y][r] (the bra slots of
        // GAM_ee), C2[x][y] = sum_r HB[x][r] Bbar[y][r] (those of GAM_2s); each
        // pair of slots gives the same sum, so FE4(x,y) = 2 (C1 + C2)[x][y]
        rc = gemm(st, ne, ne, ne3, x->HA4, ne3, 1, x->Ab4, 1, ne3, x->FE4, x->zpart, x->nz_fe);
        if (rc) return rc;
        rc = gemm(st, ne, ne, ne3, x->HB4, ne3, 1, x->Bb4, 1, ne3, x->FE4 + ne2, x->zpart, x->nz_fe);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_assemble, dim3(1), dim3(1024), 0, st, ne, np, x->nelec, x->enuc, DE, DP, x->He, x->Hp, x->FEs,
                       x->SEs, x->FEee, x->FE4, x->FPx, x->FP23, x->SP2, x->FP4, x->E4, FE, FP, comps);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

} // namespace

// ------------------------------------------------------------------
// context: integrals and workspace
// ------------------------------------------------------------------
namespace {

// Blocking copy ordered on stream st (after the work queued there).
int copy_on(hipStream_t st, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

int xchf_build(Xchf *x, i64 handle, i64 store)
{
    System *s = x->s;
    hipStream_t st = x->st;
    const i64 ne = x->ne, np = x->np, ne2 = ne * ne, np2 = np * np, ne4 = ne2 * ne2;
    const i64 ng1 = ne2 * np2, ng2 = ne4 * np2, ng3 = ne4 * ne2 * np2;
    int rc = 0;
    rc |= x->alloc(x->Se, ne2);
    rc |= x->alloc(x->Sp, np2);
    rc |= x->alloc(x->He, ne2);
    rc |= x->alloc(x->Hp, np2);
    rc |= x->alloc(x->Gep, ng1);
    rc |= x->alloc(x->Gee, ne4);
    rc |= x->alloc(x->G1, ng1);
    rc |= x->alloc(x->G1s, ng1);
    rc |= x->alloc(x->Xe, ne2);
    rc |= x->alloc(x->Xp, np2);
    if (rc) return rc;
    // class_nuc_rep (class_nuc_rep.f:1-45), on the host
    {
        const i64 nat = s->d.nat;
        std::vector<double> cat(3 * std::max<i64>(nat, 1)), zan(std::max<i64>(nat, 1));
        if (nat > 0) {
            rc = copy_on(x->st, cat.data(), s->d.cat, 3 * nat * 8, hipMemcpyDeviceToHost);
            if (rc) return rc;
            rc = copy_on(x->st, zan.data(), s->d.zan, nat * 8, hipMemcpyDeviceToHost);
            if (rc) return rc;
        }
        double e = 0.0;
        for (i64 i = 1; i < nat; ++i)
            for (i64 j = 0; j < i; ++j) {
                double r2 = 0.0;
                for (int k = 0; k < 3; ++k) r2 += (cat[3 * i + k] - cat[3 * j + k]) * (cat[3 * i + k] - cat[3 * j + k]);
                if (r2 != 0.0) e += zan[i] * zan[j] / sqrt(r2);
            }
        x->enuc = e;
    }
    hipLaunchKernelGGL(k_one_body, dim3(grid_blocks(ne2 + np2, 64)), dim3(64), 0, st, s->d, x->He, x->Se, x->Hp, x->Sp);
    hipLaunchKernelGGL(k_gam_ep, dim3(grid_blocks(ng1, 64)), dim3(64), 0, st, s->d, x->Gep);
    HIPCHK(hipGetLastError());
    rc = nrx_gam_ee_launch(s, x->Gee, st);
    if (!rc) rc = nrx_gam1_launch(s, x->G1, x->G1s, st);
    if (rc) return rc;
    // symmetric orthogonalisers (the overlaps do not change during the SCF)
    for (int k = 0; k < 2; ++k) {
        const int n = (int)(k == 0 ? ne : np);
        const size_t lds = roothan_lds(n);
        double *gw = nullptr;
        if (!lds) {
            rc = x->alloc(gw, 2 * (i64)n * n);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_orth, dim3(1), dim3(RH_BS), lds, st, n, k == 0 ? x->Se : x->Sp, k == 0 ? x->Xe : x->Xp, gw);
        HIPCHK(hipGetLastError());
    }
    if (x->nelec > 1) {
        rc = x->alloc(x->G2, ng2);
        rc |= x->alloc(x->G2s, ng2);
        if (rc) return rc;
        rc = neorx_gam2_ic1_dev(handle, x->G2, x->G2s, st);
        if (rc) return rc;
    }
    if (x->nelec > 2) {
        rc = x->alloc(x->G3, ng3);
        if (rc) return rc;
        rc = neorx_gam3_ic1_dev(handle, x->G3, st);
        if (rc) return rc;
        if (store & 2) {
            float *f = nullptr;
            rc = x->alloc(f, ng3);
            if (rc) return rc;
            hipLaunchKernelGGL(k_to_f32s, dim3(grid_blocks(ng3, 256)), dim3(256), 0, st, ng3, x->G3, f);
            x->G3f = f;
        }
        if (store & 4) {
            uint16_t *h = nullptr;
            rc = x->alloc(h, ng3);
            if (rc) return rc;
            hipLaunchKernelGGL(k_to_bf16, dim3(grid_blocks(ng3, 256)), dim3(256), 0, st, ng3, x->G3, h);
            x->G3h = h;
        }
        HIPCHK(hipGetLastError());
        if (store & 8) {
            // autotune the EG3 contraction on the stored tensors (the timing
            // does not depend on the density; the overlap stands in for it)
            HIPCHK(hipStreamSynchronize(st));
            const void *g[3] = {x->G3, x->G3f, x->G3h};
            for (int t = 0; t < 3; ++t)
                if (g[t]) {
                    int64_t best = 0;
                    double bms = 0.0, all[4] = {};
                    rc = neorx_autotune_eg3_dev(ne, np, t, x->Se, g[t], 3, &best, &bms, all);
                    if (rc) return rc;
                    if (g_verbose)
                        fprintf(stderr, "neorx: EG3 slab depth for storage %d: %lld (%.3f ms; 8/12/16/24: %.3f %.3f %.3f "
                                        "%.3f ms)\n",
                                t, (long long)best, bms, all[0], all[1], all[2], all[3]);
                }
        }
        if (!(store & 1)) {
            HIPCHK(hipStreamSynchronize(st));
            auto it = std::find(x->allocs.begin(), x->allocs.end(), (void *)x->G3);
            x->allocs.erase(it);
            HIPCHK(hipFree(x->G3));
            x->G3 = nullptr;
        }
        // the factors of the 5-particle term through underflow()
        rc = x->alloc(x->GeeU, ne4);
        rc |= x->alloc(x->G2sU, ng2);
        if (rc) return rc;
        hipLaunchKernelGGL(k_flush_copy, dim3(grid_blocks(ne4, 256)), dim3(256), 0, st, ne4, x->Gee, x->GeeU);
        hipLaunchKernelGGL(k_flush_copy, dim3(grid_blocks(ng2, 256)), dim3(256), 0, st, ng2, x->G2s, x->G2sU);
        HIPCHK(hipGetLastError());
    }
    // workspace
    x->nz_fe = 64;
    x->nch_fp = std::max<i64>(1, std::min<i64>(64, (ne4 + 4095) / 4096));
    x->nz_dot = 256;
    rc |= x->alloc(x->part2, x->nelec > 1 ? np2 * ne * 4 * ne : 1);
    rc |= x->alloc(x->part3, x->nelec > 2 ? np2 * ne * (ne + 1) : 1);
    rc |= x->alloc(x->part3ref, x->nelec > 2 ? np2 * ne * (ne + 1) : 1);
    rc |= x->alloc(x->DEref, ne2);
    rc |= x->alloc(x->DXi, ne2);
    rc |= x->alloc(x->DYi, ne2);
    rc |= x->alloc(x->FPx, 3 * np2);
    rc |= x->alloc(x->FEee, ne2);
    const i64 n4 = x->nelec > 2 ? ne4 : 1;
    rc |= x->alloc(x->A4, n4);
    rc |= x->alloc(x->R4, 4 * ne2);
    rc |= x->alloc(x->HA4, n4);
    rc |= x->alloc(x->Bb4, n4);
    rc |= x->alloc(x->GPb4, n4);
    rc |= x->alloc(x->t4, n4);
    rc |= x->alloc(x->GP4, n4);
    rc |= x->alloc(x->HB4, n4);
    rc |= x->alloc(x->B4, n4);
    rc |= x->alloc(x->Ab4, n4);
    rc |= x->alloc(x->FP4, np2);
    rc |= x->alloc(x->FE4, 2 * ne2);
    rc |= x->alloc(x->E4, 1);
    rc |= x->alloc(x->zpart, x->nz_fe * ne2);
    rc |= x->alloc(x->rows, std::max<i64>(np2 * x->nch_fp, x->nz_dot));
    rc |= x->alloc(x->FEs, ne2);
    rc |= x->alloc(x->SEs, ne2);
    rc |= x->alloc(x->FP23, 3 * np2);
    rc |= x->alloc(x->SP2, np2);
    if (rc) return rc;
    // one-electron systems never write the e-e-p partials: keep them zero
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, (i64)(3 * np2), 0.0, x->FP23);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, np2, 0.0, x->SP2);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, ne2, 0.0, x->FEee);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return NEORX_OK;
}

int scf_state(Xchf *x)
{
    if (x->DE) return NEORX_OK;
    const i64 ne2 = x->ne * x->ne, np2 = x->np * x->np, nmax = std::max(x->ne, x->np);
    int rc = 0;
    rc |= x->alloc(x->DE, ne2);
    rc |= x->alloc(x->DP, np2);
    rc |= x->alloc(x->DE0, ne2);
    rc |= x->alloc(x->DP0, np2);
    rc |= x->alloc(x->FE, ne2);
    rc |= x->alloc(x->FP, np2);
    rc |= x->alloc(x->FEx, std::max(ne2, np2));
    rc |= x->alloc(x->CE, ne2);
    rc |= x->alloc(x->CP, np2);
    rc |= x->alloc(x->evE, x->ne);
    rc |= x->alloc(x->evP, x->np);
    rc |= x->alloc(x->comps, 32);
    rc |= x->alloc(x->stat, 64);
    rc |= x->alloc(x->rwork, 2 * nmax * nmax);
    rc |= x->alloc(x->Fh, DIIS_MAX * ne2);
    rc |= x->alloc(x->Eh, DIIS_MAX * ne2);
    rc |= x->alloc(x->Bd, DIIS_MAX * DIIS_MAX);
    rc |= x->alloc(x->hist, 2);
    rc |= x->alloc(x->dw1, ne2);
    rc |= x->alloc(x->dw2, ne2);
    rc |= x->alloc(x->VwE, ne2);
    rc |= x->alloc(x->VwP, np2);
    rc |= x->alloc(x->vflag, 2);
    if (rc) return rc;
    HIPCHK(hipHostMalloc((void **)&x->hstat, 64 * sizeof(double), hipHostMallocDefault));
    return NEORX_OK;
}

int launch_roothan(Xchf *x, int n, const double *X, const double *F, int first, int last, double prefac, double *C,
                   double *ev, double *D, double *Vw, double *vflag)
{
    const size_t lds = roothan_lds4(n);
    hipLaunchKernelGGL(k_roothan, dim3(1), dim3(RH_BS), lds, x->st, n, X, F, first, last, prefac, C, ev, D,
                       lds ? nullptr : x->rwork, Vw, vflag);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

// stat layout: [0] diffE [1] diffP [2] DIIS error [3..16] comps of the
// electronic Fock build [17..30] comps of the nuclear one
int scf_iteration(Xchf *x, int dtype, bool diis, int nucst, int pass)
{
    hipStream_t st = x->st;
    const int ne = (int)x->ne, np = (int)x->np;
    const int nocc = x->nelec > 1 ? x->nelec / 2 : 1;
    const double pre = x->nelec > 1 ? 2.0 : 1.0;
    int rc = phase_a(x, x->DE, dtype, pass);
    if (!rc) rc = phase_b(x, x->DE, x->DP, x->FEx, x->FP, x->stat + 17);
    if (!rc) rc = launch_roothan(x, np, x->Xp, x->FP, nucst - 1, nucst - 1, 1.0, x->CP, x->evP, x->DP, x->VwP, x->vflag + 1);
    if (!rc) rc = phase_b(x, x->DE, x->DP, x->FE, x->FEx, x->stat + 3);
    if (rc) return rc;
    const double *F = x->FE;
    if (diis) {
        const size_t dl = diis_lds(ne);
        hipLaunchKernelGGL(k_diis, dim3(1), dim3(256), dl, st, ne, x->FE, x->DE, x->Se, x->Xe, x->Fh, x->Eh, x->Bd,
                           x->hist, x->FEx, x->stat + 2, x->dw1, x->dw2, dl ? 1 : 0);
        HIPCHK(hipGetLastError());
        F = x->FEx;
    }
    rc = launch_roothan(x, ne, x->Xe, F, 0, nocc - 1, pre, x->CE, x->evE, x->DE, x->VwE, x->vflag);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dendif, dim3(1), dim3(256), 0, st, (i64)ne * ne, x->DE, x->DE0, x->stat);
    hipLaunchKernelGGL(k_dendif, dim3(1), dim3(256), 0, st, (i64)np * np, x->DP, x->DP0, x->stat + 1);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

} // namespace

// ------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------
extern "C" int neorx_xchf_create(int64_t handle, int64_t nelec, int64_t store, int64_t *ctx)
{
    int rc = ensure_init();
    if (rc) return rc;
    System *s = get_sys(handle);
    if (!s) return fail(NEORX_E_ARG, "unknown system handle %lld", (long long)handle);
    if (!ctx || nelec < 1) return fail(NEORX_E_ARG, "neorx_xchf_create: nelec >= 1 and a handle pointer are required");
    if (s->d.nebf1 != s->d.nebf) return fail(NEORX_E_ARG, "neorx_xchf_create takes a one-basis handle");
    if (nelec > 2 && !(store & 7)) return fail(NEORX_E_ARG, "store must keep GAM_3 in at least one storage type");
    if (s->maxl > 2)
        return fail(NEORX_E_UNSUPPORTED, "the device XCHF SCF supports l <= 2 (this basis has l = %d); the integral "
                                         "builders and contractions take f functions",
                    s->maxl);
    if (s->d.nebf > RH_MAXN || s->d.npbf > RH_MAXN)
        return fail(NEORX_E_UNSUPPORTED, "ROOTHAN on the device supports at most %d basis functions", RH_MAXN);
    Xchf *x = new Xchf();
    x->s = s;
    x->ne = s->d.nebf;
    x->np = s->d.npbf;
    x->nelec = (int)nelec;
    // Non-blocking: a context's stream takes no implicit dependency on the
    // legacy stream, so another thread's legacy-stream work cannot
    // invalidate a graph capture on it (SCFs of several contexts run
    // concurrently from host threads).  Every host copy of the context's
    // buffers is therefore ordered on this stream (copy_on).
    if (hipStreamCreateWithFlags(&x->st, hipStreamNonBlocking) != hipSuccess) {
        delete x;
        return fail(NEORX_E_HIP, "hipStreamCreateWithFlags failed");
    }
    rc = xchf_build(x, handle, store ? store : 1);
    if (rc) {
        delete x;
        return rc;
    }
    std::lock_guard<std::mutex> lk(g_xmu);
    const i64 h = g_xnext++;
    g_xchf[h] = x;
    *ctx = h;
    return NEORX_OK;
}

extern "C" int neorx_xchf_destroy(int64_t ctx)
{
    Xchf *x = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_xmu);
        auto it = g_xchf.find(ctx);
        if (it == g_xchf.end()) return fail(NEORX_E_ARG, "unknown XCHF context %lld", (long long)ctx);
        x = it->second;
        g_xchf.erase(it);
    }
    (void)hipStreamSynchronize(x->st);
    delete x;
    return NEORX_OK;
}

extern "C" int neorx_xchf_integrals(int64_t ctx, double *enuc, double *se, double *sp, double *gam_ecore,
                                    double *gam_pcore, double *gam_ep, double *gam_ee)
{
    Xchf *x = get_x(ctx);
    if (!x) return fail(NEORX_E_ARG, "unknown XCHF context %lld", (long long)ctx);
    const i64 ne2 = x->ne * x->ne, np2 = x->np * x->np;
    int rc = NEORX_OK;
    HIPCHK(hipStreamSynchronize(x->st));
    if (enuc) *enuc = x->enuc;
    if (se && (rc = copy_on(x->st, se, x->Se, ne2 * 8, hipMemcpyDeviceToHost))) return rc;
    if (sp && (rc = copy_on(x->st, sp, x->Sp, np2 * 8, hipMemcpyDeviceToHost))) return rc;
    if (gam_ecore && (rc = copy_on(x->st, gam_ecore, x->He, ne2 * 8, hipMemcpyDeviceToHost))) return rc;
    if (gam_pcore && (rc = copy_on(x->st, gam_pcore, x->Hp, np2 * 8, hipMemcpyDeviceToHost))) return rc;
    if (gam_ep && (rc = copy_on(x->st, gam_ep, x->Gep, ne2 * np2 * 8, hipMemcpyDeviceToHost))) return rc;
    if (gam_ee && (rc = copy_on(x->st, gam_ee, x->Gee, ne2 * ne2 * 8, hipMemcpyDeviceToHost))) return rc;
    return NEORX_OK;
}

extern "C" int neorx_xchf_fock_dev(int64_t ctx, const double *d_de, const double *d_dp, int64_t prec,
                                   double *d_focke, double *d_fockp, double *d_comps, void *stream)
{
    Xchf *x = get_x(ctx);
    if (!x) return fail(NEORX_E_ARG, "unknown XCHF context %lld", (long long)ctx);
    if (!d_de || !d_dp || !d_focke || !d_fockp || !d_comps) return fail(NEORX_E_ARG, "null argument");
    const int dtype = (int)(prec & 15), pass = (int)(prec >> 4);
    if (prec < 0 || dtype > 2 || pass > 2)
        return fail(NEORX_E_ARG, "prec %lld: storage not in 0..2 or pass not in 0..2", (long long)prec);
    if (x->nelec > 2 && pass == G3_REFRESH && dtype != NRX_F64)
        return fail(NEORX_E_ARG, "a reference build reads GAM_3 in FP64 (prec 0x10)");
    hipStream_t keep = x->st;
    if (stream) x->st = (hipStream_t)stream;
    int rc = phase_a(x, d_de, dtype, pass);
    if (!rc) rc = phase_b(x, d_de, d_dp, d_focke, d_fockp, d_comps);
    x->st = keep;
    return rc;
}

extern "C" int neorx_convert_bf16_dev(int64_t n, const double *d_src, uint16_t *d_dst, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!d_src || !d_dst))) return fail(NEORX_E_ARG, "bad conversion arguments");
    if (n == 0) return NEORX_OK;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n, d_src, d_dst);
    HIPCHK(hipGetLastError());
    return NEORX_OK;
}

extern "C" int neorx_gemm_f64_dev(int64_t M, int64_t N, int64_t K, const double *d_a, int64_t sam, int64_t sak,
                                  const double *d_b, int64_t sbk, int64_t sbn, double *d_c, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (M < 1 || N < 1 || K < 1 || !d_a || !d_b || !d_c) return fail(NEORX_E_ARG, "GEMM: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    const i64 zcap = 64;
    double *z = nullptr;
    HIPCHK(hipMallocAsync((void **)&z, (size_t)(zcap * M * N) * 8, st));
    rc = gemm(st, M, N, K, d_a, sam, sak, d_b, sbk, sbn, d_c, z, zcap);
    (void)hipFreeAsync(z, st);
    return rc;
}

extern "C" int neorx_roothan_dev(int64_t n, const double *d_s, const double *d_f, int64_t kstart, int64_t klast,
                                 double prefac, double *d_c, double *d_ev, double *d_d, void *stream)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (n < 1 || n > RH_MAXN) return fail(NEORX_E_UNSUPPORTED, "ROOTHAN: n=%lld not in 1..%d", (long long)n, RH_MAXN);
    if (kstart < 1 || klast < kstart || klast > n) return fail(NEORX_E_ARG, "ROOTHAN: occupied range %lld..%lld",
                                                               (long long)kstart, (long long)klast);
    if (!d_s || !d_f || !d_c || !d_ev || !d_d) return fail(NEORX_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    double *w = nullptr;
    const size_t lds = roothan_lds((int)n), lds4 = roothan_lds4((int)n);
    HIPCHK(hipMallocAsync((void **)&w, (size_t)(3 * n * n) * 8, st));
    hipLaunchKernelGGL(k_orth, dim3(1), dim3(RH_BS), lds, st, (int)n, d_s, w, lds ? nullptr : w + n * n);
    hipLaunchKernelGGL(k_roothan, dim3(1), dim3(RH_BS), lds4, st, (int)n, w, d_f, (int)kstart - 1, (int)klast - 1,
                       prefac, d_c, d_ev, d_d, lds4 ? nullptr : w + n * n, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    (void)hipFreeAsync(w, st);
    if (e != hipSuccess) return fail(NEORX_E_HIP, "ROOTHAN launch failed: %s", hipGetErrorString(e));
    return NEORX_OK;
}

extern "C" int neorx_xchf_scf(int64_t ctx, const double *ce_guess, int64_t nucst, int64_t maxit, double tole,
                              double tolp, int64_t mode, int64_t flags, double sw_hi, double sw_lo, double *comps_init,
                              double *comps, double *ce, double *cp, double *ev_e, double *ev_p, double *trace,
                              int64_t *niter)
{
    Xchf *x = get_x(ctx);
    if (!x) return fail(NEORX_E_ARG, "unknown XCHF context %lld", (long long)ctx);
    const int ne = (int)x->ne, np = (int)x->np;
    if (nucst < 1 || nucst > np) return fail(NEORX_E_ARG, "NUCST=%lld not in 1..%d", (long long)nucst, np);
    if (maxit < 1) return fail(NEORX_E_ARG, "maxit < 1");
    if (mode < 0 || mode > 3) return fail(NEORX_E_ARG, "mode %lld not in 0..3", (long long)mode);
    const bool use_graph = flags & 1, diis = flags & 2;
    if (x->nelec > 2) {
        if (!x->G3 && mode == 0) return fail(NEORX_E_ARG, "FP64 mode needs GAM_3 in FP64 (store bit 1)");
        if ((mode == 1 || mode == 2) && !x->G3f) return fail(NEORX_E_ARG, "mixed mode needs GAM_3 in FP32 (store bit 2)");
        if (mode >= 2 && !x->G3h) return fail(NEORX_E_ARG, "mode %lld needs GAM_3 in BF16 (store bit 4)", (long long)mode);
        if (mode >= 1 && !x->G3) return fail(NEORX_E_ARG, "mixed mode switches to FP64: store bit 1 is needed");
    }
    int rc = scf_state(x);
    if (rc) return rc;
    hipStream_t st = x->st;
    const int nocc = x->nelec > 1 ? x->nelec / 2 : 1;
    const double pre = x->nelec > 1 ? 2.0 : 1.0;
    const i64 ne2 = (i64)ne * ne, np2 = (i64)np * np;
    // every SCF starts its diagonalisations from the identity
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (i64)2, 0.0, x->vflag);
    HIPCHK(hipGetLastError());
    // initial electronic density: read_elec_density (GAM_IO.f:247-330) from
    // the guess orbitals, or guess_elec (xcscf.f, ROOTHAN on GAM_ecore)
    if (ce_guess) {
        HIPCHK(hipMemcpyAsync(x->CE, ce_guess, ne2 * 8, hipMemcpyHostToDevice, st));
        std::vector<double> D(ne2, 0.0);
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) {
                double v = 0.0;
                for (int k = 0; k < nocc; ++k) v += pre * ce_guess[k * ne + i] * ce_guess[k * ne + j];
                D[j * ne + i] = v;
            }
        HIPCHK(hipMemcpyAsync(x->DE, D.data(), ne2 * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    } else {
        // GAM_ecore as a column-major Fock matrix: fock(i,j) = GAM_ecore[i][j]
        std::vector<double> h(ne2), f(ne2);
        rc = copy_on(st, h.data(), x->He, ne2 * 8, hipMemcpyDeviceToHost);
        if (rc) return rc;
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) f[j * ne + i] = h[i * ne + j];
        rc = copy_on(st, x->FEx, f.data(), ne2 * 8, hipMemcpyHostToDevice);
        if (rc) return rc;
        rc = launch_roothan(x, ne, x->Xe, x->FEx, 0, nocc - 1, pre, x->CE, x->evE, x->DE, x->VwE, x->vflag);
        if (rc) return rc;
    }
    // guess_prot (xcscf.f): ROOTHAN on GAM_pcore, orbital NUCST occupied
    {
        std::vector<double> h(np2), f(np2);
        rc = copy_on(st, h.data(), x->Hp, np2 * 8, hipMemcpyDeviceToHost);
        if (rc) return rc;
        for (int i = 0; i < np; ++i)
            for (int j = 0; j < np; ++j) f[j * np + i] = h[i * np + j];
        rc = copy_on(st, x->FEx, f.data(), np2 * 8, hipMemcpyHostToDevice);
        if (rc) return rc;
        rc = launch_roothan(x, np, x->Xp, x->FEx, (int)nucst - 1, (int)nucst - 1, 1.0, x->CP, x->evP, x->DP, x->VwP,
                            x->vflag + 1);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, ne2, 0.0, x->DE0);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, np2, 0.0, x->DP0);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (i64)2, 0.0, x->hist);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (i64)64, 0.0, x->stat);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));

    // precision of GAM_3 per iteration: mode 0 FP64; mode 1 FP32 until the
    // switch signal (DIIS error with DIIS, else max|dDE|) < sw_lo; mode 2
    // BF16 until < sw_hi, then FP32 until < sw_lo; FP64 from then on.
    // Mode 3 (incremental): BF16 builds until the signal < sw_hi; then a
    // reference build (from the FP32 copy if one is held, else FP64) whose
    // rows are kept, and BF16-stored increments from it (T(DE - DEref,
    // DE + DEref), whose storage error shrinks with the distance to the
    // reference); an FP64 reference once the signal < sw_lo, and a last FP64
    // build when the increments meet the convergence test.  Every mode
    // converges only on an FP64 build.
    int dtype = mode == 0 ? NRX_F64 : mode == 1 ? NRX_F32 : NRX_BF16;
    int pass = G3_FULL;
    double ref_sig = 0.0; // the switch signal when the last reference was asked for
    double eold = 0.0;
    i64 it = 0;
    bool conv = false;
    for (it = 1; it <= maxit && !conv; ++it) {
        const auto t0 = std::chrono::steady_clock::now();
        const int key = ((dtype * 3 + pass) * 2) + (diis ? 1 : 0);
        if (use_graph) {
            int slot = -1;
            for (int g = 0; g < Xchf::NGRAPH; ++g)
                if (x->graph_key[g] == key) slot = g;
            if (slot < 0) {
                for (int g = 0; g < Xchf::NGRAPH && slot < 0; ++g)
                    if (x->graph_key[g] < 0) slot = g;
                if (slot < 0) slot = 0;
                if (x->graph[slot]) (void)hipGraphExecDestroy(x->graph[slot]);
                x->graph[slot] = nullptr;
                x->graph_key[slot] = -1;
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                rc = scf_iteration(x, dtype, diis, (int)nucst, pass);
                hipError_t e = hipStreamEndCapture(st, &g);
                if (rc) return rc;
                if (e != hipSuccess) return fail(NEORX_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
                e = hipGraphInstantiate(&x->graph[slot], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) return fail(NEORX_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
                x->graph_key[slot] = key;
            }
            HIPCHK(hipGraphLaunch(x->graph[slot], st));
        } else {
            rc = scf_iteration(x, dtype, diis, (int)nucst, pass);
            if (rc) return rc;
        }
        HIPCHK(hipMemcpyAsync(x->hstat, x->stat, 32 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const double *h = x->hstat;
        const double etot = h[3], dife = h[0], difp = h[1], derr = h[2];
        if (it == 1 && comps_init) std::copy(h + 17, h + 31, comps_init);
        if (trace) {
            double *r = trace + (it - 1) * 8;
            r[0] = etot;
            r[1] = etot - eold;
            r[2] = dife;
            r[3] = difp;
            r[4] = diis ? derr : dife;
            r[5] = dtype;
            r[6] = ms;
            r[7] = pass;
        }
        eold = etot;
        // converged (xcscf.f:515-519) only on an iteration with FP64 GAM_3
        const bool crit = (np > 1) ? (dife < tole && difp < tolp) : (dife < tole);
        conv = crit && (x->nelec <= 2 || (dtype == NRX_F64 && pass != G3_INC));
        const double sig = diis ? derr : dife;
        if (mode == 3) {
            // next build: a reference (FP64) or an increment (BF16)
            const bool want_ref = pass == G3_INC ? (crit || (ref_sig >= sw_lo && sig < sw_lo))
                                                 : (pass == G3_FULL && dtype == NRX_BF16 && (sig < sw_hi || crit));
            if (want_ref) {
                // the first reference (leaving the full BF16 builds above
                // sw_hi) reads the FP32 copy when one is held: its rounding
                // (2^-24 relative) is far below the DIIS error at that point.
                // Every later reference, and so every build that may
                // converge, reads FP64: increments follow any reference, and
                // an increment meeting the criteria asks for a reference.
                const bool f32_ref = pass == G3_FULL && !crit && x->G3f;
                ref_sig = sig;
                dtype = f32_ref ? NRX_F32 : NRX_F64;
                pass = G3_REFRESH;
            } else if (pass == G3_REFRESH || pass == G3_INC) {
                dtype = NRX_BF16;
                pass = G3_INC;
            }
        } else {
            if (dtype == NRX_BF16 && sig < sw_hi) dtype = NRX_F32;
            if (dtype == NRX_F32 && sig < sw_lo) dtype = NRX_F64;
        }
    }
    if (niter) *niter = it - 1;
    HIPCHK(hipStreamSynchronize(st));
    if (comps) std::copy(x->hstat + 3, x->hstat + 17, comps);
    if (ce && (rc = copy_on(x->st, ce, x->CE, ne2 * 8, hipMemcpyDeviceToHost))) return rc;
    if (cp && (rc = copy_on(x->st, cp, x->CP, np2 * 8, hipMemcpyDeviceToHost))) return rc;
    if (ev_e && (rc = copy_on(x->st, ev_e, x->evE, ne * 8, hipMemcpyDeviceToHost))) return rc;
    if (ev_p && (rc = copy_on(x->st, ev_p, x->evP, np * 8, hipMemcpyDeviceToHost))) return rc;
    if (!conv) return fail(NEORX_E_ARG, "SCF did not converge in %lld iterations", (long long)maxit);
    return NEORX_OK;
}
