/*
 * neorx.h -- C ABI of libneorx_hip.so, the MI355X (gfx950) engine for the
 * geminal-integral and Fock-build path of NEO-RXCHF.
 *
 * Conventions (identical to the reference's Fortran interface, built with
 * -integer-size 64, source/comp:7):
 *   - every integer is int64_t, every real is double;
 *   - matrices are column-major: ELCAM(npebf,3) is elcam[k*npebf+i],
 *     cat(3,nat) is cat[3*iat+k], DE(nebf,nebf) is de[j*nebf+i];
 *   - index arrays KPESTR/KPEEND/AMPEB2C hold 1-based primitive/contracted
 *     indices exactly as the reference builds them (make_KPE,
 *     source/xcneo_HYBRID.f:1269-1307);
 *   - contracted tensors use the reference packings
 *     (source/CS_Integrals/Spot_Check/pack_GAM.f:11-240).
 * Host-pointer entry points copy inputs to HBM, run, and copy results back;
 * "_dev" entry points take device pointers and a hipStream_t (as void*), and
 * return without synchronising.
 *
 * Every function returns 0 on success and a negative NEORX_E_* code on
 * failure; neorx_last_error() returns the message.  Nothing falls back to the
 * CPU: without a usable gfx950 device every entry point fails.
 */
#ifndef NEORX_H
#define NEORX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEORX_OK 0
#define NEORX_E_NODEVICE (-1)
#define NEORX_E_ARG (-2)
#define NEORX_E_HIP (-3)
#define NEORX_E_NOMEM (-4)
#define NEORX_E_UNSUPPORTED (-5)

/* ---- library --------------------------------------------------------- */
int neorx_init(int64_t device);
int neorx_finalize(void);
const char *neorx_last_error(void);
/* name of the device the library runs on (e.g. "gfx950") */
const char *neorx_device_arch(void);

/* Resident tensors (no counterpart in the reference: the reference keeps
 * its in-core tensors in host memory).  The host-pointer builders
 * neorx_gam2_ic1, neorx_gam3_ic1 and neorx_gam4_icr fill the caller's array
 * and keep their device copy, keyed by the array's address and size; a later
 * host-pointer call given the same array as an input tensor (EG2IC1, EG3IC1,
 * EG4ICR, the XCUHF contractions, GAM4_ICR) reads that copy instead of
 * copying the tensor to the device again.  4096 entries sampled over the
 * array are compared with the values kept at registration on every use, and
 * a difference drops the copy.  Copies use at most 60 % of the device memory
 * (the oldest are released first).  NEORX_RESIDENT=0 in the environment
 * turns this off.  neorx_release_resident frees every copy (neorx_finalize
 * does too).  neorx_transfer_bytes reports the bytes of tensors these entry
 * points copied host->device and device->host since the process started. */
int neorx_release_resident(void);
int neorx_transfer_bytes(int64_t *h2d_bytes, int64_t *d2h_bytes);

/* ---- system handle: basis, nuclei and geminal resident in HBM ---------- */
/* Arguments are the basis and system arrays that the reference passes to
 * GAM2_IC1 (source/gam2_IC1_OMP.f:2-37); the handle keeps device copies. */
int neorx_system_create(int64_t nebf, int64_t npebf, int64_t npbf, int64_t nat, int64_t ngtg, double pmass,
                        const double *cat, const double *zan, const double *bcoef, const double *gamma,
                        const int64_t *kpestr, const int64_t *kpeend, const int64_t *ampeb2c, const double *agebfcc,
                        const double *agnbfcc, const double *elcex, const double *nucex, const int64_t *elcam,
                        const int64_t *nucam, const double *elcbfc, const double *nucbfc, int64_t *handle);
int neorx_system_destroy(int64_t handle);

/* ---- primitive tuples -------------------------------------------------- */
/* One fully geminal-summed primitive tuple value per entry:
 *   cls 1 (e-p)     : xcalc_GAM1_MD, source/gam_1_OMP.f:398   -> ansE, ansS
 *   cls 2 (e-e-p)   : xcalc_GAM2_MD, source/gam_2_OMP.f:546   -> ansE, ansS
 *   cls 3 (e-e-e-p) : xcalc_GAM3_MD, source/gam_3_OMP.f:617   -> ansE
 *   cls 4 (e-e-p)   : RXCHFne_xcalc_GAM2_MD, rxchfne/rxchfne_gam_2_OMP.f:2 -> ansE1, ansE2
 *   cls 9 (e-e-p)   : RXCHF_xcalc_GAM2_MD, rxchf/rxchf_gam_2_OMP.f:546 (with neorx_prim_list3_dev)
 *                     -> ansE1, ansS, ansE2
 *   cls 10 (e-e-e-p): RXCHF_xcalc_GAM3_MD, rxchf/rxchf_gam_3_OMP.f:617 -> ans1, ans2
 * d_idx holds 0-based primitive indices per tuple, 2*(n+1) int64 each for n
 * electrons (n = cls, 2 for cls 4), in the order (ie1..ie_n, ip, je1..je_n, jp). */
int neorx_prim_list_dev(int64_t handle, int64_t cls, int64_t ntup, const int64_t *d_idx, double *d_ansE,
                        double *d_ansS, void *stream);
/* Same values for the contiguous slice [start, start+count) of the full
 * primitive enumeration, ordered C-row-major as (ip,jp,ie1,je1,ie2,je2,...)
 * over primitives re-ordered by angular-momentum class and exponent; the
 * re-ordering is returned by neorx_system_order(). */
int neorx_prim_range_dev(int64_t handle, int64_t cls, int64_t start, int64_t count, double *d_ansE, double *d_ansS,
                         void *stream);
/* eorder[npebf], porder[npbf]: position k of the enumeration holds original
 * primitive eorder[k] (0-based). */
int neorx_system_order(int64_t handle, int64_t *eorder, int64_t *porder);

/* ---- contracted builders (host arrays, reference semantics) ------------ */
/* GAM1_OMP_MD, source/gam_1_OMP.f:2: GAM_1, GAM_1s in pack_4D(nebf,nebf,npbf)
 * order (returned here instead of being written to GAM_1.ufm/GAM_1s.ufm). */
int neorx_gam1(int64_t handle, double *gam1, double *gam1s);
/* GAM2_IC1, source/gam2_IC1_OMP.f:2: GM2ICR, GM2sICR (ng2 = nebf^4 npbf^2) */
int neorx_gam2_ic1(int64_t handle, double *gm2icr, double *gm2sicr);
/* GAM3_IC1, source/gam3_IC1_OMP.f:2: GM3ICR (ng3 = nebf^6 npbf^2) */
int neorx_gam3_ic1(int64_t handle, double *gm3icr);
/* calc_GAM_ee, source/gam_ee.f:2: GAM_ee (ngee = nebf^4), returned instead of
 * GAM_ee.ufm */
int neorx_gam_ee(int64_t handle, double *gam_ee);
/* GAM4_ICR, source/gam4_ICR_OMP.f:2: GM4ICR (ng4 = nebf^8 npbf^2) from
 * GAM_2s (= GM2sICR) and GAM_ee */
int neorx_gam4_icr(int64_t nebf, int64_t npbf, const double *gam_2s, const double *gam_ee, double *gm4icr);

/* Device-resident builders: outputs are device arrays of the same sizes. */
int neorx_gam2_ic1_dev(int64_t handle, double *d_gm2icr, double *d_gm2sicr, void *stream);
int neorx_gam3_ic1_dev(int64_t handle, double *d_gm3icr, void *stream);

/* ---- Fock contractions (reference semantics) ---------------------------- */
/* EG2IC1, source/gam2_IC1_OMP.f:339: FE, FP accumulate; SE, SP overwritten;
 * E_gam2, S_gam2 overwritten. */
int neorx_eg2ic1(int64_t nebf, int64_t npbf, const double *de, const double *dp, const double *gm2icr,
                 const double *gm2sicr, double *fe, double *se, double *fp, double *sp, double *e_gam2,
                 double *s_gam2);
/* EG3IC1, source/gam3_IC1_OMP.f:433: focke, fockp accumulate; E overwritten */
int neorx_eg3ic1(int64_t nebf, int64_t npbf, const double *de, const double *dp, const double *gm3icr, double *focke,
                 double *fockp, double *e_gam3);
/* EG4ICR, source/gam4_ICR_OMP.f:273 */
int neorx_eg4icr(int64_t nebf, int64_t npbf, const double *de, const double *dp, const double *gm4icr, double *focke,
                 double *fockp, double *e_gam4);

/* Device-resident contractions (all pointers device memory; d_e is one
 * double, d_se/d_sp/d_s may be NULL when gm2s is NULL). */
int neorx_eg2ic1_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const double *d_gm2,
                     const double *d_gm2s, double *d_fe, double *d_se, double *d_fp, double *d_sp, double *d_e,
                     double *d_s, void *stream);
int neorx_eg3ic1_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const double *d_gm3,
                     double *d_fe, double *d_fp, double *d_e, void *stream);

/* neorx_eg3ic1_dev without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a static device array of 2^21 doubles (16 MiB per
 * device, held until the process ends), so nebf and npbf are limited to
 * npbf^2*nebf*(nebf+1) + npbf^2 <= 2^21 (NEORX_E_ARG beyond).  Same results
 * as neorx_eg3ic1_dev; enqueued on the stream and not synchronised.  The
 * array is shared by every call on a device: calls that overlap on one device
 * race.  The device must have been set up by neorx_init. */
int neorx_eg3ic1_noalloc_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp,
                             const double *d_gm3, double *d_fe, double *d_fp, double *d_e, void *stream);

/* EG4ICR on device arrays. */
int neorx_eg4icr_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const double *d_gm4,
                     double *d_fe, double *d_fp, double *d_e, void *stream);

/* ---- mixed precision ------------------------------------------------------
 * No counterpart in the reference, whose Fock build is FP64 throughout.  The
 * in-core tensors are stored in FP32 (half the HBM bytes of every Fock build);
 * products and sums stay FP64, so the only rounding is that of the stored
 * values.  Same semantics as the FP64 entry points above. */
int neorx_convert_f32_dev(int64_t n, const double *d_src, float *d_dst, void *stream);
/* BF16 copy of an FP64 tensor: the upper 16 bits of the FP32 value, rounded
 * to nearest even (for the storage types that take dtype 2). */
int neorx_convert_bf16_dev(int64_t n, const double *d_src, uint16_t *d_dst, void *stream);
int neorx_eg2ic1_f32_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const float *d_gm2,
                         const float *d_gm2s, double *d_fe, double *d_se, double *d_fp, double *d_sp, double *d_e,
                         double *d_s, void *stream);
int neorx_eg3ic1_f32_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const float *d_gm3,
                         double *d_fe, double *d_fp, double *d_e, void *stream);
int neorx_eg4icr_f32_dev(int64_t nebf, int64_t npbf, const double *d_de, const double *d_dp, const float *d_gm4,
                         double *d_fe, double *d_fp, double *d_e, void *stream);
/* Switch back to FP64 (no counterpart in the reference).  The reference's SCF
 * exposes no DIIS error to these routines; the signal the library sees on
 * every Fock build is the density itself, so the switch compares successive
 * electronic densities: once max|DE - DE_previous| < threshold the caller
 * should use the FP64 entry points, and *use_fp64 stays 1 from then on.
 * d_state: device array of nebf*nebf + 2 doubles, zeroed before the first
 * call; *delta (may be NULL) receives max|DE - DE_previous| (+inf on the
 * first call).  Synchronises the stream. */
int neorx_precision_switch_dev(int64_t nebf, const double *d_de, double *d_state, double threshold,
                               int64_t *use_fp64, double *delta, void *stream);

/* ---- XCUHF contractions (reference semantics, source/xcuhf/) ------------ *
 * nae, nbe: alpha / beta electron counts; dae, dbe: alpha / beta densities.
 * FAE_* add into fae (the alpha Fock matrix; the beta one is the same call
 * with the spins swapped, xcuhf_fock.f:221, :241); UFP_* add into fp;
 * sae2 / sp2 and the energies are overwritten. */
/* FAE_OMG2, source/xcuhf/xcuhf_omg2.f:2 */
int neorx_fae_omg2(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                   const double *dp, const double *gm2, const double *gm2s, double *fae, double *sae2, double *e,
                   double *s);
/* UFP_OMG2, source/xcuhf/xcuhf_omg2.f:302 */
int neorx_ufp_omg2(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                   const double *dp, const double *gm2, const double *gm2s, double *fp, double *sp2, double *e,
                   double *s);
/* FAE_OMG3 / UFP_OMG3, source/xcuhf/xcuhf_omg3.f:2, :284 */
int neorx_fae_omg3(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                   const double *dp, const double *gm3, double *fae, double *e);
int neorx_ufp_omg3(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                   const double *dp, const double *gm3, double *fp, double *e);
/* FAE_OMG4 / UFP_OMG4, source/xcuhf/xcuhf_omg4.f:2, :512 */
int neorx_fae_omg4(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                   const double *dp, const double *gm4, double *fae, double *e);
int neorx_ufp_omg4(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *dae, const double *dbe,
                   const double *dp, const double *gm4, double *fp, double *e);
/* OMG4_ICR, source/xcuhf/xcuhf_omg4.f:1080: the unsymmetrised 5-particle
 * tensor (calc_GAM_4_integral per entry) from GAM_2s and GAM_ee */
int neorx_omg4_icr(int64_t nebf, int64_t npbf, const double *gam_2s, const double *gam_ee, double *gm4);
/* device-array forms (synchronise the stream) */
int neorx_fae_omg2_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                       const double *d_dp, const double *d_gm2, const double *d_gm2s, double *d_fae, double *d_sae2,
                       double *d_e, double *d_s, void *stream);
int neorx_ufp_omg2_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                       const double *d_dp, const double *d_gm2, const double *d_gm2s, double *d_fp, double *d_sp2,
                       double *d_e, double *d_s, void *stream);
int neorx_fae_omg3_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                       const double *d_dp, const double *d_gm3, double *d_fae, double *d_e, void *stream);
int neorx_ufp_omg3_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                       const double *d_dp, const double *d_gm3, double *d_fp, double *d_e, void *stream);
int neorx_fae_omg4_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                       const double *d_dp, const double *d_gm4, double *d_fae, double *d_e, void *stream);
int neorx_ufp_omg4_dev(int64_t nae, int64_t nbe, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                       const double *d_dp, const double *d_gm4, double *d_fp, double *d_e, void *stream);

/* ---- RXCHF-ne (reference semantics, source/rxchfne/) ---------------------
 * The RXCHF-ne e-e-p tuple RXCHFne_xcalc_GAM2_MD is class 4 of
 * neorx_prim_list_dev / neorx_prim_range_dev (outputs ansE1, ansE2 in the E
 * and S arrays; electron 1 is the special electron). */
/* RXCHFne_GAM2_IC1, source/rxchfne/rxchfne_gam2_IC1_OMP.f:2: GM2ICR
 * (contracted ansE1, not symmetrised) and GM2exICR ((x12 + x21)/2 of ansE2) */
int neorx_rxchfne_gam2_ic1(int64_t handle, double *gm2icr, double *gm2exicr);
int neorx_rxchfne_gam2_ic1_dev(int64_t handle, double *d_gm2icr, double *d_gm2exicr, void *stream);
/* RXCHFne_FAE_OMG2 / _FP_OMG2 / _FBE_OMG2, source/rxchfne/rxchfne_omg2.f:2,
 * :212, :409 (dae: regular, dbe: special electron density); F accumulates,
 * E overwritten */
int neorx_rxchfne_fae_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                           const double *gm2, double *fae, double *e);
int neorx_rxchfne_fp_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                          const double *gm2, double *fp, double *e);
int neorx_rxchfne_fbe_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                           const double *gm2, double *fbe, double *e);
/* The three closed-shell RXCHF-ne contractions above in one pass over GM2
 * (device pointers), with GM2 stored as dtype 0 (FP64), 1 (FP32, FP64 sums)
 * or 2 (BF16, FP32 sums; neorx_convert_bf16_dev).  d_fae (nebf^2), d_fp
 * (npbf^2) and d_fbe (nebf^2) are added into, each may be null; d_e = the
 * energy of all three (overwritten).  Stream-ordered scratch; not
 * synchronised.  No counterpart of its own in the reference: it computes
 * what RXCHFne_FAE_OMG2, _FP_OMG2 and _FBE_OMG2 compute (rxchfne_omg2.f:2,
 * :212, :409), from one read of the tensor, in the storage type asked for. */
int neorx_rxchfne_omg2_dev(int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                           const double *d_dp, const void *d_gm2, int dtype, double *d_fae, double *d_fp,
                           double *d_fbe, double *d_e, void *stream);
/* RXCUHFne_FE_OMG2 / _FP_OMG2 / _FBE_OMG2, source/rxchfne/rxcuhfne_omg2.f:2,
 * :239, :445 (daalp: regular alpha, datot: regular total density) */
int neorx_rxcuhfne_fe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot, const double *dbe,
                           const double *dp, const double *gm2, const double *gm2ex, double *faalp, double *fabet,
                           double *e);
int neorx_rxcuhfne_fp_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot, const double *dbe,
                           const double *dp, const double *gm2, const double *gm2ex, double *fp, double *e);
int neorx_rxcuhfne_fbe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot, const double *dbe,
                            const double *dp, const double *gm2, const double *gm2ex, double *fbe, double *e);
/* RXCHFne_FAE_GAMee, source/rxchfne/rxchfne_hfterms.f:2 */
int neorx_rxchfne_fae_gamee(int64_t nebf, const double *gam_ee, const double *dae, double *fae, double *e);

/* ---- RXCHF, one special electron (reference semantics, source/rxchf/) -----
 * The RXCHF e-e-p tuple RXCHF_xcalc_GAM2_MD (rxchf/rxchf_gam_2_OMP.f:546) is
 * class 9 of neorx_prim_list3_dev: outputs (ansE1, ansS, ansE2); electron 1
 * is the special electron. */
/* RXCHF_GAM2_IC1, source/rxchf/rxchf_gam2_IC1_OMP.f:2: GM2_1ICR (contracted
 * ansE1), GM2_2ICR and GM2sICR ((x12 + x21)/2 of the contracted ansE2, ansS) */
int neorx_rxchf_gam2_ic1(int64_t handle, double *gm2_1, double *gm2_2, double *gm2s);
int neorx_rxchf_gam2_ic1_dev(int64_t handle, double *d_gm2_1, double *d_gm2_2, double *d_gm2s, void *stream);
/* RXCHF_FAE_OMG2 / _FP_OMG2 / _FBE_OMG2, source/rxchf/rxchf_omg2.f:2, :231,
 * :449 (dae: regular, dbe: special electron density): F and SF accumulate,
 * e and s (the energy and overlap terms) overwritten */
int neorx_rxchf_fae_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                         const double *gm2_1, const double *gm2_2, const double *gm2s, double *fae, double *sae,
                         double *e, double *s);
int neorx_rxchf_fp_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                        const double *gm2_1, const double *gm2_2, const double *gm2s, double *fp, double *sp,
                        double *e, double *s);
int neorx_rxchf_fbe_omg2(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                         const double *gm2_1, const double *gm2_2, const double *gm2s, double *fbe, double *sbe,
                         double *e, double *s);
/* the same on device arrays: which 0 FAE, 1 FP, 2 FBE; d_es[0] = E,
 * d_es[1] = S; stream-ordered scratch, not synchronised */
int neorx_rxchf_omg2_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                         const double *d_dp, const double *d_gm2_1, const double *d_gm2_2, const double *d_gm2s,
                         double *d_f, double *d_sf, double *d_es, void *stream);
/* The RXCHF e-e-e-p tuple RXCHF_xcalc_GAM3_MD (rxchf/rxchf_gam_3_OMP.f:617)
 * is class 10 of neorx_prim_list3_dev: outputs (ans1, ans2).
 * RXCHF_GAM3_IC1, source/rxchf/rxchf_gam3_IC1_OMP.f:2: GM3_1ICR
 * ((x123 + x132)/2 of the contracted ans1) and GM3_2ICR (the contracted ans2) */
int neorx_rxchf_gam3_ic1(int64_t handle, double *gm3_1, double *gm3_2);
int neorx_rxchf_gam3_ic1_dev(int64_t handle, double *d_gm3_1, double *d_gm3_2, void *stream);
/* RXCHF_FAE_OMG3 / _FP_OMG3 / _FBE_OMG3, source/rxchf/rxchf_omg3.f:2, :263,
 * :486: F accumulates, e overwritten */
int neorx_rxchf_fae_omg3(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                         const double *gm3_1, const double *gm3_2, double *fae, double *e);
int neorx_rxchf_fp_omg3(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                        const double *gm3_1, const double *gm3_2, double *fp, double *e);
int neorx_rxchf_fbe_omg3(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                         const double *gm3_1, const double *gm3_2, double *fbe, double *e);
/* the same on device arrays: which 0 FAE, 1 FP, 2 FBE; *d_e = E.  Stream-
 * ordered scratch (3 npbf^2 nebf^4 doubles); not synchronised. */
int neorx_rxchf_omg3_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                         const double *d_dp, const double *d_gm3_1, const double *d_gm3_2, double *d_f, double *d_e,
                         void *stream);
/* RXCHF_FP_OMG1 / RXCHF_FBE_OMG1, source/rxchf/rxchf_omg1.f:2, :84, over the
 * GAM_1 and GAM_1s arrays (the contents of GAM_1.ufm / GAM_1s.ufm, which the
 * reference reads itself): F accumulates, SF (SP / SBE) overwritten, e and s
 * overwritten.  RXCHF_GAM1_OMP_MD (rxchf/rxchf_gam_1_OMP.f:2) computes what
 * RXCHFmult_GAM1_MD computes: neorx_rxchfmult_gam1. */
int neorx_rxchf_fp_omg1(int64_t nebf, int64_t npbf, const double *dbe, const double *dp, const double *gam1,
                        const double *gam1s, double *fp, double *sp, double *e, double *s);
int neorx_rxchf_fbe_omg1(int64_t nebf, int64_t npbf, const double *dbe, const double *dp, const double *gam1,
                         const double *gam1s, double *fbe, double *sbe, double *e, double *s);
/* RXCUHF (open-shell regular electrons): RXCUHF_FE_OMG2 / _FP_OMG2 /
 * _FBE_OMG2, source/rxchf/rxcuhf_omg2.f:2, :253, :469, over the RXCHF_GAM2_IC1
 * tensors; F, FAbetE and the S matrices accumulate, e and s overwritten.  FE
 * takes DAbetE as the reference does and does not read it. */
int neorx_rxcuhf_fe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet, const double *datot,
                         const double *dbe, const double *dp, const double *gm2_1, const double *gm2_2,
                         const double *gm2s, double *faalp, double *fabet, double *saalp, double *e, double *s);
int neorx_rxcuhf_fp_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot, const double *dbe,
                         const double *dp, const double *gm2_1, const double *gm2_2, const double *gm2s, double *fp,
                         double *sp, double *e, double *s);
int neorx_rxcuhf_fbe_omg2(int64_t nebf, int64_t npbf, const double *daalp, const double *datot, const double *dbe,
                          const double *dp, const double *gm2_1, const double *gm2_2, const double *gm2s, double *fbe,
                          double *sbe, double *e, double *s);
/* device arrays: which 0 FE (d_f = FAalpE, d_f2 = FAbetE, d_sf = SAalpE),
 * 1 FP, 2 FBE; d_es[0] = E, d_es[1] = S */
int neorx_rxcuhf_omg2_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_daalp, const double *d_datot,
                          const double *d_dbe, const double *d_dp, const double *d_gm2_1, const double *d_gm2_2,
                          const double *d_gm2s, double *d_f, double *d_f2, double *d_sf, double *d_es, void *stream);
/* RXCUHF_FE_OMG3 / _FP_OMG3 / _FBE_OMG3, source/rxchf/rxcuhf_omg3.f:2, :270,
 * :499, over the RXCHF_GAM3_IC1 tensors: F (and FAbetE) accumulate, e
 * overwritten */
int neorx_rxcuhf_fe_omg3(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet, const double *datot,
                         const double *dbe, const double *dp, const double *gm3_1, const double *gm3_2, double *faalp,
                         double *fabet, double *e);
int neorx_rxcuhf_fp_omg3(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet, const double *datot,
                         const double *dbe, const double *dp, const double *gm3_1, const double *gm3_2, double *fp,
                         double *e);
int neorx_rxcuhf_fbe_omg3(int64_t nebf, int64_t npbf, const double *daalp, const double *dabet, const double *datot,
                          const double *dbe, const double *dp, const double *gm3_1, const double *gm3_2, double *fbe,
                          double *e);
/* device arrays: which 0 FE (d_f = FAalpE, d_f2 = FAbetE), 1 FP, 2 FBE;
 * stream-ordered scratch (3 npbf^2 nebf^4 doubles); not synchronised */
int neorx_rxcuhf_omg3_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_daalp, const double *d_dabet,
                          const double *d_datot, const double *d_dbe, const double *d_dp, const double *d_gm3_1,
                          const double *d_gm3_2, double *d_f, double *d_f2, double *d_e, void *stream);
/* the same on device arrays: which 1 FP, 2 FBE; d_es[0] = E, d_es[1] = S */
int neorx_rxchf_omg1_dev(int64_t which, int64_t nebf, int64_t npbf, const double *d_dbe, const double *d_dp,
                         const double *d_gam1, const double *d_gam1s, double *d_f, double *d_sf, double *d_es,
                         void *stream);

/* ---- direct-SCF routines (reference semantics) -------------------------- */
/* GAM2_DSCF, source/gam2_DSCF_OMP.f:2: contracted e-e-p values rebuilt per
 * call (every primitive value through underflow(), as contract_omega2 does),
 * symmetrised and contracted like EG2IC1; focke, fockp, se2, sp2 accumulate
 * (add2fock), E_gam2, S_gam2 overwritten.  The handle supplies the basis and
 * system arrays the reference routine receives. */
int neorx_gam2_dscf(int64_t handle, const double *de, const double *dp, double *focke, double *fockp, double *se2,
                    double *sp2, double *e_gam2, double *s_gam2);
/* GAM3_DSCF, source/gam3_DSCF_OMP.f:2: focke, fockp accumulate; E overwritten */
int neorx_gam3_dscf(int64_t handle, const double *de, const double *dp, double *focke, double *fockp,
                    double *e_gam3);
/* GAM4_DSCF, source/gam4_DSCF_OMP.f:2: the 5-particle values are formed per
 * entry from GAM_2s and GAM_ee (passed in; the reference reads GAM_2s.ufm and
 * GAM_ee.ufm) and contracted as EG4ICR does; no ng4 tensor is stored. */
int neorx_gam4_dscf(int64_t nebf, int64_t npbf, const double *gam_2s, const double *gam_ee, const double *de,
                    const double *dp, double *focke, double *fockp, double *e_gam4);

/* Device-resident direct-SCF routines (all arrays device memory; d_e, d_s one
 * double each; synchronise the stream before returning). */
int neorx_gam2_dscf_dev(int64_t handle, const double *d_de, const double *d_dp, double *d_fe, double *d_fp,
                        double *d_se2, double *d_sp2, double *d_e, double *d_s, void *stream);
int neorx_gam3_dscf_dev(int64_t handle, const double *d_de, const double *d_dp, double *d_fe, double *d_fp,
                        double *d_e, void *stream);
int neorx_gam4_dscf_dev(int64_t nebf, int64_t npbf, const double *d_gam2s, const double *d_gamee, const double *d_de,
                        const double *d_dp, double *d_fe, double *d_fp, double *d_e, void *stream);

/* ---- RXCHFmult (reference semantics, source/rxchfmult/) -------------------
 * More than one special electron: a regular electron basis (nebf functions,
 * npebf primitives) and a special-electron basis (nebfBE, npebfBE).  Tensors
 * use the RXCHFmult packings, rxchfmult/rxchfmult_basis.f:2-134: C row-major
 * [ip][jp][ie1][je1] over the regular basis, then [ie][je] over the special
 * basis for every further electron.  The routines that read regular-basis
 * slots with special-basis indices (XCHF_GAM2, GAM4_INT) need
 * nebfBE <= nebf, as the reference's arrays do. */
/* System handle over both bases; the arguments are those of RXCHFmult_GAM2,
 * rxchfmult/rxchfmult_gam2.f:2-12 (the _be arrays: special-electron basis). */
int neorx_mult_system_create(int64_t nebf, int64_t npebf, int64_t nebfbe, int64_t npebfbe, int64_t npbf, int64_t nat,
                             int64_t ngtg, double pmass, const double *cat, const double *zan, const double *bcoef,
                             const double *gamma, const int64_t *kpestr, const int64_t *kpeend,
                             const int64_t *ampeb2c, const double *agebfcc, const double *elcex,
                             const int64_t *elcam, const double *elcbfc, const int64_t *kpestr_be,
                             const int64_t *kpeend_be, const int64_t *ampeb2c_be, const double *agebfcc_be,
                             const double *elcex_be, const int64_t *elcam_be, const double *elcbfc_be,
                             const double *agnbfcc, const double *nucex, const int64_t *nucam,
                             const double *nucbfc, int64_t *handle);
/* Primitive tuples of a two-basis handle: primitive indices are over the
 * regular primitives followed by the special ones (0-based).
 *   cls 5 (e-e-p)   : RXCHFmult_GAM2_MD, rxchfmult/rxchfmult_gam2.f:627 -> ansE1, ansS, ansE2
 *   cls 6 (e-e-e-p) : RXCHFmult_INT_GAM3_MD, rxchfmult/rxchfmult_gam3-int.f:661 -> ans, and
 *                     ans3 = 2 gVEEg2, ans4 = gVEEg1 of RXCHFmult_xcalc_GAM3_MDex
 *                     (rxchfmult/rxchfmult_gam3ex.f:712) in the S and E2 arrays
 * and of any handle:
 *   cls 7 (e-e-e-p) : RXCHFmult_XCHF_GAM3_MD, rxchfmult/rxchfmult_gam3-xchf.f:630 -> ans
 *                     (electron labels as in xcalc_GAM3_MD; equal after the
 *                     six-permutation average of RXCHFmult_GAM3_XCHF)
 *   cls 8 (e-p)     : RXCHFmult_xcalc_GAM1_MD, rxchfmult/rxchfmult_gam1.f:399 -> ansE, ansS
 * Index rows as for cls 2, 3, 3 and 1.  neorx_prim_list_dev also accepts
 * them (ansE2 dropped); d_ansE2 receives zero for the classes other than 5. */
int neorx_prim_list3_dev(int64_t handle, int64_t cls, int64_t ntup, const int64_t *d_idx, double *d_ansE,
                         double *d_ansS, double *d_ansE2, void *stream);
/* RXCHFmult_GAM2, rxchfmult/rxchfmult_gam2.f:2: INT_GAM2 (npbf^2 nebf^2
 * nebfBE^2, contracted ansE2, unsymmetrised), XCHF_GAM2 and XCHF_GAM2s
 * (npbf^2 nebfBE^4, index_GAM_2PK packing: (x12 + x21)/2 of ansE1 and ansS) */
int neorx_rxchfmult_gam2(int64_t handle, double *int_gam2, double *xchf_gam2, double *xchf_gam2s);
int neorx_rxchfmult_gam2_dev(int64_t handle, double *d_int_gam2, double *d_xchf_gam2, double *d_xchf_gam2s,
                             void *stream);
/* RXCHFmult_GAM1_MD, rxchfmult/rxchfmult_gam1.f:2, over the basis of a
 * one-basis handle (the driver passes the special-electron basis): GAM_1,
 * GAM_1s as neorx_gam1 returns them (the reference writes GAM_1.ufm,
 * GAM_1s.ufm). */
int neorx_rxchfmult_gam1(int64_t handle, double *gam1, double *gam1s);
/* RXCHFmult_GAM3_XCHF, rxchfmult/rxchfmult_gam3-xchf.f:2, over the basis of
 * a one-basis handle: XCHF_GAM3 (nebf^6 npbf^2, index_GAM_3PK packing).
 * RXCHFmult_GAM4_XCHF (rxchfmult/rxchfmult_gam4-xchf.f:2) is GAM4_ICR over
 * the special basis with the leading nebfBE^4 block of GAM_ee: call
 * neorx_gam4_icr(nebfBE, npbf, XCHF_GAM2s, GAM_eeBE, ...). */
int neorx_rxchfmult_gam3_xchf(int64_t handle, double *xchf_gam3);
int neorx_rxchfmult_gam3_xchf_dev(int64_t handle, double *d_xchf_gam3, void *stream);
/* RXCHFmult_GAM3_INT, rxchfmult/rxchfmult_gam3-int.f:2: INT_GAM3
 * (npbf^2 nebf^2 nebfBE^4), averaged over the exchange of electrons 2 and 3 */
int neorx_rxchfmult_gam3_int(int64_t handle, double *int_gam3);
int neorx_rxchfmult_gam3_int_dev(int64_t handle, double *d_int_gam3, void *stream);
/* RXCHFmult_GAM4_INT, rxchfmult/rxchfmult_gam4-int.f:2: INT_GAM4
 * (npbf^2 nebf^2 nebfBE^6) from GAM_2s (= XCHF_GAM2s) and GAM_ee (nebf^4;
 * the reference reads GAM_ee.ufm) */
int neorx_rxchfmult_gam4_int(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *gam_2s, const double *gam_ee,
                             double *gam4);
int neorx_rxchfmult_gam4_int_dev(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *d_gam_2s,
                                 const double *d_gam_ee, double *d_gam4, void *stream);
/* RXCHFmult_FAE/FP/FBE_OMG2, rxchfmult/rxchfmult_omg2.f:2, :175, :348;
 * _OMG3, rxchfmult_omg3.f:2, :201, :400; _OMG4, rxchfmult_omg4.f:2, :207,
 * :412.  dae (nebf^2), dbe (nebfBE^2), dp (npbf^2); gm the INT tensor of
 * the order; f (FAE, FP or FBE) accumulates (add2fock), e overwritten. */
int neorx_rxchfmult_fae_omg2(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                             const double *dp, const double *gm, double *fae, double *e);
int neorx_rxchfmult_fp_omg2(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                            const double *dp, const double *gm, double *fp, double *e);
int neorx_rxchfmult_fbe_omg2(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                             const double *dp, const double *gm, double *fbe, double *e);
int neorx_rxchfmult_fae_omg3(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                             const double *dp, const double *gm, double *fae, double *e);
int neorx_rxchfmult_fp_omg3(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                            const double *dp, const double *gm, double *fp, double *e);
int neorx_rxchfmult_fbe_omg3(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                             const double *dp, const double *gm, double *fbe, double *e);
int neorx_rxchfmult_fae_omg4(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                             const double *dp, const double *gm, double *fae, double *e);
int neorx_rxchfmult_fp_omg4(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                            const double *dp, const double *gm, double *fp, double *e);
int neorx_rxchfmult_fbe_omg4(int64_t nebf, int64_t nebfbe, int64_t npbf, const double *dae, const double *dbe,
                             const double *dp, const double *gm, double *fbe, double *e);
/* ---- RXCHFmult with exchange (LADDEXCH), one basis ------------------------
 * The routines the driver calls with LADDEXCH set take the regular basis for
 * every electron (nebfBE = nebf): the builders take a two-basis handle whose
 * two bases are the same, the contractions one basis size. */
/* RXCHFmult_GAM2_IC1ex, rxchfmult/rxchfmult_gam2ex.f:2: GM2_1ICR (XCHF_GAM2),
 * GM2_2ICR (INT_GAM2), GM2_3ICR (INT_GAM2ex), GM2sICR (XCHF_GAM2s), ng2 each */
int neorx_rxchfmult_gam2_ic1ex(int64_t handle, double *gm2_1, double *gm2_2, double *gm2_3, double *gm2s);
int neorx_rxchfmult_gam2_ic1ex_dev(int64_t handle, double *d_gm2_1, double *d_gm2_2, double *d_gm2_3,
                                   double *d_gm2s, void *stream);
/* RXCHFmult_GAM3_IC1ex, rxchfmult/rxchfmult_gam3ex.f:2: INT_GAM3, INT_GAM3ex1,
 * INT_GAM3ex2 (ng3 each) */
int neorx_rxchfmult_gam3_ic1ex(int64_t handle, double *int_gam3, double *int_gam3ex1, double *int_gam3ex2);
int neorx_rxchfmult_gam3_ic1ex_dev(int64_t handle, double *d_int_gam3, double *d_ex1, double *d_ex2, void *stream);
/* RXCHFmult_FAE/FP/FBE_OMG2ex, rxchfmult/rxchfmult_omg2ex.f:2, :181, :360;
 * _OMG3ex, rxchfmult_omg3ex.f:2, :217, :432.  f accumulates, e overwritten. */
int neorx_rxchfmult_fae_omg2ex(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                               const double *gm2, const double *gm2ex, double *fae, double *e);
int neorx_rxchfmult_fp_omg2ex(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                              const double *gm2, const double *gm2ex, double *fp, double *e);
int neorx_rxchfmult_fbe_omg2ex(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                               const double *gm2, const double *gm2ex, double *fbe, double *e);
int neorx_rxchfmult_fae_omg3ex(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                               const double *gm3, const double *gm3ex1, const double *gm3ex2, double *fae, double *e);
int neorx_rxchfmult_fp_omg3ex(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                              const double *gm3, const double *gm3ex1, const double *gm3ex2, double *fp, double *e);
int neorx_rxchfmult_fbe_omg3ex(int64_t nebf, int64_t npbf, const double *dae, const double *dbe, const double *dp,
                               const double *gm3, const double *gm3ex1, const double *gm3ex2, double *fbe, double *e);
/* The three exchange-form contractions of order 2 (OMG2ex: d_gx1 = GM2exICR)
 * or 3 (OMG3ex: d_gx1, d_gx2 = GM3ex1ICR, GM3ex2ICR) on device arrays; any of
 * d_fae, d_fp, d_fbe may be NULL. */
int neorx_rxchfmult_omgex_dev(int64_t order, int64_t nebf, int64_t npbf, const double *d_dae, const double *d_dbe,
                              const double *d_dp, const double *d_gm, const double *d_gx1, const double *d_gx2,
                              double *d_fae, double *d_fp, double *d_fbe, double *d_e, void *stream);
/* The three contractions of one order k (1: OMG2, 2: OMG3, 3: OMG4) in one
 * pass over the device tensor d_gm; any of d_fae, d_fp, d_fbe may be NULL.
 * d_e receives the energy all three reference routines return. */
int neorx_rxchfmult_omg_dev(int64_t k, int64_t nebf, int64_t nebfbe, int64_t npbf, const double *d_dae,
                            const double *d_dbe, const double *d_dp, const double *d_gm, double *d_fae,
                            double *d_fp, double *d_fbe, double *d_e, void *stream);

/* ---- autotuner (no counterpart in the reference) ---------------------------
 * Times the EG3IC1 slab kernel with 8, 12, 16 and 24 loads in flight per lane
 * on the caller's tensor (dtype 0 FP64, 1 FP32, 2 BF16 storage; reps timed
 * launches each) and records the fastest for (dtype, nebf): every later EG3
 * contraction of that storage type and basis size in this process uses it.
 * best_unroll, best_ms and all_ms[4] (per candidate) may be NULL. */
int neorx_autotune_eg3_dev(int64_t nebf, int64_t npbf, int64_t dtype, const double *d_de, const void *d_gm3,
                           int64_t reps, int64_t *best_unroll, double *best_ms, double *all_ms);

/* ---- NEO-XCHF RHF on the device (source/xchf1_independent/) --------------
 * A context holds every integral of the XCHF energy in HBM, built on the
 * device from a one-basis system handle in the order of
 * xcneo_HYBRID.f:663-1002: class_nuc_rep (class_nuc_rep.f:1), elec_ovlap /
 * nuc_ovlap (pure_ovlap.f:2, :107), calc_GAM_epcore (gam_epcore.f:1),
 * calc_GAM_ep (gam_ep.f:1), calc_GAM_ee, GAM1_OMP_MD, GAM2_IC1 and GAM3_IC1
 * (the last two only for nelec > 1, > 2).  GAM_4 (GAM4_ICR) is never stored:
 * the 5-particle contraction is evaluated in factorized form from GAM_ee and
 * GAM_2s, with the same result as EG4ICR on the in-core tensor.
 * store: bit mask of the storage types GAM_3 is kept in (1 FP64, 2 FP32,
 * 4 BF16); 0 means FP64; bit 8 also autotunes the EG3 contraction for each
 * stored type (neorx_autotune_eg3_dev).  The integrals are only built here:
 * these entry points read no .ufm file.
 * Each context owns a non-blocking stream; neorx_xchf_scf on different
 * contexts may run from different host threads at once (graph capture is
 * thread-local).  Calls on one context must not overlap. */
int neorx_xchf_create(int64_t handle, int64_t nelec, int64_t store, int64_t *ctx);
int neorx_xchf_destroy(int64_t ctx);
/* Host copies of the context's integrals in the reference packings (any
 * pointer may be NULL): E_nuc (ENUCRP.dat), S_e (eovlap.ufm), S_p
 * (novlap.ufm), GAM_ecore (GAM_ecr.ufm), GAM_pcore (GAM_pcr.ufm), GAM_ep
 * (GAM_ep.ufm, pack_4D(nebf,nebf,npbf)), GAM_ee (GAM_ee.ufm). */
int neorx_xchf_integrals(int64_t ctx, double *enuc, double *se, double *sp, double *gam_ecore, double *gam_pcore,
                         double *gam_ep, double *gam_ee);
/* xchf1_fock, source/xchf1_independent/xchf1_fock.f:2, on device arrays, for
 * the in-core path of xcneo_HYBRID (LG2IC1, LG3IC1, LG4IC): d_focke, d_fockp
 * receive the Fock matrices after XCfock_correction (xchf1_fock.f:728-784),
 * d_comps 14 doubles: E_total, E_nuc, E_ecore, E_pcore, E_ep, E_ee, E_gam1,
 * E_gam2, E_gam3, E_gam4, S_total, S_gam1, S_gam2, H_expect.  prec & 15
 * selects the stored GAM_3 read (0 FP64, 1 FP32, 2 BF16); prec >> 4 the pass
 * over it: 0 the rows of DE; 1 the same, kept as the reference rows (FP64
 * only, prec 0x10); 2 the reference rows plus the increment
 * T(DE - DEref, DE + DEref) read in the selected storage, where T is GAM_3's
 * symmetric bilinear form in its two contracted densities (its storage error
 * scales with |DE - DEref|).  All arithmetic is FP64.  Enqueued on stream,
 * not synchronised. */
int neorx_xchf_fock_dev(int64_t ctx, const double *d_de, const double *d_dp, int64_t prec, double *d_focke,
                        double *d_fockp, double *d_comps, void *stream);
/* The FP64 MFMA GEMM of the factorized 5-particle contraction (no
 * counterpart in the reference): C (M x N, row-major) = A B with
 * A(m,k) = d_a[m*sam + k*sak], B(k,n) = d_b[k*sbk + n*sbn]; split over K
 * (stream-ordered scratch) when C has few 16x16 tiles.  Not synchronised. */
int neorx_gemm_f64_dev(int64_t M, int64_t N, int64_t K, const double *d_a, int64_t sam, int64_t sak, const double *d_b,
                       int64_t sbk, int64_t sbn, double *d_c, void *stream);
/* ROOTHAN, source/xchf1_independent/xcscf.f:766, on device arrays (n <= 128):
 * X = S^-1/2, diagonalisation of X^T F X (eigenvalues ascending), C = X C',
 * D = prefac sum_{k = kstart..klast} C(:,k) C(:,k)^T (1-based k; the
 * reference uses 1..nelec/2 with prefac 2 for electrons and NUCST..NUCST
 * with 1 for the nucleus).  Stream-ordered scratch; not synchronised. */
int neorx_roothan_dev(int64_t n, const double *d_s, const double *d_f, int64_t kstart, int64_t klast, double prefac,
                      double *d_c, double *d_ev, double *d_d, void *stream);
/* xcscf, source/xchf1_independent/xcscf.f:2-560: the NEO-XCHF RHF SCF on the
 * device.  Per iteration, as xcscf: the Fock build with (DE, DP), ROOTHAN for
 * the nucleus, the build with (DE, new DP), ROOTHAN for the electrons, DENDIF;
 * converged when max|dDE| < tole (and max|dDP| < tolp if npbf > 1).  The
 * reference's second-order step (SOSCF) is not restated; flags bit 2 turns on
 * DIIS extrapolation of the electronic Fock matrix instead.  The two builds
 * of one iteration share one pass over GM3/GM2 (all terms are linear in DP).
 *   ce_guess: nebf x nebf column-major orbitals (read_elec_density of
 *             guessCE.inp), or NULL for the core guess (guess_elec);
 *   nucst:    occupied nuclear orbital (guess_prot, ROOTHAN);
 *   mode:     0 FP64 GAM_3; 1 FP32 until the switch signal < sw_lo, then
 *             FP64; 2 BF16 until < sw_hi, then FP32 until < sw_lo, then FP64
 *             (signal: the DIIS error max|X^T(FDS - SDF)X| with DIIS, else
 *             max|dDE|); 3 (incremental) BF16 until < sw_hi, then a build
 *             taken as the reference (from the FP32 copy if the context
 *             holds one, else FP64) and BF16-stored increments from it, an
 *             FP64 reference once < sw_lo, and an FP64 build to confirm
 *             convergence; convergence is only declared on an FP64
 *             (non-incremental) build;
 *   flags:    bit 1: capture one iteration as a hipGraph and replay it;
 *             bit 2: DIIS;
 *   comps_init / comps: 14 components (as neorx_xchf_fock_dev) of the first
 *             build of iteration 1 (the reference's INITIAL GUESS ENERGETICS)
 *             and of the last electronic build (the final energy);
 *   ce, cp, ev_e, ev_p: final orbitals and eigenvalues (FinalCE.dat, ...);
 *   trace:    maxit x 8 doubles per iteration: E_total, dE, max|dDE|,
 *             max|dDP|, switch signal, GAM_3 storage (0/1/2), wall ms, GAM_3
 *             pass (0 full, 1 full kept as the reference, 2 increment).
 * Returns NEORX_E_ARG (after filling the outputs) if not converged. */
int neorx_xchf_scf(int64_t ctx, const double *ce_guess, int64_t nucst, int64_t maxit, double tole, double tolp,
                   int64_t mode, int64_t flags, double sw_hi, double sw_lo, double *comps_init, double *comps,
                   double *ce, double *cp, double *ev_e, double *ev_p, double *trace, int64_t *niter);

#ifdef __cplusplus
}
#endif
#endif
