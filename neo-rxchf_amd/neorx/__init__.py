"""Host-side binding of libneorx_hip.so (the MI355X geminal-integral engine).

The names mirror the reference's Fortran entry points for this path
(GAM2_IC1, EG2IC1, ... in source/gam2_IC1_OMP.f, source/gam3_IC1_OMP.f,
source/gam4_ICR_OMP.f, source/gam_1_OMP.f, source/gam_ee.f) and take the same
arrays: Fortran-layout numpy arrays on the host, or raw device pointers for
the ``*_dev`` variants.  There is no CPU path: if the HIP library or a gfx950
device is missing, every call raises ``NeorxError``.
"""
import ctypes as _C
import os as _os

import numpy as _np

from . import synth  # noqa: F401  (synthetic systems of the benchmark configurations)
from . import xchf  # noqa: F401  (NEO-XCHF input files, orbital files and the device SCF)

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIBDIR = _os.path.join(_os.path.dirname(_HERE), "lib")
LIBPATH = _os.path.join(LIBDIR, "libneorx_hip.so")

_lib = None


class NeorxError(RuntimeError):
    pass


_I64 = _C.c_int64
_D = _C.c_double
_P = _C.c_void_p


def _sig(name, args):
    f = getattr(_lib, name)
    f.restype = _C.c_int
    f.argtypes = args
    return f


def load():
    """Load libneorx_hip.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not _os.path.exists(LIBPATH):
        raise NeorxError(f"{LIBPATH} not built: run __graft_entry__.build()")
    _lib = _C.CDLL(LIBPATH)
    _lib.neorx_last_error.restype = _C.c_char_p
    _lib.neorx_device_arch.restype = _C.c_char_p
    _sig("neorx_init", [_I64])
    _sig("neorx_finalize", [])
    _sig("neorx_release_resident", [])
    _sig("neorx_rxchf_gam2_ic1", [_I64, _P, _P, _P])
    for nm in ("neorx_rxchf_fae_omg2", "neorx_rxchf_fp_omg2", "neorx_rxchf_fbe_omg2"):
        _sig(nm, [_I64, _I64] + [_P] * 10)
    _sig("neorx_rxchf_gam3_ic1", [_I64, _P, _P])
    for nm in ("neorx_rxchf_fae_omg3", "neorx_rxchf_fp_omg3", "neorx_rxchf_fbe_omg3"):
        _sig(nm, [_I64, _I64] + [_P] * 7)
    for nm in ("neorx_rxchf_fp_omg1", "neorx_rxchf_fbe_omg1"):
        _sig(nm, [_I64, _I64] + [_P] * 8)
    _sig("neorx_rxcuhf_fe_omg2", [_I64, _I64] + [_P] * 13)
    for nm in ("neorx_rxcuhf_fp_omg2", "neorx_rxcuhf_fbe_omg2"):
        _sig(nm, [_I64, _I64] + [_P] * 11)
    _sig("neorx_rxcuhf_fe_omg3", [_I64, _I64] + [_P] * 10)
    for nm in ("neorx_rxcuhf_fp_omg3", "neorx_rxcuhf_fbe_omg3"):
        _sig(nm, [_I64, _I64] + [_P] * 9)
    _sig("neorx_transfer_bytes", [_P, _P])
    _sig("neorx_system_create", [_I64, _I64, _I64, _I64, _I64, _D] + [_P] * 15 + [_P])
    _sig("neorx_system_destroy", [_I64])
    _sig("neorx_system_order", [_I64, _P, _P])
    _sig("neorx_prim_list_dev", [_I64, _I64, _I64, _P, _P, _P, _P])
    _sig("neorx_prim_range_dev", [_I64, _I64, _I64, _I64, _P, _P, _P])
    _sig("neorx_gam1", [_I64, _P, _P])
    _sig("neorx_gam2_ic1", [_I64, _P, _P])
    _sig("neorx_gam3_ic1", [_I64, _P])
    _sig("neorx_gam_ee", [_I64, _P])
    _sig("neorx_gam4_icr", [_I64, _I64, _P, _P, _P])
    _sig("neorx_gam2_ic1_dev", [_I64, _P, _P, _P])
    _sig("neorx_gam3_ic1_dev", [_I64, _P, _P])
    _sig("neorx_eg2ic1", [_I64, _I64] + [_P] * 10)
    _sig("neorx_eg3ic1", [_I64, _I64] + [_P] * 6)
    _sig("neorx_eg4icr", [_I64, _I64] + [_P] * 6)
    _sig("neorx_eg2ic1_dev", [_I64, _I64] + [_P] * 10 + [_P])
    _sig("neorx_eg3ic1_dev", [_I64, _I64] + [_P] * 6 + [_P])
    _sig("neorx_eg3ic1_noalloc_dev", [_I64, _I64] + [_P] * 6 + [_P])
    _sig("neorx_eg4icr_dev", [_I64, _I64] + [_P] * 6 + [_P])
    _sig("neorx_convert_f32_dev", [_I64, _P, _P, _P])
    _sig("neorx_convert_bf16_dev", [_I64, _P, _P, _P])
    _sig("neorx_rxchfne_omg2_dev", [_I64, _I64] + [_P] * 4 + [_C.c_int] + [_P] * 5)
    _sig("neorx_eg2ic1_f32_dev", [_I64, _I64] + [_P] * 10 + [_P])
    _sig("neorx_eg3ic1_f32_dev", [_I64, _I64] + [_P] * 6 + [_P])
    _sig("neorx_eg4icr_f32_dev", [_I64, _I64] + [_P] * 6 + [_P])
    _sig("neorx_precision_switch_dev", [_I64, _P, _P, _D, _P, _P, _P])
    _sig("neorx_fae_omg2", [_I64] * 4 + [_P] * 9)
    _sig("neorx_ufp_omg2", [_I64] * 4 + [_P] * 9)
    for n in ("fae_omg3", "ufp_omg3", "fae_omg4", "ufp_omg4"):
        _sig("neorx_" + n, [_I64] * 4 + [_P] * 6)
    _sig("neorx_omg4_icr", [_I64, _I64, _P, _P, _P])
    _sig("neorx_rxchfne_gam2_ic1", [_I64, _P, _P])
    _sig("neorx_rxchfne_gam2_ic1_dev", [_I64, _P, _P, _P])
    for n in ("fae", "fp", "fbe"):
        _sig(f"neorx_rxchfne_{n}_omg2", [_I64, _I64] + [_P] * 6)
    _sig("neorx_rxcuhfne_fe_omg2", [_I64, _I64] + [_P] * 9)
    for n in ("fp", "fbe"):
        _sig(f"neorx_rxcuhfne_{n}_omg2", [_I64, _I64] + [_P] * 8)
    _sig("neorx_rxchfne_fae_gamee", [_I64] + [_P] * 4)
    _sig("neorx_gam2_dscf", [_I64] + [_P] * 8)
    _sig("neorx_gam3_dscf", [_I64] + [_P] * 5)
    _sig("neorx_gam4_dscf", [_I64, _I64] + [_P] * 7)
    _sig("neorx_gam2_dscf_dev", [_I64] + [_P] * 8 + [_P])
    _sig("neorx_gam3_dscf_dev", [_I64] + [_P] * 5 + [_P])
    _sig("neorx_gam4_dscf_dev", [_I64, _I64] + [_P] * 7 + [_P])
    _sig("neorx_mult_system_create", [_I64] * 7 + [_D] + [_P] * 22 + [_P])
    _sig("neorx_prim_list3_dev", [_I64, _I64, _I64, _P, _P, _P, _P, _P])
    _sig("neorx_rxchfmult_gam2", [_I64, _P, _P, _P])
    _sig("neorx_rxchfmult_gam2_dev", [_I64, _P, _P, _P, _P])
    _sig("neorx_rxchfmult_gam3_int", [_I64, _P])
    _sig("neorx_rxchfmult_gam1", [_I64, _P, _P])
    _sig("neorx_rxchfmult_gam3_xchf", [_I64, _P])
    _sig("neorx_rxchfmult_gam3_xchf_dev", [_I64, _P, _P])
    _sig("neorx_rxchfmult_gam3_int_dev", [_I64, _P, _P])
    _sig("neorx_rxchfmult_gam4_int", [_I64, _I64, _I64, _P, _P, _P])
    _sig("neorx_rxchfmult_gam4_int_dev", [_I64, _I64, _I64, _P, _P, _P, _P])
    for k in (2, 3, 4):
        for n in ("fae", "fp", "fbe"):
            _sig(f"neorx_rxchfmult_{n}_omg{k}", [_I64] * 3 + [_P] * 6)
    _sig("neorx_rxchfmult_omg_dev", [_I64] * 4 + [_P] * 8 + [_P])
    _sig("neorx_rxchfmult_gam2_ic1ex", [_I64] + [_P] * 4)
    _sig("neorx_rxchfmult_gam2_ic1ex_dev", [_I64] + [_P] * 4 + [_P])
    _sig("neorx_rxchfmult_gam3_ic1ex", [_I64] + [_P] * 3)
    _sig("neorx_rxchfmult_gam3_ic1ex_dev", [_I64] + [_P] * 3 + [_P])
    for n in ("fae", "fp", "fbe"):
        _sig(f"neorx_rxchfmult_{n}_omg2ex", [_I64, _I64] + [_P] * 7)
        _sig(f"neorx_rxchfmult_{n}_omg3ex", [_I64, _I64] + [_P] * 8)
    _sig("neorx_rxchfmult_omgex_dev", [_I64] * 3 + [_P] * 10 + [_P])
    _sig("neorx_autotune_eg3_dev", [_I64, _I64, _I64, _P, _P, _I64, _P, _P, _P])
    _sig("neorx_xchf_create", [_I64, _I64, _I64, _P])
    _sig("neorx_xchf_destroy", [_I64])
    _sig("neorx_xchf_integrals", [_I64] + [_P] * 7)
    _sig("neorx_xchf_fock_dev", [_I64, _P, _P, _I64, _P, _P, _P, _P])
    _sig("neorx_roothan_dev", [_I64, _P, _P, _I64, _I64, _D, _P, _P, _P, _P])
    _sig("neorx_gemm_f64_dev", [_I64, _I64, _I64, _P, _I64, _I64, _P, _I64, _I64, _P, _P])
    _sig("neorx_xchf_scf", [_I64, _P, _I64, _I64, _D, _D, _I64, _I64, _D, _D] + [_P] * 8)
    return _lib


def _check(rc):
    if rc != 0:
        raise NeorxError(f"neorx error {rc}: {_lib.neorx_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def init(device=0):
    load()
    _check(_lib.neorx_init(device))
    return _lib.neorx_device_arch().decode()


class System:
    """Basis, nuclei and geminal resident in HBM (the arguments of GAM2_IC1,
    source/gam2_IC1_OMP.f:2-37).  ``sysd`` is a dict produced by
    :mod:`neorx.synth` (Fortran-layout numpy arrays).  A dict that also holds a
    special-electron basis (``nebfBE`` and the ``*_be`` arrays, as from
    ``synth.mult``) gives a two-basis handle for the RXCHFmult routines
    (the arguments of RXCHFmult_GAM2, source/rxchfmult/rxchfmult_gam2.f:2-12)."""

    def __init__(self, sysd):
        load()
        s = sysd
        self.s = s
        if "nebfBE" in s:
            self._init_mult(s)
            return
        self._keep = [
            _np.ascontiguousarray(s[k])
            for k in ("cat", "zan", "bcoef", "gamma", "kpestr", "kpeend", "ampeb2c", "agebfcc", "agnbfcc",
                      "elcex", "nucex")
        ]
        self._keep += [_np.asfortranarray(s[k]) for k in ("elcam", "nucam", "elcbfc", "nucbfc")]
        h = _C.c_int64()
        _check(_lib.neorx_system_create(s["nebf"], s["npebf"], s["npbf"], s["nat"], s["ngtg"], s["pmass"],
                                        *[_p(a) for a in self._keep], _C.byref(h)))
        self.handle = h.value

    def _init_mult(self, s):
        c, f = _np.ascontiguousarray, _np.asfortranarray
        self._keep = [c(s[k]) for k in ("cat", "zan", "bcoef", "gamma", "kpestr", "kpeend", "ampeb2c", "agebfcc",
                                        "elcex")]
        self._keep += [f(s["elcam"]), f(s["elcbfc"])]
        self._keep += [c(s[k + "_be"]) for k in ("kpestr", "kpeend", "ampeb2c", "agebfcc", "elcex")]
        self._keep += [f(s["elcam_be"]), f(s["elcbfc_be"])]
        self._keep += [c(s["agnbfcc"]), c(s["nucex"]), f(s["nucam"]), f(s["nucbfc"])]
        h = _C.c_int64()
        _check(_lib.neorx_mult_system_create(s["nebf"], s["npebf"], s["nebfBE"], s["npebfBE"], s["npbf"], s["nat"],
                                             s["ngtg"], s["pmass"], *[_p(a) for a in self._keep], _C.byref(h)))
        self.handle = h.value

    def close(self):
        if self.handle:
            _check(_lib.neorx_system_destroy(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def order(self):
        e = _np.zeros(self.s["npebf"], dtype=_np.int64)
        p = _np.zeros(self.s["npbf"], dtype=_np.int64)
        _check(_lib.neorx_system_order(self.handle, _p(e), _p(p)))
        return e, p

    # --- primitive tuples (device pointers) ---
    def prim_list_dev(self, cls, ntup, d_idx, d_e, d_s, stream=0):
        _check(_lib.neorx_prim_list_dev(self.handle, cls, ntup, d_idx, d_e, d_s, stream))

    def prim_list3_dev(self, cls, ntup, d_idx, d_e, d_s, d_e2, stream=0):
        _check(_lib.neorx_prim_list3_dev(self.handle, cls, ntup, d_idx, d_e, d_s, d_e2, stream))

    # --- RXCHFmult builders (two-basis handle) ---
    def RXCHFmult_GAM2(self):
        """INT_GAM2, XCHF_GAM2, XCHF_GAM2s (rxchfmult/rxchfmult_gam2.f:2)."""
        na, nb, npb = self.s["nebf"], self.s["nebfBE"], self.s["npbf"]
        gi = _np.zeros(npb * npb * na * na * nb * nb)
        gx = _np.zeros(npb * npb * nb ** 4)
        gs = _np.zeros_like(gx)
        _check(_lib.neorx_rxchfmult_gam2(self.handle, _p(gi), _p(gx), _p(gs)))
        return gi, gx, gs

    def RXCHFmult_GAM1_MD(self):
        """GAM_1, GAM_1s of RXCHFmult_GAM1_MD (rxchfmult/rxchfmult_gam1.f:2)
        over this handle's basis (one-basis handle)."""
        ne, npb = self.s["nebf"], self.s["npbf"]
        g = _np.zeros(ne * ne * npb * npb)
        gs = _np.zeros_like(g)
        _check(_lib.neorx_rxchfmult_gam1(self.handle, _p(g), _p(gs)))
        return g, gs

    def RXCHFmult_GAM3_XCHF(self):
        """XCHF_GAM3 (rxchfmult/rxchfmult_gam3-xchf.f:2) over this handle's
        basis (one-basis handle)."""
        ne, npb = self.s["nebf"], self.s["npbf"]
        g = _np.zeros(ne ** 6 * npb * npb)
        _check(_lib.neorx_rxchfmult_gam3_xchf(self.handle, _p(g)))
        return g

    def RXCHFmult_GAM2_IC1ex(self):
        """GM2_1ICR, GM2_2ICR, GM2_3ICR, GM2sICR (rxchfmult/rxchfmult_gam2ex.f:2);
        the two bases of the handle must be the same."""
        na, npb = self.s["nebf"], self.s["npbf"]
        out = [_np.zeros(npb * npb * na ** 4) for _ in range(4)]
        _check(_lib.neorx_rxchfmult_gam2_ic1ex(self.handle, *[_p(a) for a in out]))
        return out

    def RXCHFmult_GAM3_IC1ex(self):
        """INT_GAM3, INT_GAM3ex1, INT_GAM3ex2 (rxchfmult/rxchfmult_gam3ex.f:2)."""
        na, npb = self.s["nebf"], self.s["npbf"]
        out = [_np.zeros(npb * npb * na ** 6) for _ in range(3)]
        _check(_lib.neorx_rxchfmult_gam3_ic1ex(self.handle, *[_p(a) for a in out]))
        return out

    def RXCHFmult_GAM3_INT(self):
        """INT_GAM3 (rxchfmult/rxchfmult_gam3-int.f:2)."""
        na, nb, npb = self.s["nebf"], self.s["nebfBE"], self.s["npbf"]
        g = _np.zeros(npb * npb * na * na * nb ** 4)
        _check(_lib.neorx_rxchfmult_gam3_int(self.handle, _p(g)))
        return g

    def prim_range_dev(self, cls, start, count, d_e, d_s, stream=0):
        _check(_lib.neorx_prim_range_dev(self.handle, cls, start, count, d_e, d_s, stream))

    # --- contracted builders (host arrays) ---
    def GAM1(self):
        ne, npb = self.s["nebf"], self.s["npbf"]
        g = _np.zeros(ne * ne * npb * npb)
        gs = _np.zeros_like(g)
        _check(_lib.neorx_gam1(self.handle, _p(g), _p(gs)))
        return g, gs

    def GAM2_IC1(self):
        ne, npb = self.s["nebf"], self.s["npbf"]
        g = _np.zeros(ne ** 4 * npb * npb)
        gs = _np.zeros_like(g)
        _check(_lib.neorx_gam2_ic1(self.handle, _p(g), _p(gs)))
        return g, gs

    def GAM3_IC1(self):
        ne, npb = self.s["nebf"], self.s["npbf"]
        g = _np.zeros(ne ** 6 * npb * npb)
        _check(_lib.neorx_gam3_ic1(self.handle, _p(g)))
        return g

    def GAM_ee(self):
        ne = self.s["nebf"]
        g = _np.zeros(ne ** 4)
        _check(_lib.neorx_gam_ee(self.handle, _p(g)))
        return g

    def RXCHF_GAM2_IC1(self):
        """RXCHF_GAM2_IC1 (source/rxchf/rxchf_gam2_IC1_OMP.f:2): (GM2_1ICR,
        GM2_2ICR, GM2sICR)."""
        ne, npb = self.s["nebf"], self.s["npbf"]
        g1 = _np.zeros(ne ** 4 * npb * npb)
        g2, gs = _np.zeros_like(g1), _np.zeros_like(g1)
        _check(_lib.neorx_rxchf_gam2_ic1(self.handle, _p(g1), _p(g2), _p(gs)))
        return g1, g2, gs

    def RXCHF_GAM3_IC1(self):
        """RXCHF_GAM3_IC1 (source/rxchf/rxchf_gam3_IC1_OMP.f:2): (GM3_1ICR,
        GM3_2ICR)."""
        ne, npb = self.s["nebf"], self.s["npbf"]
        g1 = _np.zeros(ne ** 6 * npb * npb)
        g2 = _np.zeros_like(g1)
        _check(_lib.neorx_rxchf_gam3_ic1(self.handle, _p(g1), _p(g2)))
        return g1, g2

    def RXCHFne_GAM2_IC1(self):
        """RXCHFne_GAM2_IC1 (source/rxchfne/rxchfne_gam2_IC1_OMP.f:2):
        returns (GM2ICR, GM2exICR)."""
        ne, npb = self.s["nebf"], self.s["npbf"]
        g = _np.zeros(ne ** 4 * npb * npb)
        gx = _np.zeros_like(g)
        _check(_lib.neorx_rxchfne_gam2_ic1(self.handle, _p(g), _p(gx)))
        return g, gx

    # --- direct-SCF routines (host arrays) ---
    def GAM2_DSCF(self, DE, DP, focke, fockp, se2, sp2):
        """GAM2_DSCF (source/gam2_DSCF_OMP.f:2): focke, fockp, se2, sp2
        (Fortran-ordered float64) accumulate in place; returns (E_gam2, S_gam2)."""
        DE = _np.asfortranarray(DE, dtype=_np.float64)
        DP = _np.asfortranarray(DP, dtype=_np.float64)
        e = _np.zeros(2)
        _check(_lib.neorx_gam2_dscf(self.handle, _p(DE), _p(DP), _p(focke), _p(fockp), _p(se2), _p(sp2), _p(e),
                                    _p(e[1:])))
        return float(e[0]), float(e[1])

    def GAM3_DSCF(self, DE, DP, focke, fockp):
        """GAM3_DSCF (source/gam3_DSCF_OMP.f:2): returns E_gam3."""
        DE = _np.asfortranarray(DE, dtype=_np.float64)
        DP = _np.asfortranarray(DP, dtype=_np.float64)
        e = _np.zeros(1)
        _check(_lib.neorx_gam3_dscf(self.handle, _p(DE), _p(DP), _p(focke), _p(fockp), _p(e)))
        return float(e[0])


def GAM4_ICR(nebf, npbf, gam_2s, gam_ee):
    load()
    g = _np.zeros(nebf ** 8 * npbf * npbf)
    _check(_lib.neorx_gam4_icr(nebf, npbf, _p(_np.ascontiguousarray(gam_2s)), _p(_np.ascontiguousarray(gam_ee)),
                               _p(g)))
    return g


def GAM4_DSCF(nebf, npbf, gam_2s, gam_ee, DE, DP, focke, fockp):
    """GAM4_DSCF (source/gam4_DSCF_OMP.f:2) with GAM_2s and GAM_ee passed in
    (the reference reads them from GAM_2s.ufm / GAM_ee.ufm); returns E_gam4."""
    load()
    DE = _np.asfortranarray(DE, dtype=_np.float64)
    DP = _np.asfortranarray(DP, dtype=_np.float64)
    e = _np.zeros(1)
    _check(_lib.neorx_gam4_dscf(nebf, npbf, _p(_np.ascontiguousarray(gam_2s)), _p(_np.ascontiguousarray(gam_ee)),
                                _p(DE), _p(DP), _p(focke), _p(fockp), _p(e)))
    return float(e[0])


def RXCHF_OMG2(which, nebf, npbf, DAE, DBE, DP, G1, G2, Gs, F, SF):
    """RXCHF_FAE_OMG2 / _FP_OMG2 / _FBE_OMG2 (source/rxchf/rxchf_omg2.f:2,
    :231, :449) for which = 0, 1, 2: F and SF accumulate in place; returns
    (E, S)."""
    load()
    fn = (_lib.neorx_rxchf_fae_omg2, _lib.neorx_rxchf_fp_omg2, _lib.neorx_rxchf_fbe_omg2)[which]
    e, sv = _np.zeros(1), _np.zeros(1)
    _check(fn(nebf, npbf, _p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)), _p(_np.ascontiguousarray(G1)),
              _p(_np.ascontiguousarray(G2)), _p(_np.ascontiguousarray(Gs)), _p(F), _p(SF), _p(e), _p(sv)))
    return float(e[0]), float(sv[0])


def RXCHF_OMG3(which, nebf, npbf, DAE, DBE, DP, G1, G2, F):
    """RXCHF_FAE_OMG3 / _FP_OMG3 / _FBE_OMG3 (source/rxchf/rxchf_omg3.f:2,
    :263, :486) for which = 0, 1, 2: F accumulates in place; returns E."""
    load()
    fn = (_lib.neorx_rxchf_fae_omg3, _lib.neorx_rxchf_fp_omg3, _lib.neorx_rxchf_fbe_omg3)[which]
    e = _np.zeros(1)
    _check(fn(nebf, npbf, _p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)), _p(_np.ascontiguousarray(G1)),
              _p(_np.ascontiguousarray(G2)), _p(F), _p(e)))
    return float(e[0])


def RXCHF_OMG1(which, nebf, npbf, DBE, DP, G1, G1s, F, SF):
    """RXCHF_FP_OMG1 / RXCHF_FBE_OMG1 (source/rxchf/rxchf_omg1.f:2, :84) for
    which = 1, 2 over the GAM_1 / GAM_1s arrays: F accumulates, SF is
    overwritten; returns (E, S)."""
    load()
    fn = _lib.neorx_rxchf_fp_omg1 if which == 1 else _lib.neorx_rxchf_fbe_omg1
    e, sv = _np.zeros(1), _np.zeros(1)
    _check(fn(nebf, npbf, _p(_fort(DBE)), _p(_fort(DP)), _p(_np.ascontiguousarray(G1)),
              _p(_np.ascontiguousarray(G1s)), _p(F), _p(SF), _p(e), _p(sv)))
    return float(e[0]), float(sv[0])


def RXCUHF_OMG2(which, nebf, npbf, DAalp, DAbet, DAtot, DBE, DP, G1, G2, Gs, F, SF, F2=None):
    """RXCUHF_FE_OMG2 / _FP_OMG2 / _FBE_OMG2 (source/rxchf/rxcuhf_omg2.f:2,
    :253, :469) for which = 0, 1, 2 (FE: F = FAalpE, F2 = FAbetE, SF =
    SAalpE); F, F2, SF accumulate in place; returns (E, S)."""
    load()
    e, sv = _np.zeros(1), _np.zeros(1)
    g = [_p(_np.ascontiguousarray(x)) for x in (G1, G2, Gs)]
    if which == 0:
        _check(_lib.neorx_rxcuhf_fe_omg2(nebf, npbf, _p(_fort(DAalp)), _p(_fort(DAbet)), _p(_fort(DAtot)),
                                         _p(_fort(DBE)), _p(_fort(DP)), *g, _p(F), _p(F2), _p(SF), _p(e), _p(sv)))
    else:
        fn = _lib.neorx_rxcuhf_fp_omg2 if which == 1 else _lib.neorx_rxcuhf_fbe_omg2
        _check(fn(nebf, npbf, _p(_fort(DAalp)), _p(_fort(DAtot)), _p(_fort(DBE)), _p(_fort(DP)), *g, _p(F), _p(SF),
                  _p(e), _p(sv)))
    return float(e[0]), float(sv[0])


def RXCUHF_OMG3(which, nebf, npbf, DAalp, DAbet, DAtot, DBE, DP, G1, G2, F, F2=None):
    """RXCUHF_FE_OMG3 / _FP_OMG3 / _FBE_OMG3 (source/rxchf/rxcuhf_omg3.f:2,
    :270, :499) for which = 0, 1, 2 (FE: F = FAalpE, F2 = FAbetE); F and F2
    accumulate in place; returns E."""
    load()
    e = _np.zeros(1)
    a = [_p(_fort(x)) for x in (DAalp, DAbet, DAtot, DBE, DP)] + [_p(_np.ascontiguousarray(G1)),
                                                                  _p(_np.ascontiguousarray(G2))]
    if which == 0:
        _check(_lib.neorx_rxcuhf_fe_omg3(nebf, npbf, *a, _p(F), _p(F2), _p(e)))
    else:
        fn = _lib.neorx_rxcuhf_fp_omg3 if which == 1 else _lib.neorx_rxcuhf_fbe_omg3
        _check(fn(nebf, npbf, *a, _p(F), _p(e)))
    return float(e[0])


def transfer_bytes():
    """(host->device, device->host) bytes of in-core tensors copied by the
    host-pointer entry points so far (neorx_transfer_bytes)."""
    load()
    h, d = _C.c_int64(), _C.c_int64()
    _check(_lib.neorx_transfer_bytes(_C.byref(h), _C.byref(d)))
    return h.value, d.value


def release_resident():
    """Free the device copies kept for host tensors (neorx_release_resident)."""
    load()
    _check(_lib.neorx_release_resident())


def EG2IC1(nebf, npbf, DE, DP, GM2ICR, GM2SICR, FE, FP):
    """EG2IC1 (source/gam2_IC1_OMP.f:339): FE, FP accumulate in place; returns
    (SE, SP, E_gam2, S_gam2)."""
    load()
    DE = _np.asfortranarray(DE, dtype=_np.float64)
    DP = _np.asfortranarray(DP, dtype=_np.float64)
    SE = _np.zeros((nebf, nebf), order="F")
    SP = _np.zeros((npbf, npbf), order="F")
    e = _np.zeros(1)
    s = _np.zeros(1)
    _check(_lib.neorx_eg2ic1(nebf, npbf, _p(DE), _p(DP), _p(_np.ascontiguousarray(GM2ICR)),
                             _p(_np.ascontiguousarray(GM2SICR)), _p(FE), _p(SE), _p(FP), _p(SP), _p(e), _p(s)))
    return SE, SP, float(e[0]), float(s[0])


def EG3IC1(nebf, npbf, DE, DP, GM3ICR, FE, FP):
    load()
    DE = _np.asfortranarray(DE, dtype=_np.float64)
    DP = _np.asfortranarray(DP, dtype=_np.float64)
    e = _np.zeros(1)
    _check(_lib.neorx_eg3ic1(nebf, npbf, _p(DE), _p(DP), _p(_np.ascontiguousarray(GM3ICR)), _p(FE), _p(FP), _p(e)))
    return float(e[0])


def EG4ICR(nebf, npbf, DE, DP, GM4ICR, FE, FP):
    load()
    DE = _np.asfortranarray(DE, dtype=_np.float64)
    DP = _np.asfortranarray(DP, dtype=_np.float64)
    e = _np.zeros(1)
    _check(_lib.neorx_eg4icr(nebf, npbf, _p(DE), _p(DP), _p(_np.ascontiguousarray(GM4ICR)), _p(FE), _p(FP), _p(e)))
    return float(e[0])


def precision_switch_dev(nebf, d_de, d_state, threshold=1e-5, stream=0):
    """Mixed-precision switch: returns (use_fp64, max|DE - DE_previous|);
    d_state is a zeroed device array of nebf*nebf + 2 doubles."""
    load()
    u = _C.c_int64()
    d = _C.c_double()
    _check(_lib.neorx_precision_switch_dev(nebf, d_de, d_state, threshold, _C.byref(u), _C.byref(d), stream))
    return bool(u.value), d.value


def call_dev(name, *args):
    """Call a ``neorx_*_dev`` entry point with raw device pointers / ints."""
    load()
    _check(getattr(_lib, name)(*args))


def _fort(a):
    return _np.asfortranarray(a, dtype=_np.float64)


def FAE_OMG2(nae, nbe, nebf, npbf, DAE, DBE, DP, GM2ICR, GM2SICR, FAE):
    """FAE_OMG2 (source/xcuhf/xcuhf_omg2.f:2): FAE accumulates in place;
    returns (SAE2, E, S)."""
    load()
    SAE2 = _np.zeros((nebf, nebf), order="F")
    e = _np.zeros(2)
    _check(_lib.neorx_fae_omg2(nae, nbe, nebf, npbf, _p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)),
                               _p(_np.ascontiguousarray(GM2ICR)), _p(_np.ascontiguousarray(GM2SICR)), _p(FAE),
                               _p(SAE2), _p(e), _p(e[1:])))
    return SAE2, float(e[0]), float(e[1])


def UFP_OMG2(nae, nbe, nebf, npbf, DAE, DBE, DP, GM2ICR, GM2SICR, FP):
    """UFP_OMG2 (source/xcuhf/xcuhf_omg2.f:302): FP accumulates; returns
    (SP2, E, S)."""
    load()
    SP2 = _np.zeros((npbf, npbf), order="F")
    e = _np.zeros(2)
    _check(_lib.neorx_ufp_omg2(nae, nbe, nebf, npbf, _p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)),
                               _p(_np.ascontiguousarray(GM2ICR)), _p(_np.ascontiguousarray(GM2SICR)), _p(FP),
                               _p(SP2), _p(e), _p(e[1:])))
    return SP2, float(e[0]), float(e[1])


def _uhf(name, nae, nbe, nebf, npbf, DAE, DBE, DP, G, F):
    load()
    e = _np.zeros(1)
    _check(getattr(_lib, name)(nae, nbe, nebf, npbf, _p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)),
                               _p(_np.ascontiguousarray(G)), _p(F), _p(e)))
    return float(e[0])


def FAE_OMG3(*a):
    """FAE_OMG3 (source/xcuhf/xcuhf_omg3.f:2): returns E."""
    return _uhf("neorx_fae_omg3", *a)


def UFP_OMG3(*a):
    """UFP_OMG3 (source/xcuhf/xcuhf_omg3.f:284): returns E."""
    return _uhf("neorx_ufp_omg3", *a)


def FAE_OMG4(*a):
    """FAE_OMG4 (source/xcuhf/xcuhf_omg4.f:2): returns E."""
    return _uhf("neorx_fae_omg4", *a)


def UFP_OMG4(*a):
    """UFP_OMG4 (source/xcuhf/xcuhf_omg4.f:512): returns E."""
    return _uhf("neorx_ufp_omg4", *a)


def OMG4_ICR(nebf, npbf, gam_2s, gam_ee):
    """OMG4_ICR (source/xcuhf/xcuhf_omg4.f:1080) with GAM_ee passed in."""
    load()
    g = _np.zeros(nebf ** 8 * npbf * npbf)
    _check(_lib.neorx_omg4_icr(nebf, npbf, _p(_np.ascontiguousarray(gam_2s)), _p(_np.ascontiguousarray(gam_ee)),
                               _p(g)))
    return g


def RXCHFne_OMG2(which, nebf, npbf, DAE, DBE, DP, GM2ICR, F):
    """RXCHFne_{FAE,FP,FBE}_OMG2 (source/rxchfne/rxchfne_omg2.f), which = 'fae',
    'fp' or 'fbe'; F accumulates; returns E."""
    load()
    e = _np.zeros(1)
    _check(getattr(_lib, f"neorx_rxchfne_{which}_omg2")(nebf, npbf, _p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)),
                                                          _p(_np.ascontiguousarray(GM2ICR)), _p(F), _p(e)))
    return float(e[0])


def RXCUHFne_OMG2(which, nebf, npbf, DAalp, DAtot, DBE, DP, GM2ICR, GM2exICR, F, F2=None):
    """RXCUHFne_{FE,FP,FBE}_OMG2 (source/rxchfne/rxcuhfne_omg2.f), which = 'fe'
    (F = FAalpE, F2 = FAbetE), 'fp' or 'fbe'; returns E."""
    load()
    e = _np.zeros(1)
    a = [nebf, npbf, _p(_fort(DAalp)), _p(_fort(DAtot)), _p(_fort(DBE)), _p(_fort(DP)),
         _p(_np.ascontiguousarray(GM2ICR)), _p(_np.ascontiguousarray(GM2exICR)), _p(F)]
    if which == "fe":
        a.append(_p(F2))
    _check(getattr(_lib, f"neorx_rxcuhfne_{which}_omg2")(*a, _p(e)))
    return float(e[0])


def RXCHFne_FAE_GAMee(nebf, GAM_ee, DAE, FAE):
    """RXCHFne_FAE_GAMee (source/rxchfne/rxchfne_hfterms.f:2); returns E."""
    load()
    e = _np.zeros(1)
    _check(_lib.neorx_rxchfne_fae_gamee(nebf, _p(_np.ascontiguousarray(GAM_ee)), _p(_fort(DAE)), _p(FAE), _p(e)))
    return float(e[0])


# ---------------------------------------------------------------------------
# RXCHFmult (source/rxchfmult/)
# ---------------------------------------------------------------------------
def RXCHFmult_GAM4_INT(nebf, nebfBE, npbf, gam_2s, gam_ee):
    """INT_GAM4 (rxchfmult/rxchfmult_gam4-int.f:2) from XCHF_GAM2s and GAM_ee."""
    g = _np.zeros(npbf * npbf * nebf * nebf * nebfBE ** 6)
    load()
    _check(_lib.neorx_rxchfmult_gam4_int(nebf, nebfBE, npbf, _p(_np.ascontiguousarray(gam_2s)),
                                         _p(_np.ascontiguousarray(gam_ee)), _p(g)))
    return g


def RXCHFmult_GAM4_XCHF(nebf, nebfBE, npbf, gam_2s, gam_ee):
    """XCHF_GAM4 (rxchfmult/rxchfmult_gam4-xchf.f:2): GAM4_ICR over the special
    basis with the leading nebfBE^4 block of GAM_ee (regular basis)."""
    ee = _np.ascontiguousarray(_np.asarray(gam_ee).reshape((nebf,) * 4)[:nebfBE, :nebfBE, :nebfBE, :nebfBE])
    return GAM4_ICR(nebfBE, npbf, gam_2s, ee.ravel())


def RXCHFmult_OMG(which, k, nebf, nebfBE, npbf, DAE, DBE, DP, GM, F):
    """RXCHFmult_{FAE,FP,FBE}_OMG{2,3,4} (rxchfmult/rxchfmult_omg{2,3,4}.f):
    which 0/1/2 = FAE/FP/FBE, k = 2, 3, 4.  F (Fortran layout) accumulates;
    returns the energy."""
    load()
    e = _np.zeros(1)
    name = f"neorx_rxchfmult_{('fae', 'fp', 'fbe')[which]}_omg{k}"
    _check(getattr(_lib, name)(nebf, nebfBE, npbf, _p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)),
                               _p(_np.ascontiguousarray(GM)), _p(F), _p(e)))
    return float(e[0])


def RXCHFmult_OMGex(which, order, nebf, npbf, DAE, DBE, DP, GM, GX1, GX2, F):
    """RXCHFmult_{FAE,FP,FBE}_OMG{2,3}ex (rxchfmult/rxchfmult_omg{2,3}ex.f):
    which 0/1/2, order 2 or 3 (GX2 only for 3).  F accumulates; returns E."""
    load()
    e = _np.zeros(1)
    name = f"neorx_rxchfmult_{('fae', 'fp', 'fbe')[which]}_omg{order}ex"
    args = [_p(_fort(DAE)), _p(_fort(DBE)), _p(_fort(DP)), _p(_np.ascontiguousarray(GM)),
            _p(_np.ascontiguousarray(GX1))]
    if order == 3:
        args.append(_p(_np.ascontiguousarray(GX2)))
    _check(getattr(_lib, name)(nebf, npbf, *args, _p(F), _p(e)))
    return float(e[0])


def autotune_eg3_dev(nebf, npbf, dtype, d_de, d_gm3, reps=3):
    """neorx_autotune_eg3_dev: (best loads-in-flight, its ms, ms of the 8, 12,
    16 and 24 variants) for storage dtype 0 FP64, 1 FP32, 2 BF16."""
    load()
    b, ms, allms = _C.c_int64(), _C.c_double(), (_C.c_double * 4)()
    _check(_lib.neorx_autotune_eg3_dev(nebf, npbf, dtype, d_de, d_gm3, reps, _C.byref(b), _C.byref(ms), allms))
    return b.value, ms.value, list(allms)
