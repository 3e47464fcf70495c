"""NEO-XCHF RHF on the device: the reference's input and orbital files, and
the SCF context of libneorx_hip.so.

* ``read_input`` parses the reference's input file (namelists SYSINFO,
  GEMINAL, the ATOMS / EBASIS / PBASIS blocks; source/xcneo_HYBRID.f:280-460)
  into the system dict of :mod:`neorx.synth`, with the reference's unit
  conversion (angstrom -> bohr, a2bohr = 1/0.529177249, xcneo_HYBRID.f:97)
  and normalisations: ELCNORM3 (xchf1_independent/pure_ovlap.f:245-363) for
  the electronic contraction coefficients and NUCNORM
  (xcneo_HYBRID.f:1309-1347) for the nuclear ones.
* ``write_input`` writes a system dict back in that format (coordinates in
  angstrom, contracted functions on contiguous primitive lines as make_KPE
  requires).
* ``read_mos`` / ``write_mos`` read and write orbital files (guessCE.inp,
  FinalCE.dat, ...) in the PUSQLF layout: per orbital, lines of
  ``FORMAT(I2,I3,1P,5E15.8)`` (GAM_IO.f:247-330, :471-520).
* ``XCHF`` wraps the device context: every integral of the XCHF energy in
  HBM, the Fock build of xchf1_fock (xchf1_fock.f:2), ROOTHAN (xcscf.f:766)
  and the SCF loop of xcscf (xcscf.f:2).
"""
import ctypes as _C
import math
import re

import numpy as np

from . import synth

A2BOHR = 1.0 / 0.529177249

# energy components of xchf1_fock, in the order of the C ABI's comps array
COMPS = ("E_TOTAL", "E_NUC", "E_ECORE", "E_PCORE", "E_EP", "E_EE", "E_GAM1", "E_GAM2", "E_GAM3", "E_GAM4",
         "S_TOTAL", "S_GAM1", "S_GAM2", "H_EXPECT")

PREC = {"f64": 0, "f32": 1, "bf16": 2}


def _num(s):
    return float(s.replace("d", "e").replace("D", "e"))


def _namelist(text, name):
    m = re.search(r"&" + name + r"\b(.*?)/", text, re.S | re.I)
    if not m:
        raise ValueError(f"namelist &{name} not found")
    out = {}
    for k, v in re.findall(r"([A-Za-z_][A-Za-z0-9_]*(?:\(\d+\))?)\s*=\s*([^,\s/]+)", m.group(1)):
        out[k.lower()] = v
    return out


def _block(lines, tag, n):
    for i, l in enumerate(lines):
        if l.strip().upper() == tag:
            rows = [l2.split() for l2 in lines[i + 1:i + 1 + n]]
            if len(rows) < n or any(not r for r in rows):
                raise ValueError(f"{tag}: expected {n} lines")
            return rows
    raise ValueError(f"{tag} block not found")


def _ovl1d(l1, a, A, l2, b, B):
    """1-D overlap of (x-A)^l1 exp(-a(x-A)^2) and (x-B)^l2 exp(-b(x-B)^2)."""
    p = a + b
    P = (a * A + b * B) / p
    pre = math.exp(-a * b / p * (A - B) ** 2) * math.sqrt(math.pi / p)
    tot = 0.0
    for i in range(l1 + 1):
        for j in range(l2 + 1):
            k = i + j
            if k % 2:
                continue
            mom = synth._dfact(k - 1) / (2 * p) ** (k // 2)
            tot += math.comb(l1, i) * math.comb(l2, j) * (P - A) ** (l1 - i) * (P - B) ** (l2 - j) * mom
    return pre * tot


def gfovlap(l1, a, A, l2, b, B):
    """Overlap of two Cartesian Gaussians (gfovlap, AC_Integrals/ovlap.f)."""
    return _ovl1d(l1[0], a, A[0], l2[0], b, B[0]) * _ovl1d(l1[1], a, A[1], l2[1], b, B[1]) * _ovl1d(
        l1[2], a, A[2], l2[2], b, B[2])


def read_input(path):
    """System dict of a reference input file (lengths converted to bohr,
    coefficients normalised as ELCNORM3 / NUCNORM do); adds nelec, nucst and
    the input's name."""
    text = open(path).read()
    lines = text.splitlines()
    si = _namelist(text, "SYSINFO")
    nat, nebf, npebf, npbf = (int(si[k]) for k in ("nat", "nebf", "npebf", "npbf"))
    ngtg, nelec = int(si["ngtg"]), int(si["nelec"])
    nucst = int(si.get("nucst", "1"))
    pmass = _num(si["pmass"])
    gm = _namelist(text, "GEMINAL")
    b = np.array([_num(gm[f"bgem({k + 1})"]) for k in range(ngtg)])
    g = np.array([_num(gm[f"ggem({k + 1})"]) for k in range(ngtg)])
    at = _block(lines, "&ATOMS", nat)
    zan = np.array([_num(r[0]) for r in at])
    cat = np.array([[_num(x) * A2BOHR for x in r[1:4]] for r in at])
    eb = _block(lines, "&EBASIS", npebf)
    ampeb2c = np.array([int(r[1]) for r in eb], dtype=np.int64)
    elcam = np.array([[int(x) for x in r[2:5]] for r in eb], dtype=np.int64)
    elcex = np.array([_num(r[5]) for r in eb])
    agebfcc = np.array([_num(r[6]) for r in eb])
    elcbfc = np.array([[_num(x) * A2BOHR for x in r[7:10]] for r in eb])
    pb = _block(lines, "&PBASIS", npbf)
    nucam = np.array([[int(x) for x in r[2:5]] for r in pb], dtype=np.int64)
    nucex = np.array([_num(r[5]) for r in pb])
    agnbfcc = np.array([_num(r[6]) for r in pb])
    nucbfc = np.array([[_num(x) * A2BOHR for x in r[7:10]] for r in pb])
    # make_KPE (xcneo_HYBRID.f:1269-1307): contiguous primitives per function
    kpestr = np.zeros(nebf, dtype=np.int64)
    kpeend = np.zeros(nebf, dtype=np.int64)
    for i, c in enumerate(ampeb2c):
        if kpestr[c - 1] == 0:
            kpestr[c - 1] = i + 1
        kpeend[c - 1] = i + 1
    # ELCNORM3: each primitive to unit self-overlap, then each contracted
    # function to unit norm over its primitives
    for i in range(npebf):
        agebfcc[i] /= math.sqrt(gfovlap(elcam[i], elcex[i], elcbfc[i], elcam[i], elcex[i], elcbfc[i]))
    kappa = np.zeros(nebf)
    for i in range(npebf):
        for j in range(npebf):
            if ampeb2c[i] == ampeb2c[j]:
                kappa[ampeb2c[i] - 1] += agebfcc[i] * agebfcc[j] * gfovlap(elcam[i], elcex[i], elcbfc[i], elcam[j],
                                                                           elcex[j], elcbfc[j])
    agebfcc = agebfcc / np.sqrt(kappa[ampeb2c - 1])
    # NUCNORM: each nuclear primitive normalised at the origin
    for i in range(npbf):
        agnbfcc[i] /= math.sqrt(gfovlap(nucam[i], nucex[i], (0, 0, 0), nucam[i], nucex[i], (0, 0, 0)))
    return dict(
        nebf=nebf, npebf=npebf, npbf=npbf, nat=nat, ngtg=ngtg, pmass=pmass,
        cat=np.ascontiguousarray(cat), zan=zan, bcoef=b, gamma=g,
        kpestr=kpestr, kpeend=kpeend, ampeb2c=ampeb2c, agebfcc=agebfcc, agnbfcc=agnbfcc,
        elcex=elcex, nucex=nucex, elcam=np.asfortranarray(elcam), nucam=np.asfortranarray(nucam),
        elcbfc=np.asfortranarray(elcbfc), nucbfc=np.asfortranarray(nucbfc),
        nelec=nelec, nucst=nucst, name=str(path),
    )


def write_input(path, sysd, raw_ecoef=None, raw_ncoef=None):
    """Write a system dict as a reference input file.  The reader normalises
    the coefficients again, so pass the raw (unnormalised) coefficients when
    they are known; by default the stored ones are written, which ELCNORM3 /
    NUCNORM leave unchanged up to rounding only for already-normalised sets."""
    s = sysd
    ec = s["agebfcc"] if raw_ecoef is None else raw_ecoef
    nc = s["agnbfcc"] if raw_ncoef is None else raw_ncoef
    out = ["&SYSINFO", f" nat={s['nat']},", f" nebf={s['nebf']},", f" npebf={s['npebf']},", f" npbf={s['npbf']},",
           f" ngtg={s['ngtg']},", f" pmass={s['pmass']:.10e},".replace("e", "d"), f" nelec={s['nelec']},",
           f" nucst={s.get('nucst', 1)}/", "", "&GEMINAL"]
    gl = [f" bgem({k + 1})={s['bcoef'][k]:.12f}, ggem({k + 1})={s['gamma'][k]:.12f}" for k in range(s["ngtg"])]
    out.append(",\n".join(gl) + "/")
    out += ["", "&ATOMS"]
    for z, c in zip(s["zan"], np.asarray(s["cat"]).reshape(-1, 3)):
        out.append(f" {z:.1f} {c[0] / A2BOHR:.12f} {c[1] / A2BOHR:.12f} {c[2] / A2BOHR:.12f} X")
    out += ["", "&EBASIS"]
    for i in range(s["npebf"]):
        am, c = s["elcam"][i], s["elcbfc"][i] / A2BOHR
        out.append(f"{i + 1} {s['ampeb2c'][i]} {am[0]} {am[1]} {am[2]} {s['elcex'][i]:.12f} {ec[i]:.12f} "
                   f"{c[0]:.12f} {c[1]:.12f} {c[2]:.12f}")
    out += ["", "&PBASIS"]
    for i in range(s["npbf"]):
        am, c = s["nucam"][i], s["nucbfc"][i] / A2BOHR
        out.append(f"{i + 1} {i + 1} {am[0]} {am[1]} {am[2]} {s['nucex'][i]:.12f} {nc[i]:.12f} "
                   f"{c[0]:.12f} {c[1]:.12f} {c[2]:.12f}")
    open(path, "w").write("\n".join(out) + "\n")


def read_mos(path, n):
    """n x n orbital matrix (column-major order: column j is orbital j) from a
    PUSQLF-format file, checking the orbital and line labels as READMO does."""
    C = np.zeros((n, n), order="F")
    with open(path) as f:
        lines = [l.rstrip("\n") for l in f if l.strip()]
    pos = 0
    for j in range(n):
        for ic, i0 in enumerate(range(0, n, 5)):
            l = lines[pos]
            pos += 1
            jj, icc = int(l[0:2]), int(l[2:5])
            if jj != (j + 1) % 100 or icc != (ic + 1) % 1000:
                raise ValueError(f"{path}: orbital {j + 1} line {ic + 1} labelled {jj} {icc}")
            body = l[5:]
            for k in range(min(5, n - i0)):
                C[i0 + k, j] = _num(body[15 * k:15 * k + 15])
    return C


def write_mos(path, C):
    """PUSQLF (GAM_IO.f:471-520): FORMAT(I2,I3,1P,5E15.8) per orbital."""
    C = np.asarray(C)
    n = C.shape[0]
    out = []
    for j in range(C.shape[1]):
        for ic, i0 in enumerate(range(0, n, 5)):
            vals = "".join(f"{v:15.8E}" for v in C[i0:i0 + 5, j])
            out.append(f"{(j + 1) % 100:2d}{(ic + 1) % 1000:3d}{vals}")
    open(path, "w").write("\n".join(out) + "\n")


class XCHF:
    """Device context of a NEO-XCHF RHF calculation (neorx_xchf_create):
    builds and keeps in HBM the integrals of xcneo_HYBRID.f:663-1002 for the
    System ``S`` -- E_nuc, the overlaps, GAM_ecore/pcore, GAM_ep, GAM_ee,
    GAM_1/1s, GM2ICR/GM2sICR and GM3ICR (in the storage types of ``store``:
    a set of "f64", "f32", "bf16").  GAM_4 is not stored; its contraction is
    evaluated in factorized form."""

    def __init__(self, S, nelec, store=("f64",), autotune=False):
        from . import _check, _lib, load
        load()
        self._check, self._lib = _check, _lib
        bits = sum({"f64": 1, "f32": 2, "bf16": 4}[k] for k in store) | (8 if autotune else 0)
        h = _C.c_int64()
        _check(_lib.neorx_xchf_create(S.handle, nelec, bits, _C.byref(h)))
        self.ctx = h.value
        self.ne, self.np, self.nelec = S.s["nebf"], S.s["npbf"], nelec

    def close(self):
        if self.ctx:
            self._check(self._lib.neorx_xchf_destroy(self.ctx))
            self.ctx = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def integrals(self):
        """E_nuc, S_e, S_p, GAM_ecore, GAM_pcore, GAM_ep, GAM_ee (reference
        packings: pack_2D, pack_4D)."""
        ne, npb = self.ne, self.np
        e = _C.c_double()
        a = [np.zeros(n) for n in (ne * ne, npb * npb, ne * ne, npb * npb, ne * ne * npb * npb, ne ** 4)]
        self._check(self._lib.neorx_xchf_integrals(self.ctx, _C.byref(e), *[x.ctypes.data_as(_C.c_void_p) for x in a]))
        return (e.value, *a)

    def fock_dev(self, d_de, d_dp, prec, d_fe, d_fp, d_comps, stream=0, g3pass=0):
        """xchf1_fock (xchf1_fock.f:2) on device arrays: FE, FP (after
        XCfock_correction) and the energy components (COMPS order).  g3pass:
        0 full; 1 full, kept as the reference (FP64); 2 the reference plus the
        increment from it, read in the storage prec."""
        p = PREC[prec] if isinstance(prec, str) else prec
        self._check(self._lib.neorx_xchf_fock_dev(self.ctx, d_de, d_dp, p | (g3pass << 4), d_fe, d_fp, d_comps,
                                                  stream))

    def scf(self, ce_guess=None, nucst=1, maxit=100, tole=1e-6, tolp=1e-4, mode="f64", graph=True, diis=False,
            sw_hi=1e-3, sw_lo=1e-5):
        """xcscf (xcscf.f:2): returns a dict with the final and initial-guess
        energy components, the orbitals, eigenvalues and the per-iteration
        trace [E_total, dE, max|dDE|, max|dDP|, switch signal, GAM_3 storage
        (0 FP64, 1 FP32, 2 BF16), ms, GAM_3 pass (0 full, 1 full kept as the
        reference, 2 increment)].  mode: "f64"; "f32" (FP32 then FP64);
        "bf16" (BF16, FP32, FP64); "inc" (BF16, then BF16-stored increments
        from reference builds: the first from the FP32 copy when the context
        holds one, the later ones and the converging build from FP64)."""
        ne, npb = self.ne, self.np
        m = {"f64": 0, "f32": 1, "bf16": 2, "inc": 3}[mode]
        ce0 = None if ce_guess is None else np.asfortranarray(ce_guess, dtype=np.float64)
        ci, cf = np.zeros(14), np.zeros(14)
        CE, CP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
        evE, evP = np.zeros(ne), np.zeros(npb)
        tr = np.zeros((maxit, 8))
        nit = _C.c_int64()
        p = lambda a: None if a is None else a.ctypes.data_as(_C.c_void_p)  # noqa: E731
        self._check(self._lib.neorx_xchf_scf(self.ctx, p(ce0), nucst, maxit, tole, tolp, m,
                                             (1 if graph else 0) | (2 if diis else 0), sw_hi, sw_lo, p(ci), p(cf),
                                             p(CE), p(CP), p(evE), p(evP), p(tr), _C.byref(nit)))
        n = nit.value
        return dict(comps=dict(zip(COMPS, cf)), comps_init=dict(zip(COMPS, ci)), CE=CE, CP=CP, evE=evE, evP=evP,
                    trace=tr[:n, :8], niter=n)


def scf_batch(contexts, workers=16, **kw):
    """The SCF of several XCHF contexts at once: up to ``workers`` host
    threads, each driving one context's own stream and captured graphs
    (capture is per thread), so the small kernels of many instances run on
    the GPU side by side.  Keyword arguments go to every XCHF.scf call;
    returns their results in input order.  Contexts made with autotune=True
    are tuned when they are built, never during the SCF, so concurrent runs
    share nothing but read-only tables."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        return list(ex.map(lambda c: c.scf(**kw), contexts))


def roothan_dev(n, d_s, d_f, kstart, klast, prefac, d_c, d_ev, d_d, stream=0):
    """ROOTHAN (xcscf.f:766) on device arrays: orbitals kstart..klast (1-based)
    occupied with weight prefac."""
    from . import _check, _lib, load
    load()
    _check(_lib.neorx_roothan_dev(n, d_s, d_f, kstart, klast, prefac, d_c, d_ev, d_d, stream))
