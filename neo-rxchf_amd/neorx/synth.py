"""Seeded synthetic systems in the reference's array layout.

A system is a dict with the arrays that GAM2_IC1 receives
(source/gam2_IC1_OMP.f:2-37): Cartesian primitives ELCAM/ELCEX/ELCBFC,
contraction maps KPESTR/KPEEND/AMPEB2C (1-based, contiguous primitives per
contracted function as make_KPE requires, source/xcneo_HYBRID.f:1269-1307),
contraction coefficients AGEBFCC/AGNBFCC that already include primitive
normalisation, classical nuclei cat(3,nat)/zan, the geminal bcoef/gamma and
the proton mass.  Lengths are in bohr.
"""
import math

import numpy as np

PMASS = 1836.15267245  # proton mass in electron masses

_SHELL = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2),
        (1, 1, 1)],
}


def _dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def prim_norm(a, lx, ly, lz):
    """Normalisation of a Cartesian Gaussian x^lx y^ly z^lz exp(-a r^2)."""
    l = lx + ly + lz
    return (2 * a / math.pi) ** 0.75 * (4 * a) ** (l / 2) / math.sqrt(
        _dfact(2 * lx - 1) * _dfact(2 * ly - 1) * _dfact(2 * lz - 1))


def _ovl_same_centre(a, b, l3):
    s = (math.pi / (a + b)) ** 1.5
    for l in l3:
        s *= _dfact(2 * l - 1) / (2 * (a + b)) ** l
    return s


def _finish(prims, contr, nuc, cat, zan, bcoef, gamma, pmass=PMASS, ncoef=None):
    """prims: list of (l3, exponent, centre, coef) in contracted order;
    contr: number of primitives of each contracted function."""
    npebf = len(prims)
    nebf = len(contr)
    elcam = np.zeros((npebf, 3), dtype=np.int64)
    elcex = np.zeros(npebf)
    elcbfc = np.zeros((npebf, 3))
    agebfcc = np.zeros(npebf)
    kpestr = np.zeros(nebf, dtype=np.int64)
    kpeend = np.zeros(nebf, dtype=np.int64)
    ampeb2c = np.zeros(npebf, dtype=np.int64)
    k = 0
    for c, n in enumerate(contr):
        kpestr[c] = k + 1
        kpeend[c] = k + n
        for j in range(k, k + n):
            l3, a, cen, cf = prims[j]
            elcam[j] = l3
            elcex[j] = a
            elcbfc[j] = cen
            agebfcc[j] = cf * prim_norm(a, *l3)
            ampeb2c[j] = c + 1
        # renormalise the contracted function (same centre and powers)
        s = 0.0
        for i in range(k, k + n):
            for j in range(k, k + n):
                s += agebfcc[i] * agebfcc[j] * _ovl_same_centre(elcex[i], elcex[j], elcam[i])
        agebfcc[k:k + n] /= math.sqrt(s)
        k += n
    npbf = len(nuc)
    nucam = np.array([p[0] for p in nuc], dtype=np.int64).reshape(npbf, 3)
    nucex = np.array([p[1] for p in nuc], dtype=np.float64)
    nucbfc = np.array([p[2] for p in nuc], dtype=np.float64).reshape(npbf, 3)
    agnbfcc = np.array([prim_norm(p[1], *p[0]) for p in nuc]) if ncoef is None else np.asarray(ncoef, float)
    cat = np.asarray(cat, dtype=np.float64).reshape(-1, 3)
    return dict(
        nebf=nebf, npebf=npebf, npbf=npbf, nat=len(zan), ngtg=len(bcoef), pmass=float(pmass),
        cat=np.ascontiguousarray(cat), zan=np.asarray(zan, dtype=np.float64),
        bcoef=np.asarray(bcoef, dtype=np.float64), gamma=np.asarray(gamma, dtype=np.float64),
        kpestr=kpestr, kpeend=kpeend, ampeb2c=ampeb2c, agebfcc=agebfcc, agnbfcc=agnbfcc,
        elcex=elcex, nucex=nucex,
        elcam=np.asfortranarray(elcam), nucam=np.asfortranarray(nucam),
        elcbfc=np.asfortranarray(elcbfc), nucbfc=np.asfortranarray(nucbfc),
    )


def _logu(rng, lo, hi, n=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))


def config1(seed=1):
    """Plain NEO-XCHF RHF system: classical C (Z=6, z=0) and N (Z=7, z=2.1),
    quantum proton at z=4.3; 36 s/p electronic primitives contracted to 24
    functions (per classical centre 3 s functions from 3/1/2 primitives and
    3 p shells from 2/1/1 primitives), exponents log-uniform in [0.1, 50],
    coefficients uniform in [0.1, 1]; 8 nuclear primitives (s, p shell, four
    d components xx,yy,zz,xy) on the proton, exponents log-uniform in [4, 32];
    3-term geminal, b in [0.1, 0.5], gamma log-uniform in [1, 30]."""
    rng = np.random.default_rng(seed)
    cents = [np.array([0.0, 0.0, 0.0]), np.array([0.0, 0.0, 2.1])]
    prims, contr = [], []
    for c in cents:
        for n in (3, 1, 2):
            ex = np.sort(_logu(rng, 0.1, 50.0, n))[::-1]
            cf = rng.uniform(0.1, 1.0, n)
            for a, f in zip(ex, cf):
                prims.append(((0, 0, 0), a, c, f))
            contr.append(n)
        for n in (2, 1, 1):
            ex = np.sort(_logu(rng, 0.1, 50.0, n))[::-1]
            cf = rng.uniform(0.1, 1.0, n)
            for l3 in _SHELL[1]:
                for a, f in zip(ex, cf):
                    prims.append((l3, a, c, f))
                contr.append(n)
    hpos = np.array([0.0, 0.0, 4.3])
    nl = [(0, 0, 0)] + _SHELL[1] + _SHELL[2][:4]
    nex = _logu(rng, 4.0, 32.0, len(nl))
    nuc = [(l3, a, hpos) for l3, a in zip(nl, nex)]
    b = rng.uniform(0.1, 0.5, 3)
    g = _logu(rng, 1.0, 30.0, 3)
    sysd = _finish(prims, contr, nuc, [c for c in cents], [6.0, 7.0], b, g)
    sysd["nelec"] = 14
    sysd["name"] = "C1 NEO-XCHF RHF (C/N/H+ line, 24/36 e-functions, 8 p-functions, 3 geminals)"
    return sysd


def config2(seed=2):
    """Integral-kernel system: 60 electronic primitives on 4 random centres
    in a 6-bohr box (per centre 3 s, 2 p shells and 1 d shell, exponents
    log-uniform in [0.1, 50], each primitive its own contracted function), 16
    nuclear primitives on centre 0 (4 s, 2 p shells, 1 d shell; exponents
    log-uniform in [4, 32]), a 4-term geminal, and a classical nucleus of
    charge 1 at each of the 4 centres."""
    rng = np.random.default_rng(seed)
    cents = [rng.uniform(0.0, 6.0, 3) for _ in range(4)]
    prims, contr = [], []
    for c in cents:
        shells = [0, 0, 0, 1, 1, 2]
        for l in shells:
            a = float(_logu(rng, 0.1, 50.0))
            for l3 in _SHELL[l]:
                prims.append((l3, a, c, 1.0))
                contr.append(1)
    nuc = []
    for l in [0, 0, 0, 0, 1, 1, 2]:
        a = float(_logu(rng, 4.0, 32.0))
        for l3 in _SHELL[l]:
            nuc.append((l3, a, cents[0]))
    b = rng.uniform(0.1, 0.5, 4)
    g = _logu(rng, 1.0, 30.0, 4)
    sysd = _finish(prims, contr, nuc, cents, [1.0] * 4, b, g)
    sysd["nelec"] = 4
    sysd["name"] = "C2 integral kernels (60 e-prims s/p/d on 4 centres, 16 p-prims l<=2, 4 geminals)"
    return sysd


def config3(seed=3):
    """RXCHF-ne closed-shell system with 2 special electrons: 20 regular plus
    2 special electrons; 48 electronic functions (one primitive each) on 5
    random centres in a 6-bohr box, per centre 3 s shells and 1 p shell, and
    a Cartesian d shell on each of the first 3 centres (5 x 6 + 3 x 6 = 48),
    exponents log-uniform in [0.08, 80]; 12 nuclear functions (3 s, 1 p and
    1 d shell, exponents log-uniform in [4, 32]) on centre 0; 3-term geminal,
    b in [0.1, 0.5], gamma log-uniform in [1, 30]; a classical nucleus of
    charge 4 at each centre."""
    rng = np.random.default_rng(seed)
    cents = [rng.uniform(0.0, 6.0, 3) for _ in range(5)]
    prims, contr = [], []
    for k, c in enumerate(cents):
        for l in [0, 0, 0, 1] + ([2] if k < 3 else []):
            a = float(_logu(rng, 0.08, 80.0))
            for l3 in _SHELL[l]:
                prims.append((l3, a, c, 1.0))
                contr.append(1)
    nuc = []
    for l in [0, 0, 0, 1, 2]:
        a = float(_logu(rng, 4.0, 32.0))
        for l3 in _SHELL[l]:
            nuc.append((l3, a, cents[0]))
    b = rng.uniform(0.1, 0.5, 3)
    g = _logu(rng, 1.0, 30.0, 3)
    sysd = _finish(prims, contr, nuc, cents, [4.0] * 5, b, g)
    sysd["nelec"] = 22
    sysd["nbe"] = 2
    sysd["name"] = "C3 RXCHF-ne closed shell (48 e-functions s/p/d on 5 centres, 12 p-functions, 3 geminals)"
    return sysd


def config5(k, n=64, seed=5):
    """Instance k of n of the batched many-instance NEO-XCHF RHF workload, at
    reduced basis sizes: a linear frame with classical nuclei of charge 2 at
    z = 0 and charge 1 at z = 3.0 bohr, 4 electrons, and the quantum proton's
    basis centre scanned over n points in [-0.6, 0.6] bohr about z = 1.5
    along the bond.  Uneven sizes per instance (drawn from the seed and k):
    on each classical centre 2 or 3 s functions with even-tempered exponents
    (0.35, 1.4, 5.6, each scaled by a factor in [0.9, 1.1]) and optionally a
    p shell (exponent 0.9, same scaling), 4 to 12 functions of one primitive
    each; nuclear basis an s shell and optionally a p shell (1 or 4
    functions), exponents in [8, 16]; a 3-term geminal shared by all
    instances."""
    rng = np.random.default_rng(seed * 1000 + k)
    zc = [np.array([0.0, 0.0, 0.0]), np.array([0.0, 0.0, 3.0])]
    zp = np.array([0.0, 0.0, 1.5 - 0.6 + 1.2 * k / max(1, n - 1)])
    prims, contr = [], []
    for c in zc:
        ns = int(rng.integers(2, 4))
        f = rng.uniform(0.9, 1.1)
        for a in (0.35, 1.4, 5.6)[:ns]:
            prims.append(((0, 0, 0), a * f, c, 1.0))
            contr.append(1)
        if rng.random() < 0.5:
            for l3 in _SHELL[1]:
                prims.append((l3, 0.9 * f, c, 1.0))
                contr.append(1)
    nuc = []
    for l in [0] + ([1] if rng.random() < 0.5 else []):
        a = float(rng.uniform(8.0, 16.0))
        for l3 in _SHELL[l]:
            nuc.append((l3, a, zp))
    g_rng = np.random.default_rng(seed)
    b = g_rng.uniform(0.1, 0.5, 3)
    g = _logu(g_rng, 1.0, 30.0, 3)
    sysd = _finish(prims, contr, nuc, zc, [2.0, 1.0], b, g)
    sysd["nelec"] = 4
    sysd["name"] = f"C5 instance {k}: proton at z = {zp[2]:.4f} bohr"
    return sysd


def small(seed=0, n_e_centres=2, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1), ngtg=3, nat=2, maxl=None):
    """Small random system for parity tests: s/p/d/f shells on random
    centres, contracted functions of 1-2 primitives."""
    rng = np.random.default_rng(seed)
    cents = [rng.uniform(-1.5, 1.5, 3) for _ in range(n_e_centres)]
    prims, contr = [], []
    for c in cents:
        for l in e_shells:
            n = int(rng.choice(contract))
            ex = np.sort(_logu(rng, 0.2, 8.0, n))[::-1]
            cf = rng.uniform(0.2, 1.0, n)
            for l3 in _SHELL[l]:
                for a, f in zip(ex, cf):
                    prims.append((l3, a, c, f))
                contr.append(n)
    ppos = rng.uniform(-0.5, 0.5, 3)
    nuc = []
    for l in p_shells:
        a = float(_logu(rng, 4.0, 32.0))
        for l3 in _SHELL[l]:
            nuc.append((l3, a, ppos))
    catv = [rng.uniform(-1.5, 1.5, 3) for _ in range(nat)]
    zan = rng.integers(1, 4, nat).astype(float)
    b = rng.uniform(0.1, 0.5, ngtg)
    g = _logu(rng, 1.0, 30.0, ngtg)
    return _finish(prims, contr, nuc, catv, zan, b, g)


BE_KEYS = ("kpestr", "kpeend", "ampeb2c", "agebfcc", "elcex", "elcam", "elcbfc")


def mult(seed=0, e_shells=(0, 1), be_shells=(0, 1), be_centres=1, **kw):
    """Two-basis system for the RXCHFmult routines: the regular electron basis
    of small(seed, ...) plus a special-electron basis (keys nebfBE, npebfBE and
    <key>_be) drawn from another seed on be_centres centres, sharing nuclei,
    the proton basis and the geminal."""
    a = small(seed=seed, e_shells=e_shells, **kw)
    b = small(seed=seed + 7919, e_shells=be_shells, n_e_centres=be_centres,
              **{k: v for k, v in kw.items() if k != "n_e_centres"})
    out = dict(a)
    out["nebfBE"], out["npebfBE"] = b["nebf"], b["npebf"]
    for k in BE_KEYS:
        out[k + "_be"] = b[k]
    return out


def be_part(sysd):
    """The special-electron basis of a two-basis system as a one-basis system
    (for the XCHF routines over that basis: GAM1_MD, GAM_ee, ...)."""
    out = dict(sysd)
    out["nebf"], out["npebf"] = sysd["nebfBE"], sysd["npebfBE"]
    for k in BE_KEYS:
        out[k] = sysd[k + "_be"]
    return out


def combined(sysd):
    """The primitive list of a two-basis handle as a one-basis dict: regular
    primitives first, then the special-electron ones (contracted functions and
    KPESTR/KPEEND/AMPEB2C offset to match).  Primitive-tuple indices of the
    RXCHFmult classes refer to this list."""
    out = dict(sysd)
    na, pa = sysd["nebf"], sysd["npebf"]
    out["nebf"], out["npebf"] = na + sysd["nebfBE"], pa + sysd["npebfBE"]
    out["kpestr"] = np.concatenate([sysd["kpestr"], sysd["kpestr_be"] + pa])
    out["kpeend"] = np.concatenate([sysd["kpeend"], sysd["kpeend_be"] + pa])
    out["ampeb2c"] = np.concatenate([sysd["ampeb2c"], sysd["ampeb2c_be"] + na])
    for k in ("agebfcc", "elcex"):
        out[k] = np.concatenate([sysd[k], sysd[k + "_be"]])
    for k in ("elcam", "elcbfc"):
        out[k] = np.asfortranarray(np.concatenate([sysd[k], sysd[k + "_be"]], axis=0))
    for k in list(out):
        if k.endswith("_be") or k in ("nebfBE", "npebfBE"):
            del out[k]
    return out


def density(n, trace, seed=0):
    """Random symmetric positive semidefinite n x n matrix with given trace."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    m = a @ a.T
    return np.asfortranarray(m * (trace / np.trace(m)))
