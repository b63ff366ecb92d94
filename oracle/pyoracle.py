"""ctypes bindings for the oracle (liborad_oracle.so) and the reference-kernel
library (oracle/_ref/libref.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by openradioss_amd/.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborad_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

# Column layout of one element's state row (orad_bstate_t).
ST_SIG = slice(0, 6)
ST_EINT, ST_RHO, ST_VOL0, ST_PLA, ST_QVIS, ST_OFFG, ST_EPSD, ST_DMG = range(6, 14)
NST = 14


class Law2(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("rho0", "shear", "bulk", "ca", "cb", "cn", "epmx", "sigmx",
                 "cc", "epdr", "fisokin", "mexp")] + \
               [(n, C.c_int) for n in ("icc", "vp", "israte", "pad")]


class SProp(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("qa", "qb", "hvis", "dtfac")]


class BOut(C.Structure):
    _fields_ = [("f", C.c_double * 24), ("sti", C.c_double), ("dtx", C.c_double),
                ("ehou", C.c_double), ("voln", C.c_double), ("mstop", C.c_int),
                ("pad", C.c_int)]


class Int7(C.Structure):
    """orad_int7_t: one /INTER/TYPE7 interface."""
    _fields_ = [("nsn", C.c_int), ("nrtm", C.c_int), ("nsv", C.c_void_p),
                ("irect", C.c_void_p), ("stfm", C.c_void_p), ("stfns", C.c_void_p)] + \
               [(n, C.c_double) for n in ("gap", "fric", "visc", "viscf", "econt", "econtv",
                                          "econtd")] + \
               [("npen", C.c_int), ("nsky", C.c_int)]


class OModel(C.Structure):
    _fields_ = [("numnod", C.c_int), ("numels", C.c_int),
                ("x", C.c_void_p), ("v", C.c_void_p), ("d", C.c_void_p),
                ("ms", C.c_void_p), ("icodt", C.c_void_p), ("ixs", C.c_void_p),
                ("adsky", C.c_void_p), ("iads", C.c_void_p), ("st", C.c_void_p),
                ("fsky", C.c_void_p), ("a", C.c_void_p), ("stifn", C.c_void_p),
                ("mat", Law2), ("prop", SProp),
                ("tt", C.c_double), ("dt1", C.c_double), ("dt2", C.c_double),
                ("dt12", C.c_double), ("dt2old", C.c_double), ("dtmx", C.c_double),
                ("ehour", C.c_double), ("ncycle", C.c_int), ("mstop", C.c_int),
                ("nthreads", C.c_int), ("pad", C.c_int),
                ("fv_dir", C.c_void_p), ("fv_val", C.c_void_p),
                ("fv_mem", C.c_void_p), ("wfext", C.c_double),
                ("encin", C.c_double), ("enrot", C.c_double), ("contact", C.c_void_p)]


class ONodes(C.Structure):
    """orad_onodes_t: the nodal part of one cycle (oracle_nodal_update)."""
    _fields_ = [("numnod", C.c_int), ("rot", C.c_int)] + \
               [(n, C.c_void_p) for n in ("x", "v", "d", "a", "ms", "icodt", "vr", "ar",
                                          "inr", "icodr", "fv_dir", "fv_val", "fv_mem")] + \
               [(n, C.c_double) for n in ("tt", "dt1", "dt2", "dt12", "dt2old", "dtmx",
                                          "wfext")] + \
               [("ncycle", C.c_int), ("nthreads", C.c_int),
                ("encin", C.c_double), ("enrot", C.c_double)]


# ---- shells: one element's state row (orad_shstate_t), npt <= 5 ----
SH_FOR = slice(0, 5)
SH_MOM = slice(5, 8)
SH_THK = 8
SH_EINT = slice(9, 11)
SH_EPSD = 11
SH_HOUR = slice(12, 17)
SH_OFFG = 17
SH_SIG0 = 18          # sig[pt][c] at 18 + 5*pt + c
SH_PLA0 = 43          # pla[pt]
SH_EPSDP0 = 48        # epsdp[pt]
SH_HOURX = slice(53, 60)   # QEPH: HOURG(6:12)
SH_NST = 60
SH_NST_BT = 53        # the Belytschko-Tsay state ends before HOURG(6:12)
# QBAT (Ishell=12): columns 0..17 as above (HOUR unused), then per Gauss point g
# (0..3, the order of CBAFORC3's IS/IR loops) and through-thickness point p:
QB_SIG0 = 18          # sig[g][p][c] at 18 + 5*(5*g + p) + c
QB_PLA0 = 118         # pla[g][p] at 118 + 5*g + p
QB_EPSDP0 = 138       # epsdp[g][p]
QB_FORPG0 = 158       # forpg[g][c] at 158 + 5*g + c (FORPG)
QB_MOMPG0 = 178       # mompg[g][c] at 178 + 3*g + c (MOMPG)
QB_NST = 190


class ShMat(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("rho0", "young", "nu", "shear", "a11", "a12", "ssp", "ca", "cb", "cn",
                 "epmx", "ymax", "cc", "epdr", "mexp", "asrate_pm9")] + \
               [(n, C.c_int) for n in ("icc", "vp", "israte", "ipla")]


class ShProp(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin")] + \
               [(n, C.c_int) for n in ("ish", "npt", "ishell", "pad")] + \
               [(n, C.c_double) for n in ("dn", "cvis")]


class ShOut(C.Structure):
    _fields_ = [("f", C.c_double * 12), ("m", C.c_double * 12), ("sti", C.c_double),
                ("stir", C.c_double), ("dt", C.c_double), ("ehou", C.c_double),
                ("alive", C.c_int), ("pad", C.c_int), ("stic", C.c_double * 4)]


class OShell(C.Structure):
    _fields_ = [("numnod", C.c_int), ("numelc", C.c_int)] + \
               [(n, C.c_void_p) for n in ("x", "v", "vr", "d", "ms", "inr", "icodt", "icodr",
                                          "fv_dir", "fv_val", "ixc", "thke", "adsky", "iads",
                                          "st", "qst", "fsky", "a", "ar", "stifn", "stifr")] + \
               [("mat", ShMat), ("prop", ShProp)] + \
               [(n, C.c_double) for n in ("tt", "dt1", "dt2", "dt12", "dt2old", "dtmx",
                                          "ehour")] + \
               [("ncycle", C.c_int), ("nthreads", C.c_int),
                ("fv_mem", C.c_void_p), ("wfext", C.c_double),
                ("encin", C.c_double), ("enrot", C.c_double), ("contact", C.c_void_p)]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _lib.oracle_brick_element.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                             C.POINTER(Law2), C.POINTER(SProp), C.POINTER(BOut)]
        _lib.oracle_brick_volume.argtypes = [C.c_void_p]
        _lib.oracle_brick_volume.restype = C.c_double
        _lib.oracle_build_pon.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]
        _lib.oracle_cycle.argtypes = [C.POINTER(OModel)]
        _lib.oracle_energies.argtypes = [C.POINTER(OModel), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]
        _lib.oracle_shell_element.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                              C.c_int, C.c_void_p, C.c_double,
                                              C.POINTER(ShMat), C.POINTER(ShProp),
                                              C.POINTER(ShOut)]
        _lib.oracle_qbat_element.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                             C.c_void_p, C.c_double, C.POINTER(ShMat),
                                             C.POINTER(ShProp), C.POINTER(ShOut)]
        _lib.oracle_qeph_element.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                             C.c_void_p, C.c_double, C.POINTER(ShMat),
                                             C.POINTER(ShProp), C.POINTER(ShOut)]
        _lib.oracle_shell_tables.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.oracle_shell_constants.argtypes = [C.c_void_p]
        _lib.oracle_build_pon_nc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.oracle_shell_cycle.argtypes = [C.POINTER(OShell)]
        _lib.oracle_shell_energies.argtypes = [C.POINTER(OShell), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double)]
        _lib.oracle_nodal_update.argtypes = [C.POINTER(ONodes), C.c_double]
        _lib.oracle_int7_candidates.argtypes = [C.POINTER(Int7), C.c_void_p, C.c_double,
                                                C.c_int, C.c_void_p, C.c_void_p]
        _lib.oracle_int7_forces.argtypes = [C.POINTER(Int7), C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_double, C.POINTER(C.c_double), C.c_void_p,
                                            C.c_void_p]
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        _ref.ref_brick_group.argtypes = [C.c_int] + [C.c_void_p] * 11 + \
            [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
             C.c_void_p, C.c_void_p, C.c_void_p]
        _ref.ref_constants.argtypes = [C.c_void_p]
        _ref.ref_shell_group.argtypes = [C.c_int] + [C.c_void_p] * 15 + \
            [C.c_double] + [C.c_void_p] * 6
        _ref.ref_qeph_group.argtypes = [C.c_int] + [C.c_void_p] * 14 + \
            [C.c_double] + [C.c_void_p] * 7
        _ref.ref_qbat_group.argtypes = [C.c_int] + [C.c_void_p] * 15 + \
            [C.c_double] + [C.c_void_p] * 7
        _ref.ref_shell_constants.argtypes = [C.c_void_p]
        _ref.ref_shell_tables.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _ref.ref_nodal_cycle.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 10 + \
            [C.c_int] + [C.c_void_p] * 4 + [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        _ref.ref_int7_batch.argtypes = [C.c_int] + [C.c_void_p] * 3 + [C.c_int, C.c_void_p,
                                                                     C.c_int] + \
            [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 8
    return _ref


def law2(p):
    """Law2 struct from a parameter dict (see openradioss_amd.deck.law2_params)."""
    return Law2(p["rho0"], p["shear"], p["bulk"], p["ca"], p["cb"], p["cn"], p["epmx"],
                p["sigmx"], p["cc"], p["epdr"], 0.0, p["mexp"], p["icc"], p["vp"],
                p["israte"], 0)


def sprop(p):
    return SProp(p["qa"], p["qb"], p["hvis"], p["dtfac"])


def brick_element(xg, vg, state, dt1, mat, prop):
    """One element through the oracle chain. xg, vg: (3,8); state: (14,) updated in place."""
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    vg = np.ascontiguousarray(vg, dtype=np.float64)
    assert xg.shape == (3, 8) and vg.shape == (3, 8) and state.shape == (NST,)
    assert state.dtype == np.float64 and state.flags.c_contiguous
    out = BOut()
    lib().oracle_brick_element(_ptr(xg), _ptr(vg), _ptr(state), dt1,
                               C.byref(law2(mat)), C.byref(sprop(prop)), C.byref(out))
    f = np.array(out.f, dtype=np.float64).reshape(3, 8)
    return dict(f=f, sti=out.sti, dtx=out.dtx, ehou=out.ehou, voln=out.voln, mstop=out.mstop)


def brick_volume(xg):
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    return lib().oracle_brick_volume(_ptr(xg))


def build_pon(numnod, ixs, uid):
    """PARITH/ON CSR: ixs (numels,8) 0-based, uid (numels,). Returns adsky, iads."""
    ixs = np.ascontiguousarray(ixs, dtype=np.int32)
    uid = np.ascontiguousarray(uid, dtype=np.int32)
    ne = ixs.shape[0]
    adsky = np.zeros(numnod + 1, np.int32)
    iads = np.zeros((ne, 8), np.int32)
    rc = lib().oracle_build_pon(numnod, ne, _ptr(ixs), _ptr(uid), _ptr(adsky), _ptr(iads))
    if rc != 0:
        raise ValueError(f"oracle_build_pon failed ({rc})")
    return adsky, iads


def ref_brick_group(xg, vg, state, dt1, mat, prop):
    """Reference kernels on a group: xg, vg (nel,3,8); state (nel,14) updated in place.
    Returns dict with fsky (nel,8,8) [slot comps 1..8], dt2t, ehou, voln, mstop."""
    nel = xg.shape[0]
    assert nel <= 128
    xg = np.ascontiguousarray(xg, dtype=np.float64)
    vg = np.ascontiguousarray(vg, dtype=np.float64)
    cols = [np.ascontiguousarray(state[:, i]) for i in range(NST)]
    sig = np.ascontiguousarray(state[:, ST_SIG].T)       # Fortran SIG(NEL,6)
    mpar = np.zeros(16)
    mpar[:15] = [mat["rho0"], mat["shear"], mat["bulk"], mat["ca"], mat["cb"], mat["cn"],
                 mat["epmx"], mat["sigmx"], mat["cc"], mat["epdr"], mat["mexp"],
                 prop["qa"], prop["qb"], prop["dtfac"], prop["hvis"]]
    ipar = np.array([mat["icc"], mat["vp"], mat["israte"], 0], np.int32)
    fsky = np.zeros((nel * 8, 8))
    dt2t = C.c_double(0.0)
    ehou = np.zeros(nel)
    voln = np.zeros(nel)
    mstop = C.c_int(0)
    ref().ref_brick_group(nel, _ptr(xg), _ptr(vg), _ptr(sig), _ptr(cols[ST_EINT]),
                          _ptr(cols[ST_RHO]), _ptr(cols[ST_VOL0]), _ptr(cols[ST_PLA]),
                          _ptr(cols[ST_QVIS]), _ptr(cols[ST_OFFG]), _ptr(cols[ST_EPSD]),
                          _ptr(cols[ST_DMG]), dt1, _ptr(mpar), _ptr(ipar), _ptr(fsky),
                          C.byref(dt2t), _ptr(ehou), _ptr(voln), C.byref(mstop))
    state[:, ST_SIG] = sig.T
    for i in range(6, NST):
        state[:, i] = cols[i]
    return dict(fsky=fsky.reshape(nel, 8, 8), dt2t=dt2t.value, ehou=ehou, voln=voln,
                mstop=mstop.value)


def ref_constants():
    c = np.zeros(16)
    ref().ref_constants(_ptr(c))
    return c


class Model:
    """Whole-model oracle (FP64, RESOL order) over numpy arrays it keeps alive."""

    def __init__(self, deck, nthreads=0):
        self.deck = deck
        self.x = np.ascontiguousarray(deck["x"], dtype=np.float64).copy()
        self.v = np.ascontiguousarray(deck["v"], dtype=np.float64).copy()
        self.d = np.zeros_like(self.x)
        self.ms = np.ascontiguousarray(deck["ms"], dtype=np.float64)
        self.icodt = np.ascontiguousarray(deck["icodt"], dtype=np.int32)
        self.ixs = np.ascontiguousarray(deck["ixs"], dtype=np.int32)
        self.adsky, self.iads = build_pon(self.x.shape[0], self.ixs, deck["uid"])
        ne = self.ixs.shape[0]
        self.st = np.zeros((ne, NST))
        self.st[:, ST_RHO] = deck["mat"]["rho0"]
        self.st[:, ST_VOL0] = deck["vol0"]
        self.st[:, ST_OFFG] = 1.0
        self.fsky = np.zeros((8 * ne, 4))
        self.a = np.zeros_like(self.x)
        self.stifn = np.zeros(self.x.shape[0])
        m = OModel()
        m.numnod, m.numels = self.x.shape[0], ne
        for name in ("x", "v", "d", "ms", "icodt", "ixs", "adsky", "iads", "st", "fsky",
                     "a", "stifn"):
            setattr(m, name, _ptr(getattr(self, name)))
        m.mat = law2(deck["mat"])
        m.prop = sprop(deck["prop"])
        if deck.get("fv_dir") is not None:
            self.fv_dir = np.ascontiguousarray(deck["fv_dir"], dtype=np.int32)
            self.fv_val = np.ascontiguousarray(deck["fv_val"], dtype=np.float64)
            m.fv_dir, m.fv_val = _ptr(self.fv_dir), _ptr(self.fv_val)
        self.fv_mem = np.zeros(self.x.shape[0])
        m.fv_mem = _ptr(self.fv_mem)
        m.tt, m.dt1, m.dt2, m.dt12 = 0.0, 0.0, 0.0, 0.0
        m.dt2old = 1.0e6 / 1.1          # starter contrl.F:1108-1110: DTINI/ONEP1
        m.dtmx = 1.0e20
        m.nthreads = nthreads
        self.m = m
        self.contact = None
        if deck.get("contact") is not None:
            self.contact = Contact7(deck["contact"])
            m.contact = C.cast(C.pointer(self.contact.it), C.c_void_p)

    def cycle(self, n=1):
        for _ in range(n):
            first = self.m.ncycle == 0
            lib().oracle_cycle(C.byref(self.m))
            if first:
                self.entot0 = self.balance(first=True)

    def balance(self, first=False):
        """ECRIT's energy balance after the last cycle (output/ecrit.F:500-508):
        ENTOT = ENCIN + ENINT + ENROT, ERR = ENTOT / (ENTOT0 + WFEXT) - 1 with ENTOT0 =
        ENTOT - WFEXT at the first cycle."""
        return _balance(self, lib().oracle_energies, first)

    def energies(self):
        """(internal, kinetic, cumulative hourglass) energy."""
        ei, ek = C.c_double(), C.c_double()
        lib().oracle_energies(C.byref(self.m), C.byref(ei), C.byref(ek))
        return ei.value, ek.value, self.m.ehour

    @property
    def dt2(self):
        return self.m.dt2

    @property
    def tt(self):
        return self.m.tt


def _balance(model, energies, first):
    ei, ek = C.c_double(), C.c_double()
    energies(C.byref(model.m), C.byref(ei), C.byref(ek))
    m = model.m
    entot = m.encin + ei.value + m.enrot
    if first:
        return entot - m.wfext
    den = model.entot0 + m.wfext
    err = entot / den - 1.0 if abs(den) > 1e-20 else 0.0
    return dict(encin=m.encin, enrot=m.enrot, enint=ei.value, ehour=m.ehour, wfext=m.wfext,
                entot=entot, entot0=model.entot0, err=err, tt=m.tt, ncycle=m.ncycle)


# ============================== shells ==============================

def shmat(p):
    """ShMat from a shell material dict (openradioss_amd.deck.shell_mat_params)."""
    return ShMat(*[p[k] for k in ("rho0", "young", "nu", "shear", "a11", "a12", "ssp", "ca",
                                  "cb", "cn", "epmx", "ymax", "cc", "epdr", "mexp",
                                  "asrate_pm9")],
                 p["icc"], p["vp"], p["israte"], p["ipla"])


def shprop(p):
    return ShProp(*[p[k] for k in ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin")],
                  p["ish"], p["npt"], p.get("ishell", 1), 0, p.get("dn", 0.015),
                  p.get("cvis", 1.0))


def shell_state(ne, thk, offg=1.0):
    st = np.zeros((ne, SH_NST))
    st[:, SH_THK] = thk
    st[:, SH_OFFG] = offg
    return st


def shell_element(xg, vg, vrg, thke, state, dt1, mat, prop, tri=False):
    """One shell through the oracle chain. xg, vg, vrg: (3,4); state (53,) in place."""
    xg, vg, vrg = (np.ascontiguousarray(a, dtype=np.float64) for a in (xg, vg, vrg))
    assert xg.shape == (3, 4) and state.shape == (SH_NST,) and state.flags.c_contiguous
    out = ShOut()
    lib().oracle_shell_element(_ptr(xg), _ptr(vg), _ptr(vrg), float(thke), int(tri),
                               _ptr(state), dt1, C.byref(shmat(mat)), C.byref(shprop(prop)),
                               C.byref(out))
    return dict(f=np.array(out.f).reshape(3, 4), m=np.array(out.m).reshape(3, 4),
                sti=out.sti, stir=out.stir, dt=out.dt, ehou=out.ehou, alive=out.alive)


def qeph_element(xg, vg, vrg, thke, state, dt1, mat, prop):
    """One QEPH quadrilateral (Ishell=24) through the oracle chain; state (60,) in
    place.  stic: FSKY(7) of each corner."""
    xg, vg, vrg = (np.ascontiguousarray(a, dtype=np.float64) for a in (xg, vg, vrg))
    assert xg.shape == (3, 4) and state.shape == (SH_NST,) and state.flags.c_contiguous
    out = ShOut()
    lib().oracle_qeph_element(_ptr(xg), _ptr(vg), _ptr(vrg), float(thke), _ptr(state), dt1,
                              C.byref(shmat(mat)), C.byref(shprop(prop)), C.byref(out))
    return dict(f=np.array(out.f).reshape(3, 4), m=np.array(out.m).reshape(3, 4),
                sti=out.sti, stir=out.stir, stic=np.array(out.stic), dt=out.dt,
                ehou=out.ehou, alive=out.alive)


def qbat_element(xg, vg, vrg, thke, state, dt1, mat, prop):
    """One QBAT quadrilateral (Ishell=12) through the oracle chain; state (190,) in
    place.  stic: FSKY(7) of each corner; ehou: CBAVISC's damping energy."""
    xg, vg, vrg = (np.ascontiguousarray(a, dtype=np.float64) for a in (xg, vg, vrg))
    assert xg.shape == (3, 4) and state.shape == (QB_NST,) and state.flags.c_contiguous
    out = ShOut()
    lib().oracle_qbat_element(_ptr(xg), _ptr(vg), _ptr(vrg), float(thke), _ptr(state), dt1,
                              C.byref(shmat(mat)), C.byref(shprop(prop)), C.byref(out))
    return dict(f=np.array(out.f).reshape(3, 4), m=np.array(out.m).reshape(3, 4),
                sti=out.sti, stir=out.stir, stic=np.array(out.stic), dt=out.dt,
                ehou=out.ehou, alive=out.alive)


def shell_tables(npt):
    z, w, m = np.zeros(npt), np.zeros(npt), np.zeros(npt)
    lib().oracle_shell_tables(npt, _ptr(z), _ptr(w), _ptr(m))
    return z, w, m


def shell_constants():
    c = np.zeros(16)
    lib().oracle_shell_constants(_ptr(c))
    return c


def ref_shell_constants():
    c = np.zeros(16)
    ref().ref_shell_constants(_ptr(c))
    return c


def ref_shell_tables(npt):
    z, w, m = np.zeros(npt), np.zeros(npt), np.zeros(npt)
    ref().ref_shell_tables(npt, _ptr(z), _ptr(w), _ptr(m))
    return z, w, m


def build_pon_nc(numnod, conn, uid):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    uid = np.ascontiguousarray(uid, dtype=np.int32)
    ne, nc = conn.shape
    adsky = np.zeros(numnod + 1, np.int32)
    iads = np.zeros((ne, nc), np.int32)
    rc = lib().oracle_build_pon_nc(numnod, ne, nc, _ptr(conn), _ptr(uid), _ptr(adsky),
                                   _ptr(iads))
    if rc != 0:
        raise ValueError(f"oracle_build_pon_nc failed ({rc})")
    return adsky, iads


def ref_shell_group(xg, vg, vrg, thke, state, dt1, mat, prop, tri=None):
    """Reference routines on a group: xg, vg, vrg (nel,3,4); state (nel,53) updated in
    place.  Returns fsky (nel,4,8), per-element dt, ehou."""
    nel = xg.shape[0]
    assert nel <= 128
    npt = prop["npt"]
    fx = lambda a: np.ascontiguousarray(np.transpose(a, (0, 1, 2)), dtype=np.float64)
    xg, vg, vrg = fx(xg), fx(vg), fx(vrg)          # Fortran XG(4,3,NEL) == C [nel][3][4]
    thke = np.ascontiguousarray(thke, dtype=np.float64)
    itri = np.zeros(nel, np.int32) if tri is None else np.ascontiguousarray(tri, np.int32)
    forc = np.asfortranarray(state[:, SH_FOR])
    mom = np.asfortranarray(state[:, SH_MOM])
    thk = np.ascontiguousarray(state[:, SH_THK])
    eint = np.asfortranarray(state[:, SH_EINT])
    epsd = np.ascontiguousarray(state[:, SH_EPSD])
    hour = np.asfortranarray(state[:, SH_HOUR])
    offg = np.ascontiguousarray(state[:, SH_OFFG])
    sig = np.zeros((nel, 5, npt), order="F")
    pla = np.zeros((nel, npt), order="F")
    epsdp = np.zeros((nel, npt), order="F")
    for p_ in range(npt):
        sig[:, :, p_] = state[:, SH_SIG0 + 5 * p_: SH_SIG0 + 5 * p_ + 5]
        pla[:, p_] = state[:, SH_PLA0 + p_]
        epsdp[:, p_] = state[:, SH_EPSDP0 + p_]
    mpar = np.array([mat[k] for k in ("rho0", "young", "nu", "shear", "a11", "a12", "ssp",
                                      "ca", "cb", "cn", "epmx", "ymax", "cc", "epdr", "mexp",
                                      "asrate_pm9")], np.float64)
    ipar = np.array([mat["icc"], mat["vp"], mat["israte"], mat["ipla"], npt, prop["ish"]],
                    np.int32)
    ppar = np.array([prop[k] for k in ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin")],
                    np.float64)
    fsky = np.zeros((4 * nel, 8))
    dte = np.zeros(nel)
    ehou = np.zeros(nel)
    ref().ref_shell_group(nel, _ptr(xg), _ptr(vg), _ptr(vrg), _ptr(thke), _ptr(itri),
                          _ptr(forc), _ptr(mom), _ptr(thk), _ptr(eint), _ptr(epsd), _ptr(hour),
                          _ptr(offg), _ptr(sig), _ptr(pla), _ptr(epsdp), dt1, _ptr(mpar),
                          _ptr(ipar), _ptr(ppar), _ptr(fsky), _ptr(dte), _ptr(ehou))
    state[:, SH_FOR] = forc
    state[:, SH_MOM] = mom
    state[:, SH_THK] = thk
    state[:, SH_EINT] = eint
    state[:, SH_EPSD] = epsd
    state[:, SH_HOUR] = hour
    state[:, SH_OFFG] = offg
    for p_ in range(npt):
        state[:, SH_SIG0 + 5 * p_: SH_SIG0 + 5 * p_ + 5] = sig[:, :, p_]
        state[:, SH_PLA0 + p_] = pla[:, p_]
        state[:, SH_EPSDP0 + p_] = epsdp[:, p_]
    return dict(fsky=fsky.reshape(nel, 4, 8), dt=dte, ehou=ehou)


def ref_qeph_group(xg, vg, vrg, thke, state, dt1, mat, prop):
    """Reference QEPH routines on a group: xg, vg, vrg (nel,3,4); state (nel,60)
    updated in place.  Returns fsky (nel,4,8), per-element dt, ehou and DT2T."""
    nel = xg.shape[0]
    assert nel <= 128 and state.shape == (nel, SH_NST)
    npt = prop["npt"]
    xg, vg, vrg = (np.ascontiguousarray(a, dtype=np.float64) for a in (xg, vg, vrg))
    thke = np.ascontiguousarray(thke, dtype=np.float64)
    forc = np.asfortranarray(state[:, SH_FOR])
    mom = np.asfortranarray(state[:, SH_MOM])
    thk = np.ascontiguousarray(state[:, SH_THK])
    eint = np.asfortranarray(state[:, SH_EINT])
    epsd = np.ascontiguousarray(state[:, SH_EPSD])
    hourg = np.asfortranarray(np.concatenate([state[:, SH_HOUR], state[:, SH_HOURX]], axis=1))
    offg = np.ascontiguousarray(state[:, SH_OFFG])
    sig = np.zeros((nel, 5, npt), order="F")
    pla = np.zeros((nel, npt), order="F")
    epsdp = np.zeros((nel, npt), order="F")
    for p_ in range(npt):
        sig[:, :, p_] = state[:, SH_SIG0 + 5 * p_: SH_SIG0 + 5 * p_ + 5]
        pla[:, p_] = state[:, SH_PLA0 + p_]
        epsdp[:, p_] = state[:, SH_EPSDP0 + p_]
    mpar = np.array([mat[k] for k in ("rho0", "young", "nu", "shear", "a11", "a12", "ssp",
                                      "ca", "cb", "cn", "epmx", "ymax", "cc", "epdr", "mexp",
                                      "asrate_pm9")], np.float64)
    ipar = np.array([mat["icc"], mat["vp"], mat["israte"], mat["ipla"], npt, prop["ish"]],
                    np.int32)
    ppar = np.array([prop[k] for k in ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin",
                                       "dn", "cvis")], np.float64)
    fsky = np.zeros((4 * nel, 8))
    dte = np.zeros(nel)
    ehou = np.zeros(nel)
    dt2t = np.zeros(1)
    ref().ref_qeph_group(nel, _ptr(xg), _ptr(vg), _ptr(vrg), _ptr(thke), _ptr(forc),
                         _ptr(mom), _ptr(thk), _ptr(eint), _ptr(epsd), _ptr(hourg), _ptr(offg),
                         _ptr(sig), _ptr(pla), _ptr(epsdp), dt1, _ptr(mpar), _ptr(ipar),
                         _ptr(ppar), _ptr(fsky), _ptr(dte), _ptr(ehou), _ptr(dt2t))
    state[:, SH_FOR] = forc
    state[:, SH_MOM] = mom
    state[:, SH_THK] = thk
    state[:, SH_EINT] = eint
    state[:, SH_EPSD] = epsd
    state[:, SH_HOUR] = hourg[:, 0:5]
    state[:, SH_HOURX] = hourg[:, 5:12]
    state[:, SH_OFFG] = offg
    for p_ in range(npt):
        state[:, SH_SIG0 + 5 * p_: SH_SIG0 + 5 * p_ + 5] = sig[:, :, p_]
        state[:, SH_PLA0 + p_] = pla[:, p_]
        state[:, SH_EPSDP0 + p_] = epsdp[:, p_]
    return dict(fsky=fsky.reshape(nel, 4, 8), dt=dte, ehou=ehou, dt2t=float(dt2t[0]))


def qbat_state(ne, thk, offg=1.0):
    st = np.zeros((ne, QB_NST))
    st[:, SH_THK] = thk
    st[:, SH_OFFG] = offg
    return st


def ref_qbat_group(xg, vg, vrg, thke, state, dt1, mat, prop):
    """Reference QBAT routines (CBAFORC3 order, oracle/ref/harness_qbat.F) on a group:
    xg, vg, vrg (nel,3,4); state (nel,190) updated in place.  Returns fsky (nel,4,8),
    per-element dt, the viscous-damping energy increment (PARTSAV(8)) and DT2T."""
    nel = xg.shape[0]
    assert nel <= 128 and state.shape == (nel, QB_NST)
    npt = prop["npt"]
    xg, vg, vrg = (np.ascontiguousarray(a, dtype=np.float64) for a in (xg, vg, vrg))
    thke = np.ascontiguousarray(thke, dtype=np.float64)
    forc = np.asfortranarray(state[:, SH_FOR])
    mom = np.asfortranarray(state[:, SH_MOM])
    forpg = np.zeros((nel, 5, 4), order="F")
    mompg = np.zeros((nel, 3, 4), order="F")
    thk = np.ascontiguousarray(state[:, SH_THK])
    eint = np.asfortranarray(state[:, SH_EINT])
    epsd = np.ascontiguousarray(state[:, SH_EPSD])
    offg = np.ascontiguousarray(state[:, SH_OFFG])
    sig = np.zeros((nel, 5, 5, 4), order="F")
    pla = np.zeros((nel, 5, 4), order="F")
    epsdp = np.zeros((nel, 5, 4), order="F")
    for g in range(4):
        forpg[:, :, g] = state[:, QB_FORPG0 + 5 * g: QB_FORPG0 + 5 * g + 5]
        mompg[:, :, g] = state[:, QB_MOMPG0 + 3 * g: QB_MOMPG0 + 3 * g + 3]
        for p_ in range(5):
            c0 = QB_SIG0 + 5 * (5 * g + p_)
            sig[:, :, p_, g] = state[:, c0:c0 + 5]
            pla[:, p_, g] = state[:, QB_PLA0 + 5 * g + p_]
            epsdp[:, p_, g] = state[:, QB_EPSDP0 + 5 * g + p_]
    mpar = np.array([mat[k] for k in ("rho0", "young", "nu", "shear", "a11", "a12", "ssp",
                                      "ca", "cb", "cn", "epmx", "ymax", "cc", "epdr", "mexp",
                                      "asrate_pm9")], np.float64)
    ipar = np.array([mat["icc"], mat["vp"], mat["israte"], mat["ipla"], npt, prop["ish"]],
                    np.int32)
    ppar = np.array([prop[k] for k in ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin",
                                       "dn", "cvis")], np.float64)
    fsky = np.zeros((4 * nel, 8))
    dte = np.zeros(nel)
    evis = np.zeros(nel)
    dt2t = np.zeros(1)
    ref().ref_qbat_group(nel, _ptr(xg), _ptr(vg), _ptr(vrg), _ptr(thke), _ptr(forc),
                         _ptr(mom), _ptr(forpg), _ptr(mompg), _ptr(thk), _ptr(eint),
                         _ptr(epsd), _ptr(offg), _ptr(sig), _ptr(pla), _ptr(epsdp), dt1,
                         _ptr(mpar), _ptr(ipar), _ptr(ppar), _ptr(fsky), _ptr(dte),
                         _ptr(evis), _ptr(dt2t))
    state[:, SH_FOR] = forc
    state[:, SH_MOM] = mom
    state[:, SH_THK] = thk
    state[:, SH_EINT] = eint
    state[:, SH_EPSD] = epsd
    state[:, SH_OFFG] = offg
    for g in range(4):
        state[:, QB_FORPG0 + 5 * g: QB_FORPG0 + 5 * g + 5] = forpg[:, :, g]
        state[:, QB_MOMPG0 + 3 * g: QB_MOMPG0 + 3 * g + 3] = mompg[:, :, g]
        for p_ in range(npt):
            c0 = QB_SIG0 + 5 * (5 * g + p_)
            state[:, c0:c0 + 5] = sig[:, :, p_, g]
            state[:, QB_PLA0 + 5 * g + p_] = pla[:, p_, g]
            state[:, QB_EPSDP0 + 5 * g + p_] = epsdp[:, p_, g]
    return dict(fsky=fsky.reshape(nel, 4, 8), dt=dte, evis=evis, dt2t=float(dt2t[0]))


class ShellModel:
    """Whole-model shell oracle (FP64, RESOL order, 6 DOF)."""

    def __init__(self, deck, nthreads=0):
        self.deck = deck
        nn = deck["x"].shape[0]
        self.x = np.ascontiguousarray(deck["x"], dtype=np.float64).copy()
        self.v = np.ascontiguousarray(deck["v"], dtype=np.float64).copy()
        self.vr = np.ascontiguousarray(deck.get("vr", np.zeros((nn, 3))), np.float64).copy()
        self.d = np.zeros_like(self.x)
        self.ms = np.ascontiguousarray(deck["ms"], dtype=np.float64)
        self.inr = np.ascontiguousarray(deck["in"], dtype=np.float64)
        self.icodt = np.ascontiguousarray(deck["icodt"], dtype=np.int32)
        self.icodr = np.ascontiguousarray(deck["icodr"], dtype=np.int32)
        self.fv_dir = np.ascontiguousarray(deck.get("fv_dir", np.zeros(nn, np.int32)), np.int32)
        self.fv_val = np.ascontiguousarray(deck.get("fv_val", np.zeros(nn)), np.float64)
        self.ixc = np.ascontiguousarray(deck["ixc"], dtype=np.int32)
        self.thke = np.ascontiguousarray(deck["thk"], dtype=np.float64)
        self.adsky, self.iads = build_pon_nc(nn, self.ixc, deck["uid"])
        ne = self.ixc.shape[0]
        qbat = int(deck["prop"].get("ishell", 1)) == 12
        self.st = qbat_state(ne, self.thke) if qbat else shell_state(ne, self.thke)
        self.qst = self.st
        self.fsky = np.zeros((4 * ne, 8))
        self.a = np.zeros_like(self.x)
        self.ar = np.zeros_like(self.x)
        self.stifn = np.zeros(nn)
        self.stifr = np.zeros(nn)
        m = OShell()
        m.numnod, m.numelc = nn, ne
        for name in ("x", "v", "vr", "d", "ms", "inr", "icodt", "icodr", "fv_dir", "fv_val",
                     "ixc", "thke", "adsky", "iads", "st", "qst", "fsky", "a", "ar",
                     "stifn", "stifr"):
            setattr(m, name, _ptr(getattr(self, name)))
        m.mat = shmat(deck["mat"])
        m.prop = shprop(deck["prop"])
        self.fv_mem = np.zeros(nn)
        m.fv_mem = _ptr(self.fv_mem)
        m.dt2old = 1.0e6 / 1.1
        m.dtmx = 1.0e20
        m.nthreads = nthreads
        self.m = m
        self.contact = None
        if deck.get("contact") is not None:
            self.contact = Contact7(deck["contact"])
            m.contact = C.cast(C.pointer(self.contact.it), C.c_void_p)

    def cycle(self, n=1):
        for _ in range(n):
            first = self.m.ncycle == 0
            lib().oracle_shell_cycle(C.byref(self.m))
            if first:
                self.entot0 = self.balance(first=True)

    def balance(self, first=False):
        """ECRIT's energy balance after the last cycle (see Model.balance)."""
        return _balance(self, lib().oracle_shell_energies, first)

    def energies(self):
        ei, ek = C.c_double(), C.c_double()
        lib().oracle_shell_energies(C.byref(self.m), C.byref(ei), C.byref(ek))
        return ei.value, ek.value, self.m.ehour

    @property
    def dt2(self):
        return self.m.dt2

    @property
    def tt(self):
        return self.m.tt


# ============================ nodal update ============================

TIME_KEYS = ("tt", "dt1", "dt2", "dt12", "dt2old", "dtmx")


def _nodal_arrays(nodes):
    """Contiguous copies of a node set dict: x, v, d, a (forces), ms, icodt and, with
    rotations, vr, ar (moments), in, icodr; fv_dir, fv_val, fv_mem (per node)."""
    nn = nodes["x"].shape[0]
    f = lambda k, dt=np.float64, shape=None: np.ascontiguousarray(
        nodes[k] if k in nodes else np.zeros(shape if shape else (nn, 3), dt), dt).copy()
    out = dict(x=f("x"), v=f("v"), d=f("d"), a=f("a"), ms=f("ms", shape=(nn,)),
               icodt=f("icodt", np.int32, (nn,)), vr=f("vr"), ar=f("ar"),
               inr=f("in", shape=(nn,)), icodr=f("icodr", np.int32, (nn,)),
               fv_dir=f("fv_dir", np.int32, (nn,)), fv_val=f("fv_val", shape=(nn,)),
               fv_mem=f("fv_mem", shape=(nn,)))
    return out


def nodal_update(nodes, time, dt2t, rot=False, nthreads=1):
    """Oracle nodal update (oracle_nodal_update) of one cycle.  nodes: dict of arrays
    (see _nodal_arrays; a/ar = assembled forces/moments); time: dict of TIME_KEYS plus
    wfext and ncycle.  Returns (new node arrays, new time dict); a/ar then hold the
    accelerations VELOCITY consumed."""
    o = _nodal_arrays(nodes)
    m = ONodes()
    m.numnod, m.rot = o["x"].shape[0], int(rot)
    for k in ("x", "v", "d", "a", "ms", "icodt", "vr", "ar", "inr", "icodr", "fv_dir",
              "fv_val", "fv_mem"):
        setattr(m, k, _ptr(o[k]))
    for k in TIME_KEYS:
        setattr(m, k, float(time[k]))
    m.wfext = float(time.get("wfext", 0.0))
    m.ncycle = int(time.get("ncycle", 0))
    m.nthreads = nthreads
    lib().oracle_nodal_update(C.byref(m), float(dt2t))
    t = {k: getattr(m, k) for k in TIME_KEYS}
    t["wfext"], t["ncycle"] = m.wfext, m.ncycle
    t["encin"], t["enrot"] = m.encin, m.enrot
    return o, t


def ref_nodal_cycle(nodes, time, dt2t, rot=False):
    """The reference's own ACCELE/BCS10/FIXVEL/VELOCITY/DEPLA (oracle/_ref, harness_nodal.F)
    on the same inputs as nodal_update.  FIXVEL entries are the imposed nodes in increasing
    node order; their energy memory travels in fv_mem."""
    o = _nodal_arrays(nodes)
    nn = o["x"].shape[0]
    fxn = np.nonzero(o["fv_dir"] > 0)[0]
    fxnode = np.ascontiguousarray(fxn + 1, np.int32)
    fxdir = np.ascontiguousarray(o["fv_dir"][fxn], np.int32)
    fxval = np.ascontiguousarray(o["fv_val"][fxn], np.float64)
    fxmem = np.zeros((len(fxn), 2))
    fxmem[:, 0] = o["fv_mem"][fxn]
    times = np.array([float(time[k]) for k in TIME_KEYS])
    ncyc = C.c_int(int(time.get("ncycle", 0)))
    wf = C.c_double(float(time.get("wfext", 0.0)))
    if len(fxn) == 0:
        fxnode, fxdir, fxval, fxmem = (np.zeros(1, np.int32), np.zeros(1, np.int32),
                                       np.zeros(1), np.zeros((1, 2)))
    ref().ref_nodal_cycle(nn, int(rot), _ptr(o["x"]), _ptr(o["v"]), _ptr(o["vr"]),
                          _ptr(o["d"]), _ptr(o["a"]), _ptr(o["ar"]), _ptr(o["ms"]),
                          _ptr(o["inr"]), _ptr(o["icodt"]), _ptr(o["icodr"]), len(fxn),
                          _ptr(fxnode), _ptr(fxdir), _ptr(fxval), _ptr(fxmem), float(dt2t),
                          _ptr(times), C.byref(ncyc), C.byref(wf))
    if len(fxn):
        o["fv_mem"][fxn] = fxmem[:, 0]
    t = dict(zip(TIME_KEYS, times.tolist()))
    t["wfext"], t["ncycle"] = wf.value, ncyc.value
    return o, t


def ref_int7(x, v, ms, nsv, irect, stfm, stfns, cand_n, cand_e, gap, fric, visc, viscf, dt1,
             dt12, dt2t=1.0e30):
    """The reference's own /INTER/TYPE7 routines (I7COR3, I7DST3, I7FOR3, I7ASS2 in
    I7MAINF order, oracle/_ref, harness_int7.F) over a candidate list.  nsv, irect,
    cand_n and cand_e are 0-based here (secondary node numbers, segment node numbers,
    index into nsv, index into irect).  Returns the skyline entries in FSKYI order
    (fsky[k] = FX, FY, FZ, STIF of node isky[k], 0-based), the lowered DT2T, the
    elastic, friction and damping contact energies of the call, the number of
    penetrating pairs and, per candidate, whether I7DST3 kept it."""
    x = np.ascontiguousarray(x, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    ms = np.ascontiguousarray(ms, np.float64)
    nsv1 = np.ascontiguousarray(np.asarray(nsv) + 1, np.int32)
    irect1 = np.ascontiguousarray(np.asarray(irect) + 1, np.int32)
    stfm = np.ascontiguousarray(stfm, np.float64)
    stfns = np.array(stfns, np.float64)          # a copy: deactivation negates entries
    cn = np.ascontiguousarray(np.asarray(cand_n) + 1, np.int32)
    ce = np.ascontiguousarray(np.asarray(cand_e) + 1, np.int32)
    nc = len(cn)
    params = np.array([gap, fric, visc, viscf], np.float64)
    dts = np.array([dt1, dt12], np.float64)
    fsky = np.zeros((max(5 * nc, 1), 4))
    isky = np.zeros(max(5 * nc, 1), np.int32)
    out = np.array([dt2t, 0.0, 0.0, 0.0, 0.0])
    pen = np.zeros(max(nc, 1), np.int32)
    nsky = C.c_int(0)
    ref().ref_int7_batch(x.shape[0], _ptr(x), _ptr(v), _ptr(ms), len(nsv1), _ptr(nsv1),
                         irect1.shape[0], _ptr(irect1), _ptr(stfm), _ptr(stfns), nc,
                         _ptr(cn if nc else np.zeros(1, np.int32)),
                         _ptr(ce if nc else np.zeros(1, np.int32)), _ptr(params), _ptr(dts),
                         C.byref(nsky), _ptr(fsky), _ptr(isky), _ptr(out), _ptr(pen))
    n = nsky.value
    return dict(fsky=fsky[:n].copy(), isky=isky[:n] - 1, dt2t=out[0], econt=out[1],
                econtv=out[2], econtd=out[3], npen=int(out[4]), pen=pen[:nc].astype(bool),
                stfns=stfns)


class Contact7:
    """One /INTER/TYPE7 interface for the oracle (orad_int7_t over arrays it keeps
    alive).  `c` is a dict: nsv (secondary nodes), irect (nrtm, 4) segment nodes,
    0-based; stfm (per segment), stfns (per secondary node); gap, fric, visc, viscf."""

    def __init__(self, c):
        self.nsv = np.ascontiguousarray(c["nsv"], np.int32)
        self.irect = np.ascontiguousarray(c["irect"], np.int32).reshape(-1, 4)
        self.stfm = np.ascontiguousarray(c["stfm"], np.float64)
        self.stfns = np.ascontiguousarray(c["stfns"], np.float64).copy()
        it = Int7()
        it.nsn, it.nrtm = len(self.nsv), self.irect.shape[0]
        it.nsv, it.irect = _ptr(self.nsv), _ptr(self.irect)
        it.stfm, it.stfns = _ptr(self.stfm), _ptr(self.stfns)
        it.gap, it.fric = float(c["gap"]), float(c.get("fric", 0.0))
        it.visc, it.viscf = float(c.get("visc", 0.05)), float(c.get("viscf", 0.0))
        self.it = it

    def candidates(self, x, tz=None):
        x = np.ascontiguousarray(x, np.float64)
        tz = self.it.gap * 1.000001 if tz is None else tz
        n = lib().oracle_int7_candidates(C.byref(self.it), _ptr(x), tz, 0, None, None)
        cn = np.zeros(max(n, 1), np.int32)
        ce = np.zeros(max(n, 1), np.int32)
        lib().oracle_int7_candidates(C.byref(self.it), _ptr(x), tz, n, _ptr(cn), _ptr(ce))
        return cn[:n], ce[:n]

    def forces(self, x, v, ms, cand_n, cand_e, dt1, dt2t=1.0e30):
        """Skyline entries of one call (see ref_int7): fsky (n, 4), isky, dt2t."""
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        ms = np.ascontiguousarray(ms, np.float64)
        cn = np.ascontiguousarray(cand_n, np.int32)
        ce = np.ascontiguousarray(cand_e, np.int32)
        nc = len(cn)
        fsky = np.zeros((max(5 * nc, 1), 4))
        isky = np.zeros(max(5 * nc, 1), np.int32)
        d = C.c_double(dt2t)
        e0 = (self.it.econt, self.it.econtv, self.it.econtd)
        n = lib().oracle_int7_forces(C.byref(self.it), _ptr(x), _ptr(v), _ptr(ms), nc,
                                     _ptr(cn if nc else np.zeros(1, np.int32)),
                                     _ptr(ce if nc else np.zeros(1, np.int32)), dt1,
                                     C.byref(d), _ptr(fsky), _ptr(isky))
        return dict(fsky=fsky[:n].copy(), isky=isky[:n].copy(), dt2t=d.value,
                    econt=self.it.econt - e0[0], econtv=self.it.econtv - e0[1],
                    econtd=self.it.econtd - e0[2], npen=self.it.npen,
                    stfns=self.stfns.copy())
