"""Host-side mirror of the reference's brick-path interface over the C ABI
(include/orad.h), through ctypes.  No CPU fallback: every compute call goes to
liborad.so, and constructing a solver fails loudly if it is missing.

Names follow the reference: sforc3 (per-group element forces, SFORC3),
asspar4 (PARITH/ON assembly, ASSPAR4), run_cycles (RESOL's explicit cycle).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborad.so")

ORAD_OK, ORAD_EINVAL, ORAD_EHIP, ORAD_ESTATE, ORAD_ENEGVOL = range(5)
FP64, MIXED = 0, 1
NIXS = 11

EXPORTS = (
    "orad_create", "orad_destroy", "orad_last_error", "orad_set_nodes", "orad_set_bricks",
    "orad_set_law2", "orad_set_solid_prop", "orad_get_pon", "orad_set_time", "orad_get_time",
    "orad_sforc3", "orad_sforc3_async", "orad_cforc3_async", "orad_group_wait",
    "orad_update_xv", "orad_asspar4", "orad_run_cycles", "orad_profile_cycles", "orad_get_nodes",
    "orad_get_brick_state", "orad_set_brick_state", "orad_get_fsky", "orad_energies",
    "orad_energies_noalloc", "orad_energy_balance",
    "orad_set_fixvel", "orad_set_nodes_rot", "orad_set_shells", "orad_set_shell_law2",
    "orad_set_shell_prop", "orad_set_shell_ishell", "orad_cforc3", "orad_asspar4_rot", "orad_get_nodes_rot",
    "orad_get_shell_state", "orad_set_shell_state", "orad_set_shell_models",
    "orad_get_model_time", "orad_model_energies", "orad_set_launch", "orad_get_launch",
    "orad_autotune", "orad_adjoint_forward", "orad_adjoint_reverse",
    "orad_write_restart", "orad_read_restart",
    "orad_set_contact7", "orad_contact7", "orad_get_contact7", "orad_get_contact7_sky",
)
NIXC = 7
SH_NCOL = 53          # Belytschko-Tsay state values per shell
SH_NCOL_QEPH = 60     # QEPH: HOURG(6:12) appended
SH_NCOL_QBAT = 190    # QBAT: per Gauss point layer stresses, strains and resultants
SH_PLA0 = 43          # state column of PLA of point 1 (PLA(1..5) at 43..47)


class TimeState(C.Structure):
    """orad_time_t: RESOL's time scalars."""
    _fields_ = [("tt", C.c_double), ("dt1", C.c_double), ("dt2", C.c_double),
                ("dt12", C.c_double), ("dt2old", C.c_double), ("dtmx", C.c_double),
                ("ncycle", C.c_int), ("mstop", C.c_int)]


class Balance(C.Structure):
    """orad_balance_t: ECRIT's energy balance of the last cycle."""
    _fields_ = [(k, C.c_double) for k in ("encin", "enrot", "enint", "ehour", "wfext", "entot",
                                          "entot0", "err", "tt")] + \
               [("ncycle", C.c_int), ("pad", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class Launch(C.Structure):
    """orad_launch_t: launch configuration (element-kernel instance, residency cap, XCD
    mapping, brick element order, nodal slot tile, element VJP of the shell adjoint)."""
    _fields_ = [(k, C.c_int) for k in ("persistent", "lds_kb", "xcd_remap", "grid", "block",
                                       "waves", "order_rows", "node_tile", "adj_vjp")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OradError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"orad error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path=LIB_PATH):
    """Load liborad.so (raises if absent) and declare the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    sig = {
        "orad_create": ([C.POINTER(P), I, I], I),
        "orad_destroy": ([P], I),
        "orad_last_error": ([P], C.c_char_p),
        "orad_set_nodes": ([P, I, P, P, P, P], I),
        "orad_set_bricks": ([P, I, P, P], I),
        "orad_set_law2": ([P, P, P], I),
        "orad_set_solid_prop": ([P, D, D, D, D], I),
        "orad_get_pon": ([P, P, P], I),
        "orad_set_time": ([P, C.POINTER(TimeState)], I),
        "orad_get_time": ([P, C.POINTER(TimeState)], I),
        "orad_sforc3": ([P, I, I, D, C.POINTER(D)], I),
        "orad_sforc3_async": ([P, I, I, D], I),
        "orad_cforc3_async": ([P, I, I, D], I),
        "orad_group_wait": ([P, C.POINTER(D)], I),
        "orad_update_xv": ([P, P, P, P], I),
        "orad_asspar4": ([P, P, P], I),
        "orad_run_cycles": ([P, I, C.POINTER(TimeState)], I),
        "orad_profile_cycles": ([P, I, C.POINTER(D), C.POINTER(D)], I),
        "orad_get_nodes": ([P, P, P, P, P, P], I),
        "orad_get_brick_state": ([P, P, P, P, P, P, P, P], I),
        "orad_set_brick_state": ([P, P, P, P, P, P, P], I),
        "orad_get_fsky": ([P, P], I),
        "orad_energies": ([P, C.POINTER(D), C.POINTER(D), C.POINTER(D)], I),
        "orad_energies_noalloc": ([P, C.POINTER(D), C.POINTER(D), C.POINTER(D)], I),
        "orad_energy_balance": ([P, C.POINTER(Balance)], I),
        "orad_set_fixvel": ([P, I, P, P, P], I),
        "orad_set_nodes_rot": ([P, P, P, P], I),
        "orad_set_shells": ([P, I, P, P], I),
        "orad_set_shell_law2": ([P, P, P], I),
        "orad_set_shell_prop": ([P, P, I, I], I),
        "orad_set_shell_ishell": ([P, I, D], I),
        "orad_cforc3": ([P, I, I, D, C.POINTER(D)], I),
        "orad_asspar4_rot": ([P, P, P, P, P], I),
        "orad_get_nodes_rot": ([P, P, P], I),
        "orad_get_shell_state": ([P, P], I),
        "orad_set_shell_state": ([P, P], I),
        "orad_set_shell_models": ([P, I, P, P, P, P], I),
        "orad_get_model_time": ([P, I, C.POINTER(TimeState)], I),
        "orad_model_energies": ([P, I, C.POINTER(D), C.POINTER(D), C.POINTER(D)], I),
        "orad_set_launch": ([P, C.POINTER(Launch)], I),
        "orad_get_launch": ([P, C.POINTER(Launch)], I),
        "orad_autotune": ([P, I, C.POINTER(Launch), C.POINTER(D)], I),
        "orad_write_restart": ([P, C.c_char_p], I),
        "orad_set_contact7": ([P, I, P, I, P, P, P, P, I], I),
        "orad_contact7": ([P, D, C.POINTER(D)], I),
        "orad_get_contact7": ([P, P, P, P], I),
        "orad_get_contact7_sky": ([P, I, P, P, C.POINTER(I)], I),
        "orad_read_restart": ([P, C.c_char_p], I),
        "orad_adjoint_forward": ([P, I, I], I),
        "orad_adjoint_reverse": ([P] + [P] * 12, I),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def ixs_from_deck(deck):
    """IXS(NIXS=11, NUMELS) in the engine's layout: material, 8 nodes (1-based),
    property, user id (scoor3.F:106-121)."""
    ne = deck["ixs"].shape[0]
    ixs = np.zeros((ne, NIXS), np.int32)
    ixs[:, 0] = 1
    ixs[:, 1:9] = deck["ixs"] + 1
    ixs[:, 9] = 1
    ixs[:, 10] = deck["uid"]
    return ixs


def _set_fixvel(solver, deck):
    fv = deck.get("fv_dir")
    if fv is None:
        return
    nodes = np.flatnonzero(np.asarray(fv) > 0)
    if nodes.size == 0:
        return
    ids = np.ascontiguousarray(nodes + 1, dtype=np.int32)
    dirs = np.ascontiguousarray(np.asarray(fv)[nodes], dtype=np.int32)
    vals = np.ascontiguousarray(np.asarray(deck["fv_val"])[nodes], dtype=np.float64)
    solver._chk(solver.lib.orad_set_fixvel(solver.h, nodes.size, _p(ids), _p(dirs), _p(vals)))


class BrickSolver:
    """One model (bricks, law 2) resident on one GPU."""

    def __init__(self, deck, precision=FP64, device=0):
        self.lib = load_library()
        self.precision = precision
        h = C.c_void_p()
        rc = self.lib.orad_create(C.byref(h), device, precision)
        if rc != ORAD_OK:
            raise OradError(rc, "orad_create failed (no HIP device?)")
        self.h = h
        self.numnod = deck["x"].shape[0]
        self.numels = deck["ixs"].shape[0]
        x = _f64(deck["x"], (self.numnod, 3))
        v = _f64(deck["v"], (self.numnod, 3))
        ms = _f64(deck["ms"], (self.numnod,))
        icodt = np.ascontiguousarray(deck["icodt"], dtype=np.int32)
        self._chk(self.lib.orad_set_nodes(h, self.numnod, _p(x), _p(v), _p(ms), _p(icodt)))
        mat, prop = deck["mat"], deck["prop"]
        p = np.array([mat["rho0"], mat["shear"], mat["bulk"], mat["ca"], mat["cb"], mat["cn"],
                      mat["epmx"], mat["sigmx"], mat["cc"], mat["epdr"], mat["mexp"]], np.float64)
        ip = np.array([mat["icc"], mat["vp"], mat["israte"]], np.int32)
        self._chk(self.lib.orad_set_law2(h, _p(p), _p(ip)))
        self._chk(self.lib.orad_set_solid_prop(h, prop["qa"], prop["qb"], prop["hvis"],
                                               prop["dtfac"]))
        ixs = ixs_from_deck(deck)
        vol0 = _f64(deck["vol0"], (self.numels,))
        self._chk(self.lib.orad_set_bricks(h, self.numels, _p(ixs), _p(vol0)))
        _set_fixvel(self, deck)
        self.nsn = 0
        if deck.get("contact") is not None:
            self.set_contact7(deck["contact"], deck["contact"].get("cand_cap", 0))

    def _chk(self, rc):
        if rc != ORAD_OK:
            raise OradError(rc, self.lib.orad_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.orad_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference-shaped operations ----
    def set_launch(self, persistent=0, lds_kb=-1, xcd_remap=1, block=0, waves=0, order_rows=-1,
                   node_tile=0, grid=0, adj_vjp=0):
        """orad_set_launch; order_rows=-1 keeps the current brick element order; adj_vjp=1
        makes the shell adjoint use forward-mode dual numbers for the element VJP."""
        l = Launch(persistent, lds_kb, xcd_remap, 0, block, waves, order_rows, node_tile, adj_vjp)
        self._chk(self.lib.orad_set_launch(self.h, C.byref(l)))

    def launch(self):
        l = Launch()
        self._chk(self.lib.orad_get_launch(self.h, C.byref(l)))
        return l.as_dict()

    def autotune(self, cycles=20):
        """Time each candidate launch configuration over `cycles` cycles from the
        current state (restored after each trial); keeps the fastest."""
        l, ms = Launch(), C.c_double()
        self._chk(self.lib.orad_autotune(self.h, cycles, C.byref(l), C.byref(ms)))
        return l.as_dict(), ms.value

    def write_restart(self, path):
        """orad_write_restart: nodes, element state, per-element stiffness, time and
        energy-balance sums of this context into the file `path`."""
        self._chk(self.lib.orad_write_restart(self.h, os.fsencode(path)))

    def read_restart(self, path):
        """orad_read_restart: resume from a file written by a context of the same model."""
        self._chk(self.lib.orad_read_restart(self.h, os.fsencode(path)))

    def set_contact7(self, c, cand_cap=0):
        """orad_set_contact7 from a dict: nsv (secondary nodes), irect (nrtm, 4) 0-based;
        stfm, stfns; gap, fric, visc (VISs, default 0.05), viscf (VISf)."""
        nsv = np.ascontiguousarray(np.asarray(c["nsv"]) + 1, np.int32)
        irect = np.ascontiguousarray(np.asarray(c["irect"]).reshape(-1, 4) + 1, np.int32)
        stfm = np.ascontiguousarray(c["stfm"], np.float64)
        stfns = np.ascontiguousarray(c["stfns"], np.float64)
        par = np.array([c["gap"], c.get("fric", 0.0), c.get("visc", 0.05), c.get("viscf", 0.0)],
                       np.float64)
        self._chk(self.lib.orad_set_contact7(self.h, len(nsv), _p(nsv), irect.shape[0],
                                             _p(irect), _p(stfm), _p(stfns), _p(par),
                                             int(cand_cap)))
        self.nsn = len(nsv)

    def contact7(self, dt1, dt2t=1.0e30):
        """orad_contact7: one contact pass with DT1; returns the lowered DT2T."""
        d = C.c_double(dt2t)
        self._chk(self.lib.orad_contact7(self.h, float(dt1), C.byref(d)))
        return d.value

    def contact7_state(self):
        """(ECONT, ECONTV, ECONTD), dict of counts, STFNS of orad_get_contact7."""
        e = np.zeros(3)
        k = np.zeros(4, np.int32)
        st = np.zeros(max(self.nsn, 1))
        self._chk(self.lib.orad_get_contact7(self.h, _p(e), _p(k), _p(st)))
        return (tuple(e.tolist()), dict(ncand=int(k[0]), npen=int(k[1]), nsky=int(k[2]),
                                        overflow=int(k[3])), st[:self.nsn].copy())

    def contact7_sky(self):
        """The last pass's skyline entries (n, 4) and their nodes (0-based)."""
        n = C.c_int(0)
        self._chk(self.lib.orad_get_contact7_sky(self.h, 0, None, None, C.byref(n)))
        f = np.zeros((max(n.value, 1), 4))
        i = np.zeros(max(n.value, 1), np.int32)
        self._chk(self.lib.orad_get_contact7_sky(self.h, n.value, _p(f), _p(i), C.byref(n)))
        return f[:n.value].copy(), i[:n.value] - 1

    def pon(self):
        """(ADSKY(NUMNOD+1), IADS(8,NUMELS)) 1-based, as the starter writes them."""
        adsky = np.zeros(self.numnod + 1, np.int32)
        iads = np.zeros((self.numels, 8), np.int32)
        self._chk(self.lib.orad_get_pon(self.h, _p(adsky), _p(iads)))
        return adsky, iads

    def adjoint_forward(self, ncycle, every):
        """Run ncycle cycles keeping a checkpoint every `every` cycles."""
        self._chk(self.lib.orad_adjoint_forward(self.h, ncycle, every))

    def adjoint_nstate(self):
        """State values with an adjoint per brick: SIG(6), PLA."""
        return 7

    def adjoint_reverse(self, lam_x=None, lam_v=None, lam_d=None, lam_state=None):
        """Gradients from the adjoints of the final state: dict with g_jc (dA, dB, dn),
        g_ms and the adjoints of the initial x, v.  lam_state has 7 values per brick
        (SIG(6), PLA), in the engine's element order."""
        nn, ne = self.numnod, self.numels
        f = lambda a, shape: None if a is None else _f64(a, shape)
        ins = [f(lam_x, (nn, 3)), f(lam_v, (nn, 3)), None, f(lam_d, (nn, 3)),
               f(lam_state, (ne, 7))]
        out = dict(g_jc=np.zeros(3), g_ms=np.zeros(nn), lam_x0=np.zeros((nn, 3)),
                   lam_v0=np.zeros((nn, 3)))
        args = [None if a is None else _p(a) for a in ins] + [
            _p(out["g_jc"]), None, _p(out["g_ms"]), None, _p(out["lam_x0"]), _p(out["lam_v0"]),
            None]
        self._chk(self.lib.orad_adjoint_reverse(self.h, *args))
        return out

    def sforc3_async(self, nft, nel, dt1):
        """Enqueue the group nft+1..nft+nel (orad_sforc3_async); see group_wait."""
        self._chk(self.lib.orad_sforc3_async(self.h, nft, nel, dt1))

    def group_wait(self, dt2t=1.0e30):
        """Wait for the queued groups; returns min(dt2t, their DT2T)."""
        d = C.c_double(dt2t)
        self._chk(self.lib.orad_group_wait(self.h, C.byref(d)))
        return d.value

    def update_xv(self, x, v, vr=None):
        """orad_update_xv: the engine's new X, V (and shells' VR) into the device node records."""
        nn = self.numnod
        w = None if vr is None else _p(_f64(vr, (nn, 3)))
        self._chk(self.lib.orad_update_xv(self.h, _p(_f64(x, (nn, 3))), _p(_f64(v, (nn, 3))), w))

    def sforc3(self, nft, nel, dt1, dt2t=1.0e30):
        """Element forces of elements nft+1..nft+nel (SFORC3); returns the group DT2T."""
        d = C.c_double(dt2t)
        self._chk(self.lib.orad_sforc3(self.h, nft, nel, dt1, C.byref(d)))
        return d.value

    def asspar4(self):
        """Ordered nodal assembly of the FSKY slots: (A(NUMNOD,3), STIFN(NUMNOD))."""
        a = np.zeros((self.numnod, 3))
        s = np.zeros(self.numnod)
        self._chk(self.lib.orad_asspar4(self.h, _p(a), _p(s)))
        return a, s

    def run_cycles(self, n):
        t = TimeState()
        self._chk(self.lib.orad_run_cycles(self.h, n, C.byref(t)))
        return t

    def profile_cycles(self, n):
        e, nd = C.c_double(), C.c_double()
        self._chk(self.lib.orad_profile_cycles(self.h, n, C.byref(e), C.byref(nd)))
        return e.value, nd.value

    def time(self):
        t = TimeState()
        self._chk(self.lib.orad_get_time(self.h, C.byref(t)))
        return t

    def set_time(self, **kw):
        t = self.time()
        for k, v in kw.items():
            setattr(t, k, v)
        self._chk(self.lib.orad_set_time(self.h, C.byref(t)))

    def nodes(self):
        x = np.zeros((self.numnod, 3)); v = np.zeros_like(x); d = np.zeros_like(x)
        a = np.zeros_like(x); s = np.zeros(self.numnod)
        self._chk(self.lib.orad_get_nodes(self.h, _p(x), _p(v), _p(d), _p(a), _p(s)))
        return dict(x=x, v=v, d=d, a=a, stifn=s)

    def state(self):
        ne = self.numels
        sig = np.zeros((6, ne))
        cols = {k: np.zeros(ne) for k in ("eint", "rho", "pla", "qvis", "offg", "epsd")}
        self._chk(self.lib.orad_get_brick_state(self.h, _p(sig), _p(cols["eint"]), _p(cols["rho"]),
                                                _p(cols["pla"]), _p(cols["qvis"]),
                                                _p(cols["offg"]), _p(cols["epsd"])))
        cols["sig"] = sig.T.copy()
        return cols

    def set_state(self, sig=None, eint=None, rho=None, pla=None, qvis=None, offg=None):
        ne = self.numels
        sg = None if sig is None else _f64(np.asarray(sig).T, (6, ne))
        args = [None if a is None else _f64(a, (ne,)) for a in (eint, rho, pla, qvis, offg)]
        self._chk(self.lib.orad_set_brick_state(self.h, _p(sg), *[_p(a) for a in args]))

    def fsky(self):
        f = np.zeros((8 * self.numels, 4))
        self._chk(self.lib.orad_get_fsky(self.h, _p(f)))
        return f

    def energies(self):
        ei, ek, eh = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.lib.orad_energies(self.h, C.byref(ei), C.byref(ek), C.byref(eh)))
        return ei.value, ek.value, eh.value

    def balance(self):
        """ECRIT's energy balance of the last cycle (orad_energy_balance) as a dict."""
        b = Balance()
        self._chk(self.lib.orad_energy_balance(self.h, C.byref(b)))
        return b.as_dict()


def ixc_from_deck(deck):
    """IXC(NIXC=7, NUMELC) in the engine's layout (ccoor3.F:69-78): material,
    4 nodes (1-based), property, user id."""
    ne = deck["ixc"].shape[0]
    ixc = np.zeros((ne, NIXC), np.int32)
    ixc[:, 0] = 1
    ixc[:, 1:5] = deck["ixc"] + 1
    ixc[:, 5] = 1
    ixc[:, 6] = deck["uid"]
    return ixc


SHMAT_KEYS = ("rho0", "young", "nu", "shear", "a11", "a12", "ssp", "ca", "cb", "cn", "epmx",
              "ymax", "cc", "epdr", "mexp", "asrate_pm9")


class ShellSolver:
    """One shell model (law 2) resident on one GPU with 6-DOF nodes: the CFORC3
    path (Belytschko-Tsay, prop ishell = 1), the CZFORC3 path (QEPH, ishell = 24,
    numerical damping prop dn) or the CBAFORC3 path (QBAT, ishell = 12, membrane
    damping prop dn = GEO(13))."""

    def __init__(self, deck, precision=FP64, device=0):
        self.lib = load_library()
        self.precision = precision
        h = C.c_void_p()
        rc = self.lib.orad_create(C.byref(h), device, precision)
        if rc != ORAD_OK:
            raise OradError(rc, "orad_create failed (no HIP device?)")
        self.h = h
        nn = deck["x"].shape[0]
        self.numnod = nn
        self.numelc = deck["ixc"].shape[0]
        x = _f64(deck["x"], (nn, 3))
        v = _f64(deck["v"], (nn, 3))
        ms = _f64(deck["ms"], (nn,))
        icodt = np.ascontiguousarray(deck["icodt"], dtype=np.int32)
        self._chk(self.lib.orad_set_nodes(h, nn, _p(x), _p(v), _p(ms), _p(icodt)))
        vr = _f64(deck.get("vr", np.zeros((nn, 3))), (nn, 3))
        inr = _f64(deck["in"], (nn,))
        icodr = np.ascontiguousarray(deck["icodr"], dtype=np.int32)
        self._chk(self.lib.orad_set_nodes_rot(h, _p(vr), _p(inr), _p(icodr)))
        _set_fixvel(self, deck)
        mat, prop = deck["mat"], deck["prop"]
        p = np.array([mat[k] for k in SHMAT_KEYS], np.float64)
        ip = np.array([mat["icc"], mat["vp"], mat["israte"], mat["ipla"]], np.int32)
        self._chk(self.lib.orad_set_shell_law2(h, _p(p), _p(ip)))
        g = np.array([prop[k] for k in ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin")],
                     np.float64)
        self._chk(self.lib.orad_set_shell_prop(h, _p(g), int(prop["ish"]), int(prop["npt"])))
        self.ishell = int(prop.get("ishell", 1))
        self._chk(self.lib.orad_set_shell_ishell(h, self.ishell, float(prop.get("dn", 0.0))))
        self.ncol = {24: SH_NCOL_QEPH, 12: SH_NCOL_QBAT}.get(self.ishell, SH_NCOL)
        ixc = ixc_from_deck(deck)
        thk = _f64(deck["thk"], (self.numelc,))
        self._chk(self.lib.orad_set_shells(h, self.numelc, _p(ixc), _p(thk)))
        self.nmodel = 1
        if "node_off" in deck:
            self.set_models(deck["node_off"], deck["elem_off"], deck.get("mats"))
        self.nsn = 0
        if deck.get("contact") is not None:
            self.set_contact7(deck["contact"], deck["contact"].get("cand_cap", 0))

    def set_models(self, node_off, elem_off, mats=None):
        """Declare batched independent models (orad_set_shell_models); mats: one
        law-2 parameter dict per model, or None for the deck's material."""
        no = np.ascontiguousarray(node_off, dtype=np.int32)
        eo = np.ascontiguousarray(elem_off, dtype=np.int32)
        nm = no.size - 1
        p = ip = None
        if mats is not None:
            p = np.ascontiguousarray([[m[k] for k in SHMAT_KEYS] for m in mats], np.float64)
            ip = np.ascontiguousarray([[m["icc"], m["vp"], m["israte"], m["ipla"]] for m in mats],
                                      np.int32)
        self._chk(self.lib.orad_set_shell_models(self.h, nm, _p(no), _p(eo),
                                                 _p(p) if p is not None else None,
                                                 _p(ip) if ip is not None else None))
        self.nmodel = nm

    def adjoint_forward(self, ncycle, every):
        """Run ncycle cycles keeping a checkpoint every `every` cycles."""
        self._chk(self.lib.orad_adjoint_forward(self.h, ncycle, every))

    def adjoint_nstate(self):
        """State values with an adjoint per element: SIG(5) of points 1..5, PLA(1..5),
        then HOUR(1:3) for Belytschko-Tsay (33) or HOURG(1:12) for QEPH (42)."""
        return 42 if self.ishell == 24 else 33

    def adjoint_reverse(self, lam_x=None, lam_v=None, lam_vr=None, lam_d=None, lam_state=None):
        """Gradients from the adjoints of the final state: dict with g_jc (dA, dB, dn),
        g_thk, g_ms, g_in and the adjoints of the initial x, v, vr.  lam_state has
        adjoint_nstate() values per element."""
        nn, ne = self.numnod, self.numelc
        f = lambda a, shape: None if a is None else _f64(a, shape)
        ins = [f(lam_x, (nn, 3)), f(lam_v, (nn, 3)), f(lam_vr, (nn, 3)), f(lam_d, (nn, 3)),
               f(lam_state, (ne, self.adjoint_nstate()))]
        out = dict(g_jc=np.zeros(3), g_thk=np.zeros(ne), g_ms=np.zeros(nn), g_in=np.zeros(nn),
                   lam_x0=np.zeros((nn, 3)), lam_v0=np.zeros((nn, 3)), lam_vr0=np.zeros((nn, 3)))
        args = [None if a is None else _p(a) for a in ins] + [_p(out[k]) for k in
                ("g_jc", "g_thk", "g_ms", "g_in", "lam_x0", "lam_v0", "lam_vr0")]
        self._chk(self.lib.orad_adjoint_reverse(self.h, *args))
        return out

    def model_time(self, m):
        t = TimeState()
        self._chk(self.lib.orad_get_model_time(self.h, m, C.byref(t)))
        return t

    def model_energies(self, m):
        ei, ek, eh = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.lib.orad_model_energies(self.h, m, C.byref(ei), C.byref(ek), C.byref(eh)))
        return ei.value, ek.value, eh.value

    _chk = BrickSolver._chk
    set_launch = BrickSolver.set_launch
    group_wait = BrickSolver.group_wait
    update_xv = BrickSolver.update_xv
    launch = BrickSolver.launch
    autotune = BrickSolver.autotune
    write_restart = BrickSolver.write_restart
    read_restart = BrickSolver.read_restart
    set_contact7 = BrickSolver.set_contact7
    contact7 = BrickSolver.contact7
    contact7_state = BrickSolver.contact7_state
    contact7_sky = BrickSolver.contact7_sky
    close = BrickSolver.close
    __del__ = BrickSolver.__del__
    run_cycles = BrickSolver.run_cycles
    profile_cycles = BrickSolver.profile_cycles
    time = BrickSolver.time
    set_time = BrickSolver.set_time
    energies = BrickSolver.energies
    balance = BrickSolver.balance

    def pon(self):
        adsky = np.zeros(self.numnod + 1, np.int32)
        iadc = np.zeros((self.numelc, 4), np.int32)
        self._chk(self.lib.orad_get_pon(self.h, _p(adsky), _p(iadc)))
        return adsky, iadc

    def cforc3_async(self, nft, nel, dt1):
        """Enqueue the shell group nft+1..nft+nel (orad_cforc3_async); see group_wait."""
        self._chk(self.lib.orad_cforc3_async(self.h, nft, nel, dt1))

    def cforc3(self, nft, nel, dt1, dt2t=1.0e30):
        """Forces of shells nft+1..nft+nel (CFORC3); returns the group DT2T."""
        d = C.c_double(dt2t)
        self._chk(self.lib.orad_cforc3(self.h, nft, nel, dt1, C.byref(d)))
        return d.value

    def asspar4(self):
        a = np.zeros((self.numnod, 3)); ar = np.zeros_like(a)
        s = np.zeros(self.numnod); sr = np.zeros(self.numnod)
        self._chk(self.lib.orad_asspar4_rot(self.h, _p(a), _p(ar), _p(s), _p(sr)))
        return a, ar, s, sr

    def nodes(self):
        x = np.zeros((self.numnod, 3)); v = np.zeros_like(x); d = np.zeros_like(x)
        a = np.zeros_like(x); s = np.zeros(self.numnod)
        vr = np.zeros_like(x); ar = np.zeros_like(x)
        self._chk(self.lib.orad_get_nodes(self.h, _p(x), _p(v), _p(d), _p(a), _p(s)))
        self._chk(self.lib.orad_get_nodes_rot(self.h, _p(vr), _p(ar)))
        return dict(x=x, v=v, d=d, a=a, stifn=s, vr=vr, ar=ar)

    def state(self):
        rows = np.zeros((self.numelc, self.ncol))
        self._chk(self.lib.orad_get_shell_state(self.h, _p(rows)))
        return rows

    def set_state(self, rows):
        rows = _f64(rows, (self.numelc, self.ncol))
        self._chk(self.lib.orad_set_shell_state(self.h, _p(rows)))

    def fsky(self):
        f = np.zeros((4 * self.numelc, 8))
        self._chk(self.lib.orad_get_fsky(self.h, _p(f)))
        return f


DECK_MAGIC = 0x4F524144   # "ORAD"


def write_deck(deck, path):
    """Binary deck for the Fortran driver (openradioss_amd/fortran/resol_brick)."""
    nn, ne = deck["x"].shape[0], deck["ixs"].shape[0]
    mat, prop = deck["mat"], deck["prop"]
    with open(path, "wb") as f:
        np.array([DECK_MAGIC, 1, nn, ne], np.int32).tofile(f)
        np.ascontiguousarray(deck["x"], np.float64).tofile(f)
        np.ascontiguousarray(deck["v"], np.float64).tofile(f)
        np.ascontiguousarray(deck["ms"], np.float64).tofile(f)
        np.ascontiguousarray(deck["icodt"], np.int32).tofile(f)
        ixs_from_deck(deck).tofile(f)
        np.ascontiguousarray(deck["vol0"], np.float64).tofile(f)
        np.array([mat[k] for k in ("rho0", "shear", "bulk", "ca", "cb", "cn", "epmx", "sigmx",
                                   "cc", "epdr", "mexp")], np.float64).tofile(f)
        np.array([mat["icc"], mat["vp"], mat["israte"], 0], np.int32).tofile(f)
        np.array([prop[k] for k in ("qa", "qb", "hvis", "dtfac")], np.float64).tofile(f)


def write_shell_deck(deck, path):
    """Binary shell deck for the Fortran driver (openradioss_amd/fortran/resol_shell)."""
    nn, ne = deck["x"].shape[0], deck["ixc"].shape[0]
    mat, prop = deck["mat"], deck["prop"]
    with open(path, "wb") as f:
        np.array([DECK_MAGIC, 2, nn, ne], np.int32).tofile(f)
        for k in ("x", "v"):
            np.ascontiguousarray(deck[k], np.float64).tofile(f)
        np.ascontiguousarray(deck.get("vr", np.zeros((nn, 3))), np.float64).tofile(f)
        np.ascontiguousarray(deck["ms"], np.float64).tofile(f)
        np.ascontiguousarray(deck["in"], np.float64).tofile(f)
        np.ascontiguousarray(deck["icodt"], np.int32).tofile(f)
        np.ascontiguousarray(deck["icodr"], np.int32).tofile(f)
        ixc_from_deck(deck).tofile(f)
        np.ascontiguousarray(deck["thk"], np.float64).tofile(f)
        np.array([mat[k] for k in SHMAT_KEYS], np.float64).tofile(f)
        np.array([mat["icc"], mat["vp"], mat["israte"], mat["ipla"]], np.int32).tofile(f)
        np.array([prop[k] for k in ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin")],
                 np.float64).tofile(f)
        np.array([prop["ish"], prop["npt"], prop.get("ishell", 1), 0], np.int32).tofile(f)
        np.array([prop.get("dn", 0.0)], np.float64).tofile(f)


def read_shell_driver_output(path, numnod):
    a = np.fromfile(path, np.float64).reshape(4, numnod, 3)
    return a[0], a[1], a[2], a[3]


def read_driver_output(path, numnod):
    a = np.fromfile(path, np.float64)
    x, v, d = a[:3 * numnod], a[3 * numnod:6 * numnod], a[6 * numnod:9 * numnod]
    return x.reshape(numnod, 3), v.reshape(numnod, 3), d.reshape(numnod, 3)
