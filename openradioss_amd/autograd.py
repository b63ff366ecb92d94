"""torch.autograd binding of the shell adjoint (orad_adjoint_forward / orad_adjoint_reverse).

ShellRun maps the law-2 parameters A, B, n, the element thicknesses THKE and the nodal
masses and rotational inertias of a shell deck to the final displacements D(NUMNOD, 3)
and plastic strains PLA(NUMELC, 5) after `ncycle` cycles; its backward pass runs the
library's adjoint sweep (checkpoint every `every` cycles, segment recomputation).  The
time-step sequence is taken as fixed, so pin it with DTMX (`dtmx`) when the gradient
must be that of the run.  The forward run is the library's own (FP64 or mixed); the
adjoint is always the FP64 chain's derivative on the forward states.

cinmas gives the nodal masses and inertias as torch functions of the thicknesses, as
the starter computes them for isotropic shells (cinmas.F:1294-1296,1322), so that mass
sensitivities reach the thickness gradient.
"""
import numpy as np
import torch
from torch.autograd.function import once_differentiable

from . import engine as en


def cinmas(x, ixc, thk, rho0):
    """Nodal masses MS, inertias IN and element areas of a shell mesh: x (NUMNOD, 3)
    and thk (NUMELC,) tensors, ixc (NUMELC, 4) 0-based node ids (int64 tensor)."""
    p = x[ixc]
    area = 0.5 * torch.linalg.norm(torch.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 1], dim=1),
                                   dim=1)
    ems = rho0 * thk * area * 0.25
    xi = ems * (area / 9.0 + thk * thk / 12.0)
    nn = x.shape[0]
    ms = torch.zeros(nn, dtype=x.dtype)
    inr = torch.zeros(nn, dtype=x.dtype)
    for k in range(4):
        ms = ms.index_add(0, ixc[:, k], ems)
        inr = inr.index_add(0, ixc[:, k], xi)
    return ms, inr, area


class ShellRun(torch.autograd.Function):
    """(A, B, n, thk, ms, in) -> (D, PLA) after `ncycle` cycles of a shell deck (a dict
    as openradioss_amd.deck builds it); every tensor input is float64 on the CPU."""

    @staticmethod
    def forward(ctx, A, B, n, thk, ms, inr, deck, precision, ncycle, every, dtmx):
        d = dict(deck)
        d["mat"] = dict(deck["mat"], ca=float(A), cb=float(B), cn=float(n))
        d["thk"] = thk.detach().numpy()
        d["ms"] = ms.detach().numpy()
        d["in"] = inr.detach().numpy()
        s = en.ShellSolver(d, precision)
        try:
            if dtmx is not None:
                s.set_time(dtmx=dtmx)
            s.adjoint_forward(ncycle, every)
            nodes, st = s.nodes(), s.state()
        except Exception:
            s.close()
            raise
        if any(ctx.needs_input_grad[:6]):
            ctx.solver = s
        else:
            # no input needs a gradient: release the device checkpoints at once
            s.close()
            ctx.solver = None
        return (torch.from_numpy(nodes["d"]),
                torch.from_numpy(st[:, en.SH_PLA0:en.SH_PLA0 + 5].copy()))

    @staticmethod
    @once_differentiable
    def backward(ctx, gd, gpla):
        s = ctx.solver
        if s is None:
            raise RuntimeError("ShellRun.backward: the adjoint sweep of this run has already "
                               "been taken (its checkpoints are released after one backward "
                               "pass), or no input required a gradient")
        ctx.solver = None
        try:
            lam_state = np.zeros((s.numelc, s.adjoint_nstate()))
            lam_state[:, 25:30] = gpla.numpy()          # PLA(1..5) in the adjoint state
            g = s.adjoint_reverse(lam_d=gd.numpy(), lam_state=lam_state)
        finally:
            s.close()
        t = torch.tensor
        return (t(g["g_jc"][0]), t(g["g_jc"][1]), t(g["g_jc"][2]), torch.from_numpy(g["g_thk"]),
                torch.from_numpy(g["g_ms"]), torch.from_numpy(g["g_in"]), None, None, None, None,
                None)


def shell_run(A, B, n, thk, ms, inr, deck, precision=en.FP64, ncycle=1000, every=100,
              dtmx=None):
    """Functional form of ShellRun.apply."""
    return ShellRun.apply(A, B, n, thk, ms, inr, deck, precision, ncycle, every, dtmx)


# COQINI force weights through the thickness (coqini.F:45-121): single-precision DATA
# literals, widened as the library holds them.
_WF = {1: (1.0,), 2: (0.5, 0.5), 3: (0.25, 0.5, 0.25),
       4: (0.1666667, 0.3333333, 0.3333333, 0.1666667), 5: (0.125, 0.25, 0.25, 0.25, 0.125)}


def force_weights(npt):
    """The npt points' force weights WF as float64 (widened from float32)."""
    return np.array(_WF[npt], dtype=np.float32).astype(np.float64)


def plastic_work(A, B, n, thk, area0, pla, npt):
    """Total plastic work of law 2 with a hardening curve A + B pla^n: per point
    V_pt (A pla + B pla^(n+1) / (n+1)), V_pt = WF_pt * thickness * initial area.
    pla (NUMELC, 5) tensor; the points past npt hold zero."""
    pos = pla > 0
    safe = torch.where(pos, pla, torch.ones_like(pla))
    w = torch.where(pos, A * pla + B * safe ** (n + 1.0) / (n + 1.0), torch.zeros_like(pla))
    wf = np.zeros(5)
    wf[:npt] = force_weights(npt)
    vpt = torch.from_numpy(wf)[None, :] * (thk * area0)[:, None]
    return (vpt * w).sum()


def displacement_plus_plastic_work(deck, A, B, n, thk, precision, ncycle, every, dtmx, node,
                                   axis=2):
    """J = final displacement of `node` along `axis` + total plastic work after `ncycle`
    cycles of a shell deck, as a torch function of A, B, n and the thicknesses (masses
    and inertias follow the thicknesses through cinmas); J.backward() runs the adjoint."""
    x = torch.from_numpy(deck["x"])
    ixc = torch.from_numpy(deck["ixc"].astype(np.int64))
    ms, inr, area0 = cinmas(x, ixc, thk, deck["mat"]["rho0"])
    d, pla = ShellRun.apply(A, B, n, thk, ms, inr, deck, precision, ncycle, every, dtmx)
    return d[node, axis] + plastic_work(A, B, n, thk, area0, pla, deck["prop"]["npt"])
