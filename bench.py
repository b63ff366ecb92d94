"""Benchmark: explicit brick cycles on one MI355X.

Workload: the large brick deck -- 200x200x200 under-integrated hexahedra
(8,000,000 elements, 8,120,601 nodes), 1 mm spacing with +-5% seeded jitter,
law-2 Johnson-Cook steel, Isolid=1 co-rotational, bottom face fixed, top node
layer at -10 m/s.  A step is one full explicit cycle (element forces, PARITH/ON
assembly, time step, accelerations, boundary conditions, velocities,
displacements) over the whole deck, replayed from a captured HIP graph with the
deck resident in HBM.

The JSON line reports the mixed form (FP32 element arithmetic, FP64 nodes) as
`value`, the all-FP64 form beside it, the element kernel's roofline position
(algorithmic bytes per launch / mean launch time from HIP events on the
library stream), and the oracle timed on the host cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import torch  # noqa: E402

from openradioss_amd import autograd as ag  # noqa: E402
from openradioss_amd import deck as dk  # noqa: E402
from openradioss_amd import engine as en  # noqa: E402

HBM_PEAK_GBS = 8000.0


def elem_bytes(ne, nn, tsize):
    """Algorithmic HBM bytes of one element-kernel launch: every array element it
    must touch, once, at its stored width; nodes once per node (ideal reuse)."""
    reads = 8 * 4 + 8 + 12 * tsize + 8       # conn, 8 corner ranks, sig(6) eint rho pla qvis
    #                                          offg epsd, vol0
    writes = 12 * tsize + 8 * 3 * tsize      # sig(6) eint rho pla qvis epsd sti, 8 force slots
    return ne * (reads + writes) + nn * (32 + 32 + 4)   # X4, V4 records, ADSKY


def node_bytes(ne, nn, tsize):
    """Algorithmic bytes of one nodal-kernel launch: every force slot once, and per node
    ADSKY (8) + X4, V4 (32 + 32) + D (24) read, X4, V4, D written."""
    return 8 * ne * 3 * tsize + nn * (8 + 32 + 32 + 24 + 32 + 32 + 24)


def shell_elem_bytes(ne, nn, tsize, ishell=1):
    """Algorithmic bytes of one shell element-kernel launch (5 points): connectivity
    and slot addresses (4+4 int32), THKE (FP64), the 53 state columns read
    (element: FOR(5) MOM(3) THK EINT(2) EPSD HOUR(5) OFFG; points: SIG(5) PLA EPSD)
    and all but OFFG written, STI, 4 slots of 6 values, and the X4/V4/VR4 node
    records once per node.  QEPH (ishell 24) adds HOURG(6:12) (60 columns) and writes
    one STI per diagonal.  QBAT (ishell 12, 2 x 2 Gauss points) reads FOR(3), THK,
    EINT(2), EPSD, OFFG and per Gauss point the 5 layers' SIG(5), PLA, EPSD and the
    resultants FORPG(5), MOMPG(3) (178 columns), writes them back with FOR(5) and
    MOM(3) (184 columns) and one STI per diagonal."""
    if ishell == 12:
        rcol, wcol, nsti = 178, 184, 2
    else:
        ncol, nsti = (60, 2) if ishell == 24 else (53, 1)
        rcol, wcol = ncol, ncol - 1
    reads = 16 + 16 + 8 + rcol * tsize
    writes = wcol * tsize + nsti * tsize + 4 * 6 * tsize
    return ne * (reads + writes) + nn * 96


def shell_node_bytes(ne, nn, tsize):
    """Nodal kernel with rotations: every slot once, ADSKY, X4/V4/VR4/D read and
    X4/V4/VR4/D written."""
    return 4 * ne * 6 * tsize + nn * (8 + 96 + 24 + 96 + 24)


def _shell_forces(s, dt1):
    """Assembled nodal forces and moments of the current state (CFORC3 + ASSPAR4),
    leaving the element state as it was."""
    st = s.state()
    s.cforc3(0, s.numelc, dt1)
    a, ar, _, _ = s.asspar4()
    s.set_state(st)
    return a, ar


def shell_plate_leg(n, steps, err_cycles, ishell=1):
    """The 512 x 512 plate (5 points, law 2, clamped edges, central 32 x 32 patch at
    15 m/s +-10%) with Belytschko-Tsay (ishell 1), QEPH (ishell 24) or QBAT (ishell
    12) shells: element-cycles/s of both forms and the mixed form's error against the
    FP64 form after `steps` and after `err_cycles` cycles."""
    deck = dk.shell_plate(n=n, ishell=ishell)
    ne, nn = deck["ixc"].shape[0], deck["x"].shape[0]
    kind = {24: "QEPH (Ishell=24)", 12: "QBAT (Ishell=12)"}.get(ishell, "BT")
    res = {"workload": f"{kind} shell plate {n}x{n} ({ne} shells, {nn} nodes), 5 points, law 2",
           "elements": ne, "nodes": nn, "cycles_timed": steps - 2}
    snaps = {}
    for prec, name in ((en.MIXED, "mixed"), (en.FP64, "fp64")):
        s = en.ShellSolver(deck, prec)
        s.run_cycles(2)                          # graph capture outside the timed region
        t0 = time.perf_counter()
        s.run_cycles(steps - 2)
        dt = time.perf_counter() - t0
        snap = []
        for target in (steps, err_cycles):
            t = s.time()
            if t.ncycle < target:
                s.run_cycles(target - t.ncycle)
                t = s.time()
            a, ar = _shell_forces(s, t.dt2)
            snap.append((s.nodes()["d"], a, ar, t.tt))
            if target == steps:
                el, nd = s.profile_cycles(20)
        snaps[name] = snap
        s.close()
        ts = 4 if prec == en.MIXED else 8
        eb = shell_elem_bytes(ne, nn, ts, ishell)
        res[name] = dict(value=ne * (steps - 2) / dt, ms_per_step=1e3 * dt / (steps - 2),
                         elem_kernel_ms=el, node_kernel_ms=nd,
                         elem_kernel_gbs=eb / (el * 1e-3) / 1e9,
                         elem_algorithmic_bytes=eb,
                         node_kernel_gbs=shell_node_bytes(ne, nn, ts) / (nd * 1e-3) / 1e9)
        print(f"{kind} plate {name}: {res[name]['value']:.4e} el-cyc/s, elem {el:.3f} ms, "
              f"node {nd:.3f} ms", file=sys.stderr, flush=True)
    l2 = lambda x, y: float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
    res["mixed_over_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    res["mixed_vs_fp64_error"] = {}
    for i, cyc in enumerate((steps, err_cycles)):
        (d32, a32, r32, t32), (d64, a64, r64, t64) = snaps["mixed"][i], snaps["fp64"][i]
        res["mixed_vs_fp64_error"][f"after_{cyc}_cycles"] = dict(
            nodal_force_l2=l2(a32, a64), nodal_moment_l2=l2(r32, r64),
            displacement_l2=l2(d32, d64), time=abs(t32 - t64) / t64)
    return res


def contact_plate_leg(n, m, steps, warmup):
    """An m x m BT shell plate dropped at 20 m/s on a clamped n x n plate with one
    /INTER/TYPE7 interface (every upper-plate node secondary, every lower-plate shell a
    segment, GAP = thickness, VISs = 0.05, Coulomb friction 0.2 with VISf 1): mixed-form
    cycles timed from cycle `warmup` (the plates are in contact) with and without the
    interface; the difference is the contact pass (search, filter, forces, skyline)."""
    deck = dk.shell_plates_contact(n=n, m=m, fric=0.2, viscf=1.0)
    ne, nn = deck["ixc"].shape[0], deck["x"].shape[0]
    nsn, nrtm = len(deck["contact"]["nsv"]), deck["contact"]["irect"].shape[0]
    res = {"workload": f"BT shell plate {m}x{m} dropped on a clamped {n}x{n} plate "
                       f"({ne} shells, {nn} nodes), TYPE7 interface of {nsn} secondary "
                       f"nodes and {nrtm} segments",
           "elements": ne, "nodes": nn, "secondary_nodes": nsn, "segments": nrtm,
           "cycles_timed": steps}
    for key, d in (("with_contact", deck), ("without_contact", dict(deck, contact=None))):
        s = en.ShellSolver(d, en.MIXED)
        s.run_cycles(warmup)
        t0 = time.perf_counter()
        s.run_cycles(steps)
        dt = time.perf_counter() - t0
        r = dict(value=ne * steps / dt, ms_per_step=1e3 * dt / steps)
        if d.get("contact") is not None:
            (ec, ev, _), cnt, _ = s.contact7_state()
            r.update(candidates_last_cycle=cnt["ncand"], penetrating_pairs_last_cycle=cnt["npen"],
                     skyline_entries_last_cycle=cnt["nsky"], contact_energy=ec,
                     friction_energy=ev)
        res[key] = r
        s.close()
        print(f"contact plate {key}: {r['ms_per_step']:.4f} ms/cycle", file=sys.stderr,
              flush=True)
    res["contact_ms_per_cycle"] = (res["with_contact"]["ms_per_step"] -
                                   res["without_contact"]["ms_per_step"])
    return res


def _plate_objective(deck, A, B, cn, thk, prec, ncyc, dtmx, centre):
    """J = centre z displacement + plastic work of one forward run (no adjoint)."""
    x = torch.from_numpy(deck["x"])
    ixc = torch.from_numpy(deck["ixc"].astype(np.int64))
    t = torch.from_numpy(thk)
    ms, inr, area0 = ag.cinmas(x, ixc, t, deck["mat"]["rho0"])
    d = dict(deck, thk=thk, ms=ms.numpy(), **{"in": inr.numpy()})
    d["mat"] = dict(deck["mat"], ca=A, cb=B, cn=cn)
    s = en.ShellSolver(d, prec)
    s.set_time(dtmx=dtmx)
    s.run_cycles(ncyc)
    nd, st = s.nodes(), s.state()
    s.close()
    pla = torch.from_numpy(st[:, en.SH_PLA0:en.SH_PLA0 + 5].copy())
    f = lambda v: torch.tensor(v, dtype=torch.float64)
    W = ag.plastic_work(f(A), f(B), f(cn), t, area0, pla, deck["prop"]["npt"])
    return float(nd["d"][centre, 2]) + float(W)


def adjoint_leg(n=128, ncyc=2000, every=100, ishell=1, fd_params=20):
    """The adjoint study on an n x n plate (thickness x (1 + U(-0.1, 0.1)) per element,
    central patch impact, ncyc cycles, checkpoint every `every`): J = the centre node's
    final z displacement + total plastic work, FP32-element forward with the FP64
    adjoint, gradient w.r.t. A, B, n and every thickness (masses and inertias follow the
    thicknesses).  Times a plain FP32 forward run, the checkpointed forward run and the
    reverse sweep, and checks fd_params sampled parameters (A, B, n and thicknesses)
    against central differences of FP64 forward runs at h = 1e-4 and 1e-6 relative.
    The time step is pinned by DTMX at 0.8 x the initial element time step, as the
    adjoint assumes a fixed time-step sequence."""
    deck = dk.shell_plate(n=n, patch=n // 4, thk_noise=0.1, seed=20261203)
    deck["prop"] = dict(deck["prop"], ishell=ishell)
    s = en.ShellSolver(deck, en.FP64)
    dtmx = 0.8 * s.cforc3(0, s.numelc, 0.0)
    s.close()
    centre = (n // 2) * (n + 2)
    times = {}
    for key, prec in (("forward_ms", en.MIXED), ("forward_fp64_ms", en.FP64)):
        s = en.ShellSolver(deck, prec)
        s.set_time(dtmx=dtmx)
        s.run_cycles(10)                       # returns after the cycles have run
        t0 = time.perf_counter()
        s.run_cycles(ncyc)
        times[key] = 1e3 * (time.perf_counter() - t0)
        s.close()
    m = deck["mat"]
    A = torch.tensor(m["ca"], dtype=torch.float64, requires_grad=True)
    B = torch.tensor(m["cb"], dtype=torch.float64, requires_grad=True)
    cn = torch.tensor(m["cn"], dtype=torch.float64, requires_grad=True)
    thk = torch.tensor(deck["thk"], dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    J = ag.displacement_plus_plastic_work(deck, A, B, cn, thk, en.MIXED, ncyc, every, dtmx, centre)
    t1 = time.perf_counter()
    J.backward()
    t2 = time.perf_counter()
    times["checkpointed_forward_ms"] = 1e3 * (t1 - t0)
    times["reverse_ms"] = 1e3 * (t2 - t1)
    g = [float(A.grad), float(B.grad), float(cn.grad)]
    gthk = thk.grad.numpy()
    res = dict(workload=f"{'QEPH' if ishell == 24 else 'BT'} shell plate {n}x{n} "
                        f"({n * n} shells), thickness +-10%, {ncyc} cycles, checkpoint every "
                        f"{every}, J = centre node final z displacement + plastic work, "
                        f"FP32-element forward, FP64 adjoint",
               shells=n * n, cycles=ncyc, every=every, J=float(J.detach()), **times,
               reverse_over_forward=times["reverse_ms"] / times["forward_ms"],
               reverse_over_fp64_forward=times["reverse_ms"] / times["forward_fp64_ms"],
               g_jc=g)
    if fd_params:
        # the same gradient from an FP64 forward run, for the FD comparison
        th = [torch.tensor(v, dtype=torch.float64, requires_grad=True)
              for v in (m["ca"], m["cb"], m["cn"], deck["thk"])]
        J64 = ag.displacement_plus_plastic_work(deck, *th, en.FP64, ncyc, every, dtmx, centre)
        J64.backward()
        grads = {"fp32_forward": (g, gthk),
                 "fp64_forward": ([float(t.grad) for t in th[:3]], th[3].grad.numpy())}
        rng = np.random.default_rng(20261203)
        ne = n * n
        patch = np.flatnonzero(np.abs(deck["x"][deck["ixc"][:, 0], 0] - 0.5 * n) < 18.0)
        k = fd_params - 3
        elems = list(rng.choice(patch, (k + 1) // 2, replace=False)) + \
            list(rng.choice(ne, k // 2, replace=False))
        base = [m["ca"], m["cb"], m["cn"], deck["thk"].copy()]
        hs = (1e-4, 1e-6)
        errs = {(key, h): [] for key in grads for h in hs}
        t0 = time.perf_counter()
        for which, idx in [(0, 0), (1, 0), (2, 0)] + [(3, int(e)) for e in elems]:
            for h in hs:
                vals = []
                for sgn in (1.0, -1.0):
                    th = [base[0], base[1], base[2], base[3].copy()]
                    if which < 3:
                        step = h * abs(base[which])
                        th[which] = base[which] + sgn * step
                    else:
                        step = h * abs(base[3][idx])
                        th[3][idx] = base[3][idx] + sgn * step
                    vals.append(_plate_objective(deck, th[0], th[1], th[2], th[3], en.FP64, ncyc,
                                                 dtmx, centre))
                gf = (vals[0] - vals[1]) / (2.0 * step)
                for key, (gj, gt) in grads.items():
                    gadj = gj[which] if which < 3 else float(gt[idx])
                    errs[(key, h)].append(abs(gadj - gf) / max(abs(gf), 1e-30))
        res["fd_check"] = dict(
            parameters=fd_params, fd_forward="FP64", seconds=time.perf_counter() - t0,
            **{f"{key}_h={h:g}": dict(max_rel_err=max(e), median_rel_err=float(np.median(e)),
                                      max_rel_err_A_B_n=max(e[:3]))
               for (key, h), e in errs.items()})
    print(f"adjoint: forward (FP32) {times['forward_ms']:.0f} ms, forward (FP64) "
          f"{times['forward_fp64_ms']:.0f} ms, checkpointed forward "
          f"{times['checkpointed_forward_ms']:.0f} ms, reverse {times['reverse_ms']:.0f} ms"
          + (f"; FD h=1e-6 max rel err FP32-fwd "
             f"{res['fd_check']['fp32_forward_h=1e-06']['max_rel_err']:.2e} FP64-fwd "
             f"{res['fd_check']['fp64_forward_h=1e-06']['max_rel_err']:.2e}" if fd_params else ""),
          file=sys.stderr, flush=True)
    return res


def group_path_leg(deck, prec, sizes=(128, 8192), reps=3):
    """The engine-side call surface on the same deck: one element pass as the FORINT
    loop drives it, orad_sforc3_async per group of NVSIZ elements and one
    orad_group_wait, timed on the host; and orad_update_xv of the whole X, V."""
    s = en.BrickSolver(deck, prec)
    s.run_cycles(5)
    dt1 = s.time().dt2
    res = {}
    for nv in sizes:
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            for nft in range(0, s.numels, nv):
                s.sforc3_async(nft, min(nv, s.numels - nft), dt1)
            s.group_wait()
            best = min(best, time.perf_counter() - t0)
        res[f"element_pass_ms_groups_of_{nv}"] = 1e3 * best
    nd = s.nodes()
    t0 = time.perf_counter()
    s.update_xv(nd["x"], nd["v"])
    res["update_xv_ms"] = 1e3 * (time.perf_counter() - t0)
    s.close()
    print(f"group path: " + ", ".join(f"{k} {v:.1f}" for k, v in res.items()), file=sys.stderr,
          flush=True)
    return res


def brick_adjoint_leg(n=60, ncyc=1000, every=100):
    """Adjoint of an n^3 brick block run whose top nodes start downward at 50 m/s:
    J = final z displacement of a top node + sum of V0 * 1e3 * PLA, gradient w.r.t.
    A, B, n, the nodal masses and the initial X, V (FP64 adjoint of an FP32-element
    forward run).  Times a plain forward run of each form, the checkpointed forward run
    and the reverse sweep; the time step is pinned by DTMX at 0.8 x the initial one."""
    deck = dk.brick_block(n, n, n, v_top=-5.0e4)
    s = en.BrickSolver(deck, en.FP64)
    dtmx = 0.8 * s.sforc3(0, s.numels, 0.0)
    s.close()
    times = {}
    for key, prec in (("forward_ms", en.MIXED), ("forward_fp64_ms", en.FP64)):
        s = en.BrickSolver(deck, prec)
        s.set_time(dtmx=dtmx)
        s.run_cycles(10)
        t0 = time.perf_counter()
        s.run_cycles(ncyc)
        times[key] = 1e3 * (time.perf_counter() - t0)
        s.close()
    top = np.flatnonzero(deck["x"][:, 2] > deck["x"][:, 2].max() - 1e-6)
    node = int(top[len(top) // 2])
    s = en.BrickSolver(deck, en.MIXED)
    s.set_time(dtmx=dtmx)
    t0 = time.perf_counter()
    s.adjoint_forward(ncyc, every)
    times["checkpointed_forward_ms"] = 1e3 * (time.perf_counter() - t0)
    lam_d = np.zeros((s.numnod, 3))
    lam_d[node, 2] = 1.0
    lam_state = np.zeros((s.numels, 7))
    lam_state[:, 6] = deck["vol0"] * 1.0e3
    t0 = time.perf_counter()
    g = s.adjoint_reverse(lam_d=lam_d, lam_state=lam_state)
    times["reverse_ms"] = 1e3 * (time.perf_counter() - t0)
    s.close()
    ne = deck["ixs"].shape[0]
    res = dict(workload=f"brick block {n}^3 ({ne} bricks), top at 50 m/s, {ncyc} cycles, "
                        f"checkpoint every {every}, J = top node z displacement + "
                        f"V0-weighted plastic strain, FP32-element forward, FP64 adjoint",
               bricks=ne, cycles=ncyc, every=every, **times,
               reverse_over_forward=times["reverse_ms"] / times["forward_ms"],
               reverse_over_fp64_forward=times["reverse_ms"] / times["forward_fp64_ms"],
               g_jc=[float(v) for v in g["g_jc"]])
    print(f"brick adjoint: forward (FP32) {times['forward_ms']:.0f} ms, forward (FP64) "
          f"{times['forward_fp64_ms']:.0f} ms, checkpointed forward "
          f"{times['checkpointed_forward_ms']:.0f} ms, reverse {times['reverse_ms']:.0f} ms",
          file=sys.stderr, flush=True)
    return res


def tube_batch_leg(ntube, steps, solo_cycles, autotune=True):
    """256 independent square tubes (one model each, per-tube A, thickness, size and
    time step), crushed axially at 10 m/s by FIXVEL: all tubes in one batched cycle
    (autotuned launch), both forms; and, for comparison, the same tubes run one
    model at a time (one context each, graph-replayed) for solo_cycles cycles."""
    deck = dk.shell_tubes(ntube=ntube)
    ne, nn = deck["ixc"].shape[0], deck["x"].shape[0]
    res = {"workload": f"{ntube} square tubes ({ne} shells, {nn} nodes, "
                       f"{int(np.diff(deck['elem_off']).min())}-"
                       f"{int(np.diff(deck['elem_off']).max())} shells each), FIXVEL crush, law 2",
           "elements": ne, "nodes": nn, "cycles_timed": steps}
    for prec, name in ((en.MIXED, "mixed"), (en.FP64, "fp64")):
        s = en.ShellSolver(deck, prec)
        s.run_cycles(2)
        if autotune:
            best, _ = s.autotune(cycles=20)
        else:
            best = s.launch()
        s.run_cycles(2)
        t0 = time.perf_counter()
        s.run_cycles(steps)
        dt = time.perf_counter() - t0
        el, nd = s.profile_cycles(20)
        ts = 4 if prec == en.MIXED else 8
        eb = shell_elem_bytes(ne, nn, ts)
        res[name] = dict(value=ne * steps / dt, ms_per_step=1e3 * dt / steps, launch=best,
                         elem_kernel_ms=el, node_kernel_ms=nd,
                         elem_kernel_gbs=eb / (el * 1e-3) / 1e9)
        s.close()
        print(f"tubes {name}: {res[name]['value']:.4e} el-cyc/s, elem {el:.3f} ms, node {nd:.3f} ms,"
              f" launch {best}", file=sys.stderr, flush=True)
        if prec == en.MIXED:
            best_mixed = best
    # the same batch with a TYPE7 self-contact interface per tube (every node against every
    # shell of its tube), mixed form, the mixed run's launch configuration
    cdeck = dk.tubes_self_contact(dict(deck), gap=0.5)
    s = en.ShellSolver(cdeck, en.MIXED)
    s.set_launch(**best_mixed)
    s.run_cycles(2)
    t0 = time.perf_counter()
    s.run_cycles(steps)
    dt = time.perf_counter() - t0
    (ec, ev, _), cnt, _ = s.contact7_state()
    s.close()
    res["mixed_self_contact"] = dict(
        value=ne * steps / dt, ms_per_step=1e3 * dt / steps,
        contact_ms_per_cycle=1e3 * dt / steps - res["mixed"]["ms_per_step"],
        secondary_nodes=nn, segments=ne, gap=0.5, contact_energy=ec, friction_energy=ev,
        penetrating_pairs_last_cycle=cnt["npen"], candidates_last_cycle=cnt["ncand"])
    print(f"tubes mixed with self-contact: {1e3 * dt / steps:.4f} ms/cycle, "
          f"{cnt['npen']} pairs at the end", file=sys.stderr, flush=True)
    # one model at a time, mixed form
    no, eo = deck["node_off"], deck["elem_off"]
    keys = ("x", "v", "vr", "ms", "in", "icodt", "icodr", "fv_dir", "fv_val")
    total = 0.0
    for m in range(ntube):
        md = {k: deck[k][no[m]:no[m + 1]] for k in keys}
        md.update(ixc=deck["ixc"][eo[m]:eo[m + 1]] - no[m], uid=deck["uid"][eo[m]:eo[m + 1]],
                  thk=deck["thk"][eo[m]:eo[m + 1]], mat=deck["mats"][m], prop=deck["prop"])
        s = en.ShellSolver(md, en.MIXED)
        s.run_cycles(2)
        t0 = time.perf_counter()
        s.run_cycles(solo_cycles)
        total += time.perf_counter() - t0
        s.close()
    res["one_model_at_a_time_mixed"] = dict(value=ne * solo_cycles / total,
                                            cycles=solo_cycles)
    res["batched_over_one_at_a_time"] = res["mixed"]["value"] / (ne * solo_cycles / total)
    print(f"tubes one at a time: {ne * solo_cycles / total:.4e} el-cyc/s", file=sys.stderr,
          flush=True)
    return res


def run(prec, deck, steps, warmup):
    s = en.BrickSolver(deck, prec)
    s.run_cycles(warmup)
    t0 = time.perf_counter()
    t = s.run_cycles(steps)                    # synchronises on return
    t1 = time.perf_counter()
    el_ms, nd_ms = s.profile_cycles(20)
    return s, (t1 - t0), t, el_ms, nd_ms


def mixed_vs_fp64_error(s32, s64, cycles):
    """Both forms continued to the same cycle count (`cycles` in all), then the relative
    L2 distance of their assembled nodal forces (ASSPAR4 of the last cycle's slots),
    displacements and stresses."""
    for s in (s32, s64):
        left = cycles - s.time().ncycle
        if left > 0:
            s.run_cycles(left)
    if s32.time().ncycle != s64.time().ncycle:
        return None
    a32, _ = s32.asspar4()
    a64, _ = s64.asspar4()
    d32, d64 = s32.nodes()["d"], s64.nodes()["d"]
    g32, g64 = s32.state()["sig"], s64.state()["sig"]
    rl2 = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    return dict(cycles=int(s64.time().ncycle), nodal_force_rel_l2=rl2(a32, a64),
                displacement_rel_l2=rl2(d32, d64), stress_rel_l2=rl2(g32, g64))


def cpu_baseline(deck, cycles):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    cores = len(os.sched_getaffinity(0))
    m = po.Model(deck, nthreads=cores)
    m.cycle(1)
    t0 = time.perf_counter()
    m.cycle(cycles)
    dt = time.perf_counter() - t0
    ne = deck["ixs"].shape[0]
    return dict(value=ne * cycles / dt, unit="element-cycles/s", cores=cores, kind="port",
                sample=f"{cycles} cycles of the same {ne}-brick deck, FP64 C oracle, OpenMP")


def copy_calibration():
    """A plain device copy of known size (3 x 1 GiB), for scaling the rocprofv3
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = 1 << 28
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {4 * n} bytes 3 times", flush=True)
    return 0


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Achievable HBM bandwidth on this box: a plain 1 GiB device-to-device copy
    (the runtime's copy kernel; read + write counted), timed with HIP events over
    `reps` copies after a warm-up, through the HIP runtime liborad.so uses."""
    import ctypes as C
    en.load_library()
    hip = C.CDLL("libamdhip64.so", mode=C.RTLD_GLOBAL)
    P = C.c_void_p
    hip.hipMalloc.argtypes = [C.POINTER(P), C.c_size_t]
    hip.hipFree.argtypes = [P]
    hip.hipMemcpy.argtypes = [P, P, C.c_size_t, C.c_int]
    hip.hipEventCreate.argtypes = [C.POINTER(P)]
    hip.hipEventRecord.argtypes = [P, P]
    hip.hipEventSynchronize.argtypes = [P]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), P, P]
    hip.hipEventDestroy.argtypes = [P]
    D2D = 3                                      # hipMemcpyDeviceToDevice
    x, y, e0, e1 = P(), P(), P(), P()
    if hip.hipMalloc(C.byref(x), nbytes) or hip.hipMalloc(C.byref(y), nbytes):
        raise RuntimeError("copy_bandwidth: hipMalloc failed")
    try:
        for _ in range(3):
            hip.hipMemcpy(y, x, nbytes, D2D)
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpy(y, x, nbytes, D2D)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(x)
        hip.hipFree(y)
    return 2.0 * nbytes / (ms.value / reps * 1e-3) / 1e9


def measured_traffic(prefix):
    """HBM bytes per launch of the kernel whose name starts with `prefix`, from the
    committed counter summary (profiles/counters_200cube.json), or None."""
    path = os.path.join(ROOT, "profiles", "counters_200cube.json")
    try:
        with open(path) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return None
    for name, rec in kernels.items():
        if name.startswith(prefix):
            return rec["hbm_bytes_per_launch"]
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=200, help="bricks per edge of the block")
    ap.add_argument("--cpu-cycles", type=int, default=6)
    ap.add_argument("--err-cycles", type=int, default=1000,
                    help="cycles after which the mixed and FP64 forms are compared")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-copy", action="store_true",
                    help="skip the device-copy bandwidth measurement")
    ap.add_argument("--no-shell", action="store_true", help="skip the 512^2 shell plate leg")
    ap.add_argument("--shell-n", type=int, default=512)
    ap.add_argument("--shell-steps", type=int, default=1000)
    ap.add_argument("--shell-err-cycles", type=int, default=5000)
    ap.add_argument("--no-qeph", action="store_true", help="skip the 512^2 QEPH plate leg")
    ap.add_argument("--qeph-steps", type=int, default=1000)
    ap.add_argument("--qeph-err-cycles", type=int, default=1000)
    ap.add_argument("--no-qbat", action="store_true", help="skip the 512^2 QBAT plate leg")
    ap.add_argument("--qbat-steps", type=int, default=500)
    ap.add_argument("--qbat-err-cycles", type=int, default=1000)
    ap.add_argument("--no-contact", action="store_true", help="skip the contact plate leg")
    ap.add_argument("--contact-n", type=int, default=512)
    ap.add_argument("--contact-m", type=int, default=256)
    ap.add_argument("--contact-steps", type=int, default=200)
    ap.add_argument("--contact-warmup", type=int, default=40)
    ap.add_argument("--no-tubes", action="store_true", help="skip the 256-tube batch leg")
    ap.add_argument("--tubes", type=int, default=256)
    ap.add_argument("--tube-steps", type=int, default=3000)
    ap.add_argument("--tube-solo-cycles", type=int, default=200)
    ap.add_argument("--no-autotune", action="store_true",
                    help="tube leg: keep the default launch configuration")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 128^2 adjoint leg")
    ap.add_argument("--no-group-path", action="store_true",
                    help="skip timing the per-group (SFORC3 drop-in) element pass")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path in this library; "
              "run it with --gpus 1", file=sys.stderr)
        return 2
    n = args.n
    deck = dk.brick_block(n, n, n)
    ne, nn = deck["ixs"].shape[0], deck["x"].shape[0]
    print(f"deck {n}^3: {ne} bricks, {nn} nodes", file=sys.stderr, flush=True)

    s32, dt32, t32, el32, nd32 = run(en.MIXED, deck, args.steps, args.warmup)
    print(f"mixed: {ne * args.steps / dt32:.4e} el-cyc/s, elem {el32:.3f} ms, node {nd32:.3f} ms",
          file=sys.stderr, flush=True)
    fp64 = None
    err = None
    if not args.no_fp64:
        s64, dt64, t64, el64, nd64 = run(en.FP64, deck, args.steps, args.warmup)
        fp64 = dict(value=ne * args.steps / dt64, ms_per_step=1e3 * dt64 / args.steps,
                    elem_kernel_ms=el64, node_kernel_ms=nd64,
                    elem_kernel_gbs=elem_bytes(ne, nn, 8) / (el64 * 1e-3) / 1e9)
        print(f"fp64: {fp64['value']:.4e} el-cyc/s, elem {el64:.3f} ms, node {nd64:.3f} ms",
              file=sys.stderr, flush=True)
        err = mixed_vs_fp64_error(s32, s64, args.err_cycles)
        print(f"mixed vs fp64 after {args.err_cycles} cycles: {err}", file=sys.stderr,
              flush=True)
        s64.close()
    s32.close()

    group_path = None if args.no_group_path else group_path_leg(deck, en.MIXED)
    eb = elem_bytes(ne, nn, 4)
    nb = node_bytes(ne, nn, 4)
    ach = eb / (el32 * 1e-3) / 1e9
    value = ne * args.steps / dt32
    cycle_gbs = (eb + nb) * args.steps / dt32 / 1e9
    shell = None
    if not args.no_shell:
        shell = shell_plate_leg(args.shell_n, args.shell_steps, args.shell_err_cycles)
    qeph = None
    if not args.no_qeph:
        qeph = shell_plate_leg(args.shell_n, args.qeph_steps, args.qeph_err_cycles, ishell=24)
    qbat = None
    if not args.no_qbat:
        qbat = shell_plate_leg(args.shell_n, args.qbat_steps, args.qbat_err_cycles, ishell=12)
    contact = None
    if not args.no_contact:
        contact = contact_plate_leg(args.contact_n, args.contact_m, args.contact_steps,
                                    args.contact_warmup)
    tubes = None
    if not args.no_tubes:
        tubes = tube_batch_leg(args.tubes, args.tube_steps, args.tube_solo_cycles,
                               autotune=not args.no_autotune)
    adjoint = adjoint_qeph = adjoint_brick = None
    if not args.no_adjoint:
        adjoint = adjoint_leg()
        adjoint_qeph = adjoint_leg(ishell=24, fd_params=0)
        adjoint_brick = brick_adjoint_leg()
    copy_gbs = None
    if not args.no_copy:
        copy_gbs = copy_bandwidth()
        print(f"device copy: {copy_gbs:.0f} GB/s (read + write)", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(deck, args.cpu_cycles)
        print(f"cpu: {cpu['value']:.4e} el-cyc/s on {cpu['cores']} cores", file=sys.stderr,
              flush=True)
    out = {
        "metric": "element-cycles/s & % HBM peak; mixed-FP32 vs FP64 GPU at stated force error",
        "value": value,
        "unit": "element-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt32 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "element arithmetic FP32, nodal X/V/D and all sums FP64",
        "data": "synthetic",
        "config": {"workload": f"brick block {n}x{n}x{n} ({ne} hexas, {nn} nodes), law-2 "
                               "Johnson-Cook steel, Isolid=1 co-rotational, PARITH/ON, "
                               "hipGraph-replayed cycle",
                   "elements": ne, "nodes": nn, "cycles_timed": args.steps},
        "roofline": {"bound": "hbm", "kernel": "brick_forces_kernel<float>",
                     "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS,
                     "traffic": (measured_traffic("brick_forces_kernel<float")
                                 if n == 200 else None),
                     "algorithmic_bytes_per_launch": eb, "mean_launch_ms": el32,
                     "measured_copy_gbs": copy_gbs,
                     "frac_of_measured_copy": (ach / copy_gbs) if copy_gbs else None},
        "cycle_hbm_frac": cycle_gbs / HBM_PEAK_GBS,
        "node_kernel": {"mean_launch_ms": nd32,
                        "achieved_gbs": nb / (nd32 * 1e-3) / 1e9, "algorithmic_bytes": nb},
        "fp64": fp64,
        "mixed_over_fp64": (value / fp64["value"]) if fp64 else None,
        "mixed_vs_fp64_error": err,
        "shell_plate": shell,
        "qeph_plate": qeph,
        "qbat_plate": qbat,
        "contact_plate": contact,
        "tube_batch": tubes,
        "adjoint_plate": adjoint,
        "adjoint_plate_qeph": adjoint_qeph,
        "adjoint_brick": adjoint_brick,
        "group_path": group_path,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
