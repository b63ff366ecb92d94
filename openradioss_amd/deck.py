"""Synthetic decks for the explicit brick path (host-side setup, the starter's role).

Units: mm, tonne, s (stress in MPa).  Random numbers come from numpy's Philox
counter-based generator, so CPU and GPU runs of one seed see identical decks.

Material/property defaults follow the reference starter:
  law 2 (starter/source/materials/mat/mat002/hm_read_mat02_jc.F90:146-239):
    G = E/(2(1+nu)), K = E/(3(1-2nu)), EPS_max = SIG_max = 1e20 when unset,
    n = 1 when unset, ICC = 1, VP = 2, rate filtering off when C = 0,
    thermal exponent m = 1.
  /PROP/SOLID Isolid=1 (starter/source/properties/solid/hm_read_prop14.F:348-362):
    qa = 1.1, qb = 0.05, h = 0.1;  DTFAC = 0.9 for the element time step.
"""
import math

import numpy as np

SEED_BASE = 20261203


def law2_params(E=2.1e5, nu=0.3, rho0=7.85e-9, A=250.0, B=600.0, n=0.4, C=0.0,
                eps0=1.0, epmx=0.0, sigmx=0.0, mexp=0.0):
    if n == 0.0:
        n = 1.0
    if epmx == 0.0:
        epmx = 1.0e20
    if sigmx == 0.0:
        sigmx = 1.0e20
    if mexp == 0.0:
        mexp = 1.0
    if C == 0.0:
        eps0, israte = 1.0, 0
    else:
        israte = 1
    G = E / (2.0 * (1.0 + nu))
    K = E / (3.0 * (1.0 - 2.0 * nu))
    return dict(rho0=rho0, shear=G, bulk=K, ca=A, cb=B, cn=n, epmx=epmx, sigmx=sigmx,
                cc=C, epdr=eps0, mexp=mexp, icc=1, vp=2, israte=israte, young=E, nu=nu)


def solid_params(qa=1.1, qb=0.05, hvis=0.1, dtfac=0.9):
    return dict(qa=qa, qb=qb, hvis=hvis, dtfac=dtfac)


def brick_volumes(x, ixs):
    """Initial brick volume: the SDERI3 determinant in the global frame, FP64,
    with the reference's operation order (sderi3.F:104-155)."""
    p = x[ixs]                                   # (ne, 8, 3)
    X = [p[:, k, 0] for k in range(8)]
    Y = [p[:, k, 1] for k in range(8)]
    Z = [p[:, k, 2] for k in range(8)]
    X17, X28, X35, X46 = X[6] - X[0], X[7] - X[1], X[4] - X[2], X[5] - X[3]
    Y17, Y28, Y35, Y46 = Y[6] - Y[0], Y[7] - Y[1], Y[4] - Y[2], Y[5] - Y[3]
    Z17, Z28, Z35, Z46 = Z[6] - Z[0], Z[7] - Z[1], Z[4] - Z[2], Z[5] - Z[3]
    J1 = X17 + X28 - X35 - X46
    J2 = Y17 + Y28 - Y35 - Y46
    J3 = Z17 + Z28 - Z35 - Z46
    x1746, x2835 = X17 + X46, X28 + X35
    y1746, y2835 = Y17 + Y46, Y28 + Y35
    z1746, z2835 = Z17 + Z46, Z28 + Z35
    J4, J5, J6 = x1746 + x2835, y1746 + y2835, z1746 + z2835
    J7, J8, J9 = x1746 - x2835, y1746 - y2835, z1746 - z2835
    J5968 = J5 * J9 - J6 * J8
    J6749 = J6 * J7 - J4 * J9
    J4857 = J4 * J8 - J5 * J7
    return 0.015625 * (J1 * J5968 + J2 * J6749 + J3 * J4857)


def nodal_mass(numnod, ixs, vol0, rho0):
    """Lumped mass rho0*V0/8 per corner (starter smass3.F), summed per node in
    element order (deterministic)."""
    ms = np.zeros(numnod)
    corner = rho0 * vol0 * 0.125
    for k in range(8):
        np.add.at(ms, ixs[:, k], corner)
    return ms


def brick_block(nx, ny, nz, h=1.0, jitter=0.05, seed=SEED_BASE, v_top=-1.0e4,
                mat=None, prop=None):
    """nx*ny*nz hexahedra, node jitter U(-jitter,+jitter)*h on every coordinate of
    every node, bottom node layer fully fixed, top node layer with Vz = v_top."""
    mat = mat or law2_params()
    prop = prop or solid_params()
    ni, nj, nk = nx + 1, ny + 1, nz + 1
    numnod = ni * nj * nk
    i, j, k = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    nid = (i + ni * (j + nj * k)).ravel(order="F")
    x = np.zeros((numnod, 3))
    x[nid, 0] = i.ravel(order="F") * h
    x[nid, 1] = j.ravel(order="F") * h
    x[nid, 2] = k.ravel(order="F") * h
    rng = np.random.Generator(np.random.Philox(seed))
    if jitter:
        x += rng.uniform(-jitter * h, jitter * h, size=x.shape)
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ei, ej, ek = ei.ravel(order="F"), ej.ravel(order="F"), ek.ravel(order="F")

    def n(a, b, c):
        return a + ni * (b + nj * c)

    ixs = np.stack([n(ei, ej, ek), n(ei + 1, ej, ek), n(ei + 1, ej + 1, ek), n(ei, ej + 1, ek),
                    n(ei, ej, ek + 1), n(ei + 1, ej, ek + 1), n(ei + 1, ej + 1, ek + 1),
                    n(ei, ej + 1, ek + 1)], axis=1).astype(np.int32)
    uid = np.arange(1, ixs.shape[0] + 1, dtype=np.int32)
    vol0 = brick_volumes(x, ixs)
    ms = nodal_mass(numnod, ixs, vol0, mat["rho0"])
    v = np.zeros((numnod, 3))
    icodt = np.zeros(numnod, np.int32)
    kk = np.arange(numnod) // (ni * nj)
    icodt[kk == 0] = 7
    v[kk == nz, 2] = v_top
    return dict(x=x, v=v, ms=ms, icodt=icodt, ixs=ixs, uid=uid, vol0=vol0, mat=mat,
                prop=prop, dims=(nx, ny, nz), seed=seed)


# ============================ shells ============================

def shell_mat_params(young=2.1e5, nu=0.3, rho0=7.85e-9, ca=250.0, cb=600.0, cn=0.4,
                     epmx=None, sigmx=None, cc=0.0, eps0=0.0, mexp=1.0, fcut=0.0, icc=1,
                     vp=2, ipla=0):
    """Law 2 for shells with the derived PM values of the starter
    (starter/source/materials/mat/mat002/hm_read_mat02_jc.F90:166-239): G, A11,
    A12, SSP = sqrt(E/rho), the EP20 defaults, ISRATE and PM(9)."""
    ep20 = 1.0e20
    israte = 1
    if cc == 0.0:
        eps0 = 1.0
        israte = 0
    if vp == 1 and fcut == 0.0:
        fcut = 10000.0
    asrate = ep20 if fcut == 0.0 else 2.0 * math.pi * fcut
    return dict(rho0=rho0, young=young, nu=nu,
                shear=young / (2.0 * (1.0 + nu)),
                a11=young / (1.0 - nu * nu), a12=nu * (young / (1.0 - nu * nu)),
                ssp=math.sqrt(young / max(rho0, 1.0e-20)),
                ca=ca, cb=cb, cn=cn, epmx=ep20 if epmx is None else epmx,
                ymax=ep20 if sigmx is None else sigmx, cc=cc, epdr=eps0, mexp=mexp,
                asrate_pm9=asrate, icc=icc, vp=vp, israte=israte, ipla=ipla)


def shell_prop_params(npt=5, hm=0.01, hf=0.01, hr=0.01, dm=0.0, dtfac=0.9, dtmin=0.0,
                      ishell=1, dn=None, cvis=1.0):
    """/PROP/SHELL TYPE1 defaults (hm_read_prop01.F:155-244): shear factor 5/6
    (ISHEAR=0), hourglass coefficients 0.01 for Ishell=1.  Ishell=24 (QEPH) uses
    instead Dn (GEO(17) on input, 0.015 by default, moved to GEO(13)) and the
    hourglass stiffness factor CVIS = 1 (GEO(17)) (hm_read_prop01.F:196-229).
    Ishell=12 (QBAT, IHBE=11) keeps Hm in GEO(13) (0.01 by default, :155,209),
    which CNCOEF3 reads as its membrane damping coefficient (cncoef3.F:421-431);
    dn is that GEO(13) value for both."""
    if dn is None:
        dn = hm if ishell == 12 else 0.015
    return dict(hm=hm, hf=hf, hr=hr, dm=dm, fsh=5.0 / 6.0, dtfac=dtfac, dtmin=dtmin,
                ish=0, npt=npt, ishell=ishell, dn=dn, cvis=cvis)


def shell_masses(x, ixc, thk, rho0):
    """Nodal mass and rotational inertia as CINMAS distributes them for an isotropic
    shell (starter/source/elements/shell/coque/cinmas.F:1294-1296,1322):
    EMS = rho*t*A/4 and XI = EMS*(A/9 + t^2/12) to each corner."""
    p = x[ixc]
    d13 = p[:, 2] - p[:, 0]
    d24 = p[:, 3] - p[:, 1]
    area = 0.5 * np.linalg.norm(np.cross(d13, d24), axis=1)
    ems = rho0 * thk * area * 0.25
    xi = ems * (area / 9.0 + thk * thk * (1.0 / 12.0))
    nn = x.shape[0]
    ms = np.zeros(nn)
    inr = np.zeros(nn)
    for k in range(4):
        np.add.at(ms, ixc[:, k], ems)
        np.add.at(inr, ixc[:, k], xi)
    return ms, inr, area


def shell_plate(n=512, h=1.0, thk=1.0, patch=32, v0=1.5e4, noise=0.1, thk_noise=0.0,
                seed=20261203, npt=5, ipla=0, ishell=1):
    """Square shell plate (Ishell 1, 12 or 24): n x n elements of size h in the XY plane, clamped edges
    (all six DOF), the nodes of the central patch x patch elements starting at
    Vz = v0 * (1 + U(-noise, noise)); element thickness thk * (1 + U(-thk_noise,
    thk_noise)).  Counter-based (Philox) streams keyed on the seed."""
    ni = n + 1
    ii, jj = np.meshgrid(np.arange(ni), np.arange(ni), indexing="ij")
    x = np.zeros((ni * ni, 3))
    nid = (ii + ni * jj).ravel(order="F")
    x[nid, 0] = (ii * h).ravel(order="F")
    x[nid, 1] = (jj * h).ravel(order="F")
    ei, ej = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ei, ej = ei.ravel(order="F"), ej.ravel(order="F")
    n1 = ei + ni * ej
    ixc = np.stack([n1, n1 + 1, n1 + 1 + ni, n1 + ni], axis=1).astype(np.int32)
    ne = n * n
    rs = np.random.Generator(np.random.Philox(key=seed))
    t = thk * (1.0 + rs.uniform(-thk_noise, thk_noise, ne)) if thk_noise > 0 else np.full(ne, thk)
    v = np.zeros_like(x)
    lo = (n - patch) // 2
    inpatch = (ii >= lo) & (ii <= lo + patch) & (jj >= lo) & (jj <= lo + patch)
    pn = (ii + ni * jj)[inpatch]
    rv = np.random.Generator(np.random.Philox(key=seed + 1))
    v[np.sort(pn), 2] = v0 * (1.0 + rv.uniform(-noise, noise, pn.size))
    edge = (ii == 0) | (jj == 0) | (ii == n) | (jj == n)
    icodt = np.zeros(ni * ni, np.int32)
    icodt[(ii + ni * jj)[edge]] = 7
    icodr = icodt.copy()
    mat = shell_mat_params(ipla=ipla)
    ms, inr, area = shell_masses(x, ixc, t, mat["rho0"])
    return dict(x=x, v=v, vr=np.zeros_like(x), ms=ms, **{"in": inr}, icodt=icodt,
                icodr=icodr, ixc=ixc, uid=np.arange(1, ne + 1, dtype=np.int32), thk=t,
                mat=mat, prop=shell_prop_params(npt=npt, ishell=ishell))


def shell_plates_contact(n=16, m=8, h=1.0, thk=0.5, v0=-2.0e4, gap=None, fric=0.0,
                         visc=0.05, viscf=0.0, stfac=1.0, ishell=1, ipla=0, tilt=0.0,
                         seed=20261205):
    """A free m x m shell plate dropped on a clamped n x n plate, with a /INTER/TYPE7
    interface: the upper plate's nodes are secondary, the lower plate's shells the main
    segments.  The upper plate starts gap/4 outside the gap, centred, tilted by `tilt`
    (radians about the y axis) and moving at Vz = v0 (1 + U(-0.05, 0.05)).  Segment
    stiffness 0.5 Stfac E t, the documented TYPE7 stiffness of a shell segment; gap
    defaults to the mean of the two plates' half thicknesses summed (= thk)."""
    lower = shell_plate(n=n, h=h, thk=thk, patch=0, v0=0.0, npt=5, ipla=ipla, ishell=ishell,
                        seed=seed)
    lower["v"][:] = 0.0
    gap = thk if gap is None else gap
    mi = m + 1
    ii, jj = np.meshgrid(np.arange(mi), np.arange(mi), indexing="ij")
    xu = np.zeros((mi * mi, 3))
    nid = (ii + mi * jj).ravel(order="F")
    off = 0.5 * (n - m) * h + 0.37 * h      # not aligned with the lower mesh
    xu[nid, 0] = (ii * h).ravel(order="F") + off
    xu[nid, 1] = (jj * h).ravel(order="F") + off
    cx = off + 0.5 * m * h
    xu[:, 2] = 1.25 * gap + np.sin(tilt) * (xu[:, 0] - cx)
    ei, ej = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    ei, ej = ei.ravel(order="F"), ej.ravel(order="F")
    n1 = ei + mi * ej
    ixu = np.stack([n1, n1 + 1, n1 + 1 + mi, n1 + mi], axis=1).astype(np.int32)
    nl = lower["x"].shape[0]
    x = np.vstack([lower["x"], xu])
    ixc = np.vstack([lower["ixc"], ixu + nl]).astype(np.int32)
    ne = ixc.shape[0]
    t = np.full(ne, thk)
    mat = lower["mat"]
    ms, inr, _ = shell_masses(x, ixc, t, mat["rho0"])
    v = np.zeros_like(x)
    rv = np.random.Generator(np.random.Philox(key=seed + 7))
    v[nl:, 2] = v0 * (1.0 + rv.uniform(-0.05, 0.05, xu.shape[0]))
    icodt = np.concatenate([lower["icodt"], np.zeros(xu.shape[0], np.int32)])
    icodr = np.concatenate([lower["icodr"], np.zeros(xu.shape[0], np.int32)])
    nsv = np.arange(nl, nl + xu.shape[0], dtype=np.int32)
    contact = dict(nsv=nsv, irect=lower["ixc"].copy(),
                   stfm=np.full(lower["ixc"].shape[0], 0.5 * stfac * mat["young"] * thk),
                   stfns=np.ones(len(nsv)), gap=gap, fric=fric, visc=visc, viscf=viscf)
    return dict(x=x, v=v, vr=np.zeros_like(x), ms=ms, **{"in": inr}, icodt=icodt,
                icodr=icodr, ixc=ixc, uid=np.arange(1, ne + 1, dtype=np.int32), thk=t,
                mat=mat, prop=lower["prop"], contact=contact)


def brick_blocks_contact(nx=8, ny=8, nz=3, mx=4, my=4, mz=2, h=1.0, v0=-2.0e4, gap=0.2,
                         fric=0.0, visc=0.05, viscf=0.0, stfac=1.0, seed=SEED_BASE):
    """A free mx x my x mz brick block dropped on an nx x ny x nz block whose bottom
    layer is fixed, with a /INTER/TYPE7 interface: the upper block's bottom nodes are
    secondary, the lower block's top faces the main segments (outward normal +z).  The
    upper block starts gap/4 outside the gap, off the lower mesh's node lines, moving at
    Vz = v0.  Segment stiffness Stfac B A^2 / V of the brick under the face, the
    documented TYPE7 stiffness of a solid segment (B the bulk modulus)."""
    lower = brick_block(nx, ny, nz, h=h, jitter=0.0, seed=seed, v_top=0.0)
    upper = brick_block(mx, my, mz, h=h, jitter=0.0, seed=seed + 1, v_top=0.0)
    xu = upper["x"].copy()
    off = np.array([0.5 * (nx - mx) * h + 0.31 * h, 0.5 * (ny - my) * h + 0.29 * h,
                    nz * h + 1.25 * gap])
    xu += off
    nl = lower["x"].shape[0]
    x = np.vstack([lower["x"], xu])
    ixs = np.vstack([lower["ixs"], upper["ixs"] + nl]).astype(np.int32)
    mat = lower["mat"]
    vol0 = brick_volumes(x, ixs)
    ms = nodal_mass(x.shape[0], ixs, vol0, mat["rho0"])
    v = np.zeros_like(x)
    v[nl:, 2] = v0
    icodt = np.concatenate([lower["icodt"], np.zeros(xu.shape[0], np.int32)])
    # top faces of the lower block: nodes 5-8 of its top element layer, counterclockwise
    # seen from +z
    top = lower["ixs"][-nx * ny:]
    irect = top[:, [4, 5, 6, 7]].astype(np.int32)
    area = h * h
    stfm = stfac * mat["bulk"] * area * area / vol0[nx * ny * (nz - 1):nx * ny * nz]
    kk = np.arange(xu.shape[0]) // ((mx + 1) * (my + 1))
    nsv = (nl + np.nonzero(kk == 0)[0]).astype(np.int32)
    contact = dict(nsv=nsv, irect=irect, stfm=stfm, stfns=np.ones(len(nsv)), gap=gap,
                   fric=fric, visc=visc, viscf=viscf)
    return dict(x=x, v=v, ms=ms, icodt=icodt, ixs=ixs,
                uid=np.arange(1, ixs.shape[0] + 1, dtype=np.int32), vol0=vol0, mat=mat,
                prop=lower["prop"], contact=contact)


def square_tube(side, length, h, thk, mat, v_imp=-1.0e4, uid0=1):
    """One square thin-walled tube: cross-section side x side centred on the z axis,
    z from 0 to length; ceil(side/h) shells per face and ceil(length/h) along the
    axis.  The z = 0 ring is clamped (six DOF); the z = length ring has an imposed
    z velocity v_imp (FIXVEL) and starts at it."""
    ns = int(np.ceil(side / h))
    nl = int(np.ceil(length / h))
    npr = 4 * ns                                   # nodes per ring (closed perimeter)
    i = np.arange(npr)
    face, t = i // ns, (i % ns) / ns
    a = 0.5 * side
    px = np.select([face == 0, face == 1, face == 2, face == 3],
                   [-a + t * side, np.full(npr, a), a - t * side, np.full(npr, -a)])
    py = np.select([face == 0, face == 1, face == 2, face == 3],
                   [np.full(npr, -a), -a + t * side, np.full(npr, a), a - t * side])
    z = np.arange(nl + 1) * (length / nl)
    x = np.zeros(((nl + 1) * npr, 3))
    x[:, 0] = np.tile(px, nl + 1)
    x[:, 1] = np.tile(py, nl + 1)
    x[:, 2] = np.repeat(z, npr)
    jj, ii = np.meshgrid(np.arange(nl), np.arange(npr), indexing="ij")
    jj, ii = jj.ravel(), ii.ravel()
    n1 = jj * npr + ii
    n2 = jj * npr + (ii + 1) % npr
    ixc = np.stack([n1, n2, n2 + npr, n1 + npr], axis=1).astype(np.int32)
    ne, nn = ixc.shape[0], x.shape[0]
    tk = np.full(ne, float(thk))
    ms, inr, _ = shell_masses(x, ixc, tk, mat["rho0"])
    icodt = np.zeros(nn, np.int32)
    icodt[:npr] = 7
    icodr = icodt.copy()
    fv_dir = np.zeros(nn, np.int32)
    fv_val = np.zeros(nn)
    fv_dir[nl * npr:] = 3
    fv_val[nl * npr:] = v_imp
    v = np.zeros_like(x)
    v[nl * npr:, 2] = v_imp
    return dict(x=x, v=v, vr=np.zeros_like(x), ms=ms, **{"in": inr}, icodt=icodt, icodr=icodr,
                ixc=ixc, uid=np.arange(uid0, uid0 + ne, dtype=np.int32), thk=tk,
                fv_dir=fv_dir, fv_val=fv_val)


def tube_self_contact(side=20.0, length=40.0, h=2.0, thk=1.2, v_imp=-5.0e4, gap=None, fric=0.2,
                      visc=0.05, viscf=1.0, stfac=1.0, npt=5, ipla=0):
    """One square tube crushed axially (square_tube: clamped base ring, the top ring driven
    at v_imp) with a /INTER/TYPE7 self-contact interface: every node secondary, every shell
    a segment (a node is never a candidate of a segment it belongs to).  Segment stiffness
    0.5 Stfac E t; gap 0.6 t unless given."""
    mat = shell_mat_params(ipla=ipla)
    d = square_tube(side, length, h, thk, mat, v_imp=v_imp)
    d["mat"] = mat
    d["prop"] = shell_prop_params(npt=npt)
    nn, ne = d["x"].shape[0], d["ixc"].shape[0]
    d["contact"] = dict(nsv=np.arange(nn, dtype=np.int32), irect=d["ixc"].copy(),
                        stfm=np.full(ne, 0.5 * stfac * mat["young"] * thk),
                        stfns=np.ones(nn), gap=0.6 * thk if gap is None else gap, fric=fric,
                        visc=visc, viscf=viscf)
    return d


def shell_tubes(ntube=256, side=(30.0, 80.0), length=(100.0, 400.0), h=(2.0, 5.0),
                thk=(0.8, 2.5), ca=(200.0, 400.0), v_imp=-1.0e4, seed=20261203, npt=5, ipla=0):
    """Batch of independent square tubes crushed axially (one model each): side,
    length, element size, thickness and Johnson-Cook A drawn per tube from seeded
    uniform distributions (Philox stream keyed on the seed), the other law-2 values
    as shell_mat_params.  Returns one concatenated deck with node_off / elem_off
    (model ranges) and mats (one law-2 dict per tube)."""
    rs = np.random.Generator(np.random.Philox(key=seed))
    draws = rs.uniform(size=(ntube, 5))
    lo = np.array([side[0], length[0], h[0], thk[0], ca[0]])
    hi = np.array([side[1], length[1], h[1], thk[1], ca[1]])
    par = lo + draws * (hi - lo)
    parts, mats = [], []
    node_off, elem_off = [0], [0]
    uid0 = 1
    for k in range(ntube):
        s_, l_, h_, t_, a_ = par[k]
        mat = shell_mat_params(ca=float(a_), ipla=ipla)
        d = square_tube(s_, l_, h_, t_, mat, v_imp=v_imp, uid0=uid0)
        d["ixc"] = d["ixc"] + node_off[-1]
        uid0 += d["ixc"].shape[0]
        node_off.append(node_off[-1] + d["x"].shape[0])
        elem_off.append(elem_off[-1] + d["ixc"].shape[0])
        parts.append(d)
        mats.append(mat)
    cat = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    cat["node_off"] = np.array(node_off, np.int32)
    cat["elem_off"] = np.array(elem_off, np.int32)
    cat["mats"] = mats
    cat["mat"] = mats[0]
    cat["prop"] = shell_prop_params(npt=npt)
    cat["tubes"] = par
    return cat


def tubes_self_contact(d, gap, fric=0.2, visc=0.05, viscf=1.0, stfac=1.0):
    """Adds a /INTER/TYPE7 self-contact interface to a shell_tubes batch: every node
    secondary, every shell a segment of its own tube (pairs never cross tubes); segment
    stiffness 0.5 Stfac E t of the shell's tube."""
    nn, ne = d["x"].shape[0], d["ixc"].shape[0]
    young = np.empty(ne)
    for m, mat in enumerate(d["mats"]):
        young[d["elem_off"][m]:d["elem_off"][m + 1]] = mat["young"]
    d["contact"] = dict(nsv=np.arange(nn, dtype=np.int32), irect=d["ixc"].copy(),
                        stfm=0.5 * stfac * young * d["thk"], stfns=np.ones(nn), gap=gap,
                        fric=fric, visc=visc, viscf=viscf)
    return d
