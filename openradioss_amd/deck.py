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
