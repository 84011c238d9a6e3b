"""Polyhedral meshes in the form cg_mesh_create_general takes, with the geometry OpenFOAM's
primitiveMesh computes (primitiveMeshFaceCentresAndAreas.C: triangle fan around the face's
point average; primitiveMeshCellCentresAndVols.C: pyramids on the faces around the average
of the cell's face centres).  `block` builds a hex block whose lattice points may be moved
by a map, faces in OpenFOAM's order (internal faces by owner, then neighbour; boundary
faces after), so a block run through the general path can be compared with the hex path,
and a warped block exercises the non-orthogonal correction."""
import numpy as np


def face_geometry(P):
    """P [F, k, 3] face points in order -> (Cf [F, 3], Sf [F, 3])"""
    est = P.mean(axis=1)
    sumN = np.zeros_like(est)
    sumA = np.zeros(len(P))
    sumAc = np.zeros_like(est)
    k = P.shape[1]
    for i in range(k):
        a, b = P[:, i], P[:, (i + 1) % k]
        nv = np.cross(b - a, est - a)
        ar = np.linalg.norm(nv, axis=1)
        sumN += nv
        sumA += ar
        sumAc += ar[:, None] * (a + b + est)
    return sumAc / (3.0 * sumA[:, None]), 0.5 * sumN


def block(nx, ny, nz, lengths=(1.0, 1.0, 1.0), warp=None):
    """dict(n, owner, neighbour, Sf, Cf, C, V, b_owner, b_Sf, b_Cf) of an nx*ny*nz block of
    the given lengths; warp(X, Y, Z) -> (X', Y', Z') moves the lattice points."""
    Lx, Ly, Lz = lengths
    xs, ys, zs = (np.linspace(0, Lx, nx + 1), np.linspace(0, Ly, ny + 1),
                  np.linspace(0, Lz, nz + 1))
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    if warp is not None:
        X, Y, Z = warp(X, Y, Z)
    pts = np.stack([X, Y, Z], axis=-1)          # [nz+1, ny+1, nx+1, 3]
    n = nx * ny * nz
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    c = I + nx * (J + ny * K)

    def P(i, j, k):
        return pts[k, j, i]

    # internal faces of each cell towards +x, +y, +z, in the order owner, then neighbour
    faces = []
    for d, (ok, nb, quad) in enumerate((
            (I < nx - 1, c + 1, lambda i, j, k: [P(i + 1, j, k), P(i + 1, j + 1, k),
                                                  P(i + 1, j + 1, k + 1), P(i + 1, j, k + 1)]),
            (J < ny - 1, c + nx, lambda i, j, k: [P(i, j + 1, k), P(i, j + 1, k + 1),
                                                   P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k)]),
            (K < nz - 1, c + nx * ny, lambda i, j, k: [P(i, j, k + 1), P(i + 1, j, k + 1),
                                                        P(i + 1, j + 1, k + 1),
                                                        P(i, j + 1, k + 1)]))):
        sel = np.nonzero(ok)[0]
        faces.append((c[sel], nb[sel], d, np.stack(quad(I[sel], J[sel], K[sel]), axis=1)))
    owner = np.concatenate([f[0] for f in faces])
    nbr = np.concatenate([f[1] for f in faces])
    dirs = np.concatenate([np.full(len(f[0]), f[2]) for f in faces])
    quads = np.concatenate([f[3] for f in faces])
    order = np.lexsort((nbr, owner))            # by owner, then neighbour
    owner, nbr, quads = owner[order], nbr[order], quads[order]
    del dirs
    # boundary faces, outward
    b = []
    jj, kk = np.meshgrid(np.arange(ny), np.arange(nz), indexing="xy")
    jj, kk = jj.ravel(), kk.ravel()
    b.append((0 + nx * (jj + ny * kk), np.stack([P(0, jj, kk), P(0, jj, kk + 1),
                                                 P(0, jj + 1, kk + 1), P(0, jj + 1, kk)], 1)))
    b.append((nx - 1 + nx * (jj + ny * kk), np.stack([P(nx, jj, kk), P(nx, jj + 1, kk),
                                                      P(nx, jj + 1, kk + 1), P(nx, jj, kk + 1)], 1)))
    ii, kk = np.meshgrid(np.arange(nx), np.arange(nz), indexing="xy")
    ii, kk = ii.ravel(), kk.ravel()
    b.append((ii + nx * (0 + ny * kk), np.stack([P(ii, 0, kk), P(ii + 1, 0, kk),
                                                 P(ii + 1, 0, kk + 1), P(ii, 0, kk + 1)], 1)))
    b.append((ii + nx * (ny - 1 + ny * kk), np.stack([P(ii, ny, kk), P(ii, ny, kk + 1),
                                                      P(ii + 1, ny, kk + 1), P(ii + 1, ny, kk)], 1)))
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    b.append((ii + nx * jj, np.stack([P(ii, jj, 0), P(ii, jj + 1, 0), P(ii + 1, jj + 1, 0),
                                      P(ii + 1, jj, 0)], 1)))
    b.append((ii + nx * (jj + ny * (nz - 1)), np.stack([P(ii, jj, nz), P(ii + 1, jj, nz),
                                                        P(ii + 1, jj + 1, nz),
                                                        P(ii, jj + 1, nz)], 1)))
    bown = np.concatenate([x[0] for x in b])
    bquads = np.concatenate([x[1] for x in b])
    Cf, Sf = face_geometry(quads)
    bCf, bSf = face_geometry(bquads)
    cnt = np.zeros(n)
    est = np.zeros((n, 3))
    for o, cf in ((owner, Cf), (nbr, Cf), (bown, bCf)):
        for q in range(3):
            est[:, q] += np.bincount(o, weights=cf[:, q], minlength=n)
        cnt += np.bincount(o, minlength=n)
    est /= cnt[:, None]
    vol3 = np.zeros(n)
    ctr = np.zeros((n, 3))
    for o, cf, sf, sgn in ((owner, Cf, Sf, 1.0), (nbr, Cf, Sf, -1.0), (bown, bCf, bSf, 1.0)):
        pv = sgn * np.einsum("ij,ij->i", sf, cf - est[o])
        pc = 0.75 * cf + 0.25 * est[o]
        vol3 += np.bincount(o, weights=pv, minlength=n)
        for q in range(3):
            ctr[:, q] += np.bincount(o, weights=pv * pc[:, q], minlength=n)
    return dict(n=n, owner=owner.astype(np.int32), neighbour=nbr.astype(np.int32), Sf=Sf,
                Cf=Cf, C=ctr / vol3[:, None], V=vol3 / 3.0, b_owner=bown.astype(np.int32),
                b_Sf=bSf, b_Cf=bCf)


def smooth_warp(eps, lengths=(1.0, 1.0, 1.0)):
    """A displacement (eps times the box size) that vanishes on the box's walls, so the
    domain stays the box while the interior faces tilt and the cells skew."""
    Lx, Ly, Lz = lengths

    def w(X, Y, Z):
        s = np.sin
        x, y, z = X / Lx, Y / Ly, Z / Lz
        dx = eps * s(np.pi * x) * s(2 * np.pi * y) * s(np.pi * z)
        dy = eps * s(np.pi * y) * s(2 * np.pi * z) * s(np.pi * x)
        dz = eps * s(np.pi * z) * s(2 * np.pi * x) * s(np.pi * y)
        return X + dx * Lx, Y + dy * Ly, Z + dz * Lz
    return w


def mesh_args(m):
    """the arguments of cardiacfoam_amd.PolyMesh"""
    return (m["owner"], m["neighbour"], m["Sf"], m["Cf"], m["C"], m["V"], m["b_owner"],
            m["b_Sf"])
