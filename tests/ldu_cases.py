"""Seeded unstructured lduMatrix systems shared by the oracle and GPU tests."""
import numpy as np


def triangle_mesh(n_points, seed):
    """Cells = triangles of a Delaunay triangulation of seeded points in the unit square,
    faces = shared edges, in upper-triangular order (lower < upper, sorted by lower then
    upper) as lduAddressing keeps them."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    tri = Delaunay(rng.random((n_points, 2)))
    faces = []
    for c, nb in enumerate(tri.neighbors):
        for d in nb:
            if d > c:
                faces.append((c, d))
    faces.sort()
    f = np.array(faces, dtype=np.int32)
    return len(tri.simplices), f[:, 0].copy(), f[:, 1].copy()


def renumbered_hex(nx, ny, nz, seed):
    """A hex block with its cells renumbered at random (seed None: blockMesh order): the
    same operator, addressing no longer banded.  Returns (n, lower, upper, perm) with faces
    in upper-triangular order."""
    n = nx * ny * nz
    perm = (np.arange(n) if seed is None else np.random.default_rng(seed).permutation(n))
    perm = perm.astype(np.int32)
    faces = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                for d, ok in ((1, i < nx - 1), (nx, j < ny - 1), (nx * ny, k < nz - 1)):
                    if ok:
                        a, b = perm[c], perm[c + d]
                        faces.append((min(a, b), max(a, b)))
    faces.sort()
    f = np.array(faces, dtype=np.int32)
    return n, f[:, 0].copy(), f[:, 1].copy(), perm


def diffusion_coeffs(n, lower, upper, seed, mass=50.0):
    """Symmetric M-matrix of an implicit diffusion step: upper = -a_f (a_f > 0 seeded),
    diag = mass + sum of a_f over the cell's faces."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2.0, lower.size)
    diag = np.full(n, mass)
    np.add.at(diag, lower, a)
    np.add.at(diag, upper, a)
    return diag, -a


def csr(n, lower, upper, diag, up):
    import scipy.sparse as sp
    return sp.coo_matrix((np.concatenate([diag, up, up]),
                          (np.concatenate([np.arange(n), lower, upper]),
                           np.concatenate([np.arange(n), upper, lower]))),
                         shape=(n, n)).tocsr()
