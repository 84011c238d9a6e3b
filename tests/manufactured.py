"""The manufactured solution of the reference's monodomain verification
(src/verificationModels/monodomainVerification/manufacturedFDAReference.H:33-160; case
tutorials/manufacturedSolutions/monodomainPseudoECG): on the unit cube with zeroGradient
walls, sigma = diag(0.111453302, 0.121585420, 0.030396355), chi = 3, Cm = 2 and the
monodomainFDAManufactured model,
    V  = sqrt(1 + t) F,            F = cos(pi x) cos(2 pi y) cos(3 pi z)   (3D)
    u1 = (1 + t) G + sqrt(1 + t) F, G = 1 + x y^2 z^3
    u2 = (1 + t)^-1 G^-1/2,  u3 = 0
solve chi Cm dV/dt = div(sigma grad V) - chi Cm Iion/Cm exactly."""
import numpy as np

SIGMA = (0.111453302, 0.121585420, 0.030396355)
CHI, CM = 3.0, 2.0


def centres(n):
    h = 1.0 / n
    c = (np.arange(n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel(), h


def fields(x, y, z, t):
    F = np.cos(np.pi * x) * np.cos(2 * np.pi * y) * np.cos(3 * np.pi * z)
    G = 1.0 + x * y ** 2 * z ** 3
    V = np.sqrt(1.0 + t) * F
    u1 = (1.0 + t) * G + np.sqrt(1.0 + t) * F
    u2 = 1.0 / (1.0 + t) * G ** -0.5
    return V, u1, u2, np.zeros_like(V)


def single_cell(G, t):
    """u1, u2 of a cell where F = 0 (V = 0): the single-cell tutorial's one 1D cell at
    x = 0.5, G = 1 + x."""
    return (1.0 + t) * G, G ** -0.5 / (1.0 + t)
