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


# electrodes of tutorials/manufacturedSolutions/monodomainPseudoECG (ecgDomains.ECG)
ELECTRODES = ((-0.5, 0.5, 0.5), (1.5, 0.5, 0.5), (1.2, 0.23, 0.61), (1.35, 0.74, 0.28),
              (1.55, 0.41, 0.83))


def pseudo_ecg_reference(P, t=0.0, order=96):
    """The continuous pseudo-ECG of the manufactured V on the unit cube,
    int (sigma grad V) . (x - P) / |x - P|^3 dx, by Gauss-Legendre quadrature of the given
    order per direction (the reference verifier's referenceQuadratureOrder 96)."""
    x, w = np.polynomial.legendre.leggauss(order)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    c = np.sqrt(1.0 + t)
    cx, cy, cz = np.cos(np.pi * X), np.cos(2 * np.pi * Y), np.cos(3 * np.pi * Z)
    gx = -np.pi * np.sin(np.pi * X) * cy * cz * c
    gy = -2 * np.pi * cx * np.sin(2 * np.pi * Y) * cz * c
    gz = -3 * np.pi * cx * cy * np.sin(3 * np.pi * Z) * c
    rx, ry, rz = X - P[0], Y - P[1], Z - P[2]
    r = np.sqrt(rx * rx + ry * ry + rz * rz)
    return float(np.sum(W * (SIGMA[0] * gx * rx + SIGMA[1] * gy * ry + SIGMA[2] * gz * rz)
                        / r ** 3))
