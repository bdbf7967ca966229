"""Seeded random MHD states used by the golden vectors and the parity tests.

Density 1 +- amp, velocity +- amp, B = (1, 0.5, -0.3) +- amp on faces,
pressure 0.6 (1 +- amp); total energy from p/(gamma-1) + rho v^2/2 + B^2/2
with gamma = 5/3.  Arrays are float32, Fortran order, component fastest.
"""
import numpy as np


def random_state(nx, ny, nz, seed, amp=0.3):
    r = np.random.default_rng(seed)
    nc = nx * ny * nz
    u = np.zeros((nc, 5))
    b = np.zeros((nc, 3))
    u[:, 0] = 1 + amp * r.uniform(-1, 1, nc)
    v = amp * r.uniform(-1, 1, (nc, 3))
    u[:, 1:4] = u[:, 0:1] * v
    b[:] = np.array([1.0, 0.5, -0.3]) + amp * r.uniform(-1, 1, (nc, 3))
    p = 0.6 * (1 + amp * r.uniform(-1, 1, nc))
    u[:, 4] = p / (2.0 / 3.0) + 0.5 * u[:, 0] * (v ** 2).sum(1) + 0.5 * (b ** 2).sum(1)
    return u.astype(np.float32).ravel(), b.astype(np.float32).ravel()
