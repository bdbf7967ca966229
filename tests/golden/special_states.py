"""Seeded states with non-positive densities, NaN and infinite values, used
by the golden vectors and the parity tests of the CFL reduction's NaN
semantics (Heterogeneous_MHD/C/calcfl.c:55 max_2num, :93-103) and of the
exact-division fallbacks.

Each case is a random_state with a few cells overwritten; names say what.
"""
import numpy as np

from golden.states import random_state

F32 = np.float32


def special_state(name, nx, ny, nz, seed):
    """(u, b) float32 AoS arrays of the named case on an nx*ny*nz box."""
    u, b = random_state(nx, ny, nz, seed=seed, amp=0.3)
    U = u.reshape(-1, 5)
    nc = nx * ny * nz
    r = np.random.default_rng(seed + 1000)
    cells = r.choice(nc, size=4, replace=False)
    if name == "neg_density":          # one negative density: its CFL speed is NaN
        U[cells[0], 0] = F32(-0.5)
    elif name == "neg_density_many":   # several, the last one early in the box
        U[np.sort(cells)[:3] // 4, 0] = F32(-0.25)
    elif name == "zero_density":       # 0/0 and x/0 quotients
        U[cells[0], 0] = F32(0.0)
        U[cells[1], 0] = F32(0.0)
        U[cells[1], 1] = F32(0.0)
    elif name == "nan_energy":         # NaN in the middle of the box
        U[cells[0], 4] = F32(np.nan)
    elif name == "nan_last_cell":      # NaN in the last cell: the whole max is NaN
        U[nc - 1, 4] = F32(np.nan)
    elif name == "inf_density":        # inf/inf quotients in the speed
        U[cells[0], 0] = F32(np.inf)
        U[cells[0], 1] = F32(np.inf)
    elif name == "tiny_density":       # densities far below the fast-division window
        U[cells[:2], 0] = F32(2.0 ** -40)
    elif name == "huge_momentum":      # momenta, energy and speeds far above it, fluxes finite
        U[cells[0], 1] = F32(2.0 ** 61)
        U[cells[0], 4] = F32(2.0 ** 63)
        U[cells[1], 2] = F32(-(2.0 ** 60))
    elif name == "huge_field":         # face fields and momenta whose products, differences
        B = b.reshape(-1, 3)           # and limiter sums overflow FP32 in the sweeps
        B[cells[0], :] = F32(2.0 ** 64)
        B[cells[1], 1] = F32(-(2.0 ** 64))
        U[cells[0], 1:4] = F32(2.0 ** 63)
        U[cells[1], 1:4] = F32(-(2.0 ** 63))
    else:
        raise KeyError(name)
    return u, b


# states outside the committed golden vectors (pinned against the reference
# build directly)
EXTRA_CASES = ["huge_field"]

CFL_CASES = ["neg_density", "neg_density_many", "zero_density", "nan_energy", "nan_last_cell", "inf_density",
             "tiny_density", "huge_momentum"]
