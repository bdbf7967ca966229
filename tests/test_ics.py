"""Seeded initial-state generators of the workloads (host side): each one
holds the properties its workload definition fixes - zero discrete div B of
the face field, the stated rms velocity and field, plasma beta, mean density,
the reference's Alfven face average - and the fixture writer produces files
the reader (mhd_init) takes back."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

GAMMA = np.float64(np.float32(5.0 / 3.0))


def fields(u, b, n):
    U = u.reshape(n, n, n, 5).astype(np.float64)  # (z, y, x, comp)
    B = b.reshape(n, n, n, 3).astype(np.float64)
    rho = U[..., 0]
    v = U[..., 1:4] / rho[..., None]
    return rho, v, B, U[..., 4]


def divb(B):
    """Left-face divergence: b1(i+1) - b1(i) + b2(j+1) - b2(j) + b3(k+1) - b3(k)."""
    return ((np.roll(B[..., 0], -1, axis=2) - B[..., 0]) + (np.roll(B[..., 1], -1, axis=1) - B[..., 1]) +
            (np.roll(B[..., 2], -1, axis=0) - B[..., 2]))


def pressure(rho, v, B, E):
    return (E - 0.5 * rho * (v ** 2).sum(-1) - 0.5 * (B ** 2).sum(-1)) * (5.0 / 3.0 - 1.0)


def test_alfven_linear_is_the_reference_wave():
    from mhdgpu.ics import alfven_linear
    n = 32
    u, b = alfven_linear(n)
    rho, v, B, E = fields(u, b, n)
    i = np.arange(n)
    dvy = 1e-4 * np.sin(2 * np.pi * (i + 1) / n)  # OpenMP/mhd.f90:179-181, 1-based i
    assert np.allclose(v[0, 0, :, 1], dvy, atol=1e-9)
    # b2 on the faces is the average of the neighbouring cells (mhd.f90:182-183)
    assert np.allclose(B[0, 0, :, 1], -(dvy + np.roll(dvy, 1)) / 2, atol=1e-9)
    assert np.all(rho == 1) and np.all(B[..., 0] == 1)
    assert np.allclose(pressure(rho, v, B, E), 0.6, atol=1e-6)
    assert np.abs(divb(B)).max() == 0


def test_random_phase_state():
    from mhdgpu.ics import random_phase
    n = 32
    u, b = random_phase(n, seed=2)
    rho, v, B, E = fields(u, b, n)
    assert np.all(rho == 1)
    rms_v = np.sqrt(((v ** 2).sum(-1)).mean())
    assert abs(rms_v - 0.1) < 1e-3  # rms Mach 0.1 at c_s = 1
    dB = B.copy()
    dB[..., 0] -= 1.0
    assert abs(np.sqrt(((dB ** 2).sum(-1)).mean()) - 0.1) < 1e-3
    # curl A on the faces: zero divergence up to the float32 rounding of b
    assert np.abs(divb(B)).max() < 1e-6
    assert np.allclose(pressure(rho, v, B, E), 0.6, atol=1e-5)
    # seeded: the same seed gives the same state, another seed another one
    u2, _ = random_phase(n, seed=2)
    u3, _ = random_phase(n, seed=3)
    assert np.array_equal(u, u2) and not np.array_equal(u, u3)


def test_turbulence_state():
    from mhdgpu.ics import turbulence
    n = 32
    u, b = turbulence(n, seed=4)
    rho, v, B, E = fields(u, b, n)
    assert abs(rho.mean() - 1) < 1e-6
    assert abs(np.log(rho).std() - 0.5) < 0.02  # log-normal, sigma 0.5
    cs = np.sqrt(GAMMA * 0.6)
    assert abs(np.sqrt(((v ** 2).sum(-1)).mean()) / cs - 1.0) < 1e-2  # rms Mach 1
    assert abs(2 * 0.6 / (B ** 2).sum(-1).mean() - 1.0) < 1e-2  # beta = 2 p / <B^2> = 1
    assert np.abs(divb(B)).max() < 1e-5
    # solenoidal velocity: the spectral divergence is small against its magnitude
    vk = [np.fft.fftn(v[..., c]) for c in range(3)]
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    div = kx * vk[0] + ky * vk[1] + kz * vk[2]
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    mag = kk * np.sqrt(sum(np.abs(x) ** 2 for x in vk))
    assert np.abs(div).sum() < 0.2 * mag.sum()  # density weighting breaks exactness


def test_batch_sizes_and_boxes():
    from mhdgpu.ics import batch_box, batch_sizes
    sizes = batch_sizes(64, 3)
    assert len(sizes) == 64 and set(sizes) <= {32, 48, 64, 96, 128} and len(set(sizes)) == 5
    assert sizes == batch_sizes(64, 3)
    u1, _ = batch_box(32, seed=1000)
    u2, _ = batch_box(32, seed=1001)
    assert not np.array_equal(u1, u2)


def test_fixture_writer_round_trip(tmp_path):
    import mhdgpu
    from mhdgpu.ics import fixture_alfven, write_fixture
    # the 16^3 writer output equals the reference's file to float32 rounding
    path = tmp_path / "fx16.dat"
    write_fixture(path, 16)
    ref = np.fromfile(os.path.join(GOLDEN, "source_init_alfvenlinear_16cube.dat"), dtype="<f4")
    got = np.fromfile(path, dtype="<f4")
    assert got.size == ref.size and np.max(np.abs(got - ref)) < 1e-6
    # a size the reference does not ship, read back through the C-ABI reader
    n = 64
    path = tmp_path / "fx64.dat"
    write_fixture(path, n)
    u, b = mhdgpu.init(path, n, order=1)  # plain Fortran order
    rec = fixture_alfven(n).reshape(-1, 8)
    assert np.array_equal(u.reshape(-1, 5), rec[:, :5]) and np.array_equal(b.reshape(-1, 3), rec[:, 5:])
