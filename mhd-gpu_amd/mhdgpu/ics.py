"""Seeded initial states for the workloads of the build (host side, numpy).

All states are float32 AoS arrays in the reference layout (u: 5 per cell,
b: 3 per cell on the left faces, x fastest), periodic, gamma = 5/3.

* ``alfven_linear``  - the linear Alfven wave: rho = 1, p0 = 0.6, B0 = (1,0,0),
  dv_y = eps sin(2 pi (i+1)/n) with the reference's 1-based i and
  b_y = -(dv_y(i) + dv_y(i-1))/2 on the faces (OpenMP/mhd.f90:168-184).
* ``fixture_alfven`` - the reference's committed fixture content
  (data/source_init_alfvenlinear_*cube.dat, OpenMP/mhd.f90:174-184, with its
  literal 3.14159 for pi), in plain x-fastest order; ``write_fixture``
  writes it in the fixture file format.
* ``random_phase``   - the Alfven background plus a Gaussian random-phase
  velocity field, P(k) ~ k^-2 for 1 <= |k| <= 8, rms |dv| = 0.1 (Mach 0.1 for
  c_s = 1), and dB = curl A on the faces from a random vector potential of the
  same spectrum, rms |dB| = 0.1 B0, so the discrete div B is zero up to the
  float32 rounding of b.
* ``turbulence``     - decaying-turbulence start: log-normal density
  (sigma 0.5, mean 1), solenoidal velocity P(k) ~ k^-11/3 for 2 <= k <= 32 at
  rms Mach 1, B = curl A with plasma beta 2 p0/<B^2> = 1.
"""
import os

import numpy as np

try:
    import scipy.fft as _fft
    _FFT_KW = {"workers": min(16, os.cpu_count() or 1)}
except ImportError:  # numpy's single-threaded FFT
    _fft = np.fft
    _FFT_KW = {}

GAMMA = 5.0 / 3.0


def _noop(*a):
    pass


def _pack(rho, v, bface, p):
    """(rho[n3], v[3,n3], b[3,n3] faces, p[n3]) in (z,y,x) index order -> AoS float32."""
    n3 = rho.size
    u = np.empty((n3, 5))
    u[:, 0] = rho.ravel()
    for c in range(3):
        u[:, 1 + c] = (rho * v[c]).ravel()
    b = np.stack([bface[c].ravel() for c in range(3)], axis=1)
    vv = sum(v[c].ravel() ** 2 for c in range(3))
    bb = (b ** 2).sum(1)
    u[:, 4] = p.ravel() / (GAMMA - 1.0) + 0.5 * rho.ravel() * vv + 0.5 * bb
    return u.astype(np.float32).ravel(), b.astype(np.float32).ravel()


def alfven_linear(n, eps=1e-4, p0=0.6, phase=0.0, ny=None, nz=None):
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    i = np.arange(n)
    dvy = eps * np.sin(2 * np.pi * (i + 1) / n + phase)
    by = -(dvy + np.roll(dvy, 1)) / 2.0
    shape = (nz, ny, n)
    rho = np.ones(shape)
    v = np.zeros((3,) + shape)
    v[1] = dvy[None, None, :]
    bf = np.zeros((3,) + shape)
    bf[0] = 1.0
    bf[1] = by[None, None, :]
    return _pack(rho, v, bf, np.full(shape, p0))


def fixture_alfven(n):
    """The content of the reference's fixture files in x-fastest order (8 floats per cell)."""
    i = np.arange(1, n + 1, dtype=np.float32)
    u3 = (np.float32(0.1) * np.sin(np.float32(2 * 3.14159) * i / np.float32(n))).astype(np.float32)
    b2 = ((-u3 + -np.roll(u3, 1)) / np.float32(2)).astype(np.float32)
    u5 = (np.float32(0.001) + (np.float32(1) + b2 * b2) / np.float32(2) + u3 * u3 / np.float32(2)).astype(np.float32)
    rec = np.zeros((n, n, n, 8), np.float32)
    rec[..., 0] = 1
    rec[..., 2] = u3[None, None, :]
    rec[..., 4] = u5[None, None, :]
    rec[..., 5] = 1
    rec[..., 6] = b2[None, None, :]
    return rec.ravel()


def write_fixture(path, n):
    """Write the reference's fixture format for an n^3 box (Fortran stream of
    8 float32 per cell, x fastest: OpenMP_offloadMIC/work.f90:127-132) with
    the fixture content of fixture_alfven; sizes the reference does not ship
    (64^3, 128^3) come out the same way as the 16^3 and 32^3 files."""
    fixture_alfven(n).astype("<f4").tofile(path)


def _kgrid(n):
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    return kx, ky, kz, np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)


def _gaussian_field(n, rng, kmin, kmax, slope, ncomp=3, solenoidal=False):
    """Real random field(s) with power spectrum P(k) ~ k^slope on kmin <= |k| <= kmax."""
    kx, ky, kz, kk = _kgrid(n)
    shell = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[shell] = kk[shell] ** (slope / 2.0)
    # P(k) is the 3-D spectrum density per mode shell: amplitude per mode
    # ~ sqrt(P(k) / (4 pi k^2))
    amp[shell] /= kk[shell]
    fh = []
    for _ in range(ncomp):
        re = rng.standard_normal(kk.shape)
        im = rng.standard_normal(kk.shape)
        fh.append(amp * (re + 1j * im))
    if solenoidal:
        k2 = np.where(kk > 0, kk ** 2, 1.0)
        kd = (kx * fh[0] + ky * fh[1] + kz * fh[2]) / k2
        fh = [fh[0] - kx * kd, fh[1] - ky * kd, fh[2] - kz * kd]
    return np.stack([np.real(_fft.ifftn(f, **_FFT_KW)) for f in fh])


def _curl_faces(A):
    """Left-face B = curl A with A on cell edges; arrays indexed (z, y, x).
    b_x(i,j,k) = dA_z/dy - dA_y/dz etc. with forward differences, so the
    left-face divergence vanishes identically."""
    Ax, Ay, Az = A
    d = lambda a, ax: np.roll(a, -1, axis=ax) - a  # noqa: E731  forward difference
    bx = d(Az, 1) - d(Ay, 0)
    by = d(Ax, 0) - d(Az, 2)
    bz = d(Ay, 2) - d(Ax, 1)
    return np.stack([bx, by, bz])


def _rms(v):
    return np.sqrt(np.mean(np.sum(v ** 2, axis=0)))


def random_phase(n, seed, mach=0.1, db=0.1, p0=0.6):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    v = _gaussian_field(n, rng, 1, 8, -2.0)
    v *= mach / _rms(v)  # c_s = sqrt(gamma p0 / rho0) = 1
    A = _gaussian_field(n, rng, 1, 8, -2.0)
    dB = _curl_faces(A)
    dB *= db / _rms(dB)
    bf = dB.copy()
    bf[0] += 1.0
    return _pack(np.ones(shape), v, bf, np.full(shape, p0))


def turbulence(n, seed, mach=1.0, beta=1.0, sigma=0.5, p0=0.6, progress=_noop):
    rng = np.random.default_rng(seed)
    progress("turbulence: density")
    g = _gaussian_field(n, rng, 1, n // 4, -11.0 / 3.0, ncomp=1)[0]
    g = g / g.std()
    rho = np.exp(sigma * g)
    rho /= rho.mean()
    progress("turbulence: velocity")
    v = _gaussian_field(n, rng, 2, min(32, n // 2), -11.0 / 3.0, solenoidal=True)
    cs = np.sqrt(GAMMA * p0)
    v *= mach * cs / _rms(v)
    progress("turbulence: vector potential")
    A = _gaussian_field(n, rng, 2, min(32, n // 2), -11.0 / 3.0)
    progress("turbulence: field")
    bf = _curl_faces(A)
    bf *= np.sqrt(2 * p0 / beta) / _rms(bf)
    return _pack(rho, v, bf, np.full(rho.shape, p0))


def batch_sizes(nbox=64, seed=3, choices=(32, 48, 64, 96, 128)):
    """Box sizes of the batched workload: nbox draws, uniform over choices."""
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.choice(np.array(choices), size=nbox)]


def batch_box(n, seed):
    """One box of the batched workload: the linear Alfven wave with a seeded
    phase and a seeded +-10% amplitude."""
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0, 2 * np.pi)
    eps = 1e-4 * (1 + 0.1 * rng.uniform(-1, 1))
    return alfven_linear(n, eps=eps, phase=phase)
