"""Seeded synthetic particle workloads (host side, numpy).

Counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random
numbers: as easy as 1, 2, 3", SC'11) so that the host generator here and the
device generator in csrc/ppmg_gen.hip produce the same uniforms for the same
(seed, stream, index).  Uniform doubles use 53 bits from two 32-bit words
(genrand_res53 construction), which is exact on both sides.

Workloads:
  jittered_lattice  -- one particle per cell at the cell centre, each axis
                       jittered by U(-0.3h, 0.3h); particle i = ix + N*(iy + N*iz)
  fourier_modes     -- 8 random Fourier modes k (integer, 1 <= |k| <= 8),
                       amplitudes A ~ N(0,1)^3, phases U[0, 2pi)
  divfree_velocity  -- u(x) = sum_m (A_m x k_m) cos(2pi k_m.x + phi_m), which is
                       divergence free since (A x k).k = 0
  advect            -- x <- wrap(x + dt u(x)), dt = CFL * h / max|u|
  gaussian_blobs    -- 80 % of particles in 4 Gaussian blobs (sigma 0.05,
                       centres U[0,1)^3, wrapped), 20 % uniform
"""
import numpy as np

PHILOX_M0 = 0xD2511F53
PHILOX_M1 = 0xCD9E8D57
PHILOX_W0 = 0x9E3779B9
PHILOX_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)

# stream ids (third counter word) used by the workloads
STREAM_POS = 0
STREAM_POS_Z = 1
STREAM_STRENGTH = 16      # + component index
STREAM_BLOB = 64
STREAM_MODES = 128


def philox4x32_10(ctr, key):
    """ctr: (4, n) uint32-valued array, key: (k0, k1).  Returns (4, n) uint32."""
    c = [np.asarray(ctr[i], dtype=np.uint64) & _MASK for i in range(4)]
    k0 = np.uint64(int(key[0]) & 0xFFFFFFFF)
    k1 = np.uint64(int(key[1]) & 0xFFFFFFFF)
    m0, m1 = np.uint64(PHILOX_M0), np.uint64(PHILOX_M1)
    for r in range(10):
        if r:
            k0 = np.uint64((int(k0) + PHILOX_W0) & 0xFFFFFFFF)
            k1 = np.uint64((int(k1) + PHILOX_W1) & 0xFFFFFFFF)
        p0 = m0 * c[0]
        p1 = m1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack([x.astype(np.uint32) for x in c])


def _res53(a, b):
    a = (a.astype(np.uint64) >> np.uint64(5)).astype(np.float64)
    b = (b.astype(np.uint64) >> np.uint64(6)).astype(np.float64)
    return (a * 67108864.0 + b) / 9007199254740992.0


def uniforms(seed, stream, n, offset=0):
    """Two uniform [0,1) doubles per index: returns (u_a, u_b), each (n,)."""
    i = np.arange(offset, offset + n, dtype=np.uint64)
    ctr = [i & _MASK, i >> np.uint64(32), np.full(n, stream, np.uint64), np.zeros(n, np.uint64)]
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    o = philox4x32_10(ctr, key)
    return _res53(o[0], o[1]), _res53(o[2], o[3])


def normals(seed, stream, n, offset=0):
    """One N(0,1) per index (Box-Muller, cosine branch)."""
    ua, ub = uniforms(seed, stream, n, offset)
    return np.sqrt(-2.0 * np.log1p(-ua)) * np.cos(2.0 * np.pi * ub)


def jittered_lattice(N, seed, dim=3, jitter=0.3, L=1.0, xmin=0.0):
    """One particle per cell, cell centre + U(-jitter h, jitter h) per axis."""
    h = L / N
    n = N ** dim
    ux, uy = uniforms(seed, STREAM_POS, n)
    us = [ux, uy]
    if dim == 3:
        uz, _ = uniforms(seed, STREAM_POS_Z, n)
        us.append(uz)
    idx = np.arange(n, dtype=np.int64)
    cells = [idx % N, (idx // N) % N, idx // (N * N)][:dim]
    xp = np.empty((n, dim), dtype=np.float64)
    for d in range(dim):
        t = (2.0 * jitter) * (us[d] - 0.5)
        xp[:, d] = xmin + ((cells[d].astype(np.float64) + 0.5) + t) * h
    return xp


def strengths(seed, n, lda):
    """u ~ N(0,1), shape (n, lda)."""
    return np.stack([normals(seed, STREAM_STRENGTH + c, n) for c in range(lda)], axis=1)


def fourier_modes(seed, nmodes=8, kmax=8):
    """Integer wave vectors with 1 <= |k| <= kmax, amplitudes N(0,1)^3, phases U[0,2pi)."""
    ks = []
    j = 0
    while len(ks) < nmodes:
        ua, ub = uniforms(seed, STREAM_MODES, 1, offset=j)
        uc, _ = uniforms(seed, STREAM_MODES + 1, 1, offset=j)
        j += 1
        k = np.floor(np.array([ua[0], ub[0], uc[0]]) * (2 * kmax + 1)).astype(np.int64) - kmax
        r = np.sqrt((k * k).sum())
        if 1.0 <= r <= kmax:
            ks.append(k)
    k = np.array(ks, dtype=np.float64)
    A = np.stack([normals(seed, STREAM_MODES + 2 + d, nmodes) for d in range(3)], axis=1)
    phi = 2.0 * np.pi * uniforms(seed, STREAM_MODES + 5, nmodes)[0]
    return k, A, phi


def divfree_velocity(x, modes):
    """u(x) = sum_m (A_m x k_m) cos(2pi k_m.x + phi_m); x (n,3)."""
    k, A, phi = modes
    B = np.cross(A, k)
    u = np.zeros_like(x)
    for m in range(k.shape[0]):
        c = np.cos(2.0 * np.pi * (x @ k[m]) + phi[m])
        u += c[:, None] * B[m][None, :]
    return u


def wrap(x, xmin=0.0, xmax=1.0):
    """ppm_impose_part_bc semantics: x < min -> x + L, then x >= max -> x - L."""
    L = xmax - xmin
    x = np.where(x < xmin, x + L, x)
    x = np.where(x >= xmax, x - L, x)
    return x


def advect(xp, modes, N, cfl=0.5):
    u = divfree_velocity(xp, modes)
    umax = np.sqrt((u * u).sum(axis=1)).max()
    dt = cfl * (1.0 / N) / umax
    return wrap(xp + dt * u), dt


def gaussian_blobs(n, seed, nblobs=4, sigma=0.05, frac=0.8):
    """frac of the particles from nblobs Gaussian blobs, the rest uniform; wrapped into [0,1)^3."""
    cu = [uniforms(seed, STREAM_BLOB + d, nblobs)[0] for d in range(3)]
    centres = np.stack(cu, axis=1)
    nb = int(round(frac * n))
    idx = np.arange(n)
    which = idx % nblobs
    g = np.stack([normals(seed, STREAM_BLOB + 8 + d, n) for d in range(3)], axis=1)
    uni = np.stack([uniforms(seed, STREAM_BLOB + 16 + d, n)[0] for d in range(3)], axis=1)
    x = np.where((idx < nb)[:, None], centres[which] + sigma * g, uni)
    return wrap(wrap(x))
