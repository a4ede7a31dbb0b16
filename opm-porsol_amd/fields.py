"""Seeded synthetic rock fields (bit-reproducible, counter-based).

Uniform deviates come from SplitMix64 applied to (seed, stream, counter), so
any sub-range of a field can be regenerated independently; normals use the
Box-Muller transform.  Permeability units: 1 mD = 9.869233e-16 m^2 (opm-core
Units, used through ReservoirPropertyCommon).
"""
from __future__ import annotations

import hashlib

import numpy as np

MILLIDARCY = 9.869233e-16
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on a uint64 array."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n deviates in (0, 1) from counters offset..offset+n-1."""
    with np.errstate(over="ignore"):
        key = splitmix64(np.array([np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream)],
                                  dtype=np.uint64))[0]
        ctr = np.arange(offset, offset + n, dtype=np.uint64) + key
    bits = splitmix64(ctr) >> np.uint64(11)                  # 53 random bits
    return (bits.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normal deviates (Box-Muller on pairs of uniforms)."""
    m = (n + 1) // 2
    u1 = uniform(seed, 2 * stream, m)
    u2 = uniform(seed, 2 * stream + 1, m)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:n]


def gaussian_filtered(dims, corr_len: float, seed: int, stream: int = 0) -> np.ndarray:
    """Standardised Gaussian random field on an (nx, ny, nz) grid.

    White noise convolved with the separable kernel exp(-r^2/(2 l^2)), truncated
    at 3 l, with periodic wrap; then shifted/scaled to zero mean, unit std.
    Returned in cell order (x fastest).
    """
    nx, ny, nz = dims
    w = normal(seed, stream, nx * ny * nz).reshape(nz, ny, nx)
    rad = int(np.ceil(3.0 * corr_len))
    offs = np.arange(-rad, rad + 1)
    ker = np.exp(-offs.astype(np.float64) ** 2 / (2.0 * corr_len ** 2))
    for axis in (2, 1, 0):
        acc = np.zeros_like(w)
        for o, kv in zip(offs, ker):
            acc += kv * np.roll(w, int(o), axis=axis)
        w = acc
    w = w.reshape(-1)
    return (w - w.mean()) / w.std()


def lognormal_iid(n: int, mean_log: float, std_log: float, seed: int, stream: int = 0) -> np.ndarray:
    return np.exp(mean_log + std_log * normal(seed, stream, n))


def diag_perm(kx: np.ndarray, ky: np.ndarray | None = None, kz: np.ndarray | None = None) -> np.ndarray:
    """Row-major 3x3 tensors [nc, 9] with the given diagonal."""
    ky = kx if ky is None else ky
    kz = kx if kz is None else kz
    K = np.zeros((kx.shape[0], 9))
    K[:, 0], K[:, 4], K[:, 8] = kx, ky, kz
    return K


def checksum(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


# --- the synthetic rock of each benchmark configuration -------------------------
def perm_c1(nc: int, seed: int = 1) -> np.ndarray:
    """20x20x10 plain case: ln k ~ N(ln 100 mD, 1.0) i.i.d., isotropic."""
    k = lognormal_iid(nc, np.log(100 * MILLIDARCY), 1.0, seed)
    return diag_perm(k)


def perm_c2(dims, seed: int = 2, corr_len: float = 8.0) -> np.ndarray:
    """128x128x64 case: ln kx = ln 100 mD + 1.5 z, z Gaussian-filtered (l = 8);
    K = diag(kx, kx, 0.1 kx)."""
    z = gaussian_filtered(dims, corr_len, seed)
    kx = np.exp(np.log(100 * MILLIDARCY) + 1.5 * z)
    return diag_perm(kx, kx, 0.1 * kx)


def perm_c4(nc: int, seed: int = 4, max_dip_deg: float = 20.0, anisotropy: float = 0.1) -> np.ndarray:
    """Faulted corner-point case: per cell K = R diag(k, k, anisotropy k) R',
    ln k ~ N(ln 100 mD, 1.0) i.i.d., R a rotation by a dip uniform on
    [0, max_dip_deg] about a horizontal axis of uniform azimuth (all seeded).
    The principal values differ (kz/kx = anisotropy), so the rotated tensors
    are full."""
    k = lognormal_iid(nc, np.log(100 * MILLIDARCY), 1.0, seed)
    dip = np.deg2rad(max_dip_deg) * uniform(seed, 1, nc)
    az = 2.0 * np.pi * uniform(seed, 2, nc)
    a = np.stack([np.cos(az), np.sin(az), np.zeros(nc)], axis=1)     # rotation axis
    c, s = np.cos(dip), np.sin(dip)
    # Rodrigues: R = c I + s [a]x + (1 - c) a a'
    ax = np.zeros((nc, 3, 3))
    ax[:, 0, 2], ax[:, 1, 2] = a[:, 1], -a[:, 0]
    ax[:, 2, 0], ax[:, 2, 1] = -a[:, 1], a[:, 0]
    R = (c[:, None, None] * np.eye(3) + s[:, None, None] * ax
         + (1.0 - c)[:, None, None] * np.einsum("ni,nj->nij", a, a))
    D = np.zeros((nc, 3, 3))
    D[:, 0, 0] = D[:, 1, 1] = k
    D[:, 2, 2] = anisotropy * k
    K = np.einsum("nij,njk,nlk->nil", R, D, R)
    K = 0.5 * (K + K.transpose(0, 2, 1))
    return K.reshape(nc, 9)
