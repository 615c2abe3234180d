"""Seeded initial states for the benchmark and parity workloads.

A portable counter-based generator (splitmix64 + Box-Muller) so the same
seed gives the same state on every machine.  The state mapping follows the
reference's make_input.py:131-139: L = Sigma * r x v, E = Pi/(gam-1) +
Sigma |v|^2 / 2, with r the unit face centre.
"""
import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """splitmix64 finaliser applied element-wise to uint64 counters."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n):
    """n uniforms in (0, 1] from counter (seed, stream, i)."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        ctr = (base + np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) & _M
    r = splitmix64(ctr) >> np.uint64(11)
    return (r.astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def gaussian(seed, stream, n):
    """n standard normals (Box-Muller, cosine branch)."""
    u1 = uniform(seed, 2 * stream, n)
    u2 = uniform(seed, 2 * stream + 1, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def unit_centres(face_centres):
    fc = np.asarray(face_centres, dtype=np.float64)
    return fc / np.linalg.norm(fc, axis=1)[:, None]


def adiabatic_state(face_centres, gam, seed, sigma_amp=0.1, pi_frac=0.1, omega=(0.0, 0.0, 0.0),
                    noise=0.01, sigma=None, pressure=None):
    """dim-5 state: Sigma = 1 + sigma_amp*N(0,1) (or the given array),
    Pi = pi_frac*Sigma (or the given array), v = omega x r + noise * P_t(N^3)."""
    r = unit_centres(face_centres)
    n = r.shape[0]
    if sigma is None:
        sigma = 1.0 + sigma_amp * gaussian(seed, 0, n)
    sigma = np.asarray(sigma, dtype=np.float64)
    pi = pi_frac * sigma if pressure is None else np.asarray(pressure, dtype=np.float64)
    v = np.cross(np.asarray(omega, dtype=np.float64)[None, :], r)
    if noise:
        g = np.stack([gaussian(seed, 1 + c, n) for c in range(3)], axis=1)
        g -= r * np.sum(g * r, axis=1)[:, None]
        v = v + noise * g
    L = sigma[:, None] * np.cross(r, v)
    E = pi / (gam - 1.0) + sigma * np.sum(v * v, axis=1) / 2.0
    return np.column_stack([sigma, L, E])


def isothermal_state(face_centres, seed, sigma_amp=0.05):
    """dim-4 state at rest: Sigma = 1 + sigma_amp*N(0,1), L = 0."""
    n = np.asarray(face_centres).shape[0]
    sigma = 1.0 + sigma_amp * gaussian(seed, 0, n)
    return np.column_stack([sigma, np.zeros((n, 3))])


def real_sph_harm_sum(face_centres, seed, n_modes=8, lmax=16, amplitude=0.05):
    """Sum of n_modes random real spherical harmonics (l <= lmax), each scaled
    to unit peak then by `amplitude`; modes drawn from the seed."""
    from scipy.special import sph_harm_y
    r = unit_centres(face_centres)
    theta = np.arccos(np.clip(r[:, 2], -1.0, 1.0))
    phi = np.arctan2(r[:, 1], r[:, 0])
    u = uniform(seed, 99, 3 * n_modes)
    out = np.zeros(r.shape[0])
    for k in range(n_modes):
        l = 1 + int(u[3 * k] * lmax) % lmax
        m = int(u[3 * k + 1] * (2 * l + 1)) - l
        y = sph_harm_y(l, abs(m), theta, phi)
        y = y.real if m >= 0 else y.imag
        peak = np.max(np.abs(y))
        sign = 1.0 if u[3 * k + 2] > 0.5 else -1.0
        out += sign * amplitude * y / (peak if peak > 0 else 1.0)
    return out
