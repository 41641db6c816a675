"""Seeded synthetic inputs: splitmix64 -> uniform doubles -> Box-Muller normals.

All random coefficient fields are generated once on the host with this generator and copied to
the device, so the GPU path and the CPU oracle see bit-identical inputs.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n consecutive outputs of splitmix64 started at `seed` (skipping `offset` outputs)."""
    with np.errstate(over="ignore"):
        idx = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, offset=0):
    """doubles in [0, 1) from the top 53 bits"""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    """standard normals by Box-Muller on pairs of uniforms (u1 in (0, 1])"""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:n]


def lognormal(seed, n, sigma=1.0):
    """kappa_e = exp(sigma g_e), g ~ N(0, 1) i.i.d."""
    return np.exp(sigma * normal(seed, n))


def channel_field(seed, nx, ny, fraction=0.2, low=1e-4, high=1.0):
    """kappa = high except whole cell rows (horizontal channels) of kappa = low covering `fraction`
    of the rows, chosen by a seeded permutation; element index = row * nx + column"""
    rows = int(round(fraction * ny))
    order = np.argsort(uniform(seed, ny), kind="stable")
    kappa = np.full((ny, nx), high)
    kappa[np.sort(order[:rows]), :] = low
    return kappa.reshape(-1)


def batch_sizes(seed, count, lo=32, hi=512):
    """grid sizes n_i uniform in {lo, ..., hi}"""
    return lo + (splitmix64(seed, count) % np.uint64(hi - lo + 1)).astype(np.int64)
