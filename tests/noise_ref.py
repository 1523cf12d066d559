"""CPU restatement of the Brownian noise stream (test infrastructure).

Philox4x64-10 (Salmon, Moraes, Dror & Shaw, SC'11) keyed by (seed, 'UMCP')
at counter (vertex, step, 0, 0); the four 64-bit outputs become uniforms
((x >> 11) + 1/2) 2^-53 and, by Box-Muller, the normals
  z0 = sqrt(-2 ln u0) cos(2 pi u1), z1 = sqrt(-2 ln u0) sin(2 pi u1),
  z2 = sqrt(-2 ln u2) cos(2 pi u3).
``philox`` is pinned against numpy's own Philox bit generator in
tests/test_oracle_restraints.py.
"""
import math

import numpy as np

M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
KEY1 = 0x554D4350
MASK = (1 << 64) - 1


def philox(ctr, key):
    x = [int(c) & MASK for c in ctr]
    k0, k1 = int(key[0]) & MASK, int(key[1]) & MASK
    for _ in range(10):
        p0, p1 = M0 * x[0], M1 * x[2]
        x = [(p1 >> 64) ^ x[1] ^ k0, p1 & MASK, (p0 >> 64) ^ x[3] ^ k1, p0 & MASK]
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return x


def normals(seed: int, step: int, n: int) -> np.ndarray:
    """3n deviates of (seed, step) for vertices 0..n-1."""
    out = np.empty(3 * n)
    k53 = 2.0 ** -53
    for v in range(n):
        x = philox((v, step, 0, 0), (seed, KEY1))
        u = [((xx >> 11) + 0.5) * k53 for xx in x]
        r0 = math.sqrt(-2.0 * math.log(u[0]))
        r1 = math.sqrt(-2.0 * math.log(u[2]))
        out[3 * v] = r0 * math.cos(2 * math.pi * u[1])
        out[3 * v + 1] = r0 * math.sin(2 * math.pi * u[1])
        out[3 * v + 2] = r1 * math.cos(2 * math.pi * u[3])
    return out
