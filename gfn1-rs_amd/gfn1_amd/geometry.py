"""Seeded synthetic molecular clusters for the benchmark and parity workloads.

Rigid H2O / CH4 / NH3 monomers with uniformly random orientations
(Shoemake quaternions) placed uniformly in a cube with a minimum heavy-atom
distance.  The random stream is SplitMix64-seeded xoshiro256**, so the same
seed gives the same cluster in every implementation.

Workloads (all positions returned in bohr):
  water_cluster(48, 11.3, seed=0)     48 waters, 144 atoms, 384 AO
  water_cluster(1000, 31.0, seed=1)   1000 waters, 3000 atoms, 8000 AO
  water_cluster(2000, 39.1, seed=4)   2000 waters, 6000 atoms, 16000 AO
  mixed_batch(4096)                   4096 clusters of 2-40 molecules drawn
                                      from {H2O, CH4, NH3} at 0.9 g/cm^3,
                                      instance i seeded with 1000 + i
  noisy_snapshots(20000)              learned-guess training data: mixed
                                      clusters (seed 10^6 + j) displaced by
                                      N(0, 0.05 A) per coordinate
"""
import math

import numpy as np

ANGSTROM_TO_BOHR = 1.8897261246257702
MASK = (1 << 64) - 1
AVOGADRO_SCALE = 0.602214076  # g/mol / (g/cm^3) -> A^3 / 1e24 ... V[A^3] = M / (rho * 0.602214076)


class Xoshiro256:
    """xoshiro256** seeded through SplitMix64 (Blackman & Vigna)."""

    def __init__(self, seed):
        x = seed & MASK
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & MASK
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
            s.append(z ^ (z >> 31))
        self.s = s

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & MASK

    def next_u64(self):
        s = self.s
        result = (self._rotl((s[1] * 5) & MASK, 7) * 9) & MASK
        t = (s[1] << 17) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self):
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def normal(self):
        """Standard normal variate (Box-Muller, one value per call)."""
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)


def _water():
    half = math.radians(104.52) / 2.0
    x, y = 0.9572 * math.sin(half), 0.9572 * math.cos(half)
    return [8, 1, 1], np.array([[0.0, 0.0, 0.0], [x, y, 0.0], [-x, y, 0.0]])


def _methane():
    d = 1.09 / math.sqrt(3.0)
    return [6, 1, 1, 1, 1], np.array([[0, 0, 0], [d, d, d], [d, -d, -d], [-d, d, -d], [-d, -d, d]], dtype=float)


def _ammonia():
    hnh = math.radians(106.7)
    cos_t = math.sqrt((2.0 * math.cos(hnh) + 1.0) / 3.0)
    sin_t = math.sqrt(1.0 - cos_t * cos_t)
    pts = [[0.0, 0.0, 0.0]]
    for k in range(3):
        phi = 2.0 * math.pi * k / 3.0
        pts.append([1.01 * sin_t * math.cos(phi), 1.01 * sin_t * math.sin(phi), 1.01 * cos_t])
    return [7, 1, 1, 1], np.array(pts)


MONOMERS = {"water": _water(), "methane": _methane(), "ammonia": _ammonia()}
MASSES = {"water": 18.015, "methane": 16.043, "ammonia": 17.031}
MONOMER_ORDER = ("water", "methane", "ammonia")


def _rotation(rng):
    """Uniform random rotation matrix from a Shoemake unit quaternion."""
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    a = math.sqrt(1.0 - u1)
    b = math.sqrt(u1)
    x, y = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
    z, w = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def cube_side(kinds, density):
    """Cube side (Angstrom) holding the monomers at the given density (g/cm^3)."""
    mass = sum(MASSES[k] for k in kinds)
    return (mass / (density * AVOGADRO_SCALE)) ** (1.0 / 3.0)


def place(kinds, side, rng, min_dist=2.6, max_tries=2_000_000, periodic=False):
    """Random sequential placement; returns (numbers, positions in Angstrom).
    periodic: distances between centres by the minimum image of the cube."""
    centres = []
    numbers = []
    coords = []
    tries = 0
    md2 = min_dist * min_dist
    for kind in kinds:
        z, geo = MONOMERS[kind]
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("cluster placement failed (too dense)")
            c = np.array([rng.uniform() * side, rng.uniform() * side, rng.uniform() * side])
            if centres:
                d = np.asarray(centres) - c
                if periodic:
                    d -= side * np.round(d / side)
                if np.min(np.einsum("ij,ij->i", d, d)) < md2:
                    continue
            break
        rot = _rotation(rng)
        centres.append(c)
        numbers.extend(z)
        coords.append(geo @ rot.T + c)
    return np.array(numbers, dtype=np.int32), np.vstack(coords)


def water_cluster(nmol, side, seed):
    """nmol rigid waters in a cube of the given side (Angstrom); positions in bohr."""
    rng = Xoshiro256(seed)
    z, x = place(["water"] * nmol, side, rng)
    return z, x * ANGSTROM_TO_BOHR


def water_box(nmol, side, seed):
    """nmol rigid waters in a periodic cube of the given side (Angstrom): the
    centres at least 2.6 A apart under the minimum image.  Returns (numbers,
    positions in bohr, cell vectors in bohr as a 3 x 3 array of rows)."""
    rng = Xoshiro256(seed)
    z, x = place(["water"] * nmol, side, rng, periodic=True)
    return z, x * ANGSTROM_TO_BOHR, np.eye(3) * side * ANGSTROM_TO_BOHR


def mixed_cluster(seed, density=0.9, kmin=2, kmax=40):
    """One mixed H2O/CH4/NH3 cluster of k ~ U{kmin..kmax} molecules; bohr."""
    rng = Xoshiro256(seed)
    k = rng.integer(kmin, kmax)
    kinds = [MONOMER_ORDER[rng.integer(0, 2)] for _ in range(k)]
    side = cube_side(kinds, density)
    z, x = place(kinds, side, rng)
    return z, x * ANGSTROM_TO_BOHR


def mixed_batch(count, first_seed=1000, **kw):
    return [mixed_cluster(first_seed + i, **kw) for i in range(count)]


def noisy_snapshot(j, sigma_angstrom=0.05, density=0.9, kmin=2, kmax=40):
    """Snapshot j of the learned-guess data: the mixed cluster of seed 10^6 + j with
    every coordinate displaced by N(0, sigma) drawn from the same random stream
    right after placement; bohr."""
    rng = Xoshiro256(1_000_000 + j)
    k = rng.integer(kmin, kmax)
    kinds = [MONOMER_ORDER[rng.integer(0, 2)] for _ in range(k)]
    side = cube_side(kinds, density)
    z, x = place(kinds, side, rng)
    noise = np.array([rng.normal() for _ in range(x.size)]).reshape(x.shape)
    return z, (x + noise * sigma_angstrom) * ANGSTROM_TO_BOHR


def noisy_snapshots(count, first=0, **kw):
    return [noisy_snapshot(first + j, **kw) for j in range(count)]


def perturbed(z, x_bohr, seed, sigma_angstrom=0.05):
    """Gaussian displacement of every coordinate (snapshots for the learned guess)."""
    rng = Xoshiro256(seed)
    noise = np.array([rng.normal() for _ in range(x_bohr.size)]).reshape(x_bohr.shape)
    return z, x_bohr + noise * sigma_angstrom * ANGSTROM_TO_BOHR


def basis_size(numbers):
    """(shells, AOs) of the GFN1 basis: H 2 shells / 2 AO, C N O 2 shells / 4 AO."""
    nsh = 0
    nao = 0
    for z in numbers:
        if z == 1:
            nsh += 2
            nao += 2
        elif z in (6, 7, 8):
            nsh += 2
            nao += 4
        else:
            raise ValueError("basis_size covers H, C, N, O only")
    return nsh, nao
