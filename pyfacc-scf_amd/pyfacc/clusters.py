"""Seeded water-cluster generator for the benchmark and test geometries.

The reference ships no generator (it has only hand-made ``data/molecules``
files); these geometries are new inputs for the MI355X engine, defined here:

* monomer (Angstrom): O (0, 0, 0), H (+-0.757, 0.586, 0)
* OH radical (Angstrom): O (0, 0, 0), H (0.97, 0, 0)
* cluster: one molecule per cubic grid site ``a*(i, j, k)`` with a = 3.0 A,
  sites visited with i slowest and k fastest.  Each molecule gets a Haar-uniform
  rotation about its O atom (normalised Gaussian quaternion, four normal
  draws) and then a centre jitter drawn from U[-0.2, 0.2]^3 A, in that order,
  from ``numpy.random.Generator(PCG64(seed))``.
* OH substitution replaces molecule 0 by OH whose H lies along that
  molecule's rotated +x axis.
* batched instance ``idx``: seed 4096 + idx; k ~ U{1..12}; k sites drawn
  without replacement from a 3x2x2 grid; rotation and jitter per molecule as
  above; then OH substitution of molecule 0 with probability 0.1.
"""
import numpy as np

MONOMER = np.array([[0.0, 0.0, 0.0], [0.757, 0.586, 0.0], [-0.757, 0.586, 0.0]])
OH_RADICAL = np.array([[0.0, 0.0, 0.0], [0.97, 0.0, 0.0]])
SPACING = 3.0
JITTER = 0.2


def quat_to_matrix(q):
    """Rotation matrix of a unit quaternion (w, x, y, z)."""
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def _place(rng, site, oh):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    rot = quat_to_matrix(q)
    shift = rng.uniform(-JITTER, JITTER, size=3)
    ref = OH_RADICAL if oh else MONOMER
    xyz = ref @ rot.T + site + shift
    return (["O", "H"] if oh else ["O", "H", "H"]), xyz


def water_cluster(grid, seed, oh_first=False, spacing=SPACING):
    """Return ``(species, coords_angstrom)`` for a seeded cluster on ``grid`` = (nx, ny, nz)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = grid
    species, coords = [], []
    m = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                s, x = _place(rng, spacing * np.array([i, j, k], float), oh_first and m == 0)
                species += s
                coords.append(x)
                m += 1
    return species, np.vstack(coords)


def batched_instance(idx, spacing=SPACING):
    """Return ``(species, coords_angstrom, k, has_oh)`` for batched instance ``idx``."""
    rng = np.random.Generator(np.random.PCG64(4096 + idx))
    k = int(rng.integers(1, 13))
    grid = [np.array([i, j, l], float) for i in range(3) for j in range(2) for l in range(2)]
    sites = rng.choice(len(grid), size=k, replace=False)
    placed = [_place(rng, spacing * grid[s], False) for s in sites]
    has_oh = bool(rng.random() < 0.1)
    species, coords = [], []
    for m, (s, x) in enumerate(placed):
        if m == 0 and has_oh:
            # OH keeps molecule 0's O position and rotated +x axis.
            # The rotated +x axis is the direction from H2 to H1.
            o = x[0]
            ex = (x[1] - x[2]) / np.linalg.norm(x[1] - x[2])
            s, x = ["O", "H"], np.vstack([o, o + 0.97 * ex])
        species += s
        coords.append(x)
    return species, np.vstack(coords), k, has_oh


def xyz_text(species, coords, comment=""):
    lines = [str(len(species)), comment]
    lines += [f"{s:2s} {x: .10f} {y: .10f} {z: .10f}" for s, (x, y, z) in zip(species, coords)]
    return "\n".join(lines) + "\n"
