"""Drop-in mirror of the reference Python driver API.

Same names, argument meaning and error behaviour as
``src/python/helper_functions.py`` of PyFAcc-SCF:

* ``symbols_to_z``            -- helper_functions.py:21-22
* ``read_xyz``                -- helper_functions.py:25-54 (Angstrom -> bohr, coords (3, natoms))
* ``guess_multiplicity``      -- helper_functions.py:63-74
* ``compute_spin_electrons``  -- helper_functions.py:76-87
* ``run_scf``                 -- helper_functions.py:90-149 (packs ctypes args, calls ``run_scf_c``)

``run_scf`` takes the same ``lib`` argument as the reference: any
``ctypes.CDLL`` that exports ``run_scf_c``.  ``pyfacc.load_library()``
returns the MI355X library; the reference's ``libscf.so`` would work too.

Paths: the reference resolves ``./data/molecules/<name>`` relative to the
working directory.  That is kept as the first choice; an absolute path, a
path that exists as given, ``$PYFACC_DATA/molecules/<name>`` and the data
directory shipped inside this package are tried afterwards.
"""
import ctypes
import os

import numpy as np

angtobohr = 1.8897259886  # helper_functions.py:7

# Symbol -> Z for H..Xe (helper_functions.py:10-19).
_SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni "
    "Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe"
).split()
periodic_table = {s: i + 1 for i, s in enumerate(_SYMBOLS)}

# Special-cased homonuclear molecules (helper_functions.py:58-61).
known_multiplicities = {("O", 2): 3, ("N", 2): 4}

_PKG_DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data")


def data_dir():
    """Directory holding ``basis_set/`` and ``molecules/`` shipped with the package."""
    return os.environ.get("PYFACC_DATA", _PKG_DATA)


def resolve_data_file(kind, name):
    """Find ``name`` under ``./data/<kind>/`` (reference behaviour), else fall back.

    Order: ``./data/<kind>/<name>`` (CWD-relative, as the reference does),
    ``name`` itself if it exists, ``$PYFACC_DATA/<kind>/<name>``, then the
    package data directory.  Raises ``FileNotFoundError`` naming every
    candidate when none exists.
    """
    cands = [os.path.join(".", "data", kind, name), name,
             os.path.join(data_dir(), kind, name), os.path.join(_PKG_DATA, kind, name)]
    for c in cands:
        if os.path.isfile(c):
            return c
    raise FileNotFoundError(f"{kind} file {name!r} not found; tried {cands}")


def symbols_to_z(species):
    """Atomic numbers of element symbols; ``KeyError`` on unknown symbols."""
    return [periodic_table[s] for s in species]


def read_xyz(filename):
    """Read an XYZ file.

    Returns ``(natoms, coords, species)`` with ``coords`` of shape
    ``(3, natoms)`` in bohr (Fortran-ordered in memory, so the raw buffer is
    x1 y1 z1 x2 ... as the C-ABI expects) and ``species`` a numpy str array.
    A data line without exactly four fields raises ``ValueError``.
    """
    path = resolve_data_file("molecules", filename)
    with open(path, "r") as fh:
        lines = fh.readlines()
    natoms = int(lines[0].strip())
    species, coords = [], []
    for line in lines[2:2 + natoms]:
        parts = line.strip().split()
        if len(parts) != 4:
            raise ValueError(f"Invalid line in XYZ file: {line}")
        species.append(parts[0])
        coords.append([float(parts[1]), float(parts[2]), float(parts[3])])
    xyz = np.array(coords, dtype=np.float64).T * angtobohr
    return natoms, xyz, np.array(species, dtype=str)


def guess_multiplicity(atoms, charge=0):
    """1 for an even electron count, else 2; O2 -> 3 and N2 -> 4."""
    total = sum(periodic_table[a] for a in atoms) - charge
    counts = {}
    for a in atoms:
        counts[a] = counts.get(a, 0) + 1
    for (sym, cnt), mult in known_multiplicities.items():
        if counts.get(sym, 0) == cnt and len(counts) == 1:
            return mult
    return 1 if total % 2 == 0 else 2


def compute_spin_electrons(atoms, charge=0):
    """Return ``(n_alpha, n_beta, multiplicity)``; ``ValueError`` if inconsistent."""
    total = sum(periodic_table[a] for a in atoms) - charge
    mult = guess_multiplicity(atoms, charge)
    two_s = mult - 1
    if (total + two_s) % 2 != 0:
        raise ValueError("Inconsistent electron count and spin state.")
    n_alpha = (total + two_s) // 2
    return n_alpha, total - n_alpha, mult


def pack_species(species):
    """Species as ``2*natoms`` space-padded bytes, no NUL (helper_functions.py:105-107)."""
    flat = b"".join(s.ljust(2).encode("utf-8") for s in species)
    return (ctypes.c_char * len(flat))(*flat)


def run_scf(lib, atom_file, eps_scf, max_iter, final_energy, basis_set, charge=0):
    """Run an SCF through ``lib.run_scf_c`` and return the total energy (Eh).

    Same contract as helper_functions.py:90-149: ``final_energy`` is ignored
    and a fresh ``c_double`` is passed; ``run_scf_c`` is called without
    ``argtypes`` (ctypes default conversions).  RHF when n_alpha == n_beta,
    UHF otherwise.  A NaN result means the library reported an error; the
    MI355X library's message is then raised as ``RuntimeError``.
    """
    natoms, coords, species = read_xyz(atom_file)
    nalpha, nbeta, _ = compute_spin_electrons(species, charge)
    nelectrons = nalpha + nbeta
    z = np.array(symbols_to_z(species), dtype=np.int32)
    sp = pack_species(species)
    xyz = np.array(coords, dtype=np.float64)  # keeps the (3, natoms) F-order buffer
    out = ctypes.c_double(0.0)
    lib.run_scf_c(ctypes.c_int(natoms), ctypes.c_int(nelectrons), ctypes.c_int(charge),
                  ctypes.c_int(nalpha), ctypes.c_int(nbeta), sp,
                  xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                  ctypes.c_double(eps_scf), ctypes.c_int(max_iter), ctypes.byref(out),
                  basis_set.encode("utf-8"), z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    val = out.value
    if val != val and hasattr(lib, "pf_last_error"):
        lib.pf_last_error.restype = ctypes.c_char_p
        raise RuntimeError((lib.pf_last_error() or b"unknown error").decode())
    return val
