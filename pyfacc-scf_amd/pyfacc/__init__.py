"""MI355X-native Hartree-Fock SCF engine (drop-in for PyFAcc-SCF's driver API).

``pyfacc.helper_functions`` mirrors the reference's Python driver
(``src/python/helper_functions.py``); ``load_library()`` returns the ctypes
handle of ``libpyfacc.so`` which exports the reference's ``run_scf_c`` /
``run_scf_interface_py`` plus the ``pf_*`` entry points of
``include/pyfacc.h``.  There is no CPU fallback: every function here runs
the HIP kernels and raises if the library or a GPU is missing.
"""
from . import helper_functions  # noqa: F401
from ._native import (  # noqa: F401
    LIB_PATH,
    Options,
    Result,
    PyfaccError,
    basis_info,
    batch_scf,
    boys,
    cached_device_bytes,
    default_options,
    device_count,
    eigh,
    eri_packed,
    eri_quartets,
    gemm,
    jk,
    load_library,
    one_electron,
    projector,
    release_device_cache,
    scf,
    scf_profile,
    schwarz,
)
from .helper_functions import compute_spin_electrons, guess_multiplicity, read_xyz, run_scf, symbols_to_z  # noqa: F401
