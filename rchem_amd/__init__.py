"""rchem_amd: the restricted-Hartree-Fock hot path of berquist/rchem on MI355X.

The compute path is ``librchem_gpu.so`` (HIP, gfx950) behind the C ABI in
``include/rchem_gpu.h``; this package is its Python host side, shaped like
rchem's own API (see ``rchem_amd.basis``), plus input generators.
"""
from ._lib import (BOYS_COMPAT, BOYS_EXACT, PREC_FP64, PREC_MIXED, RchemError, default_opts,
                   lib)
from .basis import (Basis, Context, JK_direct, JK_inmem, S, T, V, build_I, jk_last_stats,
                    one_electron, schwarz, scf)

__all__ = ["lib", "default_opts", "RchemError", "BOYS_EXACT", "BOYS_COMPAT", "PREC_FP64",
           "PREC_MIXED", "Basis", "Context", "JK_direct", "JK_inmem", "S", "T", "V", "build_I",
           "scf", "one_electron", "schwarz", "jk_last_stats"]
