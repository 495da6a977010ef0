"""dakc_amd -- DAKC's k-mer counting hot path on MI355X (gfx950).

Host code talks to libdakc_amd.so through the C ABI in include/dakc.h; this
package holds the HIP sources (csrc/), the in-tree build (build.py), the C++
CLI (cli/) and a thin Python face (kcounter.py) used by tests and benchmarks.
"""
from .kcounter import FqReader, KmerCounter, KmerTable, gpu_available  # noqa: F401
from ._lib import DakcError, gen_params, generate_reads  # noqa: F401

__all__ = ["FqReader", "KmerCounter", "KmerTable", "DakcError", "gen_params", "generate_reads",
           "gpu_available"]
