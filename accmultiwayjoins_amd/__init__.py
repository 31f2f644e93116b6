"""MI355X-native multi-way joins (AMHJ hash join and ALFTJ leapfrog triejoin).

The compute path is libamj.so (hand-written gfx950 HIP kernels behind a C ABI, see
include/amj.h); this package is its host-side interface.
"""
from .amj import AmjError, Context, HashTable, JoinResult, Relation, Trie  # noqa: F401
from . import formats, generators, serving  # noqa: F401

__all__ = ["AmjError", "Context", "HashTable", "JoinResult", "Relation", "Trie", "formats", "generators", "serving"]
