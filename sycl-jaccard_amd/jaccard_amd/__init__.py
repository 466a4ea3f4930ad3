"""MI355X (gfx950) edge-connected Jaccard similarity.

* ``graphs``  host CSR container and deterministic generators (RMAT, Chung-Lu)
* ``api``     the reference's ``sygraph::jaccard`` / ``jaccard_list`` interface
              and kernel-level entry points over ``libjaccard_amd.so``
"""
from ._native import NativeLibraryMissing  # noqa: F401

__all__ = ["NativeLibraryMissing"]
