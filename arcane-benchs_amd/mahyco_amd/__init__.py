"""Host-side Python harness of the MI355X MaHyCo time step.

The product is the C-ABI library built from ``csrc/`` (libmahyco_gpu.so);
this package only binds it with ctypes (``abi``), builds the benchmark
workloads (``cases``) and evaluates reference solutions used by the tests
(``riemann``).  ``load_gpu()`` raises if the library is missing: there is no
CPU fallback.
"""
import os

from . import abi  # noqa: F401

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
GPU_LIB = os.path.join(ROOT_DIR, "csrc", "libmahyco_gpu.so")


def load_gpu(path=None):
    """Bind the HIP library.  Fails loudly when it has not been built."""
    p = path or GPU_LIB
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    return abi.Library(p, "mhg_")
