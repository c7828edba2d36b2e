"""fhelium_amd: an MI355X (gfx950) backend for fhelium's CKKS evaluation hot path.

Hand-written HIP kernels behind a C ABI (``include/fhelium_amd.h``), driven
from PyTorch-ROCm tensors.  The public surface mirrors the reference's
``CkksEngine`` for ``coefficient_domain_to_ntt_domain``, ``multiply``,
``relinearize``, ``rotate_by_step`` and ``rescale_to_next_level`` and registers
the reference's native operator schemas (``fhelium_rns_ops``,
``fhelium_ntt_ops``, ``fhelium_ckks_ops``) for ROCm tensors via
:mod:`fhelium_amd.native`.
"""

__version__ = "0.1.0"

from fhelium_amd.config import CkksConfig, Preset  # noqa: E402,F401


def __getattr__(name):
    if name == "CkksEngine":
        from fhelium_amd.engine import CkksEngine

        return CkksEngine
    raise AttributeError(name)
