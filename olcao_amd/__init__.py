"""MI355X-native implementation of OLCAO's integral -> H(k)/S(k) ->
eigenvalue path (two- and three-centre Gaussian integrals with lattice sums,
contraction and Bloch sums, batched generalised eigensolve) behind a C-ABI
(include/olcao_gpu.h).  `olcao_amd.gpu` binds it with the reference's
procedure names; `olcao_amd.system` and `olcao_amd.configs` describe and
generate the input systems."""
from . import configs, system  # noqa: F401

__all__ = ["configs", "system", "gpu"]
