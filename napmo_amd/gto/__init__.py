from napmo_amd.gto.basis_set import BasisSet, cartesian_powers  # noqa: F401
from napmo_amd.gto.contracted_gaussian import ContractedGaussian  # noqa: F401
from napmo_amd.gto.primitive_gaussian import PrimitiveGaussian  # noqa: F401
