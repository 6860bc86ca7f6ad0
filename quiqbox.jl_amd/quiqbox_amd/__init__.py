"""quiqbox_amd -- MI355X-native Gaussian-integral engine and Hartree-Fock Fock
build with Quiqbox.jl's call surface.

All integrals and Fock builds run in libqbx_hip.so (HIP, gfx950) through the
C-ABI of include/qbx.h; the host keeps basis construction and the SCF
control flow.
"""
from .basis import (AtomicGTOrbSetNames, CompositeOrb, FlatBasis, PrimitiveOrb, flatten_basis,
                    genGaussTypeOrb, genGaussTypeOrbSeq, get3DimPGTOrbNormFactor, splitOrb, subshellXYZs)
from .engine import Context, GPUBasis, class_table, default_context
from .hartree_fock import (DenseERI, DirectJK, HFconfig, HFfinalInfo, OrbCoeffInitialConfig, RCHartreeFock,
                           SCFconfig, UOHartreeFock, getEhf, getEhfMO, getOrthonormalization, runHartreeFock)
from .integrals import (changeOrbitalBasis, coreHamiltonian, elecKinetic, elecKinetics, elecRepulsion,
                        elecRepulsions, multipoleMoment, multipoleMoments, nucAttraction, nucAttractions, overlap,
                        overlaps)
from .particles import NuclearCluster, OccupationState, getCharge, nucRepulsion, prepareSpinConfiguration
from ._lib import QBX_PREC_FP64, QBX_PREC_MIXED, LIB_PATH

__all__ = [n for n in dir() if not n.startswith("_")]
