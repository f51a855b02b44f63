"""b3-facing surface of aspartik_amd (mirrors aspartik.b3's calculator,
tree, substitution and likelihood objects that drive the pruning path)."""
from .batch import Batch
from .calculator import Calculator, HipCalculator, gamma_rates
from .clock import Clock, RelaxedClock, StrictClock
from .likelihood import DNALikelihood, GammaLikelihood, HeteroLikelihood, deduplicate
from .parameters import ClassVector, IntVector, Real, RealVector
from .mcmc import Chain, Surrogate, ess
from .substitution import GTR, GY94, HKY, JC, K80, SENSE_CODONS
from .tree import TOMBSTONE, DatedTree
from .tuning import autotune, autotune_batch

__all__ = [
    "Batch", "Calculator", "HipCalculator", "gamma_rates", "DNALikelihood", "GammaLikelihood",
    "HeteroLikelihood", "Clock", "StrictClock", "RelaxedClock", "ClassVector", "Real", "RealVector", "IntVector",
    "deduplicate", "GTR", "GY94", "HKY", "JC", "K80", "SENSE_CODONS", "TOMBSTONE",
    "DatedTree", "Chain", "Surrogate", "ess", "autotune", "autotune_batch",
]
