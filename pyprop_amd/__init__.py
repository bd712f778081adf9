"""pyprop_amd: MI355X-native tensor-potential Krylov propagation behind
pyprop's config.ini / TensorPotential / Krylov-propagator interface.

The names a pyprop config file evaluates (``core.CombinedRepresentation_2``,
``InitBSpline``, ``PamPropagator``, ``ExpokitPropagator``, ``CayleyPropagator``, ``BasisPropagator``,
``InitialConditionType``, ``PotentialType``, numpy's ``pi`` ...) are exported
here, so ``pyprop_amd.Load(file)`` reads the reference's files unchanged.
"""
from numpy import *  # noqa: F401,F403  (config expressions use pi, sqrt, ...; pyprop/__init__.py:16)

from .config import Config, Load, LoadString, Section, ProjectNamespace  # noqa: F401
from .bspline import InitBSpline  # noqa: F401
from .core import core  # noqa: F401
from .potentials import PotentialType  # noqa: F401
from .problem import InitialConditionType, WavefunctionFileFormat, Problem  # noqa: F401
from .propagator import PropagatorBase, PamPropagator, ExpokitPropagator, CayleyPropagator  # noqa: F401
from .tensorpotential import BasisPropagator, TensorPotential  # noqa: F401
from .coupled import CoupledIndex, DefaultCoupledIndexIterator  # noqa: F401
from .solver import PiramSolver, GMRESShiftInvertSolver, BSplinePreconditioner, GmresSolver  # noqa: F401
from .helium import (RadialTwoElectronPreconditioner, RadialTwoElectronPreconditionerSuperLU,  # noqa: F401
                     RadialTwoElectronPreconditionerIfpack)
from .serialization import (SaveWavefunctionHDF, LoadWavefunctionHDF, LoadConfigFromFile,  # noqa: F401
                            CreateWavefunctionFromFile, SaveTensorPotential, LoadTensorPotential)

del abs, max, min, round, sum  # keep the builtins for code that star-imports this package
