"""Representations and wavefunctions (the ``core.*`` names config files use).

Mirrors the reference's Boost.Python classes that the tensor path touches:
CombinedRepresentation_R (core/representation/combinedrepresentation.cpp:186-474),
BSplineRepresentation (grid = B-spline index, core/representation/bspline/
bsplinerepresentation.cpp:31-35), ReducedSphericalHarmonicRepresentation
(grid = l, core/representation/reducedspherical/lrange.h:26-40) and
Wavefunction (core/wavefunction.cpp:194-207).  The wavefunction data is a
complex128 torch tensor on the GPU; overlap operations and inner products run
through the HIP library.
"""
import sys
import types

import numpy as np

from . import _native as nat
from . import bspline as _bs
from . import device as dev
from .sphharm import ReducedSphericalTools


class RepresentationBase:
    def __init__(self):
        self.BaseRank = 0

    def SetBaseRank(self, r):
        self.BaseRank = r

    def GetBaseRank(self):
        return self.BaseRank

    def SetDistributedModel(self, d):
        self.DistributedModel = d

    def GetDistributedModel(self):
        return getattr(self, "DistributedModel", None)


class BSplineRepresentation(RepresentationBase):
    def ApplyConfigSection(self, conf):
        init = getattr(conf, "init_function", _bs.InitBSpline)
        self.SetupRepresentation(init(conf))

    def SetupRepresentation(self, bsplineObject):
        self.BSplineObject = bsplineObject
        self._overlap = None

    def GetBSplineObject(self):
        return self.BSplineObject

    def GetFullShape(self):
        return [self.BSplineObject.NumberOfBSplines]

    def GetGlobalGrid(self, rank=0):
        return np.arange(self.BSplineObject.NumberOfBSplines, dtype=np.float64)

    GetLocalGrid = GetGlobalGrid

    def IsOrthogonalBasis(self):
        return False

    def GetOverlap(self):
        """Device Cholesky factor S = U^H U (OverlapMatrix, overlapmatrix.cpp:9-104)."""
        if self._overlap is None:
            b = self.BSplineObject
            self._overlap = dev.Overlap(b.overlap_upper(), b.NumberOfBSplines, b.MaxSplineOrder - 1)
        return self._overlap


class ReducedSphericalHarmonicRepresentation(RepresentationBase):
    def ApplyConfigSection(self, conf):
        self.SetupRepresentation(conf.maxl, getattr(conf, "m", 0))

    def SetupRepresentation(self, lmax, m=0):
        self.LMax, self.M = int(lmax), int(m)
        self.Tools = ReducedSphericalTools(self.LMax, self.M)

    def GetFullShape(self):
        return [self.LMax + 1 - abs(self.M)]

    def GetGlobalGrid(self, rank=0):
        return np.arange(abs(self.M), self.LMax + 1, dtype=np.float64)

    GetLocalGrid = GetGlobalGrid

    def IsOrthogonalBasis(self):
        return True


class CoupledSphericalHarmonicRepresentation(RepresentationBase):
    """Two-electron angular basis |(l1 l2) L M> (core/representation/
    coupledspherical/coupledsphericalharmonicrepresentation.h); the grid is
    the channel index, channels in the order of ``index_iterator``
    (coupledrange.h:84-140)."""

    def ApplyConfigSection(self, conf):
        self.SetupRepresentation(conf.index_iterator)

    def SetupRepresentation(self, indexIterator):
        from .coupled import CoupledIndex
        self.Channels = [CoupledIndex(*tuple(c)) for c in indexIterator]
        if len(set(self.Channels)) != len(self.Channels):
            raise ValueError("index_iterator yields a channel twice")

    def GetFullShape(self):
        return [len(self.Channels)]

    def GetGlobalGrid(self, rank=0):
        return np.arange(len(self.Channels), dtype=np.float64)

    GetLocalGrid = GetGlobalGrid

    def IsOrthogonalBasis(self):
        return True


class CombinedRepresentation(RepresentationBase):
    Rank = None

    def __init__(self):
        super().__init__()
        self.Representations = [None] * self.Rank

    def ApplyConfigSection(self, conf):
        pass

    def SetRepresentation(self, i, repr):
        self.Representations[i] = repr

    def GetRepresentation(self, i):
        return self.Representations[i]

    def RadialMatrices(self):
        """Row-major band matrices (S, D1, D2, 1/r, 1/r^2) of the B-spline
        basis shared by the radial ranks (rank-3 two-electron states)."""
        if getattr(self, "_radial", None) is None:
            from .helium import radial_matrices
            self._radial = radial_matrices(self.Representations[1].GetBSplineObject())
        return self._radial

    def GetFullShape(self):
        return [r.GetFullShape()[0] for r in self.Representations]

    GetInitialShape = GetFullShape

    def GetLocalGrid(self, i):
        return self.Representations[i].GetLocalGrid(i)

    GetGlobalGrid = GetLocalGrid

    def IsOrthogonalBasis(self, i):
        return self.Representations[i].IsOrthogonalBasis()

    def _ranks(self):
        return [i for i in range(self.Rank) if not self.IsOrthogonalBasis(i)]

    def _apply(self, op, psi):
        data = psi.GetData()
        for i in self._ranks():
            self.Representations[i].GetOverlap().apply(op, i, data)

    # combinedrepresentation.cpp:186-238
    def SolveSqrtOverlap(self, conj, psi):
        self._apply(nat.OV_SOLVE_SQRT_CONJ if conj else nat.OV_SOLVE_SQRT, psi)

    def MultiplySqrtOverlap(self, conj, psi):
        self._apply(nat.OV_MUL_SQRT_CONJ if conj else nat.OV_MUL_SQRT, psi)

    def SolveOverlap(self, psi):
        self._apply(nat.OV_SOLVE, psi)

    def MultiplyOverlap(self, psi):
        self._apply(nat.OV_MUL, psi)

    def InnerProduct(self, psi1, psi2):
        """<psi1|S|psi2> (Algo 4, combinedrepresentation.cpp:423-474): S psi2
        by the overlap kernels along every non-orthogonal rank, then the
        device zdotc (pp_inner_product)."""
        a = psi1.GetData()
        tmp = psi2.GetData().clone()
        if tuple(a.shape) != tuple(tmp.shape):
            raise ValueError("wavefunctions of different shapes")
        for i in self._ranks():
            self.Representations[i].GetOverlap().apply(nat.OV_MUL, i, tmp)
        out = np.zeros(2)
        nat.check(nat.lib().pp_inner_product(dev.ptr(a), dev.ptr(tmp), int(tmp.numel()), nat.dptr(out),
                                             dev.stream_ptr()), "pp_inner_product")
        return complex(out[0], out[1])


def _combined(rank):
    return type("CombinedRepresentation_%d" % rank, (CombinedRepresentation,), {"Rank": rank})


CombinedRepresentation_1 = _combined(1)
CombinedRepresentation_2 = _combined(2)
CombinedRepresentation_3 = _combined(3)


class Wavefunction:
    """Rank-R complex128 amplitudes on the GPU with a CombinedRepresentation."""

    def __init__(self, representation=None, data=None):
        self.Representation = representation
        self.Data = data

    def SetRepresentation(self, r):
        self.Representation = r

    def GetRepresentation(self):
        return self.Representation

    def AllocateData(self):
        t = dev.torch()
        shape = self.Representation.GetFullShape()
        if int(np.prod(shape)) * 16 > 288 * 2 ** 30:
            raise RuntimeError("wavefunction does not fit one MI355X")
        self.Data = t.zeros(shape, dtype=t.complex128, device=dev.device())

    def GetData(self):
        return self.Data

    def SetData(self, d):
        self.Data = d

    def GetRank(self):
        return len(self.Representation.GetFullShape())

    def Copy(self):
        return Wavefunction(self.Representation, self.Data.clone())

    CopyDeep = Copy

    def Clear(self):
        self.Data.zero_()

    def InnerProduct(self, other):
        return self.Representation.InnerProduct(self, other)

    def GetNorm(self):
        """sqrt(<psi|S|psi>) (wavefunction.cpp:194-198)."""
        return float(np.sqrt(self.InnerProduct(self).real))

    def Normalize(self):
        n = self.GetNorm()
        self.Data /= n
        return n


class _CoreNamespace(types.ModuleType):
    """``core.X`` lookup for config files; TensorPotentialMultiply_<ids>
    resolves to the HIP tensor multiply with those storage ids."""

    def __getattr__(self, name):
        prefix = "TensorPotentialMultiply_"
        if name.startswith(prefix):
            ids = name[len(prefix):].split("_")
            for s in ids:
                if s not in nat.STORAGE_IDS:
                    raise AttributeError(name)

            def multiply(potential, scaling, source, dest, *args):
                pairs = []
                it = iter(args)
                for s in ids:
                    pairs.append(next(it) if s in ("Simp", "Herm") else None)
                return dev.tp_multiply(ids, potential, scaling, source, dest, pairs)
            multiply.__name__ = name
            return multiply
        raise AttributeError(name)


core = _CoreNamespace("core")
for _name in ["BSplineRepresentation", "ReducedSphericalHarmonicRepresentation",
              "CoupledSphericalHarmonicRepresentation", "CombinedRepresentation_1",
              "CombinedRepresentation_2", "CombinedRepresentation_3", "Wavefunction"]:
    setattr(core, _name, globals()[_name])
core.BSpline = _bs.BSplineBasis
sys.modules.setdefault("pyprop_amd.core_namespace", core)
