"""Tensor potentials: geometry infos, the generator, TensorPotential and
BasisPropagator.

Follows pyprop/tensorpotential/GeometryInfo.py:11-464,
basis/BSpline.py:6-234, basis/ReducedSphericalHarmonic.py:5-148,
TensorPotentialGenerator.py:72-256 (evaluate on the grid, project each rank in
reverse order), TensorPotential.py:3-163 and BasisPropagator.py:1-168
(generation, consolidation, balanced-overlap matvec).  Storage layout and basis
pair ordering are the reference's, so potential arrays are interchangeable.
Projection runs on the host (numpy) once; the multiplies run on the GPU.
"""
import numpy as np

from . import device as dev
from .potentials import CreatePotentialInstance
from .propagator import PropagatorBase


class UnsupportedGeometryException(Exception):
    pass


# ---------------------------------------------------------------- geometry infos
class GeometryInfoBase:
    def GetGlobalBasisPairs(self):
        return self.GetBasisPairs()

    def GetLocalBasisPairCount(self):
        return self.GetGlobalBasisPairCount()

    def HasParallelMultiply(self):
        return False

    def GetMultiplyArguments(self, psi):
        return []


class GeometryInfoCommonIdentity(GeometryInfoBase):
    def __init__(self, useGrid):
        self.UseGrid = useGrid

    def UseGridRepresentation(self):
        return self.UseGrid

    def GetGlobalBasisPairCount(self):
        return 1

    def GetBasisPairs(self):
        return np.zeros((1, 2), dtype=np.int32)

    def GetStorageId(self):
        return "Ident"


class GeometryInfoCommonDiagonal(GeometryInfoBase):
    def __init__(self, size, useGrid):
        self.Size = size
        self.UseGrid = useGrid

    def UseGridRepresentation(self):
        return self.UseGrid

    def GetGlobalBasisPairCount(self):
        return self.Size

    def GetBasisPairs(self):
        p = np.zeros((self.Size, 2), dtype=np.int32)
        p[:, 0] = p[:, 1] = np.arange(self.Size)
        return p

    def GetStorageId(self):
        return "Diag"


class GeometryInfoCommonDense(GeometryInfoBase):
    def __init__(self, rankCount, useGrid):
        self.RankCount = rankCount
        self.UseGrid = useGrid

    def UseGridRepresentation(self):
        return self.UseGrid

    def GetGlobalBasisPairCount(self):
        return self.RankCount ** 2

    def GetBasisPairs(self):
        i, j = np.meshgrid(np.arange(self.RankCount), np.arange(self.RankCount), indexing="ij")
        return np.stack([i.ravel(), j.ravel()], axis=1).astype(np.int32)

    def GetStorageId(self):
        return "Simp"

    def GetMultiplyArguments(self, psi):
        return [self.GetBasisPairs()]


class GeometryInfoCommonBandedNonHermitian(GeometryInfoBase):
    """General band storage (GeometryInfo.py:410-464): pair (i,j) at j*(2k+1) + k+i-j;
    unused slots carry (0, N-1)."""

    def __init__(self, rankCount, bandCount, useGrid):
        self.RankCount = rankCount
        self.BandCount = bandCount
        self.UseGrid = useGrid

    def UseGridRepresentation(self):
        return self.UseGrid

    def GetGlobalBasisPairCount(self):
        return self.RankCount * (2 * self.BandCount + 1)

    def GetBasisPairs(self):
        N, k = self.RankCount, self.BandCount
        pairs = np.zeros((self.GetGlobalBasisPairCount(), 2), dtype=np.int32)
        pairs[:, 1] = N - 1
        for i in range(N):
            for j in range(i - k, i + k + 1):
                if 0 <= j < N:
                    idx = j * (2 * k + 1) + k + i - j
                    pairs[idx] = (i, j)
        return pairs

    def GetStorageId(self):
        return "BandNH"


class GeometryInfoBSplineBanded(GeometryInfoBase):
    """'banded-packed': Simp storage of all pairs |i-j|<k (basis/BSpline.py:6-55)."""

    def __init__(self, bsplineObject):
        self.BSplineObject = bsplineObject

    def UseGridRepresentation(self):
        return True

    def GetGlobalBasisPairCount(self):
        N = self.BSplineObject.NumberOfBSplines
        k = self.BSplineObject.MaxSplineOrder
        return 2 * N * k - k * (k - 1) - N

    def GetBasisPairs(self):
        N = self.BSplineObject.NumberOfBSplines
        k = self.BSplineObject.MaxSplineOrder
        pairs = []
        for i in range(N):
            for j in range(max(0, i - k + 1), min(i + k, N)):
                pairs.append((i, j))
        return np.array(pairs, dtype=np.int32)

    def GetStorageId(self):
        return "Simp"

    def GetMultiplyArguments(self, psi):
        if not hasattr(self, "BasisPairs"):
            self.BasisPairs = self.GetBasisPairs()
        return [self.BasisPairs]


class GeometryInfoReducedSphHarmSelectionRule(GeometryInfoBase):
    """dipole l <-> l+-1 (basis/ReducedSphericalHarmonic.py:5-45)."""

    def __init__(self, tools, useGrid):
        self.Tools = tools
        self.UseGrid = useGrid

    def UseGridRepresentation(self):
        return self.UseGrid

    def GetGlobalBasisPairCount(self):
        return self.Tools.GetSize() * 2

    def GetBasisPairs(self):
        count = self.Tools.GetSize() + 1
        pairs = []
        for i in range(count):
            if i > 0:
                pairs.append((i, i - 1))
            if i < count - 1:
                pairs.append((i, i + 1))
        return np.array(pairs, dtype=np.int32).reshape(-1, 2)

    def GetStorageId(self):
        return "Simp"

    def GetMultiplyArguments(self, psi):
        if not hasattr(self, "BasisPairs"):
            self.BasisPairs = self.GetBasisPairs()
        return [self.BasisPairs]


# ---------------------------------------------------------------- basis functions
class BasisfunctionBSpline:
    def SetupBasis(self, repr):
        self.Representation = repr
        self.BSplineObject = repr.GetBSplineObject()

    def GetGridPoints(self):
        return self.BSplineObject.QuadratureGridGlobal

    def GetGeometryInfo(self, geometryName):
        geom = geometryName.lower().strip()
        N = self.BSplineObject.NumberOfBSplines
        band = self.BSplineObject.MaxSplineOrder - 1
        if geom == "identity":
            return GeometryInfoCommonIdentity(True)
        if geom == "banded-nonhermitian":
            return GeometryInfoCommonBandedNonHermitian(N, band, True)
        if geom == "banded-packed":
            return GeometryInfoBSplineBanded(self.BSplineObject)
        if geom in ("dense", "hermitian"):
            return GeometryInfoCommonDense(N, True)
        raise UnsupportedGeometryException("Geometry '%s' not supported by BasisfunctionBSpline" % geometryName)

    def RepresentPotentialInBasis(self, source, rank, geometryInfo, differentiation, device=False):
        """source: potential with the grid on axis `rank`; returns the array with
        that axis replaced by the basis pairs (basis_bspline.h:11-47);
        device: the projection runs on the GPU."""
        sid = geometryInfo.GetStorageId()
        src = np.moveaxis(source, rank, -1)
        if sid == "Ident":
            out = src[..., :1]
            return np.moveaxis(out, -1, rank)
        band = self.BSplineObject.project_bandnh(src, differentiation, device=device)
        if sid == "BandNH":
            out = band
        else:
            pairs = geometryInfo.GetBasisPairs()
            k = self.BSplineObject.MaxSplineOrder
            s = k - 1
            bw = 2 * s + 1
            i, j = pairs[:, 0], pairs[:, 1]
            inband = np.abs(i - j) <= s
            idx = np.where(inband, j * bw + s + i - j, 0)
            out = np.where(inband, band[..., idx], 0.0)
        return np.moveaxis(out, -1, rank)


class BasisfunctionReducedSphericalHarmonic:
    def SetupBasis(self, repr):
        self.Representation = repr
        self.Tools = repr.Tools
        self.LMax, self.M = repr.LMax, repr.M

    def GetGridPoints(self):
        return self.Tools.ThetaGrid

    def GetGeometryInfo(self, geometryName):
        geom = geometryName.lower().strip()
        useGrid = True
        if geom.startswith("custom-"):
            useGrid = False
            geom = geom[len("custom-"):]
        L = self.LMax + 1 - abs(self.M)
        if geom == "identity":
            return GeometryInfoCommonIdentity(useGrid)
        if geom == "diagonal":
            return GeometryInfoCommonDiagonal(L, False)
        if geom == "dense":
            return GeometryInfoCommonDense(L, useGrid)
        if geom == "dipoleselectionrule":
            return GeometryInfoReducedSphHarmSelectionRule(self.Tools, useGrid)
        raise UnsupportedGeometryException("Geometry '%s' not supported by BasisfunctionReducedSpherical" % geometryName)

    def RepresentPotentialInBasis(self, source, rank, geometryInfo, differentiation):
        sid = geometryInfo.GetStorageId()
        src = np.moveaxis(source, rank, -1)
        if sid == "Ident":
            return np.moveaxis(src[..., :1], -1, rank)
        pairs = geometryInfo.GetBasisPairs()
        out = self.Tools.project(src, pairs, differentiation)
        return np.moveaxis(out, -1, rank)


def CreateBasisFromRepresentation(repr):
    from .core import BSplineRepresentation, ReducedSphericalHarmonicRepresentation
    if isinstance(repr, BSplineRepresentation):
        b = BasisfunctionBSpline()
    elif isinstance(repr, ReducedSphericalHarmonicRepresentation):
        b = BasisfunctionReducedSphericalHarmonic()
    else:
        raise NotImplementedError("Unknown representation %s" % repr)
    b.SetupBasis(repr)
    return b


# ---------------------------------------------------------------- generator
class TensorPotentialGenerator:
    def __init__(self, representation):
        self.Rank = len(representation.GetFullShape())
        self.BasisList = [CreateBasisFromRepresentation(representation.GetRepresentation(i))
                          for i in range(self.Rank)]

    def GetGeometryList(self, configSection):
        return [b.GetGeometryInfo(configSection.Get("geometry%i" % i)) for i, b in enumerate(self.BasisList)]

    def GeneratePotential(self, configSection):
        """Evaluate the potential on the grid, then project rank by rank in
        reverse order (TensorPotentialGenerator.py:72-256)."""
        geometryList = self.GetGeometryList(configSection)
        grids = []
        for i, (b, g) in enumerate(zip(self.BasisList, geometryList)):
            if g.UseGridRepresentation():
                grids.append(b.GetGridPoints())
            else:
                grids.append(b.Representation.GetGlobalGrid(i))
        pos = np.meshgrid(*grids, indexing="ij", sparse=True)
        evaluator = CreatePotentialInstance(configSection.classname)
        configSection.Apply(evaluator)
        data = np.array(evaluator.GetPotentialValue(tuple(pos)), dtype=np.complex128)
        data = np.broadcast_to(data, tuple(len(g) for g in grids)).copy()
        if getattr(configSection, "debug_potential", False):
            self.OriginalPotential = data.copy()
        # extension key (no reference counterpart): generate_on_device = True
        # projects the B-spline ranks on the GPU (pp_projector_apply)
        device = bool(getattr(configSection, "generate_on_device", False))
        for rank in reversed(range(self.Rank)):
            g = geometryList[rank]
            if g.UseGridRepresentation():
                d = getattr(configSection, "differentiation%i" % rank, 0)
                basis = self.BasisList[rank]
                if device and isinstance(basis, BasisfunctionBSpline):
                    data = basis.RepresentPotentialInBasis(data, rank, g, d, device=True)
                else:
                    data = basis.RepresentPotentialInBasis(data, rank, g, d)
        return np.ascontiguousarray(data, dtype=np.complex128)


# ---------------------------------------------------------------- TensorPotential
class TensorPotential:
    """PotentialWrapper for a tensor potential (TensorPotential.py:3-163)."""

    def __init__(self, psi):
        self.psi = psi
        self.Rank = psi.GetRank()
        self.GeometryList = None
        self.PotentialData = None
        self.Name = None
        self.OriginalPotential = None

    def ApplyConfigSection(self, configSection):
        self.DebugPotential = getattr(configSection, "debug_potential", False)
        self.IsTimeDependent = False
        if hasattr(configSection, "time_function"):
            self.IsTimeDependent = True
            self.TimeFunction = lambda t: configSection.time_function(configSection, t)
            self.OriginalTimeFunction = configSection.time_function
        generator = TensorPotentialGenerator(self.psi.GetRepresentation())
        self.GeometryList = generator.GetGeometryList(configSection)
        self.Name = configSection.name
        if hasattr(configSection, "filename"):
            # TensorPotential.py:56-71: load the data instead of generating it
            from .serialization import LoadTensorPotential
            shape = tuple(len(g.GetBasisPairs()) for g in self.GeometryList)
            self.PotentialData = np.zeros(shape, dtype=np.complex128)
            LoadTensorPotential(configSection.filename, configSection.dataset, self)
        else:
            self.PotentialData = generator.GeneratePotential(configSection)
            self.OriginalPotential = getattr(generator, "OriginalPotential", None)
        self._device = None

    def StorageIds(self):
        return [g.GetStorageId() for g in self.GeometryList]

    def SetupStep(self, timestep):
        self.BasisPairs = [g.GetBasisPairs() for g in self.GeometryList]
        from .core import core
        name = "core.TensorPotentialMultiply_" + "_".join(self.StorageIds())
        try:
            self.MultiplyFunction = eval(name, {"core": core})
        except AttributeError:
            print("ERROR: Could not find multiplyfunction for potential: %s" % name)
            raise

    def DeviceData(self):
        if self._device is None:
            self._device = dev.to_device(self.PotentialData)
        return self._device

    def GetTimeScaling(self, t):
        return float(np.real(self.TimeFunction(t))) if self.IsTimeDependent else 1.0

    def MultiplyPotential(self, srcPsi, destPsi, t, timestep):
        if self.PotentialData.size == 0:   # no pairs (the dipole term at lmax = 0): adds nothing
            return
        argList = [self.DeviceData(), self.GetTimeScaling(t), srcPsi.GetData(), destPsi.GetData()]
        for g in self.GeometryList:
            argList += g.GetMultiplyArguments(srcPsi)
        self.MultiplyFunction(*argList)

    def GetExpectationValue(self, psi, tmpPsi, t, timeStep):
        tmpPsi.Clear()
        self.MultiplyPotential(psi, tmpPsi, t, timeStep)
        psi.GetRepresentation().SolveOverlap(tmpPsi)
        return psi.InnerProduct(tmpPsi)

    def CanConsolidate(self, otherPot):
        if self.IsTimeDependent != otherPot.IsTimeDependent:
            return False
        if self.IsTimeDependent and otherPot.OriginalTimeFunction != self.OriginalTimeFunction:
            return False
        if self.DebugPotential or otherPot.DebugPotential:
            return False
        for r in range(self.Rank):
            if self.GeometryList[r].GetStorageId() != otherPot.GeometryList[r].GetStorageId():
                return False
        for r in range(self.Rank):
            a = self.GeometryList[r].GetBasisPairs()
            b = otherPot.GeometryList[r].GetBasisPairs()
            if a.shape != b.shape or np.any(a != b):
                return False
        return True


def _consolidate(existingPot, pot):
    """existingPot += pot (BasisPropagator.py:95-103); factored potentials
    merge their term lists."""
    if hasattr(existingPot, "ConsolidateWith"):
        existingPot.ConsolidateWith(pot)
    else:
        existingPot.PotentialData[:] += pot.PotentialData
        existingPot.Name += "+" + pot.Name


# ---------------------------------------------------------------- BasisPropagator
class BasisPropagator(PropagatorBase):
    """Applies tensor potentials in the basis (BasisPropagator.py:1-168)."""

    def __init__(self, psi):
        PropagatorBase.__init__(self, psi)
        self.Rank = psi.GetRank()
        self.TempPsi2 = None

    def ApplyConfig(self, config):
        PropagatorBase.ApplyConfig(self, config)
        self.GeneratePotentials(config)
        self.ConsolidatePotentials()

    def GeneratePotential(self, configSection):
        from .core import CoupledSphericalHarmonicRepresentation
        repr = self.psi.GetRepresentation()
        if self.Rank == 3 and isinstance(repr.GetRepresentation(0), CoupledSphericalHarmonicRepresentation):
            # two-electron potentials are generated in factored form
            from .helium import CoupledTensorPotential
            potential = CoupledTensorPotential(self.psi)
        else:
            potential = TensorPotential(self.psi)
        configSection.Apply(potential)
        return potential

    def GeneratePotentials(self, config):
        if hasattr(config.Propagation, "grid_potential_list"):
            for potentialName in config.Propagation.grid_potential_list:
                configSection = config.GetSection(potentialName)
                pot = self.GeneratePotential(configSection)
                for existingPot in self.PotentialList:
                    if existingPot.CanConsolidate(pot):
                        _consolidate(existingPot, pot)
                        pot = None
                        break
                if pot is not None:
                    self.PotentialList.append(pot)

    def ConsolidatePotentials(self):
        potentials = list(self.PotentialList)
        remove = []
        i = 0
        while i < len(potentials) - 1:
            cur = potentials[i]
            for other in list(potentials[i + 1:]):
                if cur.CanConsolidate(other):
                    _consolidate(cur, other)
                    potentials.remove(other)
                    remove.append(other)
            i += 1
        for p in remove:
            self.PotentialList.remove(p)

    def ApplyConfigSection(self, configSection):
        PropagatorBase.ApplyConfigSection(self, configSection)

    def SetupStep(self, dt):
        PropagatorBase.SetupStep(self, dt)

    def MultiplyHamiltonian(self, srcPsi, destPsi, t, dt):
        destPsi.GetData().zero_()
        self.MultiplyPotential(srcPsi, destPsi, t, dt)
        repr = srcPsi.GetRepresentation()
        if not all(repr.IsOrthogonalBasis(i) for i in range(self.Rank)):
            repr.SolveOverlap(destPsi)

    def MultiplyHamiltonianNoOverlap(self, srcPsi, destPsi, t, dt):
        destPsi.GetData().zero_()
        self.MultiplyPotential(srcPsi, destPsi, t, dt)

    def MultiplyHamiltonianBalancedOverlap(self, srcPsi, destPsi, t, dt):
        """U^-H H U^-1 with src restored afterwards (BasisPropagator.py:131-159)."""
        if self.TempPsi2 is None:
            self.TempPsi2 = srcPsi.Copy()
        self.TempPsi2.GetData().copy_(srcPsi.GetData())
        repr = srcPsi.GetRepresentation()
        repr.SolveSqrtOverlap(False, srcPsi)
        destPsi.GetData().zero_()
        self.MultiplyPotential(srcPsi, destPsi, t, dt)
        destPsi.GetRepresentation().SolveSqrtOverlap(True, destPsi)
        srcPsi.GetData().copy_(self.TempPsi2.GetData())

    def AdvanceStep(self, t, dt):
        raise NotImplementedError("BasisPropagator does not support AdvanceStep. Use it as a base for explicit propagators")
