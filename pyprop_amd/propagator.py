"""Propagators: PropagatorBase and the Krylov PamPropagator / ExpokitPropagator.

PropagatorBase follows pyprop/propagator/Propagator.py:5-80.  PamPropagator
keeps the reference's config keys and its balanced-overlap stepping
(pyprop/propagator/PamPropagator.py:2-83, core/krylov/pamp/pampwrapper.cpp:
51-108): psi~ = U psi, pAMP on U^-H H U^-1, psi = U^-1 psi~.  The Arnoldi loop,
the matvec callback and the Pade exponential all run inside one persistent HIP
launch; the time functions are sampled on the host at t_n for each step
(H held at the step start, Problem.py:166, TensorPotential.py:98-100).
"""
import warnings

import numpy as np

from . import _native as nat
from . import device as dev


class PropagatorBase:
    def __init__(self, psi):
        self.psi = psi
        self.PotentialList = []
        self.RenormalizeActive = False

    def ApplyConfig(self, config):
        from .problem import CreatePotential
        names = getattr(config.Propagation, "potential_evaluation", [])
        self.PotentialList = [CreatePotential(config, name, self.psi) for name in names]

    def ApplyConfigSection(self, configSection):
        self.RenormalizeActive = getattr(configSection, "renormalization", False)

    def SetupStep(self, dt):
        self.SetupPotential(dt)

    def RestartPropagation(self, timestep, startTime, propagationTime):
        pass

    def AdvanceStep(self, t, dt):
        raise NotImplementedError("AdvanceStep must be implemented by the propagator")

    def MultiplyHamiltonian(self, srcPsi, destPsi, t, dt):
        raise NotImplementedError("MultiplyHamiltonian should be implemented by inheriting class %s" % self.__class__)

    def SetupPotential(self, dt):
        for potential in self.PotentialList:
            potential.SetupStep(dt)

    def MultiplyPotential(self, srcPsi, destPsi, t, dt):
        for potential in self.PotentialList:
            potential.MultiplyPotential(srcPsi, destPsi, t, dt)

    def GetPotentialList(self):
        return self.PotentialList


class PamPropagator(PropagatorBase):
    """Fixed-m Arnoldi/Pade exponential propagator (pamp.h:613-669)."""

    def __init__(self, psi):
        PropagatorBase.__init__(self, psi)
        self.Rank = psi.GetRank()
        self.BasisSize = 20
        self.ExponentiationMethod = 0
        self.DoubleOrthogonalization = False
        self.MatvecPrecision = "fp64"
        self.Engine = None

    def ApplyConfig(self, config):
        PropagatorBase.ApplyConfig(self, config)
        self.BasePropagator = config.Propagation.base_propagator(self.psi)
        config.Apply(self.BasePropagator)
        self._apply_krylov(config.Propagation)

    def ApplyConfigSection(self, configSection):
        PropagatorBase.ApplyConfigSection(self, configSection)
        configSection.Apply(self.BasePropagator)
        self._apply_krylov(configSection)

    def _apply_krylov(self, section):
        # pampwrapper.cpp:51-70
        self.BasisSize = int(section.krylov_basis_size)
        if hasattr(section, "krylov_exponentiation_method"):
            self.ExponentiationMethod = int(section.krylov_exponentiation_method)
        if hasattr(section, "krylov_double_orthogonalization"):
            self.DoubleOrthogonalization = bool(section.krylov_double_orthogonalization)
        # extension key (no reference counterpart): "fp64" (default, the
        # drop-in) or "mixed" (one-electron states: every potential applied
        # in complex FP32; two-electron states: the electron-electron
        # multipole terms, the bulk of the apply, in FP32; Gram-Schmidt,
        # overlap solves and Pade stay FP64)
        if hasattr(section, "matvec_precision"):
            prec = str(section.matvec_precision).lower()
            if prec not in ("fp64", "mixed"):
                raise ValueError("matvec_precision must be 'fp64' or 'mixed', got %r" % prec)
            self.MatvecPrecision = prec

    def SetupStep(self, dt):
        self.BasePropagator.SetupStep(dt)
        if self.ExponentiationMethod not in (0, 1):
            raise ValueError("krylov_exponentiation_method must be 0 (Pade) or 1 (eigenvector), got %r"
                             % (self.ExponentiationMethod,))
        repr = self.psi.GetRepresentation()
        self.IsOrthogonalBasis = all(repr.IsOrthogonalBasis(i) for i in range(self.Rank))
        self.UseBalancedOverlap = hasattr(self.BasePropagator, "MultiplyHamiltonianBalancedOverlap")
        self.Engine = self._build_engine()
        # pampwrapper.cpp:56-61: 0 scaling + Pade(6), 1 the eigenvector exponential
        self.Engine.set_exponentiation(self.ExponentiationMethod)

    def _build_engine(self):
        """Map the consolidated tensor potentials onto the persistent engine:
        rank 2, angular rank 0 (orthogonal), B-spline rank 1 with BandNH storage."""
        repr = self.psi.GetRepresentation()
        pots = self.BasePropagator.PotentialList
        if self.Rank == 3:
            return self._build_engine3(repr, pots)
        if self.Rank != 2 or repr.IsOrthogonalBasis(1) is not False or not repr.IsOrthogonalBasis(0):
            raise NotImplementedError("the pAMP HIP engine handles (angular, B-spline) rank-2 states")
        # a coupling potential without angular pairs (the dipole term at
        # lmax = 0) contributes nothing and is left out of the engine
        pots = [p for p in pots if p.PotentialData.shape[0] > 0]
        L, N = repr.GetFullShape()
        overlap = self._engine_overlap(repr, L, N)
        eng = dev.PampEngine(L, N, self.BasisSize, overlap, reorth=self.DoubleOrthogonalization)
        for pot in pots:
            ids = pot.StorageIds()
            if ids[1] != "BandNH" or ids[0] not in ("Ident", "Diag", "Simp", "Herm"):
                raise NotImplementedError("storage %s is not on the pAMP HIP engine" % "_".join(ids))
            pairs = pot.GeometryList[0].GetBasisPairs() if ids[0] in ("Simp", "Herm") else None
            eng.add_term(ids[0], pot.DeviceData(), pairs=pairs)
        if self.MatvecPrecision == "mixed":
            # every term's radial bands in FP32 with FP32 band sums (the
            # complex64 operator apply); the exchange, Gram-Schmidt, the
            # overlap solves and Pade stay FP64
            eng.set_mixed(list(range(len(pots))))
        self._pots = pots
        return eng

    def _engine_overlap(self, repr, L, N):
        """The overlap factor (with its chunked solver plan) the engine uses:
        the representation's own."""
        return repr.GetRepresentation(1).GetOverlap()

    def _build_engine3(self, repr, pots):
        """Two-electron states (coupled channels x B-spline x B-spline): the
        factored potentials go to the rank-3 engine; each time-dependent
        potential gets its own time slot."""
        from .helium import CoupledTensorPotential, Pamp3Engine, radial_supports
        if not all(isinstance(p, CoupledTensorPotential) for p in pots):
            raise NotImplementedError("rank-3 states run on the two-electron engine only")
        C, N, N2 = repr.GetFullShape()
        if N != N2:
            raise NotImplementedError("both radial ranks must have the same size")
        overlap = repr.GetRepresentation(1).GetOverlap()
        eng = Pamp3Engine(C, N, self.BasisSize, overlap, reorth=self.DoubleOrthogonalization)
        eng.set_radial_supports(*radial_supports(repr.GetRepresentation(1).GetBSplineObject()))
        self._slots = []
        for pot in pots:
            slot = -1
            if pot.IsTimeDependent:
                slot = len(self._slots)
                self._slots.append(pot)
            eng.add_potential(pot, slot)
        if self.MatvecPrecision == "mixed":
            # two-electron form: the electron-electron multipole terms (the
            # bulk of the matvec) in FP32; the one-electron and laser terms
            # keep FP64
            eng.set_mixed(True)
        self._pots = pots
        return eng

    def TimeScalings(self, times):
        if self.Rank == 3:
            return np.array([[p.GetTimeScaling(t) for p in self._slots] for t in times],
                            dtype=np.float64).reshape(len(times), len(self._slots))
        return np.array([[p.GetTimeScaling(t) for p in self._pots] for t in times], dtype=np.float64)

    def MultiplyHamiltonian(self, srcPsi, destPsi, t, dt):
        self.BasePropagator.MultiplyHamiltonian(srcPsi, destPsi, t, dt)

    def AdvanceStep(self, t, dt):
        self.AdvanceSteps([t], dt)

    def AdvanceSteps(self, times, dt):
        """Consecutive AdvanceStep(t_n, dt) for each t_n in ``times`` in one launch."""
        sc = self.TimeScalings(times)
        self.Engine.propagate(self.psi.GetData(), len(times), dt, sc)

    def GetDiagnostics(self):
        """Per step of the last launch: |last Krylov coefficient| (pamp.h:666),
        basis size used, |U psi|, re-orthogonalisation passes."""
        return self.Engine.diagnostics()


def expokit_chunk(L, N, m, kd, chunk):
    """Solver chunk length for an Expokit engine over an (L, N) problem with
    Krylov dimension m: None when the overlap's own plan (`chunk`, 0 = the
    default) leaves room in 160 KB of LDS for zgexpv's Hessenberg matrix
    and exponential workspace (16 ((m+2)^2 + m+2) bytes), else the shortest
    longer chunk that does."""
    lib = nat.lib()
    hm = 16 * ((m + 2) * (m + 2) + m + 2)

    def fits(B):
        est = lib.pp_pamp_lds_estimate(L, N, m, 3, kd, B)
        return est > 0 and est + hm <= 160 * 1024

    if fits(chunk):
        return None
    for B in (32, 64, 128, 256):
        if B > chunk and fits(B):
            return B
    return None


class ExpokitPropagator(PamPropagator):
    """Expokit zgexpv propagator (pyprop/propagator/KrylovPropagator.py:1-60,
    core/krylov/expokit/expokitpropagator.cpp:9-103): config keys
    krylov_basis_size, krylov_tolerance, krylov_norm.  Same balanced-overlap
    wrapping and persistent engine as PamPropagator; the exponential is the
    adaptive zgexpv step with error control.  The Krylov time is |dt| (the
    reference's expokitpropagator.cpp:61 passes dt^2 for real dt)."""

    def __init__(self, psi):
        PamPropagator.__init__(self, psi)
        self.Tolerance = 0.0
        self.MatrixNorm = 1e6

    def _apply_krylov(self, section):
        # expokitpropagator.cpp:13-20
        self.BasisSize = int(section.krylov_basis_size)
        if hasattr(section, "krylov_tolerance"):
            self.Tolerance = float(section.krylov_tolerance)
        if hasattr(section, "krylov_norm"):
            self.MatrixNorm = float(section.krylov_norm)
        if hasattr(section, "matvec_precision"):
            prec = str(section.matvec_precision).lower()
            if prec not in ("fp64", "mixed"):
                raise ValueError("matvec_precision must be 'fp64' or 'mixed', got %r" % prec)
            self.MatvecPrecision = prec

    def _engine_overlap(self, repr, L, N):
        """zgexpv keeps its (m+2) x (m+2) Hessenberg matrix in LDS beside
        the pAMP buffers of every fiber; when the representation's solver
        plan leaves too little of the 160 KB, the engine gets its own plan
        with longer chunks (fewer scan matrices): same factor U, so the
        results agree to rounding."""
        ov = repr.GetRepresentation(1).GetOverlap()
        B = expokit_chunk(L, N, self.BasisSize, ov.kd, ov.chunk)
        if B is None:
            return ov
        b = repr.GetRepresentation(1).GetBSplineObject()
        return dev.Overlap(b.overlap_upper(), N, ov.kd, chunk=B)

    def AdvanceSteps(self, times, dt):
        sc = self.TimeScalings(times)
        self.Engine.propagate_expokit(self.psi.GetData(), len(times), dt, sc, self.Tolerance, self.MatrixNorm)

    def GetDiagnostics(self):
        """Per step of the last launch: matvecs, exponentials, substeps,
        rejected substeps, basis size of the last substep, zgexpv iflag."""
        return self.Engine.expokit_stats()


class CayleyPropagator(PropagatorBase):
    """Crank-Nicolson step (S + i dt/2 H) psi(t+dt) = (S - i dt/2 H) psi(t),
    H at the midpoint t + |dt/2|, solved by the reference's single-cycle
    GMRES with an optional left preconditioner
    (pyprop/propagator/CayleyPropagator.py:1-140): the [Propagation] keys
    base_propagator, preconditioner, krylov_basis_size, krylov_tolerance
    and krylov_double_orthogonalization keep their meaning.  psi(t) is the
    GMRES right-hand side's source and psi(t+dt) its solution (zero start).
    Rank-2 states apply H through the base propagator's tensor-potential
    kernels (preconditioner: BSplinePreconditioner); rank-3 (two-electron)
    states through the rank-3 engine's raw apply (preconditioner:
    RadialTwoElectronPreconditioner and its SuperLU / Ifpack names)."""

    def __init__(self, psi):
        PropagatorBase.__init__(self, psi)
        from .solver import GmresSolver
        self.Rank = psi.GetRank()
        self.Solver = GmresSolver(psi)
        self.Preconditioner = None
        self.LastErrorEstimate = 0.0

    def ApplyConfig(self, config):
        PropagatorBase.ApplyConfig(self, config)
        self.BasePropagator = config.Propagation.base_propagator(self.psi)
        config.Apply(self.BasePropagator)
        self.Solver.ApplyConfigSection(config.Propagation)
        name = config.Propagation.get("preconditioner", None)
        if name:
            sec = config.GetSection(name)
            self.Preconditioner = sec.type(self.psi)
            sec.Apply(self.Preconditioner)

    def ApplyConfigSection(self, configSection):
        PropagatorBase.ApplyConfigSection(self, configSection)
        configSection.Apply(self.BasePropagator)
        self.Solver.ApplyConfigSection(configSection)

    def SetupStep(self, dt):
        self.BasePropagator.SetupStep(dt)
        self.Solver.Setup(self.psi)
        self.TempPsi = self.psi.Copy()
        repr = self.psi.GetRepresentation()
        self.IsOrthogonalBasis = all(repr.IsOrthogonalBasis(i) for i in range(self.Rank))
        self._engine3 = None
        if self.Rank == 3:
            # raw H of the two-electron engine (its Krylov buffers are sized
            # by the GMRES basis but unused here)
            helper = PamPropagator(self.psi)
            helper.BasePropagator = self.BasePropagator
            helper.BasisSize = max(2, self.Solver.BasisSize)
            self._engine3 = helper._build_engine3(repr, self.BasePropagator.PotentialList)
            self._helper3 = helper
        elif self.Rank != 2:
            raise NotImplementedError("CayleyPropagator on the HIP path handles rank-2 and rank-3 states")
        if self.Preconditioner is not None:
            self.Preconditioner.SetHamiltonianScaling(1.0j * dt / 2.0)
            self.Preconditioner.SetOverlapScaling(1.0)
            self.Preconditioner.Setup(self)
        self.TempPsiMultiply = self.psi.Copy()

    def MultiplyHamiltonian(self, srcPsi, destPsi, t, dt):
        self.BasePropagator.MultiplyHamiltonian(srcPsi, destPsi, t, dt)

    def _raw_hamiltonian(self, srcPsi, destPsi, t, dt):
        if self._engine3 is not None:
            sc = self._helper3.TimeScalings([t])[0]
            self._engine3.apply(srcPsi.GetData(), destPsi.GetData(), sc, balanced=False)
        else:
            self.BasePropagator.MultiplyHamiltonianNoOverlap(srcPsi, destPsi, t, dt)

    def MultiplySpH(self, sourcePsi, destPsi, hFactor, t, dt):
        """destPsi = (S + hFactor H(t)) sourcePsi (CayleyPropagator.py:117-134)."""
        tmp = self.TempPsiMultiply
        tmp.GetData().copy_(sourcePsi.GetData())
        if not self.IsOrthogonalBasis:
            tmp.GetRepresentation().MultiplyOverlap(tmp)
        self._raw_hamiltonian(sourcePsi, destPsi, t, dt)
        destPsi.GetData().mul_(hFactor).add_(tmp.GetData())

    def SolverCallback(self, sourcePsi, destPsi, t, dt):
        self.MultiplySpH(sourcePsi, destPsi, 1.0j * dt / 2.0, t, dt)
        if self.Preconditioner is not None:
            self.Preconditioner.Solve(destPsi)

    def AdvanceStep(self, t, dt):
        tm = t + abs(dt / 2.0)
        self.TempPsi.Clear()
        self.MultiplySpH(self.psi, self.TempPsi, -1.0j * dt / 2.0, tm, dt)
        if self.Preconditioner is not None:
            self.Preconditioner.Solve(self.TempPsi)
        self.Solver.Solve(lambda s, d: self.SolverCallback(s, d, tm, dt), self.TempPsi, self.psi, False)
        self.LastErrorEstimate = self.Solver.GetErrorEstimate()
        if self.LastErrorEstimate > 1e-4:
            print("Error = %s" % self.LastErrorEstimate)
        if self.LastErrorEstimate > self.Solver.Tolerance:
            warnings.warn("CayleyPropagator: GMRES did not converge in %d vectors (residual estimate %.3e > "
                          "tolerance %.3e); the step is inaccurate" % (self.Solver.BasisSize, self.LastErrorEstimate,
                                                                       self.Solver.Tolerance),
                          RuntimeWarning, stacklevel=2)

    def AdvanceSteps(self, times, dt):
        for t in times:
            self.AdvanceStep(t, dt)
