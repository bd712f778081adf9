"""Eigenvalue solvers on the device behind pyprop's solver interface.

PiramSolver mirrors pyprop/solver/PiramSolver.py:3-142 (class name,
constructor from a Problem, the [Arpack] config keys, Solve /
GetEigenvalues / GetEigenvector / SetEigenvector / GetConvergenceEstimates)
and restates the implicitly restarted Arnoldi method of pIRAM
(core/krylov/piram/piram/piram.h): an m-step Arnoldi factorisation, the Ritz
values of the Hessenberg matrix sorted by increasing real part
(ShiftFunctorSelectRitzValues with CompareComplexLessReal, :261-267), the
unwanted m - k of them as exact shifts ordered by their error bounds
(:494-518), m - k shifted QR steps on the Hessenberg matrix and the same
unitary transformation of the basis (:540-640), until every wanted Ritz
pair's error bound |s_{m,i}| |f| passes the convergence test or the restart
limit is reached.

The n-length work runs on the device: the operator (the Hamiltonian through
the HIP tensor-potential and overlap kernels), classical Gram-Schmidt with
one re-orthogonalisation pass (pp_gram_schmidt) and the basis
recombination of a restart (pp_basis_combine).  The m x m work (eigenvalues
and QR steps of the Hessenberg matrix) runs on the host, as pIRAM does it
through LAPACK.

Default operator: the balanced Hamiltonian U^-H H U^-1 (S = U^H U), which has
the eigenvalues of the reference's default S^-1 H and is Hermitian in the
plain inner product; eigenvectors are mapped back by U^-1.  A
`matrix_vector_func` in [Arpack] replaces it with the caller's operator on
plain wavefunction data (the reference's custom operator hook).
"""
import ctypes

import numpy as np

from . import _native as nat
from . import device as dev


def _balanced_operator(prop):
    """dst = U^-H H(t) U^-1 src for the problem's propagator: one-electron
    states through BasisPropagator.MultiplyHamiltonianBalancedOverlap (the
    HIP overlap and tensor kernels); two-electron states through the rank-3
    engine's balanced apply, whose potentials are applied only as a whole."""
    if prop.psi.GetRank() == 3:
        eng = prop.Propagator.Engine

        def apply3(src, dst, t, dt):
            sc = prop.Propagator.TimeScalings([t])[0]
            eng.apply(src.GetData(), dst.GetData(), scal=sc if len(sc) else None, balanced=True)
        return apply3
    return prop.Propagator.BasePropagator.MultiplyHamiltonianBalancedOverlap


class PiramSolver:
    """Implicitly restarted Arnoldi for the lowest eigenvalues of the
    problem's Hamiltonian (pyprop/solver/PiramSolver.py)."""

    def __init__(self, prop):
        self.BaseProblem = prop
        self.Rank = prop.psi.GetRank()
        sec = prop.Config.Arpack
        self.BasisSize = int(sec.krylov_basis_size)
        self.EigenvalueCount = int(sec.krylov_eigenvalue_count)
        self.MaxRestartCount = int(sec.get("krylov_max_iteration_count", 1000))
        eps = np.finfo(np.float64).eps
        self.Tolerance = max(float(sec.get("krylov_tolerance", 0.0)), eps)
        self.UseRandomStart = bool(sec.get("krylov_use_random_start", True))
        self.Debug = bool(sec.get("krylov_debug", False))
        if 2 * self.EigenvalueCount > self.BasisSize:
            # piram.h:269-272
            raise RuntimeError("Basis size too small. should be at least 2*EigenvalueCount")
        func = sec.get("matrix_vector_func", None)
        self.InverseIterations = bool(sec.get("inverse_iterations", False))
        if self.InverseIterations and not func:
            raise Exception("Inverse Iterations must be used together with a userdefined matrix_vector_func")
        # Ritz value order (core/krylov/piram/piramsolver.cpp:58-95): smallest
        # real part first; nearest to krylov_eigenvalue_shift if given;
        # largest magnitude first with inverse iterations
        shift = sec.get("krylov_eigenvalue_shift", None)
        self.EigenvalueShift = None if shift is None else complex(shift)
        self.Balanced = not func
        if func:
            self.ApplyMatrix = func
        else:
            self.ApplyMatrix = _balanced_operator(prop)
        self.TempPsi = prop.psi.Copy()
        self.SrcPsi = prop.psi.Copy()
        self.Shape = tuple(prop.psi.GetData().shape)
        self.MatrixSize = int(np.prod(self.Shape))
        self.OperatorCount = 0
        self.RestartCount = 0

    # ------------------------------------------------------------------ device pieces
    def _op(self, src, dst):
        """dst = A src for basis rows src, dst (device views of length n)."""
        self.SrcPsi.GetData().copy_(src.view(self.Shape))
        self.ApplyMatrix(self.SrcPsi, self.TempPsi, 0, 0)
        dst.copy_(self.TempPsi.GetData().reshape(-1))
        self.OperatorCount += 1

    def _gs(self, k, w):
        """CGS2 of w against V_0..V_{k-1}: returns (h (k,), |w| after)."""
        h = np.zeros(2 * max(k, 1), dtype=np.float64)
        nrm = ctypes.c_double(0.0)
        nat.check(nat.lib().pp_gram_schmidt(dev.ptr(self.V), self.MatrixSize, int(k), dev.ptr(w), 2, nat.dptr(h),
                                            ctypes.byref(nrm), dev.stream_ptr()), "pp_gram_schmidt")
        return (h[0:2 * k:2] + 1j * h[1:2 * k:2]), nrm.value

    def _combine(self, Q, out):
        """out[p] = sum_j V_j Q[j, p] (Q: (rows of V used, np) complex)."""
        Qf = np.asfortranarray(Q, dtype=np.complex128)
        m, npv = Qf.shape
        nat.check(nat.lib().pp_basis_combine(dev.ptr(self.V), self.MatrixSize, int(m),
                                             Qf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(npv),
                                             dev.ptr(out), dev.stream_ptr()), "pp_basis_combine")

    # ------------------------------------------------------------------ Arnoldi
    def _random_orth(self, k, w):
        """w := a seeded random vector orthogonal to V_0..V_{k-1}, normalised
        (the continuation after an invariant subspace, as ARPACK's getv0)."""
        t = dev.torch()
        self._rng = getattr(self, "_rng", np.random.default_rng(7))
        w.copy_(t.from_numpy(self._rng.uniform(-1.0, 1.0, self.MatrixSize).astype(np.complex128)).to(w.device))
        _, nrm = self._gs(k, w)
        w.mul_(1.0 / nrm)

    def _extend(self, j0):
        """Arnoldi steps j0 .. m-1: V[j+1], H[:j+2, j] from A V[j].  A residual
        below 1e-12 of |A v_j| marks an invariant subspace: H[j+1, j] = 0 and
        the basis continues with a random orthogonal vector."""
        m = self.BasisSize
        for j in range(j0, m):
            w = self.V[j + 1]
            self._op(self.V[j], w)
            h, beta = self._gs(j + 1, w)
            self.H[:j + 1, j] = h
            if beta <= 1e-12 * np.sqrt(np.sum(np.abs(h) ** 2) + beta * beta):
                beta = 0.0
                self._random_orth(j + 1, w)
            else:
                w.mul_(1.0 / beta)
            self.Beta = beta
            if j + 1 < m:
                self.H[j + 1, j] = beta

    def _ritz(self):
        """Ritz values sorted by increasing real part, their vectors and
        error bounds |s_{m-1, i}| |f| (piram.h:483-532)."""
        theta, S = np.linalg.eig(self.H)
        if self.InverseIterations:
            key = -np.abs(theta)
        elif self.EigenvalueShift is not None:
            key = np.abs(theta - self.EigenvalueShift)
        else:
            key = theta.real
        order = np.argsort(key, kind="stable")
        theta, S = theta[order], S[:, order]
        S = S / np.linalg.norm(S, axis=0)
        bounds = np.abs(S[-1, :]) * self.Beta
        return theta, S, bounds

    def _converged(self, theta, bounds):
        """ARPACK's test on the wanted Ritz pairs: bound <= tol max(eps^(2/3), |theta|)."""
        eps23 = np.finfo(np.float64).eps ** (2.0 / 3.0)
        k = self.EigenvalueCount
        acc = self.Tolerance * np.maximum(eps23, np.abs(theta[:k]))
        self.ConvergenceEstimates = bounds[:k] - acc
        self.ErrorEstimates = bounds[:k].copy()
        return bool(np.all(self.ConvergenceEstimates <= 0.0))

    def _restart(self, theta, bounds):
        """m - k exact shifts (the unwanted Ritz values, by increasing error
        bound) as shifted QR steps; V_k+ = V Q[:, :k], f_k+ = V Q[:, k] H+[k, k-1]
        + f Q[m-1, k-1] (piram.h:540-640)."""
        m, k = self.BasisSize, self.EigenvalueCount
        shifts = theta[k:][np.argsort(bounds[k:], kind="stable")]
        H = self.H.copy()
        Qacc = np.eye(m, dtype=np.complex128)
        I = np.eye(m)
        for mu in shifts:
            Qi, Ri = np.linalg.qr(H - mu * I)
            H = Ri @ Qi + mu * I
            Qacc = Qacc @ Qi
        # columns 0..k-1: the new basis; column k: the new residual (row m:
        # the old residual direction V[m] with weight |f| Q[m-1, k-1])
        C = np.zeros((m + 1, k + 1), dtype=np.complex128)
        C[:m, :k] = Qacc[:, :k]
        C[:m, k] = H[k, k - 1] * Qacc[:, k]
        C[m, k] = self.Beta * Qacc[m - 1, k - 1]
        tmp = dev.torch().empty((k + 1, self.MatrixSize), dtype=self.V.dtype, device=self.V.device)
        self._combine(C, tmp)
        self.V[:k + 1].copy_(tmp)
        self.H = np.zeros((m, m), dtype=np.complex128)
        self.H[:k, :k] = H[:k, :k]
        # re-orthogonalise the residual against the kept basis (its components
        # there are rounding-level and are dropped), then normalise; a
        # vanishing residual (the kept vectors span an invariant subspace)
        # continues with a random orthogonal vector and H[k, k-1] = 0
        f = self.V[k]
        _, beta = self._gs(k, f)
        scale = np.abs(H[:k, :k]).max() if k > 0 else 1.0
        if beta <= 1e-12 * max(scale, 1e-300):
            beta = 0.0
            self._random_orth(k, f)
        else:
            f.mul_(1.0 / beta)
        self.H[k, k - 1] = beta
        self.RestartCount += 1

    def Solve(self):
        t = dev.torch()
        m, n = self.BasisSize, self.MatrixSize
        self.V = t.zeros((m + 1, n), dtype=t.complex128, device=dev.device())
        if self.UseRandomStart:
            rng = np.random.default_rng(20261203)
            start = rng.uniform(0.0, 1.0, n).astype(np.complex128)   # piram.h SetupRandomResidual: uniform [0, 1]
        else:
            start = self.BaseProblem.psi.GetData().reshape(-1).cpu().numpy()
            if self.Balanced:   # the balanced operator acts on U psi
                w = self.BaseProblem.psi.Copy()
                w.GetRepresentation().MultiplySqrtOverlap(False, w)
                start = w.GetData().reshape(-1).cpu().numpy()
        self.V[0].copy_(t.from_numpy(start).to(self.V.device))
        _, nrm = self._gs(0, self.V[0])
        self.V[0].mul_(1.0 / nrm)
        self.H = np.zeros((m, m), dtype=np.complex128)
        self.OperatorCount = self.RestartCount = 0
        self._extend(0)
        while True:
            theta, S, bounds = self._ritz()
            self.IsConverged = self._converged(theta, bounds)
            if self.Debug:
                print("pIRAM restart %d: eigenvalues %s, bounds %s" % (self.RestartCount, theta[:self.EigenvalueCount].real,
                                                                      bounds[:self.EigenvalueCount]))
            if self.IsConverged or self.RestartCount >= self.MaxRestartCount:
                break
            self._restart(theta, bounds)
            self._extend(self.EigenvalueCount)
        self.Eigenvalues = theta
        self.RitzVectors = S
        return self.IsConverged

    # ------------------------------------------------------------------ results
    def GetEigenvalues(self):
        """Real parts of the wanted Ritz values in the selection order."""
        return self.Eigenvalues[:self.EigenvalueCount].real.copy()

    def GetConvergenceEstimates(self):
        """Error bound minus the tolerance per wanted eigenvalue (<= 0: converged)."""
        return self.ConvergenceEstimates.copy()

    def GetErrorEstimates(self):
        return self.ErrorEstimates.copy()

    def GetEigenvector(self, index):
        """Eigenvector `index` as a flat numpy array, 2-norm 1 (as the
        reference's; assign it to a wavefunction and Normalize() for the
        S-norm)."""
        t = dev.torch()
        out = t.empty((1, self.MatrixSize), dtype=self.V.dtype, device=self.V.device)
        s = np.zeros((self.BasisSize + 1, 1), dtype=np.complex128)
        s[:self.BasisSize, 0] = self.RitzVectors[:, index]
        self._combine(s, out)
        if self.Balanced:   # x = U^-1 x^
            w = self.BaseProblem.psi.Copy()
            w.GetData().copy_(out.view(self.Shape))
            w.GetRepresentation().SolveSqrtOverlap(False, w)
            x = w.GetData().reshape(-1).cpu().numpy()
        else:
            x = out.reshape(-1).cpu().numpy()
        return x / np.linalg.norm(x)

    def SetEigenvector(self, psi, eigenvectorIndex, normalize=True):
        psi.GetData().copy_(dev.to_device(self.GetEigenvector(eigenvectorIndex).reshape(psi.GetData().shape)))
        if normalize:
            psi.Normalize()


class BSplinePreconditioner:
    """Per-fiber banded solve along one B-spline rank of
    M = scalingH * sum(potentials) + scalingS * S, the other ranks diagonal
    (pyprop's BSplinePreconditioner interface,
    examples/tensor/helium_stabilization_sae/example.py:275-318: config keys
    `rank` and `potential_evaluation`, SetHamiltonianScaling /
    SetOverlapScaling, Setup(prop), Solve(psi)).  Each fiber's band matrix is
    LU-factored once on the host (LAPACK zgbtrf); Solve runs the
    substitutions on the device (pp_band_lu_solve)."""

    def __init__(self, psi):
        self.Rank = psi.GetRank()
        self.psi = psi
        self.HamiltonianScaling = 1.0
        self.OverlapScaling = 0.0

    def ApplyConfigSection(self, conf):
        self.PreconditionRank = int(conf.rank)
        self.PotentialSections = [conf.Config.GetSection(s) for s in conf.potential_evaluation]

    def SetHamiltonianScaling(self, scalingH):
        self.HamiltonianScaling = scalingH

    def SetOverlapScaling(self, scalingS):
        self.OverlapScaling = scalingS

    def GetHamiltonianScaling(self):
        return self.HamiltonianScaling

    def GetOverlapScaling(self):
        return self.OverlapScaling

    def Setup(self, prop):
        import scipy.linalg.lapack as lapack
        r = self.PreconditionRank
        shape = tuple(self.psi.GetData().shape)
        if self.Rank != 2 or r != 1:
            raise NotImplementedError("BSplinePreconditioner on the device handles rank-2 states with the B-spline "
                                      "rank 1 preconditioned")
        L, N = shape
        bs = self.psi.GetRepresentation().GetRepresentation(r).GetBSplineObject()
        kd = bs.MaxSplineOrder - 1
        bw = 2 * kd + 1
        # BandNH of each fiber: element (i, j) at j*bw + kd + i - j
        band = np.zeros((L, N * bw), dtype=np.complex128)
        for conf in self.PotentialSections:
            pot = prop.BasePropagator.GeneratePotential(conf)
            ids = pot.StorageIds()
            if ids[r] != "BandNH" or ids[0] != "Diag":
                raise ValueError("Potentials must be banded-nonhermitian in precondition rank and diagonal elsewhere")
            band += self.HamiltonianScaling * pot.PotentialData
        S = bs.overlap_dense()
        for j in range(N):
            for i in range(max(0, j - kd), min(N, j + kd + 1)):
                band[:, j * bw + kd + i - j] += self.OverlapScaling * S[i, j]
        ldab = 3 * kd + 1
        lub = np.zeros((L, N, ldab), dtype=np.complex128)
        piv = np.zeros((L, N), dtype=np.int32)
        for a in range(L):
            ab = np.zeros((ldab, N), dtype=np.complex128, order="F")
            for j in range(N):
                for i in range(max(0, j - kd), min(N, j + kd + 1)):
                    ab[2 * kd + i - j, j] = band[a, j * bw + kd + i - j]
            lu, pv, info = lapack.zgbtrf(ab, kd, kd)
            if info != 0:
                raise RuntimeError("preconditioner matrix of fiber %d is singular (zgbtrf info %d)" % (a, info))
            lub[a] = lu.T
            piv[a] = pv
        self.kd, self.N, self.L = kd, N, L
        self._lub = dev.to_device(lub)
        t = dev.torch()
        self._piv = t.from_numpy(piv).to(dev.device())

    def Solve(self, psi):
        nat.check(nat.lib().pp_band_lu_solve(self.L, self.N, self.kd, self.kd, dev.ptr(self._lub),
                                             ctypes.cast(ctypes.c_void_p(self._piv.data_ptr()),
                                                         ctypes.POINTER(ctypes.c_int)),
                                             dev.ptr(psi.GetData()), dev.stream_ptr()), "pp_band_lu_solve")


class GMRESShiftInvertSolver:
    """(S^-1 H - shift) x = b solved as M^-1 (H - shift S) x = M^-1 S b by
    restarted GMRES on the device (pyprop/solver/GMRESShiftInvertSolver.py:2-136:
    the [GMRES] section's krylov_basis_size, krylov_tolerance, shift,
    preconditioner; InverseIterations as PiramSolver's matrix_vector_func).
    Arnoldi vectors, Gram-Schmidt and the solution update run on the device;
    the (m+1) x m least-squares problem on the host."""

    def __init__(self, prop):
        self.BaseProblem = prop
        self.Rank = prop.psi.GetRank()
        if self.Rank != 2:
            raise NotImplementedError("GMRESShiftInvertSolver on the device handles (angular, B-spline) rank-2 states")
        self.Config = prop.Config.GMRES
        self.psi = prop.psi
        self.Shift = float(self.Config.shift)
        self.BasisSize = int(self.Config.krylov_basis_size)
        self.Tolerance = max(float(self.Config.get("krylov_tolerance", 0.0)), 1e-12)
        self.MaxRestarts = int(self.Config.get("krylov_max_iteration_count", 100))
        self.TempPsi = prop.psi.Copy()
        self.TempPsi2 = prop.psi.Copy()
        name = self.Config.get("preconditioner", None)
        self.Preconditioner = None
        if name:
            sec = prop.Config.GetSection(name)
            self.Preconditioner = sec.type(self.psi)
            sec.Apply(self.Preconditioner)
            self.Preconditioner.SetHamiltonianScaling(1.0)
            self.Preconditioner.SetOverlapScaling(-self.Shift)
            self.Preconditioner.Setup(prop.Propagator)
        repr = self.psi.GetRepresentation()
        self.IsOrthogonalBasis = all(repr.IsOrthogonalBasis(i) for i in range(self.Rank))
        self.Shape = tuple(prop.psi.GetData().shape)
        self.MatrixSize = int(np.prod(self.Shape))
        self.Counters = {"MultiplyHamiltonian": 0, "PreconditionerSolve": 0, "Inverse Iterations": 0,
                         "GMRES iterations": 0}
        self.LastResidual = None

    def _op(self, src, dst):
        """dst = M^-1 (H - shift S) src (device rows of length n)."""
        a, b = self.TempPsi, self.TempPsi2
        a.GetData().copy_(src.view(self.Shape))
        bp = self.BaseProblem.Propagator.BasePropagator
        if self.IsOrthogonalBasis:
            bp.MultiplyHamiltonian(a, b, 0, 0)
        else:
            bp.MultiplyHamiltonianNoOverlap(a, b, 0, 0)
        self.Counters["MultiplyHamiltonian"] += 1
        a.GetRepresentation().MultiplyOverlap(a)
        b.GetData().sub_(a.GetData(), alpha=self.Shift)
        if self.Preconditioner:
            self.Preconditioner.Solve(b)
            self.Counters["PreconditionerSolve"] += 1
        dst.copy_(b.GetData().reshape(-1))

    def _gs(self, V, k, w):
        h = np.zeros(2 * max(k, 1))
        nrm = ctypes.c_double(0.0)
        nat.check(nat.lib().pp_gram_schmidt(dev.ptr(V), self.MatrixSize, int(k), dev.ptr(w), 2, nat.dptr(h),
                                            ctypes.byref(nrm), dev.stream_ptr()), "pp_gram_schmidt")
        return h[0:2 * k:2] + 1j * h[1:2 * k:2], nrm.value

    def _gmres(self, rhs, x):
        """x := solution of (M^-1 A) x = rhs (device rows); x starts at 0."""
        t = dev.torch()
        m, n = self.BasisSize, self.MatrixSize
        V = t.zeros((m + 1, n), dtype=t.complex128, device=rhs.device)
        x.zero_()
        r = rhs.clone()
        _, bnorm = self._gs(V, 0, r)
        if bnorm == 0.0:
            self.LastResidual = 0.0
            return
        for restart in range(self.MaxRestarts):
            V[0].copy_(r)
            _, beta = self._gs(V, 0, V[0])
            V[0].mul_(1.0 / beta)
            Hh = np.zeros((m + 1, m), dtype=np.complex128)
            k = 0
            for j in range(m):
                self._op(V[j], V[j + 1])
                h, hn = self._gs(V, j + 1, V[j + 1])
                Hh[:j + 1, j] = h
                Hh[j + 1, j] = hn
                k = j + 1
                self.Counters["GMRES iterations"] += 1
                e1 = np.zeros(j + 2, dtype=np.complex128)
                e1[0] = beta
                y, *_ = np.linalg.lstsq(Hh[:j + 2, :j + 1], e1, rcond=None)
                res = np.linalg.norm(Hh[:j + 2, :j + 1] @ y - e1)
                if hn <= 1e-14 * beta or res <= self.Tolerance * bnorm:
                    break
                V[j + 1].mul_(1.0 / hn)
            # x += V_k y
            dx = t.empty((1, n), dtype=t.complex128, device=rhs.device)
            Yc = np.zeros((k, 1), dtype=np.complex128)
            Yc[:, 0] = y
            Yf = np.asfortranarray(Yc)
            nat.check(nat.lib().pp_basis_combine(dev.ptr(V), n, k, Yf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                 1, dev.ptr(dx), dev.stream_ptr()), "pp_basis_combine")
            x.add_(dx[0])
            self.LastResidual = res / bnorm
            if res <= self.Tolerance * bnorm or hn <= 1e-14 * beta:
                return
            # true residual for the restart
            self._op(x, r)
            r.mul_(-1.0).add_(rhs)
        raise RuntimeError("GMRES did not reach tolerance %g in %d restarts (residual %g)"
                           % (self.Tolerance, self.MaxRestarts, self.LastResidual))

    def InverseIterations(self, srcPsi, destPsi, t, dt):
        """destPsi = (S^-1 H - shift)^-1 srcPsi (the PiramSolver operator with
        inverse_iterations = True)."""
        self.Counters["Inverse Iterations"] += 1
        c = self.TempPsi2.Copy()
        c.GetData().copy_(srcPsi.GetData())
        c.GetRepresentation().MultiplyOverlap(c)
        if self.Preconditioner:
            self.Preconditioner.Solve(c)
            self.Counters["PreconditionerSolve"] += 1
        self._gmres(c.GetData().reshape(-1), destPsi.GetData().reshape(-1))

    def PrintStatistics(self):
        print("Statistics for GMRESShiftInvertSolver")
        print(self.Counters)


class GmresSolver:
    """The reference's GMRES wrapper (core.krylov_GmresWrapper,
    core/krylov/gmres/gmreswrapper.cpp:52-111 over gmres::GMRES<cplx>,
    gmres.h:207-500) on device vectors: zero start, one cycle of at most
    krylov_basis_size Arnoldi vectors, Euclidean inner products, classical
    Gram-Schmidt with the 0.717 re-orthogonalisation rule (at most 3 extra
    passes; krylov_double_orthogonalization, on by default), tolerance
    floored at eps^(2/3), stop at happy breakdown or when the least-squares
    residual is below the tolerance.  Arnoldi vectors, Gram-Schmidt and the
    solution update run on the device (pp_gram_schmidt, pp_basis_combine);
    the (k+1) x k least-squares problem on the host."""

    def __init__(self, psi=None):
        self.BasisSize = 20
        self.Tolerance = float(np.sqrt(np.finfo(float).eps))
        self.PerformDoubleOrthogonalization = True
        self.MaxOrthogonalizationCount = 3
        self.ErrorEstimate = 0.0
        self.ErrorEstimateList = np.zeros(0)
        self.HessenbergMatrix = np.zeros((0, 0), dtype=np.complex128)
        self.OperatorCount = 0
        self.OrthogonalizationCount = 0
        self.V = None

    def ApplyConfigSection(self, conf):
        self.BasisSize = int(conf.krylov_basis_size)
        if hasattr(conf, "krylov_tolerance"):
            self.Tolerance = float(conf.krylov_tolerance)
        if hasattr(conf, "krylov_double_orthogonalization"):
            self.PerformDoubleOrthogonalization = bool(conf.krylov_double_orthogonalization)

    def Setup(self, psi):
        if self.BasisSize < 1:
            raise ValueError("krylov_basis_size must be at least 1")
        self.Tolerance = max(self.Tolerance, np.finfo(float).eps ** (2.0 / 3.0))
        self.Shape = tuple(psi.GetData().shape)
        self.MatrixSize = int(np.prod(self.Shape))
        t = dev.torch()
        self.V = t.zeros((self.BasisSize, self.MatrixSize), dtype=t.complex128, device=psi.GetData().device)
        self.Residual = t.zeros(self.MatrixSize, dtype=t.complex128, device=psi.GetData().device)
        self.ErrorEstimateList = np.zeros(self.BasisSize)
        self.ResetStatistics()

    def ResetStatistics(self):
        self.OperatorCount = 0
        self.OrthogonalizationCount = 0

    def _gs(self, k, w):
        """w -= V_k V_k^H w (one pass); returns (V_k^H w, |w|) after the pass."""
        h = np.zeros(2 * max(k, 1))
        nrm = ctypes.c_double(0.0)
        nat.check(nat.lib().pp_gram_schmidt(dev.ptr(self.V), self.MatrixSize, int(k), dev.ptr(w), 1, nat.dptr(h),
                                            ctypes.byref(nrm), dev.stream_ptr()), "pp_gram_schmidt")
        return h[0:2 * k:2] + 1j * h[1:2 * k:2], nrm.value

    def _apply(self, src, dst):
        """ApplyOperator (gmreswrapper.cpp:17-50): the callback sees the two
        wavefunctions of the Solve call holding the input / output buffers."""
        self.OperatorCount += 1
        a, b = self._psi, self._tmp
        da, db = a.GetData(), b.GetData()
        a.SetData(src.view(self.Shape))
        b.SetData(dst.view(self.Shape))
        dst.zero_()
        try:
            self._callback(a, b)
        finally:
            a.SetData(da)
            b.SetData(db)

    def Solve(self, callback, psi, tempPsi, usePypropIntegration=False):
        """tempPsi := x with A x = psi, A applied by callback(src, dst)."""
        if usePypropIntegration:
            raise NotImplementedError("GMRES with the S inner product is not on the HIP path")
        if self.V is None:
            raise RuntimeError("GmresSolver.Setup was not called")
        self._callback, self._psi, self._tmp = callback, psi, tempPsi
        m, V, r = self.BasisSize, self.V, self.Residual
        tol = self.Tolerance
        Hm = np.zeros((m + 1, m), dtype=np.complex128)
        self.ErrorEstimateList[:] = 0.0
        r.copy_(psi.GetData().reshape(-1))
        _, bnorm = self._gs(0, r)
        sol = tempPsi.GetData().reshape(-1)
        if bnorm == 0.0:
            sol.zero_()
            self.ErrorEstimate = 0.0
            self.HessenbergMatrix = Hm
            return
        r.mul_(1.0 / bnorm)
        V[0].copy_(r)
        self._apply(V[0], r)
        h, beta = self._gs(1, r)
        Hm[0, 0], Hm[1, 0] = h[0], beta
        happy = beta < tol
        y = np.array([bnorm / h[0]]) if happy else np.zeros(1, dtype=np.complex128)
        cur = 0
        err = beta
        self.ErrorEstimateList[0] = err
        while cur < m - 1 and not happy and abs(err) > tol:
            j = cur
            V[j + 1].copy_(r).mul_(complex(1.0 / Hm[j + 1, j]))
            self._apply(V[j + 1], r)
            _, orig = self._gs(0, r)
            h, res = self._gs(j + 2, r)
            count = 0
            while self.PerformDoubleOrthogonalization and res < 0.717 * orig:
                count += 1
                if count > self.MaxOrthogonalizationCount:
                    break
                self.OrthogonalizationCount += 1
                h2, res2 = self._gs(j + 2, r)
                h = h + h2
                orig, res = res, res2
            Hm[:j + 2, j + 1] = h
            beta = res
            happy = beta < tol
            Hm[j + 2, j + 1] = beta
            cur += 1
            k = cur + 1
            e1 = np.zeros(k + 1, dtype=np.complex128)
            e1[0] = bnorm
            y, *_ = np.linalg.lstsq(Hm[:k + 1, :k], e1, rcond=None)
            err = float(np.linalg.norm(Hm[:k + 1, :k] @ y - e1))
            self.ErrorEstimateList[cur] = err
        k = cur + 1
        Yf = np.asfortranarray(np.asarray(y[:k], dtype=np.complex128).reshape(k, 1))
        nat.check(nat.lib().pp_basis_combine(dev.ptr(V), self.MatrixSize, k,
                                             Yf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, dev.ptr(sol),
                                             dev.stream_ptr()), "pp_basis_combine")
        self.ErrorEstimate = err
        self.HessenbergMatrix = Hm[:k + 1, :k].copy()

    def GetErrorEstimate(self):
        return abs(self.ErrorEstimate)

    def GetErrorEstimateList(self):
        return self.ErrorEstimateList

    def GetHessenbergMatrix(self):
        return self.HessenbergMatrix

    def GetOperatorCount(self):
        return self.OperatorCount

    def GetOrthogonalizationCount(self):
        return self.OrthogonalizationCount

    def EstimateMemoryUsage(self, matrixSize, basisSize):
        """MiB of the Arnoldi workspace (gmres.h:197-203)."""
        return ((matrixSize * (basisSize + 3.0)) + basisSize * (3 * basisSize + 7.0)) * 16 / (1024.0 * 1024.0)
