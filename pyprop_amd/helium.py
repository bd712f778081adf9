"""Two-electron potentials in coupled spherical harmonics x two B-spline
radial ranks, in factored form, and their mapping onto the rank-3 engine.

The reference builds each potential of examples/tensor/helium_stabilization/
config.ini as a full rank-3 tensor (angular pairs x radial pairs x radial
pairs) by evaluating it on the product quadrature grid and projecting rank by
rank (pyprop/tensorpotential/TensorPotentialGenerator.py:72-256 with the
custom evaluators of coupledspherical.cpp, coupledvelocity.cpp and
potential.cpp).  Every potential on this path is a sum of terms

  sum_p a_p |c_p><c'_p|  (x)  A (x) B        (one-electron, dipole), or
  sum_p a_p |c_p><c'_p|  (x)  R_l            (electron-electron, one per l),

with banded radial matrices A, B (13 diagonals) and the 4-index multipole
tensor R_l[i1, j1, i2, j2] = sum_{q1,q2} w_q1 B_i1 B_j1 (q1) w_q2 B_i2 B_j2 (q2)
r<^l / r>^(l+1) on the same product grid.  Because the product quadrature
factorises, projecting a separable grid function gives exactly the
Kronecker product of the one-rank projections, so these factors carry the
reference's numbers without forming its tensors (48 GB for the electron-
electron term at 66 channels, 120 B-splines).  ``ReferenceLayout`` expands
a potential to the reference layout for small cases.

Radial band matrices are row-major: element (i, j) at [i, j - i + k - 1],
with the reference's convention for i > j (the transposed integral times
(-1)^d, bspline.cpp:380-392).
"""
import ctypes

import numpy as np

from . import coupled as cp
from . import device as dev
from . import _native as nat


# ---------------------------------------------------------------- radial pieces
def band_rowmajor(basis, f, d):
    """<B_i| f d^d |B_j> as an (N, 2k-1) row-major band (real part)."""
    k = basis.MaxSplineOrder
    N = basis.NumberOfBSplines
    s = k - 1
    bw = 2 * s + 1
    nh = np.real(basis.project_bandnh(f, d))        # BandNH: (i, j) at j*bw + s + i - j
    out = np.zeros((N, bw))
    for i in range(N):
        for dd in range(bw):
            j = i + dd - s
            if 0 <= j < N:
                out[i, dd] = nh[j * bw + s + i - j]
    return out


def radial_matrices(basis):
    """S, D1, D2, 1/r, 1/r^2 as row-major bands."""
    r = basis.QuadratureGridGlobal
    one = np.ones_like(r)
    return {"S": band_rowmajor(basis, one, 0), "D1": band_rowmajor(basis, one, 1),
            "D2": band_rowmajor(basis, one, 2), "R1": band_rowmajor(basis, 1.0 / r, 0),
            "R2": band_rowmajor(basis, 1.0 / (r * r), 0)}


def radial_supports(basis):
    """(lo, hi): the first and last breakpoint interval on which each
    B-spline is nonzero (from the basis tables at the quadrature points)."""
    k = basis.MaxSplineOrder
    N = basis.NumberOfBSplines
    iv = np.arange(len(basis.QuadratureGridGlobal)) // basis.nq
    lo = np.full(N, np.iinfo(np.int32).max, dtype=np.int64)
    hi = np.full(N, -1, dtype=np.int64)
    for r in range(k):
        i = basis.Span + r
        ok = (i >= 0) & (i < N) & (basis.B[:, r] != 0.0)
        np.minimum.at(lo, i[ok], iv[ok])
        np.maximum.at(hi, i[ok], iv[ok])
    return lo.astype(np.int32), hi.astype(np.int32)


def multipole_tensor(basis, l):
    """R_l[i1, d1, i2, d2] (j = i + d - (k-1)) of r<^l / r>^(l+1) on the
    product quadrature grid (coupledspherical.cpp:92-110 evaluated, then
    projected on both radial ranks)."""
    k = basis.MaxSplineOrder
    N = basis.NumberOfBSplines
    s = k - 1
    bw = 2 * s + 1
    r = basis.QuadratureGridGlobal
    w = basis.QuadratureWeightsGlobal
    rmin = np.minimum.outer(r, r)
    rmax = np.maximum.outer(r, r)
    F = (rmin / rmax) ** l / rmax
    Fw = F * w[:, None] * w[None, :]
    M = basis._projector(0)                          # (G, N*bw) sparse, BandNH columns
    T = np.asarray((M.T @ Fw))                       # (N*bw, G)
    Rnh = np.asarray((M.T @ T.T)).T                  # (N*bw, N*bw), BandNH x BandNH
    # BandNH index of (i, j) is j*bw + s + i - j; reorder to row-major (i, d)
    idx = np.full((N, bw), -1, dtype=np.int64)
    for i in range(N):
        for dd in range(bw):
            j = i + dd - s
            if 0 <= j < N:
                idx[i, dd] = j * bw + s + i - j
    valid = idx >= 0
    flat = np.where(valid, idx, 0).reshape(-1)
    R = Rnh[np.ix_(flat, flat)].reshape(N, bw, N, bw)
    R *= valid[:, :, None, None]
    R *= valid[None, None, :, :]
    return np.ascontiguousarray(R)


# ---------------------------------------------------------------- factored potentials
class KronTerm:
    def __init__(self, pairs, coef, A, B):
        self.pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        self.coef = np.ascontiguousarray(coef, dtype=np.complex128)
        self.A, self.B = A, B


class MultipoleTerm:
    """R: the 4-index tensor; P, Q: <B|r^l|B> and <B|r^-(l+1)|B>, the two
    factors of R wherever the radial supports do not overlap."""

    def __init__(self, pairs, coef, l, R, P=None, Q=None):
        self.pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        self.coef = np.ascontiguousarray(coef, dtype=np.complex128)
        self.l = l
        self.R = R
        self.P, self.Q = P, Q


class CoupledTensorPotential:
    """A config-section potential on (coupled channels, B-spline, B-spline),
    held as factored terms.  Mirrors TensorPotential's attributes
    (TensorPotential.py:3-163): Name, IsTimeDependent, GetTimeScaling."""

    def __init__(self, psi):
        self.psi = psi
        self.Name = None
        self.IsTimeDependent = False
        self.Terms = []

    def ApplyConfigSection(self, conf):
        repr = self.psi.GetRepresentation()
        self.Name = conf.name
        self.IsTimeDependent = hasattr(conf, "time_function")
        if self.IsTimeDependent:
            self.TimeFunction = lambda t: conf.time_function(conf, t)
            self.OriginalTimeFunction = conf.time_function
        ang = repr.GetRepresentation(0)
        b1 = repr.GetRepresentation(1).GetBSplineObject()
        b2 = repr.GetRepresentation(2).GetBSplineObject()
        if b1 is not b2 and (b1.NumberOfBSplines != b2.NumberOfBSplines
                             or not np.array_equal(b1.Knots, b2.Knots)):
            raise NotImplementedError("both radial ranks must use the same B-spline basis")
        self.Basis = b1
        self.Channels = ang.Channels
        g0 = conf.geometry0
        for r in (1, 2):
            g = getattr(conf, "geometry%i" % r).lower().strip()
            if g not in ("banded-packed", "banded-nonhermitian"):
                raise NotImplementedError("radial geometry %s is not on the rank-3 engine" % g)
        self.Pairs = np.array(cp.basis_pairs(self.Channels, g0), dtype=np.int32).reshape(-1, 2)
        self.GeometryNames = [g0, conf.geometry1, conf.geometry2]
        if hasattr(conf, "filename"):
            # TensorPotential.py:56-71: the terms come from the file, checked
            # against the angular pairs of the current representation
            from .serialization import LoadTensorPotential
            LoadTensorPotential(conf.filename, conf.dataset, self)
        else:
            self.Terms = _evaluate(conf, self.Channels, self.Pairs, self.Basis, repr.RadialMatrices())

    def GetTimeScaling(self, t):
        return float(np.real(self.TimeFunction(t))) if self.IsTimeDependent else 1.0

    def SetupStep(self, timestep):
        pass

    def MultiplyPotential(self, srcPsi, destPsi, t, timestep):
        raise NotImplementedError("two-electron potentials are applied by the rank-3 engine "
                                  "(PamPropagator's Engine.apply), not one by one")

    def CanConsolidate(self, other):
        """Same rule as TensorPotential.CanConsolidate (TensorPotential.py:
        128-163): same time dependence and time function, same storage and
        the same basis pairs on every rank."""
        if not isinstance(other, CoupledTensorPotential):
            return False
        if self.IsTimeDependent != other.IsTimeDependent:
            return False
        if self.IsTimeDependent and self.OriginalTimeFunction != other.OriginalTimeFunction:
            return False
        if [g.lower().strip() for g in self.GeometryNames[1:]] != [g.lower().strip() for g in other.GeometryNames[1:]]:
            return False
        return self.Pairs.shape == other.Pairs.shape and np.array_equal(self.Pairs, other.Pairs)

    def ConsolidateWith(self, other):
        """self += other (the factored terms are concatenated)."""
        self.Terms = list(self.Terms) + list(other.Terms)
        self.Name += "+" + other.Name

    def ReferenceLayout(self):
        """The reference's tensor (angular pairs, radial pairs, radial pairs or
        BandNH) for small cases: expands the factored terms."""
        b = self.Basis
        N = b.NumberOfBSplines
        k = b.MaxSplineOrder
        s = k - 1
        bw = 2 * s + 1
        if len(self.Pairs) * (N * bw) ** 2 > 2e7:
            raise RuntimeError("reference layout of %s is too large to form" % self.Name)
        r1 = _radial_pairs(self.GeometryNames[1], N, k)
        r2 = _radial_pairs(self.GeometryNames[2], N, k)
        out = np.zeros((len(self.Pairs), len(r1), len(r2)), dtype=np.complex128)
        pindex = {tuple(p): n for n, p in enumerate(self.Pairs)}

        def rm(M, prs):
            i, j = prs[:, 0], prs[:, 1]
            ok = np.abs(i - j) <= s
            return np.where(ok, M[i, np.clip(j - i + s, 0, bw - 1)], 0.0)
        for t in self.Terms:
            if isinstance(t, KronTerm):
                rad = np.multiply.outer(rm(t.A, r1), rm(t.B, r2))
            else:
                i1, j1 = r1[:, 0], r1[:, 1]
                i2, j2 = r2[:, 0], r2[:, 1]
                ok1 = np.abs(i1 - j1) <= s
                ok2 = np.abs(i2 - j2) <= s
                rad = t.R[i1[:, None], np.clip(j1 - i1 + s, 0, bw - 1)[:, None],
                          i2[None, :], np.clip(j2 - i2 + s, 0, bw - 1)[None, :]]
                rad = rad * ok1[:, None] * ok2[None, :]
            for p, c in zip(t.pairs, t.coef):
                out[pindex[tuple(p)]] += c * rad
        return out


def _radial_pairs(geom, N, k):
    g = geom.lower().strip()
    s = k - 1
    if g == "banded-packed":
        return np.array([(i, j) for i in range(N) for j in range(max(0, i - k + 1), min(i + k, N))], dtype=np.int64)
    bw = 2 * s + 1
    pr = np.zeros((N * bw, 2), dtype=np.int64)
    pr[:, 1] = N - 1
    for i in range(N):
        for j in range(i - s, i + s + 1):
            if 0 <= j < N:
                pr[j * bw + s + i - j] = (i, j)
    return pr


def _diag(channels, pairs, f):
    coef = np.array([f(channels[a]) for a, b in pairs], dtype=np.complex128)
    return coef


def _evaluate(conf, channels, pairs, basis, mats):
    """Factored terms of one config section, by the section's classname."""
    cls = conf.classname
    S, D1, D2, R1, R2 = mats["S"], mats["D1"], mats["D2"], mats["R1"], mats["R2"]
    ch = channels
    diag_ok = all(a == b for a, b in pairs) if len(pairs) else True
    if cls == "KineticEnergyPotential":
        # -1/(2 mass) with differentiation on one radial rank (potential.cpp:7-35)
        d1 = int(getattr(conf, "differentiation1", 0))
        d2 = int(getattr(conf, "differentiation2", 0))
        dm = {0: S, 1: D1, 2: D2}
        c = np.full(len(pairs), -1.0 / (2.0 * conf.mass), dtype=np.complex128)
        return [KronTerm(pairs, c, dm[d1], dm[d2])]
    if cls == "OverlapPotential":
        return [KronTerm(pairs, np.ones(len(pairs), dtype=np.complex128), S, S)]
    if cls == "CoupledSphericalKineticEnergyEvaluator":
        # l1(l1+1)/(2 m r1^2) + l2(l2+1)/(2 m r2^2) on matching channels
        # (coupledspherical.cpp:120-205)
        m = conf.mass
        match = np.array([ch[a] == ch[b] for a, b in pairs])
        c1 = np.array([ch[a].l1 * (ch[a].l1 + 1.0) / (2.0 * m) for a, b in pairs]) * match
        c2 = np.array([ch[a].l2 * (ch[a].l2 + 1.0) / (2.0 * m) for a, b in pairs]) * match
        return [KronTerm(pairs, c1, R2, S), KronTerm(pairs, c2, S, R2)]
    if cls == "CoupledSphericalCoulombPotential":
        # -Z/r1 - Z/r2 (potential.cpp:41-77)
        if not diag_ok:
            raise NotImplementedError("CoupledSphericalCoulombPotential on a non-diagonal angular geometry")
        c = np.full(len(pairs), -float(conf.z), dtype=np.complex128)
        return [KronTerm(pairs, c, R1, S), KronTerm(pairs, c, S, R1)]
    if cls == "CustomPotentialEvaluationR12":
        # sum_l coefficient(l) r<^l / r>^(l+1) (coupledspherical.cpp:16-114)
        terms = []
        lmax = 0
        for a, b in pairs:
            lo, hi = cp.multipole_range(ch[a], ch[b])
            lmax = max(lmax, hi)
        for l in range(lmax + 1):
            sel, coef = [], []
            for a, b in pairs:
                lo, hi = cp.multipole_range(ch[a], ch[b])
                if lo <= l <= hi:
                    v = cp.ee_coefficient(ch[a], ch[b], l)
                    if v != 0.0:
                        sel.append((a, b))
                        coef.append(v)
            if sel:
                r = basis.QuadratureGridGlobal
                terms.append(MultipoleTerm(sel, np.array(coef, dtype=np.complex128), l,
                                           multipole_tensor(basis, l), band_rowmajor(basis, r ** l, 0),
                                           band_rowmajor(basis, r ** (-(l + 1.0)), 0)))
        return terms
    if cls in ("CustomPotential_LaserVelocity", "CustomPotential_LaserVelocityDerivativeR1",
               "CustomPotential_LaserVelocityDerivativeR2"):
        cc = np.array([cp.velocity_couplings(ch[a], ch[b]) for a, b in pairs]).reshape(-1, 4)
        mi = -1j
        if cls == "CustomPotential_LaserVelocity":        # -i (c1/r1 + c2/r2)
            return [KronTerm(pairs, mi * cc[:, 0], R1, S), KronTerm(pairs, mi * cc[:, 1], S, R1)]
        if cls == "CustomPotential_LaserVelocityDerivativeR1":   # -i e1 d/dr1
            if int(getattr(conf, "differentiation1", 0)) != 1:
                raise NotImplementedError("LaserVelocityDerivativeR1 needs differentiation1 = 1")
            return [KronTerm(pairs, mi * cc[:, 2], D1, S)]
        if int(getattr(conf, "differentiation2", 0)) != 1:
            raise NotImplementedError("LaserVelocityDerivativeR2 needs differentiation2 = 1")
        return [KronTerm(pairs, mi * cc[:, 3], S, D1)]
    raise NotImplementedError("potential class %s is not on the rank-3 path" % cls)


# ---------------------------------------------------------------- engine
class Pamp3Engine:
    """pp_pamp3_*: pAMP steps of a (channels, N, N) state on the device."""

    def __init__(self, C, N, m, overlap, reorth=False):
        self.C, self.N, self.m = int(C), int(N), int(m)
        self.overlap = overlap
        h = ctypes.c_void_p()
        nat.check(nat.lib().pp_pamp3_create(self.C, self.N, self.m, int(bool(reorth)), overlap.h,
                                            ctypes.byref(h)), "pp_pamp3_create")
        self.h = h
        self._token = nat.track("pp_pamp3_destroy", h)
        self.nslots = 0
        self._last_steps = 0

    def __del__(self, _release=dev._release):   # bound now: module globals are gone at teardown
        _release(self)

    def add_kron(self, term, slot=-1):
        coef = term.coef.view(np.float64)
        A = np.ascontiguousarray(term.A, dtype=np.float64)
        B = np.ascontiguousarray(term.B, dtype=np.float64)
        nat.check(nat.lib().pp_pamp3_add_kron_term(self.h, len(term.pairs), nat.iptr(term.pairs), nat.dptr(coef),
                                                   nat.dptr(A), nat.dptr(B), int(slot)), "pp_pamp3_add_kron_term")
        self.nslots = max(self.nslots, slot + 1)

    def add_multipole(self, term, slot=-1):
        coef = term.coef.view(np.float64)
        R = np.ascontiguousarray(term.R, dtype=np.float64)
        nul = ctypes.POINTER(ctypes.c_double)()
        P = Q = None
        if term.P is not None:
            P = np.ascontiguousarray(term.P, dtype=np.float64)
            Q = np.ascontiguousarray(term.Q, dtype=np.float64)
        nat.check(nat.lib().pp_pamp3_add_multipole_term(self.h, len(term.pairs), nat.iptr(term.pairs),
                                                        nat.dptr(coef), nat.dptr(R),
                                                        nat.dptr(P) if P is not None else nul,
                                                        nat.dptr(Q) if Q is not None else nul, int(slot)),
                  "pp_pamp3_add_multipole_term")
        self.nslots = max(self.nslots, slot + 1)

    def set_radial_supports(self, lo, hi):
        """Enables the near/far split of the multipole terms."""
        lo = np.ascontiguousarray(lo, dtype=np.int32)
        hi = np.ascontiguousarray(hi, dtype=np.int32)
        self._sup = (lo, hi)
        nat.check(nat.lib().pp_pamp3_set_radial_supports(self.h, nat.iptr(lo), nat.iptr(hi)),
                  "pp_pamp3_set_radial_supports")

    def add_potential(self, pot, slot=-1):
        for t in pot.Terms:
            if isinstance(t, KronTerm):
                self.add_kron(t, slot)
            else:
                self.add_multipole(t, slot)

    def _check(self, x):
        t = dev.torch()
        if tuple(x.shape) != (self.C, self.N, self.N) or x.dtype != t.complex128 or not x.is_contiguous():
            raise ValueError("expected a contiguous complex128 (%d, %d, %d) device tensor" % (self.C, self.N, self.N))

    def _scal(self, scal, nsteps):
        if self.nslots == 0:
            return None
        sc = np.ascontiguousarray(np.asarray(scal, dtype=np.float64).reshape(nsteps, self.nslots))
        return sc

    def apply(self, src, dst, scal=None, balanced=False):
        """dst = H src (balanced: U^-H H U^-1 src); src, dst device (C, N, N)."""
        self._check(src)
        self._check(dst)
        sc = self._scal(scal if scal is not None else np.ones(max(self.nslots, 1)), 1)
        nat.check(nat.lib().pp_pamp3_apply(self.h, dev.ptr(src), dev.ptr(dst), nat.dptr(sc) if sc is not None else None,
                                           int(bool(balanced)), dev.stream_ptr()), "pp_pamp3_apply")

    def propagate(self, psi, nsteps, dt, scal=None, check=True):
        """nsteps pAMP steps of psi; check: raise on a zero wavefunction or
        a failed Pade step (reads the diagnostics, synchronises)."""
        self._check(psi)
        dt = complex(dt)
        sc = self._scal(scal, nsteps) if self.nslots else None
        nat.check(nat.lib().pp_pamp3_propagate(self.h, dev.ptr(psi), int(nsteps), dt.real, dt.imag,
                                               nat.dptr(sc) if sc is not None else None, dev.stream_ptr()),
                  "pp_pamp3_propagate")
        self._last_steps = int(nsteps)
        if check and nsteps > 0:
            dev.check_pamp_steps(self.diagnostics(), "pAMP (two-electron)")

    def set_timing(self, enable=True):
        nat.check(nat.lib().pp_pamp3_set_timing(self.h, int(bool(enable))), "pp_pamp3_set_timing")

    def set_exponentiation(self, method):
        """Exponential of the Krylov matrix: 0 scaling + Pade(6), 1 the
        eigenvector exponential (krylov_exponentiation_method)."""
        nat.check(nat.lib().pp_pamp3_set_exponentiation(self.h, int(method)), "pp_pamp3_set_exponentiation")

    def set_mixed(self, enable=True):
        """Mixed form: the multipole terms in FP32 with complex64
        intermediates (everything else FP64); needs the near/far split."""
        nat.check(nat.lib().pp_pamp3_set_mixed(self.h, int(bool(enable))), "pp_pamp3_set_mixed")

    def kernel_time(self):
        """(summed ms, launches) of the timed multipole contraction launches."""
        ms = ctypes.c_double()
        cnt = ctypes.c_long()
        nat.check(nat.lib().pp_pamp3_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(cnt)), "pp_pamp3_kernel_time")
        return ms.value, cnt.value

    def multipole_flops(self):
        f = ctypes.c_double()
        nat.check(nat.lib().pp_pamp3_multipole_flops(self.h, ctypes.byref(f)), "pp_pamp3_multipole_flops")
        return f.value

    def diagnostics(self):
        n = self._last_steps
        out = np.zeros(4 * max(n, 1))
        nat.check(nat.lib().pp_pamp3_diagnostics(self.h, n, nat.dptr(out), dev.stream_ptr()), "pp_pamp3_diagnostics")
        return out[:4 * n].reshape(n, 4)


# ---------------------------------------------------------------- preconditioner
def band_dense(band):
    """Dense N x N matrix of a row-major band (N, 2k-1), j = i + d - (k-1)."""
    N, bw = band.shape
    s = (bw - 1) // 2
    out = np.zeros((N, N))
    for d in range(bw):
        i = np.arange(max(0, s - d), min(N, N + s - d))
        out[i, i + d - s] = band[i, d]
    return out


class RadialTwoElectronPreconditioner:
    """Left preconditioner of the two-electron Cayley and shift-invert
    solves: x = M^-1 y with M = scalingS S + scalingH (P1 + P2 + ...) per
    coupled channel (pyprop's interface,
    examples/tensor/helium_stabilization/preconditioner.py:77-140: the
    section's potential_evaluation list and the config's OverlapPotential
    section, SetHamiltonianScaling / SetOverlapScaling, Setup(prop),
    Solve(psi); the potentials must be diagonal in the angular rank).

    The reference factorises each channel's N^2 x N^2 radial block (SuperLU,
    or an incomplete LU with Ifpack).  Here the block's separable part, b S
    (x) S + a (H1_c (x) S + S (x) H2_c) with H1_c, H2_c the listed terms that
    carry S on the other radial rank (radial and angular kinetic energy,
    nuclear attraction), is inverted exactly by fast diagonalisation on the
    device (pp_radial_prec_*): the S-orthonormal eigenvectors of each
    distinct H_c come from a host eigensolver once, and a solve is four dense
    N x N products per channel.  Listed terms that are not separable (the
    channel-diagonal electron-electron multipoles) are left out of M and
    named in OmittedTerms; M stays a preconditioner, GMRES solves the full
    system."""

    def __init__(self, psi):
        self.Rank = psi.GetRank()
        self.psi = psi
        self.HamiltonianScaling = 1.0
        self.OverlapScaling = 1.0
        self.OverlapSection = None
        self.PotentialSections = []
        self.OmittedTerms = []
        self._h = None

    def __del__(self, _release=dev._release):
        _release(self)

    def ApplyConfigSection(self, conf):
        cfg = conf.Config
        self.OverlapSection = getattr(cfg, "OverlapPotential", None)
        self.PotentialSections = [cfg.GetSection(s) for s in conf.potential_evaluation]

    def SetHamiltonianScaling(self, scalingH):
        self.HamiltonianScaling = scalingH

    def SetOverlapScaling(self, scalingS):
        self.OverlapScaling = scalingS

    def GetHamiltonianScaling(self):
        return self.HamiltonianScaling

    def GetOverlapScaling(self):
        return self.OverlapScaling

    def _radial_terms(self, prop, S):
        """(H1, H2): (C, N, N) real matrices of the listed potentials' terms
        A (x) S and S (x) B, summed per channel."""
        repr = self.psi.GetRepresentation()
        C, N, _ = repr.GetFullShape()
        H1 = np.zeros((C, N, N))
        H2 = np.zeros((C, N, N))
        self.OmittedTerms = []
        for conf in self.PotentialSections:
            if conf.classname == "CustomPotentialEvaluationR12" and str(conf.geometry0).strip().lower() == "diagonal":
                # only multipole terms, none separable: not generated at all
                self.OmittedTerms.append("%s (electron-electron multipoles)" % conf.name)
                continue
            pot = prop.BasePropagator.GeneratePotential(conf)
            if not isinstance(pot, CoupledTensorPotential):
                raise NotImplementedError("RadialTwoElectronPreconditioner needs two-electron potentials")
            for row, col in pot.Pairs:
                if row != col:
                    raise ValueError("%i != %i, angular rank must be diagonal for potential %s" % (row, col, pot.Name))
            for t in pot.Terms:
                if not isinstance(t, KronTerm):
                    self.OmittedTerms.append("%s (multipole l = %d)" % (pot.Name, t.l))
                    continue
                if np.any(t.coef.imag != 0.0):
                    raise NotImplementedError("potential %s has complex coefficients; the preconditioner's "
                                              "eigenvectors need real symmetric radial matrices" % pot.Name)
                ch = t.pairs[:, 0]
                if np.array_equal(t.B, S):
                    np.add.at(H1, ch, t.coef.real[:, None, None] * band_dense(t.A)[None])
                elif np.array_equal(t.A, S):
                    np.add.at(H2, ch, t.coef.real[:, None, None] * band_dense(t.B)[None])
                else:
                    self.OmittedTerms.append("%s (A (x) B term)" % pot.Name)
        return H1, H2

    def Decompose(self, prop):
        """(idx1, idx2, V, lam): the S-orthonormal generalised eigenvectors
        H V = S V diag(lam), V^T S V = I, of each distinct radial matrix of
        the listed potentials (host), and the matrix index of each channel
        on radial rank 1 and 2."""
        import scipy.linalg
        repr = self.psi.GetRepresentation()
        C, N, N2 = repr.GetFullShape()
        if self.Rank != 3 or N != N2:
            raise NotImplementedError("RadialTwoElectronPreconditioner needs (channels, N, N) two-electron states")
        S = repr.RadialMatrices()["S"]
        if self.OverlapSection is not None:
            ov = prop.BasePropagator.GeneratePotential(self.OverlapSection)
            ok = (isinstance(ov, CoupledTensorPotential) and len(ov.Terms) == 1 and isinstance(ov.Terms[0], KronTerm)
                  and np.array_equal(ov.Terms[0].A, S) and np.array_equal(ov.Terms[0].B, S)
                  and np.array_equal(ov.Terms[0].pairs[:, 0], np.arange(C))
                  and np.array_equal(ov.Terms[0].pairs[:, 1], np.arange(C)) and np.all(ov.Terms[0].coef == 1.0))
            if not ok:
                raise NotImplementedError("the OverlapPotential section must be the overlap S (x) S on every channel")
        H1, H2 = self._radial_terms(prop, S)
        Sd = band_dense(S)
        mats, idx, uniq = {}, [], []
        for H in list(H1) + list(H2):
            key = H.tobytes()
            if key not in mats:
                mats[key] = len(uniq)
                uniq.append(H)
            idx.append(mats[key])
        lam = np.zeros((len(uniq), N))
        V = np.zeros((len(uniq), N, N))
        for m, H in enumerate(uniq):
            if np.abs(H - H.T).max() > 1e-12 * max(np.abs(H).max(), 1.0):
                raise NotImplementedError("radial matrix %d of the preconditioner is not symmetric" % m)
            lam[m], V[m] = scipy.linalg.eigh(0.5 * (H + H.T), Sd)
        return np.array(idx[:C], dtype=np.int32), np.array(idx[C:], dtype=np.int32), V, lam

    def Setup(self, prop):
        """Decompose the listed potentials and set up the device solver for
        the scalings set now (the reference factorises here too,
        preconditioner.py:111-140)."""
        idx1, idx2, V, lam = self.Decompose(prop)
        C, N = len(idx1), V.shape[1]
        a = complex(self.HamiltonianScaling)
        b = complex(self.OverlapScaling)
        absD = np.abs(b + a * (lam[idx1][:, :, None] + lam[idx2][:, None, :]))
        if not np.all(np.isfinite(absD)) or absD.min() <= 1e-14 * absD.max():
            c = int(np.argmin(absD.reshape(C, -1).min(axis=1)))
            raise RuntimeError("preconditioner matrix of channel %d is singular" % c)
        dev._release(self)
        h = ctypes.c_void_p()
        V = np.ascontiguousarray(V)
        lam = np.ascontiguousarray(lam)
        nat.check(nat.lib().pp_radial_prec_create(C, N, len(V), nat.iptr(idx1), nat.iptr(idx2), nat.dptr(V),
                                                  nat.dptr(lam), ctypes.byref(h)), "pp_radial_prec_create")
        self.h = h
        self._token = nat.track("pp_radial_prec_destroy", h)
        self._a, self._b = a, b
        self._shape = (C, N, N)

    def Solve(self, psi):
        data = psi.GetData()
        if getattr(self, "_token", None) is None:
            raise RuntimeError("call Setup(prop) first")
        t = dev.torch()
        if tuple(data.shape) != self._shape or data.dtype != t.complex128 or not data.is_contiguous():
            raise ValueError("expected a contiguous complex128 %s device tensor" % (self._shape,))
        nat.check(nat.lib().pp_radial_prec_solve(self.h, dev.ptr(data), self._a.real, self._a.imag,
                                                 self._b.real, self._b.imag, dev.stream_ptr()),
                  "pp_radial_prec_solve")


class RadialTwoElectronPreconditionerSuperLU(RadialTwoElectronPreconditioner):
    """Config name of the exact variant (preconditioner.py:154-198); the
    device solve is the same."""


class RadialTwoElectronPreconditionerIfpack(RadialTwoElectronPreconditioner):
    """Config name of the incomplete-LU variant (preconditioner.py:200-244);
    its drop `cutoff` is read and kept, the device solve is exact for the
    separable part, so there is nothing to drop."""

    def ApplyConfigSection(self, conf):
        RadialTwoElectronPreconditioner.ApplyConfigSection(self, conf)
        self.Cutoff = conf.cutoff
