"""CPU oracle for pyprop's tensor-potential Krylov propagation path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  It wraps the C
restatement in pyprop_oracle.c (built by the Makefile beside it) and adds the
operator assembly the reference performs in Python:

* potential list, per-potential projection and consolidation
  (pyprop/tensorpotential/BasisPropagator.py:41-114, TensorPotential.py:128-163,
  TensorPotentialGenerator.py:72-256);
* the hydrogen / SAE evaluators of examples/tensor/helium_stabilization_sae/
  potential.cpp:7-35,112-314 and the time function LaserFunctionVelocity
  (example.py:220-230);
* the 1s start state by inverse iteration on the l=0 block (the reference
  offers shift-invert Arnoldi, pyprop/solver/PiramSolver.py:36-69; for a
  single block this is the same fixed point);
* the Cayley (Crank-Nicolson) step with the reference's single-cycle GMRES
  (`gmres_solve`, `cayley_step`; numpy, dense operators on small shapes).
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ST_IDENT, ST_DIAG, ST_SIMP, ST_HERM, ST_BANDNH, ST_BAND = range(6)


_BUILD = "liborc.so"
# builds of pyprop_oracle.c (oracle/Makefile): the checker the tests pin
# (-O2, x86-64-v2), and two -O3 builds for the CPU baseline timing
_BUILDS = {"checker": "liborc.so", "v3": "liborc_v3.so", "v4": "liborc_v4.so"}
_ISA = {"v3": {"avx2", "fma", "bmi1", "bmi2", "f16c", "movbe", "abm"},
        "v4": {"avx2", "fma", "bmi1", "bmi2", "f16c", "movbe", "abm",
               "avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}}


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def fastest_build():
    """The -O3 build whose instruction set this host has (x86-64-v4:
    AVX-512; v3: AVX2/FMA), else the checker build."""
    try:
        with open("/proc/cpuinfo") as fh:
            flags = set()
            for line in fh:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        return "checker"
    for name in ("v4", "v3"):
        if _ISA[name] <= flags and os.path.exists(os.path.join(_HERE, _BUILDS[name])):
            return name
    return "checker"


def use_build(name):
    """Select the build later lib() calls load ("checker", "v3", "v4");
    only before the first lib() call of the process."""
    global _BUILD
    if _LIB is not None and _BUILDS[name] != _BUILD:
        raise RuntimeError("the oracle library is already loaded")
    _BUILD = _BUILDS[name]
    return name


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, _BUILD)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_bspline_create.restype = P
        L.orc_bspline_create.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_bspline_destroy.argtypes = [P]
        L.orc_bspline_count.argtypes = [P]
        L.orc_bspline_grid_size.argtypes = [P]
        L.orc_bspline_grid.argtypes = [P, dp, dp]
        L.orc_bspline_value.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_bspline_value.restype = ctypes.c_double
        L.orc_bspline_integral.argtypes = [P, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_bspline_project_bandnh.argtypes = [P, dp, dp, ctypes.c_int, dp]
        L.orc_bspline_overlap_upper.argtypes = [P, dp]
        L.orc_zpbtrf_upper.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.orc_band_apply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, dp, dp]
        L.orc_angular_tables.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, dp, dp]
        L.orc_tp_multiply.argtypes = [ctypes.c_int, ip, ip, dp, ctypes.c_double, ip, dp, dp,
                                      ctypes.POINTER(ip)]
        L.orc_ham_create.restype = P
        L.orc_ham_create.argtypes = [ctypes.c_int, ip]
        L.orc_ham_set_overlap.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
        L.orc_ham_add_term.argtypes = [P, ip, ip, dp, ctypes.POINTER(ip)]
        L.orc_ham_add_factored.argtypes = [P, ctypes.c_int, ctypes.c_int, ip, dp, dp, dp, dp, ctypes.c_int]
        L.orc_ham_add_factored.restype = ctypes.c_int
        L.orc_ham_set_scaling.argtypes = [P, dp]
        L.orc_ham_destroy.argtypes = [P]
        L.orc_ham_balanced.argtypes = [P, dp, dp]
        L.orc_ham_potentials.argtypes = [P, dp, dp]
        L.orc_ham_inner.argtypes = [P, dp, dp, dp]
        L.orc_pade_exp.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.orc_scaling_pade_exp.argtypes = [ctypes.c_int, dp, dp]
        L.orc_pamp_create.restype = P
        L.orc_pamp_create.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_int]
        L.orc_pamp_destroy.argtypes = [P]
        L.orc_pamp_step.argtypes = [P, P, dp, ctypes.c_double, ctypes.c_double]
        L.orc_pamp_error_estimate.argtypes = [P]
        L.orc_pamp_error_estimate.restype = ctypes.c_double
        L.orc_propagate_pamp.argtypes = [P, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int, ctypes.c_int]
        L.orc_propagate_pamp_ex.argtypes = [P, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int, ctypes.c_int,
                                            _EXPFN]
        L.orc_zgpadm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp, ip]
        L.orc_zgexpv.argtypes = [P, ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, ip]
        L.orc_propagate_expokit.argtypes = [P, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double, ip]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def set_threads(n):
    """OpenMP threads of the oracle's parallel regions (independent outputs
    only, so results do not depend on n); returns the previous count."""
    prev = lib().orc_get_threads()
    lib().orc_set_threads(int(n))
    return prev


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


# --------------------------------------------------------------------------
# breakpoints (pyprop/bspline/breakpointsequences.py:31-41,6-28,72-93)
# --------------------------------------------------------------------------
def linear_breakpoints(rmin, rmax, n):
    return np.array([rmin + (rmax - rmin) * i / float(n - 1) for i in range(n)])


def exponential_breakpoints(rmin, rmax, n, gamma):
    h = rmax - rmin
    return np.array([rmin + h * (np.exp(gamma * float(i - 1) / float(n - 1)) - 1) / (np.exp(gamma) - 1)
                     for i in range(1, n + 1)])


def _py2_round(x):
    """Python 2 round(): halves away from zero (the reference runs on Python 2)."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def exponential_linear_breakpoints(rmin, rpartition, rmax, n, gamma):
    """breakpointsequences.py:72-90: exponential inner points, then
    pylab.frange(rpartition + h, rmax, h), which holds
    round((rmax - rpartition - h)/h) + 1 points (matplotlib.mlab.frange,
    closed interval)."""
    xi = [rmin + (rpartition - rmin) * (math.exp(gamma * float(i - 1) / float(n - 1)) - 1) / (math.exp(gamma) - 1)
          for i in range(1, n + 1)]
    h = xi[-1] - xi[-2]
    x0 = rpartition + h
    npts = int(_py2_round((rmax - x0) / h)) + 1
    for i in range(npts):
        xi.append(x0 + i * h)
    return np.array(xi)


def center_exponential_linear_breakpoints(rmin, rpartition, rmax, n, gamma):
    """breakpointsequences.py:92-121: the exponential-linear sequence from the
    centre point outwards, then mirrored through zero."""
    right = list(exponential_linear_breakpoints((rmax + rmin) / 2.0, rpartition, rmax, n, gamma))
    return np.array([-x for x in reversed(right[1:])] + right)


def quadratic_linear_breakpoints(rmin, rmax, n, joinPoint):
    """breakpointsequences.py:41-70."""
    rmin, rmax = float(rmin), float(rmax)
    r0 = (rmax * (joinPoint - 1) + rmin * (n - joinPoint)) / (2.0 * n - joinPoint - 1.0)
    alpha = (r0 - rmin) / float(joinPoint - 1) ** 2
    beta = (rmax - r0) / float(n - joinPoint)
    xi = [rmin + alpha * (i - 1) ** 2 for i in range(1, joinPoint)]
    xi += [r0 + beta * (i - joinPoint) for i in range(joinPoint, n + 1)]
    return np.array(xi)


class BSpline:
    """B-spline basis of bspline.cpp restated in C; 'zero' continuity by default."""

    def __init__(self, breakpoints, order, continuity="zero", quad_add=0):
        bp = np.ascontiguousarray(breakpoints, dtype=np.float64)
        ends = {"zero": (1, 1), "vanilla": (0, 0)}[continuity]
        self.order = order
        self.breakpoints = bp
        self.h = lib().orc_bspline_create(_dp(bp), len(bp), order, ends[0], ends[1], quad_add)
        self.N = lib().orc_bspline_count(self.h)
        G = lib().orc_bspline_grid_size(self.h)
        self.x = np.zeros(G)
        self.w = np.zeros(G)
        lib().orc_bspline_grid(self.h, _dp(self.x), _dp(self.w))

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_bspline_destroy(self.h)
            self.h = None

    def value(self, x, i, deriv=0):
        return lib().orc_bspline_value(self.h, float(x), int(i), int(deriv))

    def integral(self, f, d, i, j):
        f = np.asarray(f, dtype=np.complex128)
        re = np.ascontiguousarray(f.real)
        im = np.ascontiguousarray(f.imag)
        out = np.zeros(2)
        lib().orc_bspline_integral(self.h, _dp(re), _dp(im), d, i, j, _dp(out))
        return out[0] + 1j * out[1]

    def project_bandnh(self, f, d):
        f = np.broadcast_to(np.asarray(f, dtype=np.complex128), self.x.shape)
        re = np.ascontiguousarray(f.real)
        im = np.ascontiguousarray(f.imag)
        bw = 2 * (self.order - 1) + 1
        out = np.zeros(self.N * bw, dtype=np.complex128)
        lib().orc_bspline_project_bandnh(self.h, _dp(re), _dp(im), d, _dp(out.view(np.float64)))
        return out

    def overlap_upper(self):
        kd = self.order - 1
        out = np.zeros(self.N * (kd + 1), dtype=np.complex128)
        lib().orc_bspline_overlap_upper(self.h, _dp(out.view(np.float64)))
        return out


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def zpbtrf_upper(ab, N, kd):
    a = np.array(ab, dtype=np.complex128)
    info = lib().orc_zpbtrf_upper(N, kd, _dp(a.view(np.float64)))
    return a, info


OP_SOLVE_N, OP_SOLVE_C, OP_MUL_N = 0, 1, 2


def band_apply(op, rank, x, U, kd):
    x = np.array(x, dtype=np.complex128, order="C")
    shape = np.array(x.shape, dtype=np.int32)
    U = np.ascontiguousarray(U, dtype=np.complex128)
    lib().orc_band_apply(op, rank, x.ndim, _ip(shape), kd, _dp(U.view(np.float64)), _dp(x.view(np.float64)))
    return x


def angular_tables(lmax, m=0):
    nth = lmax + 1
    L = lmax + 1 - abs(m)
    x = np.zeros(nth)
    w = np.zeros(nth)
    P = np.zeros((nth, L))
    D = np.zeros((nth, L))
    lib().orc_angular_tables(lmax, m, _dp(x), _dp(w), _dp(P), _dp(D))
    return x, w, P, D


def dipole_pairs(lmax, m=0):
    """GeometryInfoReducedSphHarmSelectionRule.GetBasisPairs (ReducedSphericalHarmonic.py:20-37)."""
    count = lmax - abs(m) + 1
    pairs = []
    for i in range(count):
        if i > 0:
            pairs.append((i, i - 1))
        if i < count - 1:
            pairs.append((i, i + 1))
    return np.array(pairs, dtype=np.int32).reshape(-1, 2)


def tp_multiply(storage, pot, scaling, src, dst, pairs=None):
    R = src.ndim
    st = np.array(storage, dtype=np.int32)
    ps = np.array(pot.shape, dtype=np.int32)
    sh = np.array(src.shape, dtype=np.int32)
    pot = np.ascontiguousarray(pot, dtype=np.complex128)
    src = np.ascontiguousarray(src, dtype=np.complex128)
    keep = []
    arr = (ctypes.POINTER(ctypes.c_int) * R)()
    for r in range(R):
        if pairs is not None and pairs[r] is not None:
            p = np.ascontiguousarray(pairs[r], dtype=np.int32)
            keep.append(p)
            arr[r] = _ip(p)
        else:
            arr[r] = ctypes.POINTER(ctypes.c_int)()
    assert dst.flags.c_contiguous and dst.dtype == np.complex128
    rc = lib().orc_tp_multiply(R, _ip(st), _ip(ps), _dp(pot.view(np.float64)), float(scaling), _ip(sh),
                               _dp(src.view(np.float64)), _dp(dst.view(np.float64)), arr)
    if rc:
        raise RuntimeError("orc_tp_multiply failed %d" % rc)
    return dst


# --------------------------------------------------------------------------
# time function (examples/tensor/helium_stabilization_sae/example.py:220-230)
# --------------------------------------------------------------------------
def laser_velocity(t, amplitude, frequency, pulse_duration, phase=0.0):
    if 0 <= t < pulse_duration:
        f = amplitude
        f *= np.sin(t * np.pi / pulse_duration) ** 2
        f *= -np.cos(t * frequency + phase)
        return f
    return 0.0


def step_times(nsteps, dt, t0=0.0):
    """PropagatedTime sequence of Problem.AdvanceStep (Problem.py:158-173)."""
    ts = []
    t = t0
    for _ in range(nsteps):
        ts.append(t)
        t += dt
    return ts


# --------------------------------------------------------------------------
# hydrogen operator in the SAE example's form, potential list
#   RadialKineticEnergy, AngularKineticEnergy, SAEPotential,
#   LaserPotentialVelocity1..3  (config_paper.ini:41,79-140 + stabilization.py:44)
# --------------------------------------------------------------------------
class Hydrogen:
    def __init__(self, rmin=0.0, rmax=100.0, nbp=100, order=7, lmax=10, m=0, z=1.0,
                 a=(0, 0, 0, 0, 0, 0), soft=0.0, mass=1.0, charge=-1.0,
                 frequency=0.057, cycles=6, E0=0.05, phase=0.0, quad_add=0,
                 bpstype="linear", gamma=1.0):
        if bpstype == "linear":
            bp = linear_breakpoints(rmin, rmax, nbp)
        else:
            bp = exponential_breakpoints(rmin, rmax, nbp, gamma)
        self.bs = BSpline(bp, order, "zero", quad_add)
        self.N = N = self.bs.N
        self.k = order
        self.kd = order - 1
        self.bw = 2 * (order - 1) + 1
        self.lmax, self.m = lmax, m
        self.L = L = lmax + 1 - abs(m)
        self.frequency, self.cycles = frequency, cycles
        self.pulse_duration = 2 * np.pi / frequency * cycles
        self.amplitude = E0 / frequency
        self.phase = phase
        r = self.bs.x
        ls = np.arange(abs(m), lmax + 1)
        # ---- Diag x BandNH, time independent (consolidated in list order)
        kin = self.bs.project_bandnh(-1.0 / (2 * mass) * np.ones_like(r), 2)
        r2 = self.bs.project_bandnh(1.0 / (r * r), 0)
        a1, a2, a3, a4, a5, a6 = a
        sae = self.bs.project_bandnh(-(z + a1 * np.exp(-a2 * r) + a3 * r * np.exp(-a4 * r)
                                       + a5 * np.exp(-a6 * r)) / np.sqrt(r * r + soft * soft), 0)
        H0 = np.zeros((L, N * self.bw), dtype=np.complex128)
        for li, l in enumerate(ls):
            H0[li] += kin
        for li, l in enumerate(ls):
            H0[li] += l * (l + 1.0) / (2 * mass) * r2
        for li, l in enumerate(ls):
            H0[li] += sae
        self.H0 = H0
        # ---- Simp x BandNH, time dependent (LaserFunctionVelocity)
        x, w, P, D = angular_tables(lmax, m)
        pairs = dipole_pairs(lmax, m)
        self.pairs = pairs
        cos_ang = np.array([np.sum(w * P[:, a_] * x * P[:, b_]) for a_, b_ in pairs])
        der_ang = np.array([np.sum(w * P[:, a_] * D[:, b_]) for a_, b_ in pairs])
        self.cos_ang, self.der_ang = cos_ang, der_ang
        d1 = self.bs.project_bandnh(np.ones_like(r), 1)
        r1 = self.bs.project_bandnh(1.0 / r, 0)
        Hi = np.zeros((len(pairs), N * self.bw), dtype=np.complex128)
        Hi += (charge * 1j) * cos_ang[:, None] * d1[None, :]           # Velocity1: charge*i*cos, d/dr
        Hi += (-charge * 1j) * der_ang[:, None] * r1[None, :]          # Velocity2: -charge*i/r, angular deriv
        Hi += (-charge * 1j) * cos_ang[:, None] * r1[None, :]          # Velocity3: -charge*i*cos/r
        self.Hint = Hi
        # ---- overlap + Cholesky
        self.S = self.bs.overlap_upper()
        self.U, info = zpbtrf_upper(self.S, N, self.kd)
        if info:
            raise RuntimeError("overlap not positive definite (%d)" % info)
        self.shape = (L, N)

    def scaling(self, t):
        return laser_velocity(t, self.amplitude, self.frequency, self.pulse_duration, self.phase)

    def ham(self):
        return Ham(self.shape, [(ST_DIAG, ST_BANDNH, self.H0, None),
                                (ST_SIMP, ST_BANDNH, self.Hint, self.pairs)],
                   {1: (self.U, self.kd)})

    # dense helpers (for small checks)
    def band_to_dense(self, band):
        N, s, bw = self.N, self.kd, self.bw
        M = np.zeros((N, N), dtype=np.complex128)
        for j in range(N):
            for i in range(max(0, j - s), min(N, j + s + 1)):
                M[i, j] = band[j * bw + s + i - j]
        return M

    def overlap_dense(self):
        N, kd = self.N, self.kd
        S = np.zeros((N, N), dtype=np.complex128)
        for j in range(N):
            for i in range(max(0, j - kd), j + 1):
                S[i, j] = self.S[j * (kd + 1) + kd + i - j]
                S[j, i] = np.conj(S[i, j])
        return S

    def ground_state(self, l=0, shift=-0.5 + 1e-3, iterations=5):
        """1s start state: inverse iteration on the l block, S-normalised."""
        A = self.band_to_dense(self.H0[l - abs(self.m)])
        S = self.overlap_dense()
        x = np.ones(self.N, dtype=np.complex128)
        M = A - shift * S
        for _ in range(iterations):
            x = np.linalg.solve(M, S @ x)
            x /= np.sqrt(np.real(np.conj(x) @ S @ x))
        i = np.argmax(np.abs(x))
        x *= np.abs(x[i]) / x[i]
        psi = np.zeros(self.shape, dtype=np.complex128)
        psi[l - abs(self.m)] = x
        return psi

    def bound_states(self, l, emax=0.0):
        """Per-l generalized eigenproblem (analysis.py:92-140), S-normalised."""
        import scipy.linalg
        A = self.band_to_dense(self.H0[l - abs(self.m)])
        S = self.overlap_dense()
        E, V = scipy.linalg.eigh(A, S)
        sel = E < emax
        return E[sel], V[:, sel]

    def populations(self, psi, emax=0.0):
        """P_bound = sum_l sum_{E<0} |c^H S psi_l|^2 (analysis.py:344-364)."""
        S = self.overlap_dense()
        tot = 0.0
        per = []
        for li in range(self.L):
            E, V = self.bound_states(li + abs(self.m), emax)
            c = V.conj().T @ (S @ psi[li])
            p = np.abs(c) ** 2
            per.append(p)
            tot += p.sum()
        return tot, per

    def norm(self, psi):
        S = self.overlap_dense()
        return float(np.sqrt(np.real(np.sum(np.conj(psi) * (psi @ S.T)))))


class Ham:
    """Balanced Hamiltonian U^-H (sum_t s_t V_t) U^-1 held by the C oracle."""

    def __init__(self, shape, terms, overlaps):
        self.shape = tuple(shape)
        sh = np.array(shape, dtype=np.int32)
        self.h = lib().orc_ham_create(len(shape), _ip(sh))
        self.nterms = 0
        self._keep = []
        for rank, (U, kd) in overlaps.items():
            U = np.ascontiguousarray(U, dtype=np.complex128)
            lib().orc_ham_set_overlap(self.h, rank, kd, _dp(U.view(np.float64)))
        for t in terms:
            st = t[:-2]
            pot, pairs = t[-2], t[-1]
            R = len(shape)
            st = np.array(st, dtype=np.int32)
            ps = np.array(pot.shape, dtype=np.int32)
            arr = (ctypes.POINTER(ctypes.c_int) * R)()
            for r in range(R):
                arr[r] = ctypes.POINTER(ctypes.c_int)()
            if pairs is not None:
                pr = pairs if isinstance(pairs, (list, tuple)) else [pairs] + [None] * (R - 1)
                for r in range(R):
                    if pr[r] is not None:
                        p = np.ascontiguousarray(pr[r], dtype=np.int32)
                        self._keep.append(p)
                        arr[r] = _ip(p)
            pot = np.ascontiguousarray(pot, dtype=np.complex128)
            lib().orc_ham_add_term(self.h, _ip(st), _ip(ps), _dp(pot.view(np.float64)), arr)
            self.nterms += 1

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_ham_destroy(self.h)
            self.h = None

    def set_scaling(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        lib().orc_ham_set_scaling(self.h, _dp(s))

    def add_factored(self, pairs, coef, A=None, B=None, R=None, s=6):
        """Factored two-electron term on a (C, N, N) state: sum_p coef_p
        |a_p><b_p| (x) A (x) B (row-major real bands) or (x) R[i1][d1][i2][d2].
        Used where the reference layout cannot be formed (the helium
        workload's CPU baseline); returns the term index."""
        pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        cf = np.ascontiguousarray(coef, dtype=np.complex128)
        kind = 1 if R is not None else 0
        keep = [pr, cf]
        a = b = r = None
        if kind == 0:
            a = np.ascontiguousarray(A, dtype=np.float64)
            b = np.ascontiguousarray(B, dtype=np.float64)
            keep += [a, b]
        else:
            r = np.ascontiguousarray(R, dtype=np.float64)
            keep.append(r)
        nul = ctypes.POINTER(ctypes.c_double)()
        t = lib().orc_ham_add_factored(self.h, kind, len(pr), _ip(pr), _dp(cf.view(np.float64)),
                                       _dp(a) if a is not None else nul, _dp(b) if b is not None else nul,
                                       _dp(r) if r is not None else nul, int(s))
        if t < 0:
            raise RuntimeError("orc_ham_add_factored failed")
        self.nterms += 1
        return t

    def balanced(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        lib().orc_ham_balanced(self.h, _dp(x.view(np.float64)), _dp(y.view(np.float64)))
        return y

    def potentials(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        lib().orc_ham_potentials(self.h, _dp(x.view(np.float64)), _dp(y.view(np.float64)))
        return y

    def inner(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        b = np.ascontiguousarray(b, dtype=np.complex128)
        out = np.zeros(2)
        lib().orc_ham_inner(self.h, _dp(a.view(np.float64)), _dp(b.view(np.float64)), _dp(out))
        return out[0] + 1j * out[1]

    def zgexpv(self, v, t, m, tol=0.0, anorm=1e6, imag_time=False):
        """w = exp(t A) v, A = -i U^-H H U^-1 (zgexpv.cpp:32-586); returns (w, iflag, stats)."""
        v = np.ascontiguousarray(v, dtype=np.complex128)
        w = np.zeros_like(v)
        st = np.zeros(5, dtype=np.int32)
        rc = lib().orc_zgexpv(self.h, int(m), float(t), _dp(v.view(np.float64)), _dp(w.view(np.float64)),
                              float(tol), float(anorm), int(bool(imag_time)), _ip(st))
        return w, rc, st

    def propagate_expokit(self, psi, nsteps, dt, scalings, m, tol=0.0, anorm=1e6):
        psi = np.array(psi, dtype=np.complex128, order="C")
        sc = np.ascontiguousarray(np.asarray(scalings, dtype=np.float64).reshape(nsteps, self.nterms))
        st = np.zeros(5, dtype=np.int32)
        rc = lib().orc_propagate_expokit(self.h, _dp(psi.view(np.float64)), nsteps, float(dt), _dp(sc), int(m),
                                         float(tol), float(anorm), _ip(st))
        if rc:
            raise RuntimeError("Expokit returned nonzero value %d" % rc)
        return psi, st

    def propagate_pamp(self, psi, nsteps, dt, scalings, m, reorth=False, exponentiation=0):
        """exponentiation: 0 scaling + Pade (pamp.h:430-564), 1 the
        eigenvector exponential (pamp.h:567-607, `eigenvector_exponential`)."""
        psi = np.array(psi, dtype=np.complex128, order="C")
        sc = np.ascontiguousarray(np.asarray(scalings, dtype=np.float64).reshape(nsteps, self.nterms))
        if exponentiation:
            rc = lib().orc_propagate_pamp_ex(self.h, _dp(psi.view(np.float64)), nsteps, float(dt), _dp(sc),
                                             int(m), int(bool(reorth)), _EIG_EXPFN)
        else:
            rc = lib().orc_propagate_pamp(self.h, _dp(psi.view(np.float64)), nsteps, float(dt), _dp(sc),
                                          int(m), int(bool(reorth)))
        if rc:
            raise RuntimeError("oracle pAMP failed (%d)" % rc)
        return psi


def gmres_solve(apply, b, basis_size, tol=1.4901161193847656e-08, double_orth=True, max_orth=3):
    """x with apply(x) ~ b by gmres::GMRES<cplx>::SolveVector
    (core/krylov/gmres/gmres.h:430-500): zero start, one cycle (no restart)
    of at most basis_size Arnoldi vectors, Euclidean inner products (the
    BLAS integration, gmreswrapper.cpp:91-96); the initial step
    orthogonalises once (:257-290); later steps by classical Gram-Schmidt
    with the 0.717 re-orthogonalisation rule, at most max_orth extra passes
    (:293-400); tolerance sqrt(eps) unless set, floored at eps^(2/3)
    (:67, :221-225: krylov_tolerance = 0 gives eps^(2/3)); the loop stops at
    happy breakdown (|beta| < tol) or when the least-squares residual
    ||H y - ||b|| e1|| is <= tol (an absolute residual, :459-490).
    Returns (x, error estimate, operator count)."""
    b = np.asarray(b, dtype=np.complex128).reshape(-1)
    tol = max(tol, np.finfo(float).eps ** (2.0 / 3.0))     # default sqrt(eps) (gmres.h:67)
    m = int(basis_size)
    n = b.size
    V = np.zeros((m, n), dtype=np.complex128)
    Hm = np.zeros((m + 1, m), dtype=np.complex128)       # Hm[i, j] = h_ij
    bnorm = np.linalg.norm(b)
    V[0] = b / bnorm
    r = apply(V[0])
    nops = 1
    alpha = np.vdot(V[0], r)
    r = r - alpha * V[0]
    Hm[0, 0] = alpha
    beta = np.linalg.norm(r)
    happy = beta < tol
    Hm[1, 0] = beta
    cur = 0
    y = np.array([bnorm / alpha]) if happy else np.zeros(1, dtype=np.complex128)
    err = beta
    while cur < m - 1 and not happy and abs(err) > tol:
        j = cur
        V[j + 1] = r / Hm[j + 1, j]
        r = apply(V[j + 1])
        nops += 1
        orig = np.linalg.norm(r)
        h = V[:j + 2].conj() @ r
        r = r - V[:j + 2].T @ h
        res = np.linalg.norm(r)
        count = 0
        while double_orth and res < 0.717 * orig:
            count += 1
            if count > max_orth:
                break
            h2 = V[:j + 2].conj() @ r
            r = r - V[:j + 2].T @ h2
            h = h + h2
            orig, res = res, np.linalg.norm(r)
        Hm[:j + 2, j + 1] = h
        beta = np.linalg.norm(r)
        happy = beta < tol
        Hm[j + 2, j + 1] = beta
        cur += 1
        k = cur + 1
        e1 = np.zeros(k + 1, dtype=np.complex128)
        e1[0] = bnorm
        y, *_ = np.linalg.lstsq(Hm[:k + 1, :k], e1, rcond=None)
        err = np.linalg.norm(Hm[:k + 1, :k] @ y - e1)
    k = cur + 1
    return V[:k].T @ y[:k], err, nops


def cayley_step(apply_h, S_dense, M_solve, psi, dt, basis_size, tol=1.4901161193847656e-08):
    """One CayleyPropagator.AdvanceStep (pyprop/propagator/CayleyPropagator.py:73-140):
    rhs = M^-1 (S - i dt/2 H) psi, then psi' solves M^-1 (S + i dt/2 H) psi'
    = rhs by gmres_solve; H = apply_h(x) is the raw (non-balanced)
    Hamiltonian at t + |dt/2|, S the overlap: an N x N matrix applied along
    the last rank, or a callable (two-electron states: both radial ranks),
    M_solve the left preconditioner (identity when None).  Returns (psi',
    error estimate, operator count)."""
    psi = np.asarray(psi, dtype=np.complex128)
    shape = psi.shape
    prec = M_solve if M_solve is not None else (lambda x: x)
    apply_s = S_dense if callable(S_dense) else (lambda x: x @ S_dense.T)

    def sph(x, f):
        x = x.reshape(shape)
        return (apply_s(x) + f * apply_h(x)).reshape(-1)

    rhs = prec(sph(psi, -0.5j * dt).reshape(shape)).reshape(-1)
    x, err, nops = gmres_solve(lambda v: prec(sph(v, 0.5j * dt).reshape(shape)).reshape(-1), rhs, basis_size, tol)
    return x.reshape(shape), err, nops


def zgpadm(A, t, ideg=6):
    """Expokit zgpadm (zgpadm.cpp:28-285): exp(t A), returns (E, ns)."""
    A = np.array(A, dtype=np.complex128)
    m = A.shape[0]
    a = A.reshape(-1, order="F").copy()
    out = np.zeros(m * m, dtype=np.complex128)
    ns = ctypes.c_int(0)
    rc = lib().orc_zgpadm(ideg, m, float(t), _dp(a.view(np.float64)), _dp(out.view(np.float64)), ctypes.byref(ns))
    if rc:
        raise RuntimeError("zgpadm failed %d" % rc)
    return out.reshape((m, m), order="F"), ns.value


def pade_exp(A, order=6):
    A = np.asfortranarray(A, dtype=np.complex128)
    m = A.shape[0]
    out = np.zeros((m, m), dtype=np.complex128, order="F")
    info = lib().orc_pade_exp(m, _dp(A.reshape(-1, order="F").view(np.float64)),
                              _dp(out.reshape(-1, order="F").view(np.float64)), order)
    return out, info


def eigenvector_exponential(A):
    """pAMP::EigenvectorExponential (pamp.h:567-607): R diag(exp(lambda)) R^H
    with lambda, R the eigenvalues and unit right eigenvectors of A from
    LAPACK's zgeev (numpy.linalg.eig calls the same routine).  pAMP stores its
    Hessenberg matrix so that the column-major view zgeev and zgemm see is H
    itself (pamp.h:289,315,345 write H(j, j+1) = beta into a row-major
    array), so this is the matrix of A, not of its transpose."""
    A = np.asarray(A, dtype=np.complex128)
    lam, R = np.linalg.eig(A)
    return (R * np.exp(lam)[None, :]) @ R.conj().T


EIG_EXPFN_CALLS = [0]


def _eig_expfn(m, a_ptr, out_ptr):
    a = np.ctypeslib.as_array(a_ptr, shape=(2 * m * m,)).view(np.complex128).reshape((m, m), order="F")
    E = eigenvector_exponential(a)
    out = np.ctypeslib.as_array(out_ptr, shape=(2 * m * m,)).view(np.complex128)
    out[:] = E.reshape(-1, order="F")
    EIG_EXPFN_CALLS[0] += 1
    return 0


_EXPFN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double))
_EIG_EXPFN = _EXPFN(_eig_expfn)


def scaling_pade_exp(A):
    A = np.array(A, dtype=np.complex128, order="F")
    m = A.shape[0]
    a = A.reshape(-1, order="F").copy()
    out = np.zeros(m * m, dtype=np.complex128)
    info = lib().orc_scaling_pade_exp(m, _dp(a.view(np.float64)), _dp(out.view(np.float64)))
    return out.reshape((m, m), order="F"), info


# --------------------------------------------------------------------------
# helium: coupled spherical harmonics x two B-spline radial ranks
# (examples/tensor/helium_stabilization/config.ini, coupledspherical.cpp,
# coupledvelocity.cpp, potential.cpp; pyprop/CoupledSphericalHarmonics.py)
# --------------------------------------------------------------------------
def coupled_channels(lmax, Ls, Ms=(0,)):
    """DefaultCoupledIndexIterator order (CoupledSphericalHarmonics.py:52-63)."""
    out = []
    for M in Ms:
        for L in Ls:
            for l1 in range(lmax + 1):
                for l2 in range(lmax + 1):
                    if abs(l1 - l2) <= L <= l1 + l2:
                        out.append((l1, l2, L, M))
    return out


def _fact(n):
    return float(np.prod(np.arange(1, n + 1, dtype=np.float64))) if n > 1 else 1.0


def threej(j1, j2, j3, m1, m2, m3):
    """Wigner 3j by the Racah sum in floating point (GSL's
    gsl_sf_coupling_3j is what the reference calls, coupledrange.h:17-32)."""
    if m1 + m2 + m3 != 0 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    if j3 > j1 + j2 or j3 < abs(j1 - j2):
        return 0.0
    tri = _fact(j1 + j2 - j3) * _fact(j1 - j2 + j3) * _fact(-j1 + j2 + j3) / _fact(j1 + j2 + j3 + 1)
    pre = np.sqrt(tri * _fact(j1 + m1) * _fact(j1 - m1) * _fact(j2 + m2) * _fact(j2 - m2)
                  * _fact(j3 + m3) * _fact(j3 - m3))
    s = 0.0
    for k in range(0, j1 + j2 + j3 + 1):
        a = [k, j3 - j2 + k + m1, j3 - j1 + k - m2, j1 + j2 - j3 - k, j1 - k - m1, j2 - k + m2]
        if min(a) < 0:
            continue
        s += (-1.0) ** k / np.prod([_fact(x) for x in a])
    return (-1.0) ** (j1 - j2 - m3) * pre * s


def cg(l1, l2, m1, m2, L, M):
    """ClebschGordan::operator() (coupledrange.h:21-31)."""
    return (-1.0) ** (M + l1 - l2) * np.sqrt(2 * L + 1.0) * threej(l1, l2, L, m1, m2, -M)


def ee_l3sum(left, right, l3):
    """The l3 coefficient of CustomPotentialEvaluationR12::UpdatePotentialData
    (coupledspherical.cpp:53-90): left = (l1', l2', L, M), right = (l1, l2, L, M)."""
    l1p, l2p, L, M = left
    l1, l2 = right[0], right[1]
    coeff = np.sqrt((2. * l1p + 1.) * (2. * l1 + 1.) / (2. * l3 + 1.))
    coeff *= np.sqrt((2. * l2p + 1.) * (2. * l2 + 1.) / (2. * l3 + 1.))
    coeff *= 1. / (2 * l3 + 1.)
    coeff *= cg(l1p, l1, 0, 0, l3, 0)
    coeff *= cg(l2p, l2, 0, 0, l3, 0)
    s = 0.0
    lstop = max(l1, l1p, l2, l2p)
    for m1p in range(-lstop, lstop + 1):
        m2p = M - m1p
        for m1 in range(-lstop, lstop + 1):
            m2 = M - m1
            m3 = m1 - m1p
            cur = cg(l1p, l2p, m1p, m2p, L, M) * cg(l1, l2, m1, m2, L, M)
            cur *= cg(l1p, l1, -m1p, m1, l3, m3) * cg(l2p, l2, -m2p, m2, l3, -m3)
            cur *= (-1.0) ** (m1p + m2p + m3)
            s += cur
    return s * coeff


def _lh(l, m):
    C = l * np.sqrt(((l + 1 + m) * (l + 1 - m)) / ((2 * l + 1.) * (2 * l + 3)))
    D = -(l + 1) * np.sqrt(max(0.0, ((l + m) * (l - m)) / ((2 * l + 1.) * (2 * l - 1)))) if l > 0 else 0.0
    E = np.sqrt(((l + 1 + m) * (l + 1 - m)) / ((2 * l + 1.) * (2 * l + 3)))
    F = np.sqrt(max(0.0, ((l + m) * (l - m)) / ((2 * l + 1.) * (2 * l - 1)))) if l > 0 else 0.0
    return C, D, E, F


def velocity_coefficients(left, right):
    """(coupling1, coupling2) of CustomPotential_LaserVelocity and the
    couplings of the DerivativeR1 / DerivativeR2 parts
    (coupledvelocity.cpp:96-146, :258-276, and the R2 part)."""
    l1, l2, L, M = left
    l1p, l2p, Lp, Mp = right
    if M != Mp:
        return 0.0, 0.0, 0.0, 0.0
    lstop = max(l1, l1p, l2, l2p)
    c1 = c2 = d1 = d2 = 0.0
    for m in range(-lstop, lstop + 1):
        cur = cg(l1, l2, m, M - m, L, M) * cg(l1p, l2p, m, Mp - m, Lp, Mp)
        if l2 == l2p and abs(m) <= l1 and abs(m) <= l1p:
            C, D, E, F = _lh(l1p, m)
            C *= (l1 == l1p + 1); E *= (l1 == l1p + 1)
            D *= (l1 == l1p - 1); F *= (l1 == l1p - 1)
            c1 += (-(C + D) - (E + F)) * cur
            d1 += (E + F) * cur
        if l1 == l1p and abs(m) <= l2 and abs(m) <= l2p:
            C, D, E, F = _lh(l2p, M - m)
            C *= (l2 == l2p + 1); E *= (l2 == l2p + 1)
            D *= (l2 == l2p - 1); F *= (l2 == l2p - 1)
            c2 += (-(C + D) - (E + F)) * cur
            d2 += (E + F) * cur
    return c1, c2, d1, d2


class Helium:
    """Reference-layout helium operator for small cases: every potential of
    the helium config evaluated on (angular pairs, r1 grid, r2 grid) and
    projected on r2 then r1 (TensorPotentialGenerator.py:72-256)."""

    def __init__(self, lmax=1, Ls=(0, 1), nbp=8, rmax=10.0, order=7, z=2.0, mass=1.0,
                 frequency=2.0, cycles=8, E0=0.05, bpstype="linear", gamma=1.0, quad_add=0,
                 laser=True):
        bp = linear_breakpoints(0.0, rmax, nbp) if bpstype == "linear" else \
            exponential_breakpoints(0.0, rmax, nbp, gamma)
        self.bs = BSpline(bp, order, "zero", quad_add)
        self.N = N = self.bs.N
        self.k, self.kd = order, order - 1
        s = order - 1
        self.ch = ch = coupled_channels(lmax, Ls)
        self.C = C = len(ch)
        r = self.bs.x
        G = len(r)
        # basis tables at the quadrature points
        val = np.array([[self.bs.value(x, i, 0) for x in r] for i in range(N)])
        der = {d: np.array([[self.bs.value(x, i, d) for x in r] for i in range(N)]) for d in (1, 2)}
        der[0] = val
        self._tab = der
        self.packed = np.array([(i, j) for i in range(N) for j in range(max(0, i - order + 1), min(i + order, N))])
        bwid = 2 * s + 1
        nh = np.zeros((N * bwid, 2), dtype=np.int64)
        nh[:, 1] = N - 1
        for i in range(N):
            for j in range(i - s, i + s + 1):
                if 0 <= j < N:
                    nh[j * bwid + s + i - j] = (i, j)
        self.bandnh = nh
        diag = [(a, a) for a in range(C)]
        linpol = [(a, b) for a in range(C) for b in range(C)
                  if abs(ch[a][2] - ch[b][2]) == 1 and ch[a][3] == ch[b][3]
                  and ((abs(ch[a][0] - ch[b][0]) == 1 and ch[a][1] == ch[b][1])
                       or (abs(ch[a][1] - ch[b][1]) == 1 and ch[a][0] == ch[b][0]))]
        r12 = [(a, b) for a in range(C) for b in range(C)
               if ch[a][2] == ch[b][2] and ch[a][3] == ch[b][3]]
        self.pairs = {"diag": np.array(diag), "linpol": np.array(linpol).reshape(-1, 2), "r12": np.array(r12)}
        R1 = r[:, None] * np.ones((1, G))
        R2 = np.ones((G, 1)) * r[None, :]
        one = np.ones((G, G))
        terms = []
        # RadialKineticEnergy1/2 (KineticEnergyPotential, differentiation 2 on one rank)
        kin = np.full((C, G, G), -1.0 / (2 * mass))
        terms.append(("static", "diag", self._project(kin, diag, 2, 0, "packed")))
        terms.append(("static", "diag", self._project(kin, diag, 0, 2, "packed")))
        # AngularKineticEnergy (coupledspherical.cpp:160-205)
        ang = np.array([ch[a][0] * (ch[a][0] + 1.) / (2 * mass) / R1 ** 2
                        + ch[a][1] * (ch[a][1] + 1.) / (2 * mass) / R2 ** 2 for a, b in diag])
        terms.append(("static", "diag", self._project(ang, diag, 0, 0, "packed")))
        # CoupledSphericalCoulombPotential (potential.cpp:71-76)
        cou = np.array([-z / R1 - z / R2 for a, b in diag])
        terms.append(("static", "diag", self._project(cou, diag, 0, 0, "packed")))
        # CustomPotentialEvaluationR12 (coupledspherical.cpp:16-114), r2 banded-nonhermitian
        rmin = np.minimum(R1, R2)
        rmax_ = np.maximum(R1, R2)
        ee = np.zeros((len(r12), G, G))
        for n, (a, b) in enumerate(r12):
            l1p, l2p = ch[a][0], ch[a][1]
            l1, l2 = ch[b][0], ch[b][1]
            lo = max(abs(l1 - l1p), abs(l2 - l2p))
            hi = min(l1 + l1p, l2 + l2p)
            for l3 in range(lo, hi + 1):
                ee[n] += ee_l3sum(ch[a], ch[b], l3) * (rmin / rmax_) ** l3 / rmax_
        terms.append(("static", "r12", self._project(ee, r12, 0, 0, "bandnh")))
        self.laser = laser and len(linpol) > 0
        if self.laser:
            vc = np.array([velocity_coefficients(ch[a], ch[b]) for a, b in linpol])
            v0 = np.array([-1j * (vc[n, 0] / R1 + vc[n, 1] / R2) for n in range(len(linpol))])
            terms.append(("laser", "linpol", self._project(v0, linpol, 0, 0, "packed")))
            v1 = np.array([-1j * vc[n, 2] * one for n in range(len(linpol))])
            terms.append(("laser", "linpol", self._project(v1, linpol, 1, 0, "packed")))
            v2 = np.array([-1j * vc[n, 3] * one for n in range(len(linpol))])
            terms.append(("laser", "linpol", self._project(v2, linpol, 0, 1, "packed")))
        self.terms = terms
        self.frequency, self.cycles = frequency, cycles
        self.pulse_duration = 2 * np.pi / frequency * cycles
        self.amplitude = E0 / frequency
        self.S = self.bs.overlap_upper()
        self.U, info = zpbtrf_upper(self.S, N, self.kd)
        if info:
            raise RuntimeError("overlap not positive definite (%d)" % info)
        self.shape = (C, N, N)

    def _w(self, prs, d):
        """W[p, q] = w_q B_i(q) B_j^(d)(q) for i <= j, (-1)^d B_j B_i^(d) else."""
        B, D = self._tab[0], self._tab[d]
        i, j = prs[:, 0], prs[:, 1]
        lo = np.where(i <= j, i, j)
        hi = np.where(i <= j, j, i)
        sign = np.where(i <= j, 1.0, (-1.0) ** d)
        return sign[:, None] * B[lo] * D[hi] * self.bs.w[None, :]

    def _project(self, data, angpairs, d1, d2, r2layout):
        """data (P0, G, G) on the product grid -> (P0, packed pairs, r2 pairs)."""
        W1 = self._w(self.packed, d1)
        W2 = self._w(self.packed if r2layout == "packed" else self.bandnh, d2)
        t = np.einsum("pab,qb->paq", np.asarray(data, dtype=np.complex128), W2)
        return np.einsum("paq,ra->prq", t, W1)

    def scaling(self, t):
        return laser_velocity(t, self.amplitude, self.frequency, self.pulse_duration)

    def ham(self):
        out = []
        for kind, pk, pot in self.terms:
            r2 = ST_SIMP if pot.shape[2] == len(self.packed) else ST_BANDNH
            pr = [self.pairs[pk].astype(np.int32), self.packed.astype(np.int32),
                  self.packed.astype(np.int32) if r2 == ST_SIMP else None]
            out.append((ST_SIMP, ST_SIMP, r2, pot, pr))
        return Ham(self.shape, out, {1: (self.U, self.kd), 2: (self.U, self.kd)})

    def term_scalings(self, times):
        return np.array([[self.scaling(t) if kind == "laser" else 1.0 for kind, _, _ in self.terms]
                         for t in times])
