"""Hydrogen in a linearly polarised laser, velocity gauge (the single-active-
electron example of examples/tensor/helium_stabilization_sae with model "h").

Provides the example's time function, the workload configurations used by
the benchmark and tests, and the 1s start state by inverse iteration on the
l=0 block (the reference reaches it with shift-invert Arnoldi or imaginary
time, pyprop/solver/PiramSolver.py:36-69, example.py:195-205).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = os.path.join(HERE, "configs")


def LaserFunctionVelocity(conf, t):
    """A(t) = A0 sin^2(pi t/T) (-cos(w t + phase)) for 0 <= t < T (example.py:220-230)."""
    phase = getattr(conf, "phase", 0.0)
    if 0 <= t < conf.pulse_duration:
        curField = conf.amplitude
        curField *= np.sin(t * np.pi / conf.pulse_duration) ** 2
        curField *= -np.cos(t * conf.frequency + phase)
    else:
        curField = 0
    return curField


# workload name -> (config file, number of steps)
WORKLOADS = {
    "hydrogen_small": ("hydrogen_small.ini", 2000),   # order 7, 100 breakpoints, l<=10, m=20, dt=0.01
    "hydrogen_large": ("hydrogen_large.ini", 20000),  # order 7, N=800, l<=63, m=24, dt=0.005
    "hydrogen_adjoint": ("hydrogen_adjoint.ini", 4000),  # order 7, N=400, r<=800, l<=31, m=24, dt=0.005
}


def adjoint_control_samples(K, dt, seed=20261203, E0=0.05, omega=0.3, sigma=0.005):
    """Control samples of the adjoint workload: the interaction time factor
    (velocity-gauge vector potential) at t_k = k dt, a sin^2-enveloped
    carrier A0 sin^2(pi t/T) (-cos(omega t)), A0 = E0/omega, T = K dt, plus
    seeded Gaussian noise of sigma on every sample."""
    rng = np.random.default_rng(seed)
    t = dt * np.arange(K)
    T = dt * K
    return (E0 / omega) * np.sin(np.pi * t / T) ** 2 * (-np.cos(omega * t)) + sigma * rng.standard_normal(K)


def register():
    """Put the example's names into the project namespace for config evaluation."""
    import pyprop_amd
    pyprop_amd.ProjectNamespace["LaserFunctionVelocity"] = LaserFunctionVelocity
    pyprop_amd.ProjectNamespace["GroundStateInverseIteration"] = GroundStateInverseIteration


def load(name, overrides=None):
    """Config of a named workload; overrides: {(section, key): value}."""
    import pyprop_amd
    register()
    conf = pyprop_amd.Load(os.path.join(CONFIGS, WORKLOADS[name][0]))
    for (section, key), value in (overrides or {}).items():
        conf.SetValue(section, key, value)
    return conf


def batch_instance_config(draw, krylov=16, dt=0.01, precision="mixed", dr=0.5):
    """Config of one instance of the batched scan: hydrogen_base.ini with the
    drawn radial size N (linear grid, r in [0, dr*N]), lmax and pulse, Krylov
    dimension 16 and dt = 0.01; precision "mixed" or "fp64"."""
    import pyprop_amd
    register()
    k = 7
    extra = """
[AngularRepresentation]
maxl = {lmax}
[RadialRepresentation]
xmax = {rmax!r}
xsize = {xsize}
[PulseParameters]
frequency = {omega!r}
cycles = {cycles}
field_strength = {E0!r}
[Propagation]
krylov_basis_size = {krylov}
timestep = {dt!r}
matvec_precision = "{precision}"
""".format(lmax=draw["lmax"], rmax=dr * draw["N"], xsize=draw["N"] - k + 4, omega=draw["omega"],
           cycles=draw["cycles"], E0=draw["E0"], krylov=krylov, dt=dt, precision=precision)
    return pyprop_amd.Load(os.path.join(CONFIGS, "hydrogen_base.ini"), extraText=extra)


def _block(pot, l_index, N, bw):
    s = (bw - 1) // 2
    band = pot.PotentialData[l_index]
    M = np.zeros((N, N), dtype=np.complex128)
    for j in range(N):
        for i in range(max(0, j - s), min(N, j + s + 1)):
            M[i, j] = band[j * bw + s + i - j]
    return M


def GroundStateInverseIteration(psi, conf):
    """Custom initial condition: S-normalised lowest l=0 state of the
    time-independent Diag x BandNH potentials by inverse iteration,
    (H_0 - shift S) x_{k+1} = S x_k (shift/iterations from the section)."""
    raise RuntimeError("GroundStateInverseIteration needs the problem; call ground_state(problem)")


def ground_state(problem, l=0, shift=-0.5 + 1e-3, iterations=5):
    """Inverse iteration (H0_l - shift S) x_{k+1} = S x_k on the l block of the
    static Diag x BandNH potentials, S-normalised, phase of the largest
    component real (analysis.py:134-138); banded LU of width 2(k-1)."""
    import scipy.linalg
    prop = problem.Propagator.BasePropagator
    repr = problem.psi.GetRepresentation()
    L, N = repr.GetFullShape()
    bs = repr.GetRepresentation(1).GetBSplineObject()
    kd = bs.MaxSplineOrder - 1
    S = bs.overlap_dense()
    li = l - repr.GetRepresentation(0).M
    bw = 2 * kd + 1
    H = np.zeros(N * bw, dtype=np.complex128)
    for pot in prop.PotentialList:
        if pot.IsTimeDependent or pot.StorageIds() != ["Diag", "BandNH"]:
            continue
        if pot.PotentialData.shape[1] != N * bw:
            raise ValueError("radial band width differs from the overlap's")
        H += pot.PotentialData[li]
    # BandNH (j*bw + kd + i - j) -> LAPACK general band ab[kd + i - j, j]
    ab = H.reshape(N, bw).T.copy()
    Sb = np.zeros_like(ab)
    for dgl in range(-kd, kd + 1):
        i = np.arange(max(0, -dgl), min(N, N - dgl))
        Sb[kd - dgl, i + dgl] = S[i, i + dgl]
    M = ab - shift * Sb
    x = np.ones(N, dtype=np.complex128)
    for _ in range(iterations):
        x = scipy.linalg.solve_banded((kd, kd), M, S @ x)
        x /= np.sqrt(np.real(np.conj(x) @ S @ x))
    i = np.argmax(np.abs(x))
    x *= np.abs(x[i]) / x[i]
    out = np.zeros((L, N), dtype=np.complex128)
    out[li] = x
    return out


def bound_state(problem, l, index):
    """The index-th (from 0) S-normalised eigenvector of (H0_l, S) placed in
    channel l of an (L, N) array (index 0 of l = 1 is the 2p state)."""
    import scipy.linalg
    prop = problem.Propagator.BasePropagator
    repr = problem.psi.GetRepresentation()
    L, N = repr.GetFullShape()
    S = repr.GetRepresentation(1).GetBSplineObject().overlap_dense()
    li = l - repr.GetRepresentation(0).M
    Hl = np.zeros((N, N), dtype=np.complex128)
    for pot in prop.PotentialList:
        if pot.IsTimeDependent or pot.StorageIds() != ["Diag", "BandNH"]:
            continue
        Hl += _block(pot, li, N, pot.PotentialData.shape[1] // N)
    E, V = scipy.linalg.eigh(Hl, S)
    out = np.zeros((L, N), dtype=np.complex128)
    out[li] = V[:, index]
    return out, E[index]


def bound_projectors(problem, emax=0.0):
    """Per angular channel, the S-normalised eigenvectors of (H0_l, S) with
    E < emax, and S (analysis.py:92-140): what populations() projects on."""
    import scipy.linalg
    prop = problem.Propagator.BasePropagator
    repr = problem.psi.GetRepresentation()
    L, N = repr.GetFullShape()
    S = repr.GetRepresentation(1).GetBSplineObject().overlap_dense()
    out = []
    for li in range(L):
        H = np.zeros((N, N), dtype=np.complex128)
        for pot in prop.PotentialList:
            if pot.IsTimeDependent or pot.StorageIds() != ["Diag", "BandNH"]:
                continue
            H += _block(pot, li, N, pot.PotentialData.shape[1] // N)
        E, V = scipy.linalg.eigh(H, S, subset_by_value=(-np.inf, emax))
        out.append(V)
    return out, S


def populations(problem, psi, emax=0.0, projectors=None):
    """Bound populations per l: sum_{E<0} |c^H S psi_l|^2 with c the S-normalised
    eigenvectors of (H0_l, S) (analysis.py:92-140,344-364).  projectors: the
    result of bound_projectors(problem, emax), to reuse for several states."""
    Vs, S = projectors if projectors is not None else bound_projectors(problem, emax)
    per = [np.abs(V.conj().T @ (S @ psi[li])) ** 2 for li, V in enumerate(Vs)]
    return sum(p.sum() for p in per), per
