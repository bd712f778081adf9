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
}


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
    prop = problem.Propagator.BasePropagator
    repr = problem.psi.GetRepresentation()
    L, N = repr.GetFullShape()
    bs = repr.GetRepresentation(1).GetBSplineObject()
    S = bs.overlap_dense()
    H = np.zeros((N, N), dtype=np.complex128)
    li = l - repr.GetRepresentation(0).M
    for pot in prop.PotentialList:
        if pot.IsTimeDependent or pot.StorageIds() != ["Diag", "BandNH"]:
            continue
        bw = pot.PotentialData.shape[1] // N
        H += _block(pot, li, N, bw)
    x = np.ones(N, dtype=np.complex128)
    M = H - shift * S
    for _ in range(iterations):
        x = np.linalg.solve(M, S @ x)
        x /= np.sqrt(np.real(np.conj(x) @ S @ x))
    i = np.argmax(np.abs(x))
    x *= np.abs(x[i]) / x[i]
    out = np.zeros((L, N), dtype=np.complex128)
    out[li] = x
    return out


def populations(problem, psi, emax=0.0):
    """Bound populations per l: sum_{E<0} |c^H S psi_l|^2 with c the S-normalised
    eigenvectors of (H0_l, S) (analysis.py:92-140,344-364)."""
    import scipy.linalg
    prop = problem.Propagator.BasePropagator
    repr = problem.psi.GetRepresentation()
    L, N = repr.GetFullShape()
    bs = repr.GetRepresentation(1).GetBSplineObject()
    S = bs.overlap_dense()
    per = []
    for li in range(L):
        H = np.zeros((N, N), dtype=np.complex128)
        for pot in prop.PotentialList:
            if pot.IsTimeDependent or pot.StorageIds() != ["Diag", "BandNH"]:
                continue
            H += _block(pot, li, N, pot.PotentialData.shape[1] // N)
        E, V = scipy.linalg.eigh(H, S)
        V = V[:, E < emax]
        c = V.conj().T @ (S @ psi[li])
        per.append(np.abs(c) ** 2)
    return sum(p.sum() for p in per), per
