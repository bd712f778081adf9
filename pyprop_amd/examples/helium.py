"""Two-electron helium in a linearly polarised laser, velocity gauge (the
potentials of examples/tensor/helium_stabilization/config.ini, propagated
with PamPropagator and the sin^2 pulse LaserFunctionVelocity).

Provides the workload configurations and the ground state by imaginary-time
propagation with renormalisation, the reference's route
(examples/tensor/helium_stabilization/example.py:379-390, FindGroundstate:
SetupProblem(imtime=True) and Advance with a negative imaginary time step).
"""
import os

import numpy as np

from . import hydrogen

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = os.path.join(HERE, "configs")

# workload name -> (config file, number of steps)
WORKLOADS = {
    "helium": ("helium.ini", 5000),          # 66 channels, N = 120, m = 20, dt = 0.01
    "helium_small": ("helium_small.ini", 50),
}

FIELD_FREE = ["RadialKineticEnergy1", "RadialKineticEnergy2", "AngularKineticEnergy", "CoulombPotential",
              "ElectronicCouplingPotential"]


def register():
    import pyprop_amd
    pyprop_amd.ProjectNamespace["LaserFunctionVelocity"] = hydrogen.LaserFunctionVelocity


def load(name, overrides=None, field_free=False):
    """Config of a named workload; overrides: {(section, key): value}."""
    import pyprop_amd
    register()
    conf = pyprop_amd.Load(os.path.join(CONFIGS, WORKLOADS[name][0]))
    if field_free:
        conf.SetValue("Propagation", "grid_potential_list", list(FIELD_FREE))
    for (section, key), value in (overrides or {}).items():
        conf.SetValue(section, key, value)
    return conf


def start_state(problem):
    """Product of hydrogen-like 1s functions (Z = 2) in the (0, 0, 0, 0)
    channel, on the B-spline coefficients' Greville abscissae."""
    repr = problem.psi.GetRepresentation()
    b = repr.GetRepresentation(1).GetBSplineObject()
    k = b.MaxSplineOrder
    t = b.Knots
    N = b.NumberOfBSplines
    xi = np.array([t[i + 1:i + k].mean() for i in range(N)])
    f = xi * np.exp(-2.0 * xi)
    psi = np.zeros(problem.psi.GetData().shape, dtype=np.complex128)
    ch = repr.GetRepresentation(0).Channels
    c0 = [i for i, c in enumerate(ch) if tuple(c) == (0, 0, 0, 0)]
    if not c0:
        raise ValueError("the channel list has no (0, 0, 0, 0) channel")
    psi[c0[0]] = np.outer(f, f)
    return psi


def ground_state(name, steps=400, dt=0.05, overrides=None):
    """Imaginary-time pAMP with renormalisation on the field-free operator;
    returns (psi as a numpy array, energy estimate from the last step's norm
    decay)."""
    import pyprop_amd
    over = dict(overrides or {})
    over[("Propagation", "timestep")] = -1j * dt
    over[("Propagation", "renormalization")] = True
    conf = load(name, over, field_free=True)
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    prob.SetWavefunction(start_state(prob))
    prob.psi.Normalize()
    norm = 1.0
    for _ in range(steps):
        prob.Propagator.AdvanceStep(prob.PropagatedTime, prob.TimeStep)
        norm = prob.psi.Normalize()
        prob._advance_time()
    energy = -np.log(norm) / dt
    return prob.psi.GetData().cpu().numpy(), energy
