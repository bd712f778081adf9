"""Grid potential evaluators.

In the reference the potential classes named by ``classname`` are C++
templates compiled into a project's libpotential.so and evaluated point by
point by DynamicPotentialEvaluator (core/potential/dynamicpotentialevaluator.h:
129-136).  Here an evaluator is a Python class with ``ApplyConfigSection`` and
a vectorised ``GetPotentialValue(pos)``, where ``pos`` is a tuple with one
broadcastable numpy array per rank.  Lookup order: the project namespace
(pyprop_amd.config.ProjectNamespace), then the classes below, which restate
the single-active-electron example's potentials
(examples/tensor/helium_stabilization_sae/potential.cpp:7-314).
"""
import numpy as np


class PotentialType:
    """pyprop/Enum.py:2-9 (the values config files may give as numbers)."""
    Static = 1
    Dynamic = 3
    FiniteDifference = 4
    CrankNicholson = 5
    Matrix = 6
    RankOne = 7


class PotentialBase:
    def ApplyConfigSection(self, conf):
        self.conf = conf


class KineticEnergyPotential(PotentialBase):
    """-1/(2 mass) (potential.cpp:7-35); used with differentiation = 2."""

    def ApplyConfigSection(self, conf):
        self.mass = conf.mass

    def GetPotentialValue(self, pos):
        return np.full(np.broadcast(*pos).shape, -1.0 / (2.0 * self.mass))


class KineticEnergyPotential2(PotentialBase):
    """-1/(mass sqrt(r^2+soft^2)) (potential.cpp:38-72)."""

    def ApplyConfigSection(self, conf):
        self.mass = conf.mass
        self.radial_rank = conf.radial_rank
        self.soft = conf.soft

    def GetPotentialValue(self, pos):
        r = pos[self.radial_rank]
        return np.broadcast_to(-1.0 / (self.mass * np.sqrt(r * r + self.soft ** 2)), np.broadcast(*pos).shape)


class DipoleLaserPotentialLength(PotentialBase):
    """-charge r cos(theta) (potential.cpp:75-108)."""

    def ApplyConfigSection(self, conf):
        self.angular_rank = conf.angular_rank
        self.radial_rank = conf.radial_rank
        self.charge = conf.charge

    def GetPotentialValue(self, pos):
        return -self.charge * pos[self.radial_rank] * np.cos(pos[self.angular_rank])


class DipoleLaserPotentialVelocityRadialDerivative(DipoleLaserPotentialLength):
    """charge i cos(theta); used with differentiation1 = 1 (potential.cpp:112-144)."""

    def GetPotentialValue(self, pos):
        th = pos[self.angular_rank]
        return np.broadcast_to(self.charge * 1j * np.cos(th), np.broadcast(*pos).shape)


class DipoleLaserPotentialVelocityAngularDerivative(DipoleLaserPotentialLength):
    """-charge i / r; used with differentiation0 = 1 (potential.cpp:146-180)."""

    def GetPotentialValue(self, pos):
        r = pos[self.radial_rank]
        return np.broadcast_to(-self.charge * 1j / r, np.broadcast(*pos).shape)


class DipoleLaserPotentialVelocity(DipoleLaserPotentialLength):
    """-charge i cos(theta) / r (potential.cpp:182-216)."""

    def GetPotentialValue(self, pos):
        return -self.charge * 1j * np.cos(pos[self.angular_rank]) / pos[self.radial_rank]


class AngularKineticEnergyPotential(PotentialBase):
    """l(l+1)/(2 mass r^2), evaluated with l on a Diagonal angular rank (potential.cpp:218-253)."""

    def ApplyConfigSection(self, conf):
        self.angular_rank = conf.angular_rank
        self.radial_rank = conf.radial_rank
        self.mass = conf.mass

    def GetPotentialValue(self, pos):
        l = pos[self.angular_rank]
        r = pos[self.radial_rank]
        return l * (l + 1.0) / (2 * self.mass * r * r)


class SingleActiveElectronPotential(PotentialBase):
    """-(Z + a1 e^{-a2 r} + a3 r e^{-a4 r} + a5 e^{-a6 r}) / sqrt(r^2+soft^2) (potential.cpp:256-314)."""

    def ApplyConfigSection(self, conf):
        self.Z = conf.z
        self.a = [conf.a1, conf.a2, conf.a3, conf.a4, conf.a5, conf.a6]
        self.soft = conf.soft
        self.radial_rank = conf.radial_rank

    def GetPotentialValue(self, pos):
        r = np.abs(pos[self.radial_rank])
        a1, a2, a3, a4, a5, a6 = self.a
        V = -(self.Z + a1 * np.exp(-a2 * r) + a3 * r * np.exp(-a4 * r) + a5 * np.exp(-a6 * r)) / \
            np.sqrt(r * r + self.soft * self.soft)
        return np.broadcast_to(V, np.broadcast(*pos).shape)


class ComplexAbsorbingPotential(PotentialBase):
    """Polynomial-free CAP used by the SAE example's absorber section (potential.cpp:317+):
    factor * ((r - start)/length)^2 inside [start, start+length]."""

    def ApplyConfigSection(self, conf):
        self.radial_rank = conf.radial_rank
        self.start = conf.absorber_start
        self.length = conf.absorber_length
        self.factor = conf.factor_real + 1j * conf.factor_imag
        self.sreal = conf.scaling_real
        self.simag = conf.scaling_imag

    def GetPotentialValue(self, pos):
        r = pos[self.radial_rank]
        x = np.clip((r - self.start) / self.length, 0.0, None)
        V = np.where(r > self.start, self.factor.real * x ** self.sreal + 1j * self.factor.imag * x ** self.simag, 0.0)
        return np.broadcast_to(V, np.broadcast(*pos).shape)


BUILTIN = {c.__name__: c for c in [
    KineticEnergyPotential, KineticEnergyPotential2, DipoleLaserPotentialLength,
    DipoleLaserPotentialVelocityRadialDerivative, DipoleLaserPotentialVelocityAngularDerivative,
    DipoleLaserPotentialVelocity, AngularKineticEnergyPotential, SingleActiveElectronPotential,
    ComplexAbsorbingPotential]}


def CreatePotentialInstance(classname):
    """pyprop/Potential.py:23-67: resolve the evaluator class by name."""
    from . import config
    cls = config.ProjectNamespace.get(classname) or BUILTIN.get(classname)
    if cls is None:
        raise Exception("Unknown potential class '%s'" % classname)
    return cls()
