"""Problem: setup and the time loop (pyprop/Problem.py:55-476,
pyprop/CreateInstance.py:9-175)."""
import numpy as np

from . import device as dev


class InitialConditionType:
    Function = 1
    File = 2
    Class = 3
    Custom = 4


class WavefunctionFileFormat:
    """pyprop/Enum.py:21-24 (the values config files may give as numbers)."""
    Ascii = 1
    Binary = 2
    HDF = 3


def CreateRepresentation(config):
    representation = config.Representation.type()
    config.Representation.Apply(representation)
    rank = config.Representation.rank
    for i in range(rank):
        section = config.GetSection(config.Representation.Get("representation%i" % i))
        repr = section.type()
        repr.SetBaseRank(i)
        section.Apply(repr)
        representation.SetRepresentation(i, repr)
    return representation


def CreateWavefunction(config):
    from .core import Wavefunction
    representation = CreateRepresentation(config)
    psi = Wavefunction(representation)
    psi.AllocateData()
    return psi


def CreatePropagator(config, psi):
    propagator = config.Propagation.propagator(psi)
    config.Apply(propagator)
    config.Propagation.Apply(propagator)
    return propagator


def CreatePotential(config, name, psi):
    raise NotImplementedError("grid potentials (potential_evaluation) are not on the tensor path; "
                              "use grid_potential_list")


class Problem:
    def __init__(self, config):
        self.TempPsi = None
        self.Config = config
        self.Silent = getattr(config.Propagation, "silent", False)
        self.psi = CreateWavefunction(config)
        self.Propagator = CreatePropagator(config, self.psi)
        config.Propagation.Apply(self)

    def ApplyConfigSection(self, configSection):
        self.TimeStep = complex(configSection.timestep)
        self.Duration = configSection.duration
        self.PropagatedTime = 0
        self.StartTime = 0
        if hasattr(configSection, "start_time"):
            self.StartTime = configSection["start_time"]
            self.PropagatedTime = self.StartTime

    def GetGrid(self):
        repr = self.psi.GetRepresentation()
        return [repr.GetLocalGrid(i) for i in range(self.psi.GetRank())]

    def SetupStep(self, skipWavefunctionSetup=False):
        if self.Propagator is not None:
            self.Propagator.SetupStep(self.TimeStep)
        if not skipWavefunctionSetup:
            self.SetupWavefunction()

    def RestartPropagation(self, timestep, startTime, propagationTime):
        self.TimeStep = timestep
        self.PropagatedTime = startTime
        self.StartTime = startTime
        self.Duration = propagationTime
        if self.Propagator is not None:
            self.Propagator.RestartPropagation(timestep, startTime, propagationTime)

    def _advance_time(self):
        if abs(self.TimeStep.real) < 1e-10:
            self.PropagatedTime += abs(self.TimeStep)
        else:
            self.PropagatedTime += self.TimeStep.real

    def AdvanceStep(self):
        """Problem.py:158-173."""
        self.Propagator.AdvanceStep(self.PropagatedTime, self.TimeStep)
        if self.Propagator.RenormalizeActive:
            self.psi.Normalize()
        self._advance_time()

    def AdvanceSteps(self, n):
        """n consecutive AdvanceStep()s; with a batching propagator and no
        renormalisation they run in one device launch."""
        if n <= 0:
            return
        if hasattr(self.Propagator, "AdvanceSteps") and not self.Propagator.RenormalizeActive:
            times = []
            for _ in range(n):
                times.append(self.PropagatedTime)
                self._advance_time()
            self.Propagator.AdvanceSteps(times, self.TimeStep)
        else:
            for _ in range(n):
                self.AdvanceStep()

    def MultiplyHamiltonian(self, srcPsi, dstPsi):
        self.Propagator.MultiplyHamiltonian(srcPsi, dstPsi, self.PropagatedTime, self.TimeStep)

    def _stop(self, duration):
        if self.TimeStep.real == 0:
            return lambda: (self.StartTime + duration - self.PropagatedTime) > 0.5 * abs(self.TimeStep)
        endTime = self.StartTime + np.sign(self.TimeStep.real) * duration
        return lambda: abs(self.PropagatedTime - endTime) > 0.5 * abs(self.TimeStep)

    def StepCount(self, duration=None):
        """Number of AdvanceStep()s Advance() performs (the stopping rule of Problem.py:214-220)."""
        if duration is None:
            duration = self.Duration
        saved = self.PropagatedTime
        crit = self._stop(duration)
        n = 0
        while crit():
            self._advance_time()
            n += 1
        self.PropagatedTime = saved
        return n

    def Advance(self, yieldCount, duration=None, yieldEnd=False):
        """Generator over the propagation (Problem.py:182-241); steps between
        yields are launched together."""
        if duration is None:
            duration = self.Duration
        if isinstance(yieldCount, bool):
            yieldStep = 1
        else:
            yieldStep = max(1, int((duration / abs(self.TimeStep)) / yieldCount))
        total = self.StepCount(duration)
        index = 0
        prevYield = np.nan
        while index < total:
            n = min(yieldStep - index % yieldStep, total - index)
            self.AdvanceSteps(n)
            index += n
            if index % yieldStep == 0:
                yield self.PropagatedTime
                prevYield = self.PropagatedTime
        if yieldEnd and prevYield != self.PropagatedTime:
            yield self.PropagatedTime

    def GetTempPsi(self):
        if self.TempPsi is None:
            self.TempPsi = self.psi.CopyDeep()
        return self.TempPsi

    def GetEnergyExpectationValue(self):
        self.psi.Normalize()
        tempPsi = self.GetTempPsi()
        tempPsi.GetData().zero_()
        self.MultiplyHamiltonian(self.psi, tempPsi)
        return self.psi.InnerProduct(tempPsi).real

    def GetEnergy(self):
        return self.GetEnergyExpectationValue()

    # ---------------------------------------------------------- initial state
    def SetupWavefunction(self):
        ic = getattr(self.Config, "InitialCondition", None)
        if ic is None:
            return
        type_ = ic.type
        if type_ == InitialConditionType.Function:
            self.SetupWavefunctionFunction(self.Config)
        elif type_ == InitialConditionType.File:
            self.SetupWavefunctionFile(self.Config)
        elif type_ == InitialConditionType.Class:
            ic.classname(self.Config, self.psi)
        elif type_ == InitialConditionType.Custom:
            ic.function(self.psi, ic)
        elif type_ is None:
            pass
        else:
            raise Exception("Invalid InitialConditionType: %s" % type_)

    def SetupWavefunctionFunction(self, config):
        """Evaluate InitialCondition.function on the representation grids
        (l for the angular rank, the B-spline index for the radial rank)."""
        func = config.InitialCondition.function
        grids = self.GetGrid()
        pos = np.meshgrid(*grids, indexing="ij")
        vals = np.asarray(func(config.InitialCondition, pos), dtype=np.complex128)
        vals = np.broadcast_to(vals, self.psi.GetData().shape)
        self.psi.GetData().copy_(dev.to_device(np.ascontiguousarray(vals)))

    def SetWavefunction(self, data):
        self.psi.GetData().copy_(dev.to_device(np.ascontiguousarray(data, dtype=np.complex128)))

    def SetupWavefunctionFile(self, config):
        """InitialCondition.type = InitialConditionType.File: load the
        wavefunction from InitialCondition.filename (format HDF: the dataset
        InitialCondition.dataset of a wavefunction archive; the grid must be
        the one it was saved on) (Problem.py:393-421)."""
        ic = config.InitialCondition
        fmt = getattr(ic, "format", WavefunctionFileFormat.HDF)
        if fmt != WavefunctionFileFormat.HDF:
            raise NotImplementedError("wavefunction file format %r (only HDF archives are read here)" % (fmt,))
        self.LoadWavefunctionHDF(ic.filename, str(getattr(ic, "dataset", "/wavefunction")))

    # ---------------------------------------------------------- (de)serialisation
    def LoadWavefunctionData(self, newdata):
        """Problem.py:425-429: copy an array of the wavefunction's shape in."""
        newdata = np.asarray(newdata)
        if tuple(newdata.shape) != tuple(self.psi.GetData().shape):
            raise Exception("Invalid shape on loaded wavefunction, got %s expected %s"
                            % (newdata.shape, tuple(self.psi.GetData().shape)))
        self.SetWavefunction(newdata)

    def LoadWavefunctionHDF(self, filename, datasetPath):
        from . import serialization
        serialization.LoadWavefunctionHDF(filename, datasetPath, self.psi)

    def SaveWavefunctionHDF(self, filename, datasetPath):
        """Store psi with this problem's config (Problem.py:441-442)."""
        from . import serialization
        serialization.SaveWavefunctionHDF(filename, datasetPath, self.psi, conf=self.Config)
