"""Batched propagation of many independent problems in packed persistent
launches, plus the chunk-length autotuner of the band solves.

No reference counterpart: pyprop runs one problem per process (MPI ranks
split one wavefunction).  Each problem keeps its own config, operator,
overlap and wavefunction exactly as a single Problem; ``Batch.AdvanceSteps``
hands all of them to ``pp_pamp_batch_propagate``, which packs their fibers
into cooperative launches and synchronises each problem's workgroups only
among themselves.
"""
import numpy as np

from . import _native as nat
from . import device as dev


def instance_draws(count=512, seed=20261203):
    """Seeded draws of the batched hydrogen scan: N ~ U{300..600} B-splines,
    lmax ~ U{20..60}, omega ~ U[0.05, 0.6], log10 E0 ~ U[log10 0.01, log10 0.2],
    cycles ~ U{4..12} (numpy PCG64)."""
    rng = np.random.default_rng(seed)
    N = rng.integers(300, 601, size=count)
    lmax = rng.integers(20, 61, size=count)
    omega = rng.uniform(0.05, 0.6, size=count)
    E0 = 10.0 ** rng.uniform(np.log10(0.01), np.log10(0.2), size=count)
    cycles = rng.integers(4, 13, size=count)
    return [dict(N=int(N[i]), lmax=int(lmax[i]), omega=float(omega[i]), E0=float(E0[i]), cycles=int(cycles[i]))
            for i in range(count)]


def autotune_chunks(shapes, kd, m=16, nslots=3, candidates=(16, 32, 64), fibers=64, reps=10):
    """Pick the band-solver chunk length per radial size N.  shapes: (N, L)
    of the problems.  A candidate must keep the step kernel's LDS within the
    GPU's 160 KB per CU; among the candidates that give the most workgroups
    per CU (at most two: the kernel's register use) the fastest overlap solve
    pair U^-1, U^-H on a (fibers, N) tensor wins (HIP events).  Records the
    winners in device.CHUNK_TABLE; returns {N: (chunk, {chunk: us}, wg/CU)}."""
    torch = dev.torch()
    from .bspline import BSplineBasis, LinearBreakpointSequence
    lib = nat.lib()
    Lmax = {}
    for N, L in shapes:
        Lmax[int(N)] = max(Lmax.get(int(N), 0), int(L))
    result = {}
    for N in sorted(Lmax):
        occ = {}
        for B in candidates:
            lds = lib.pp_pamp_lds_estimate(Lmax[N], N, m, nslots, kd, B)
            if lds > 0 and lds <= 160 * 1024:
                occ[B] = min(2, (160 * 1024) // lds)
        if not occ:
            continue
        best_occ = max(occ.values())
        k = kd + 1
        basis = BSplineBasis(LinearBreakpointSequence(0.0, 0.5 * N, N - k + 4), k, "zero")
        S = basis.overlap_upper()
        x = torch.ones((fibers, N), dtype=torch.complex128, device=dev.device())
        times = {}
        for B in sorted(b for b in occ if occ[b] == best_occ):
            try:
                ov = dev.Overlap(S, N, kd, chunk=B)
            except nat.NativeError:
                continue
            ov.apply(nat.OV_SOLVE_SQRT, 1, x)
            ov.apply(nat.OV_SOLVE_SQRT_CONJ, 1, x)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                ov.apply(nat.OV_SOLVE_SQRT, 1, x)
                ov.apply(nat.OV_SOLVE_SQRT_CONJ, 1, x)
            e1.record()
            torch.cuda.synchronize()
            times[B] = 1e3 * e0.elapsed_time(e1) / reps
        if not times:
            continue
        best = min(times, key=times.get)
        dev.CHUNK_TABLE[(N, kd)] = best
        result[N] = (best, times, best_occ)
    return result


class Batch:
    """A set of set-up Problems whose propagators are PamPropagators on the
    HIP engine; AdvanceSteps(n) advances every problem by n steps."""

    def __init__(self, problems):
        from .propagator import ExpokitPropagator, PamPropagator
        for p in problems:
            if not isinstance(p.Propagator, PamPropagator) or isinstance(p.Propagator, ExpokitPropagator):
                raise NotImplementedError("batched propagation runs PamPropagator problems")
            if p.Propagator.Engine is None:
                raise RuntimeError("call SetupStep() on every problem first")
            if p.Propagator.RenormalizeActive:
                raise NotImplementedError("renormalisation between steps is not batched")
        self.Problems = list(problems)
        self.Graph = None

    def CaptureGraph(self, n):
        """Record the packed launches of an n-step advance as one HIP graph;
        later AdvanceSteps(n) calls replay it (other step counts launch
        directly).  Changing an engine afterwards (terms, precision) makes the
        replay raise; record again then."""
        self.Graph = None
        self.Graph = dev.BatchGraph([p.Propagator.Engine for p in self.Problems],
                                    [p.psi.GetData() for p in self.Problems], n,
                                    [p.TimeStep for p in self.Problems])

    def AdvanceSteps(self, n):
        if n <= 0:
            return
        engines, psis, dts, scs = [], [], [], []
        for p in self.Problems:
            times = []
            for _ in range(n):
                times.append(p.PropagatedTime)
                p._advance_time()
            engines.append(p.Propagator.Engine)
            psis.append(p.psi.GetData())
            dts.append(p.TimeStep)
            scs.append(p.Propagator.TimeScalings(times))
        if self.Graph is not None and self.Graph.nsteps == n:
            self.Graph.launch(scs, psis)
        else:
            dev.pamp_batch_propagate(engines, psis, n, dts, scs)

    def GetDiagnostics(self):
        return [p.Propagator.GetDiagnostics() for p in self.Problems]
