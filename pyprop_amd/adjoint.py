"""Gradient of a final-state population with respect to the sampled time
factor of the interaction term, by a hand-written adjoint sweep on the HIP
engine, behind a torch.autograd.Function.

No reference counterpart (pyprop has no derivatives).  The discrete forward
map of one step is F_k = U^-1 exp(-i dt H^_k) U with H^_k = U^-H (H0 + s_k V)
U^-1 (one PamPropagator.AdvanceStep with the interaction scaled by the control
sample s_k).  F_k is unitary in the S inner product, so with
c = <phi|S|psi_K> and P = |c|^2

    dP/ds_k = 2 Re( conj(c) <lambda_{k+1}| S |dF_k/ds_k psi_k> ),
    lambda_K = phi,  lambda_k = F_k^-1 lambda_{k+1}  (a step with -dt).

dF_k/ds_k psi = U^-1 L(A, E) U psi is the Frechet derivative of the
exponential (A = -i dt H^_k, E = -i dt V^), read off the upper block of
exp([[A, E], [0, A]]) [0; U psi]: one step of the same engine on a doubled
problem (2L fibers: the operator on both halves plus V from the lower to the
upper half).  Forward states are recomputed from checkpoints.  All steps use
the 1-norm Pade scaling (accurate to rounding), so the forward map is smooth in
the samples and its finite differences match the gradient.
"""
import numpy as np

from . import _native as nat
from . import device as dev


class PopulationAdjoint:
    """Adjoint gradients for a set-up Problem whose PamPropagator runs a
    static Diag term and one time-dependent Simp dipole term (the hydrogen
    operator).  target: the state phi (L, N) whose population |<phi|S|psi_K>|^2
    is the objective.  derivative_krylov: Krylov dimension of the doubled
    (Frechet-derivative) steps, default the propagator's."""

    def __init__(self, problem, target, checkpoint=50, derivative_krylov=None):
        prop = problem.Propagator
        if prop.Engine is None:
            raise RuntimeError("call problem.SetupStep() first")
        pots = prop._pots
        if len(pots) != 2 or pots[0].IsTimeDependent or not pots[1].IsTimeDependent:
            raise NotImplementedError("adjoint expects one static and one time-dependent potential")
        if pots[0].StorageIds() != ["Diag", "BandNH"] or pots[1].StorageIds() != ["Simp", "BandNH"]:
            raise NotImplementedError("adjoint expects Diag x BandNH and Simp x BandNH terms")
        torch = dev.torch()
        self.problem = problem
        self.engine = prop.Engine
        self.overlap = self.engine.overlap
        self.L, self.N, self.m = self.engine.L, self.engine.N, self.engine.m
        self.dt = complex(problem.TimeStep)
        self.checkpoint = int(checkpoint)
        self.psi0 = problem.psi.GetData().clone()
        self.target = dev.to_device(np.asarray(target, dtype=np.complex128))
        # the doubled problem: fibers [0, L) upper half, [L, 2L) lower half
        H0 = pots[0].DeviceData()
        V = pots[1].DeviceData()
        pairs = np.asarray(pots[1].GeometryList[0].GetBasisPairs(), dtype=np.int32)
        L = self.L
        self.mder = int(derivative_krylov) if derivative_krylov else self.m
        aug = dev.PampEngine(2 * L, self.N, self.mder, self.overlap, reorth=prop.DoubleOrthogonalization)
        aug.add_term("Diag", torch.cat([H0, H0]).contiguous())
        aug.add_term("Simp", torch.cat([V, V]).contiguous(), pairs=np.concatenate([pairs, pairs + L]))
        aug.add_term("Simp", V, pairs=np.stack([pairs[:, 0], pairs[:, 1] + L], axis=1))
        self.aug = aug
        # pAMP's largest-element Pade scaling makes a step jump by the Pade
        # error whenever the scaling exponent changes; derivatives need the
        # accurate, smooth variant on every engine of the sweep
        self.engine.set_pade_scaling(1)
        aug.set_pade_scaling(1)

    # ---------------------------------------------------------------- helpers
    def _sdot(self, a, b):
        """<a|S|b> over the radial rank of (L, N) device tensors (overlap
        kernel, then the device zdotc)."""
        tmp = b.clone()
        self.overlap.apply(nat.OV_MUL, 1, tmp)
        out = np.zeros(2)
        nat.check(nat.lib().pp_inner_product(dev.ptr(a.contiguous()), dev.ptr(tmp), int(tmp.numel()),
                                             nat.dptr(out), dev.stream_ptr()), "pp_inner_product")
        return complex(out[0], out[1])

    def _steps(self, psi, s, k0, n, sign=1.0):
        sc = np.stack([np.ones(n), np.asarray(s[k0:k0 + n], dtype=np.float64)], axis=1)
        self.engine.propagate(psi, n, sign * self.dt, sc)

    # ---------------------------------------------------------------- API
    def forward(self, s):
        """Propagate psi0 through len(s) steps with interaction factors s;
        returns (P, c, checkpoints)."""
        torch = dev.torch()
        s = np.asarray(s, dtype=np.float64)
        K = len(s)
        psi = self.psi0.clone()
        cps = []
        for k0 in range(0, K, self.checkpoint):
            cps.append(psi.clone())
            self._steps(psi, s, k0, min(self.checkpoint, K - k0))
        torch.cuda.synchronize()
        c = self._sdot(self.target, psi)
        return abs(c) ** 2, c, cps

    def population(self, s):
        return self.forward(s)[0]

    def gradient(self, s):
        """(P, dP/ds_k for all k) by the adjoint sweep."""
        torch = dev.torch()
        s = np.asarray(s, dtype=np.float64)
        K = len(s)
        P, c, cps = self.forward(s)
        lam = self.target.clone()
        grad = np.zeros(K)
        L = self.L
        z = torch.zeros((2 * L, self.N), dtype=torch.complex128, device=self.psi0.device)
        for seg in range(len(cps) - 1, -1, -1):
            k0 = seg * self.checkpoint
            n = min(self.checkpoint, K - k0)
            states = [cps[seg].clone()]
            for j in range(n - 1):
                nxt = states[-1].clone()
                self._steps(nxt, s, k0 + j, 1)
                states.append(nxt)
            for j in range(n - 1, -1, -1):
                k = k0 + j
                z.zero_()
                z[L:] = states[j]
                sc = np.array([[1.0, s[k], 1.0]])
                self.aug.propagate(z, 1, self.dt, sc)
                grad[k] = 2.0 * (np.conj(c) * self._sdot(lam, z[:L].contiguous())).real
                self._steps(lam, s, k, 1, sign=-1.0)
        torch.cuda.synchronize()
        return P, grad


def central_differences(adjoint, s, indices, h=1e-6):
    """(P(s + h e_k) - P(s - h e_k)) / 2h for each k in indices: 2 forward
    runs per index, all of them in one batched set of packed launches
    (pp_pamp_batch_propagate) from the adjoint's start state, with the same
    operator, Pade scaling and step as the adjoint's forward map."""
    torch = dev.torch()
    s = np.asarray(s, dtype=np.float64)
    K = len(s)
    eng0, pots = adjoint.engine, adjoint.problem.Propagator._pots
    pairs = pots[1].GeometryList[0].GetBasisPairs()
    engines, psis, scs = [], [], []
    for k in indices:
        for sgn in (1.0, -1.0):
            e = dev.PampEngine(eng0.L, eng0.N, eng0.m, eng0.overlap)
            e.add_term("Diag", pots[0].DeviceData())
            e.add_term("Simp", pots[1].DeviceData(), pairs=pairs)
            e.set_pade_scaling(1)
            sv = s.copy()
            sv[k] += sgn * h
            engines.append(e)
            psis.append(adjoint.psi0.clone())
            scs.append(np.stack([np.ones(K), sv], axis=1))
    dev.pamp_batch_propagate(engines, psis, K, [adjoint.dt] * len(engines), scs)
    torch.cuda.synchronize()
    pops = [abs(adjoint._sdot(adjoint.target, p)) ** 2 for p in psis]
    return np.array([(pops[2 * i] - pops[2 * i + 1]) / (2 * h) for i in range(len(indices))])


def population_function(adjoint):
    """torch.autograd.Function: samples (float64 tensor, K) -> population;
    backward runs the adjoint sweep."""
    torch = dev.torch()

    class _Population(torch.autograd.Function):
        @staticmethod
        def forward(ctx, samples):
            s = samples.detach().cpu().numpy().astype(np.float64)
            ctx.s = s
            return samples.new_tensor(adjoint.population(s))

        @staticmethod
        def backward(ctx, gP):
            _, grad = adjoint.gradient(ctx.s)
            return gP * torch.from_numpy(grad).to(gP.device, gP.dtype)

    return _Population.apply
