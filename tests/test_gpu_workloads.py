"""The headline workloads at their own sizes, against the oracle or against
size-independent properties:

* the adjoint gradient on its full workload (N = 400, l <= 31, 4000 control
  samples, checkpoints every 50 steps);
* the large hydrogen workload (N = 800, l <= 63, m = 24) against the oracle;
* the batched scan of 512 seeded instances (N 300-600, lmax 20-60, m = 16)
  through its HIP graph, sampled instances against the oracle;
* the helium workload (66 channels, N = 120) against the oracle's factored
  propagation.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O

try:
    import torch  # noqa: F401  (first import can take a minute on a fresh box)
except ImportError:  # pragma: no cover
    torch = None

pytestmark = pytest.mark.gpu


def _population_noise_floor(adj, s, count=3):
    """Spread of the final population over `count` runs from psi0 (1 + 1e-16
    eta): the rounding noise of the forward map itself, through the same
    batched launches central_differences uses."""
    from pyprop_amd import device as D
    K = len(s)
    eng0, pots = adj.engine, adj.problem.Propagator._pots
    pairs = pots[1].GeometryList[0].GetBasisPairs()
    rng = np.random.default_rng(3)
    engines, psis = [], []
    x0 = adj.psi0.cpu().numpy()
    for j in range(count + 1):
        e = D.PampEngine(eng0.L, eng0.N, eng0.m, eng0.overlap)
        e.add_term("Diag", pots[0].DeviceData())
        e.add_term("Simp", pots[1].DeviceData(), pairs=pairs)
        e.set_pade_scaling(1)
        engines.append(e)
        psis.append(D.to_device(x0 if j == 0 else x0 * (1 + 1e-16 * rng.standard_normal(x0.shape))))
    D.pamp_batch_propagate(engines, psis, K, [adj.dt] * len(engines), [np.stack([np.ones(K), s], axis=1)] * len(engines))
    pops = np.array([abs(adj._sdot(adj.target, p)) ** 2 for p in psis])
    return float(np.abs(pops[1:] - pops[0]).max())


def test_adjoint_gradient_on_the_full_workload():
    """dP(2p)/ds_k for all 4000 samples by the adjoint sweep against
    (a) central differences of the GPU forward map on 8 seeded samples over
    the whole sequence: with h = 1e-4 to 1e-6 max|g| (measured 1.8e-7);
    with h = 1e-6 to 1e-6 max|g| plus the differences' own noise,
    3 (2 dP_noise) / 2h, where dP_noise is the population's spread over
    1e-16 perturbations of psi0 (measured 2.6e-15 after 4000 steps: the
    noise term is 1.4e-5 max|g|, the differences 1.3e-5); and (b) central differences h = 1e-6 of the oracle's FP64
    propagation of the same operator on 2 samples of a 120-sample prefix
    (the oracle takes about 50 ms a step here), to 1e-6 max|g| of the prefix
    gradient."""
    import pyprop_amd
    from pyprop_amd.adjoint import PopulationAdjoint, central_differences
    from pyprop_amd.examples import hydrogen as H
    print("adjoint, full workload: setup", flush=True)
    conf = H.load("hydrogen_adjoint")
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    prob.SetWavefunction(H.ground_state(prob))
    target, _ = H.bound_state(prob, 1, 0)
    dt = prob.TimeStep.real
    K = 4000
    s = H.adjoint_control_samples(K, dt)
    adj = PopulationAdjoint(prob, target, checkpoint=50)
    t0 = time.time()
    P, g = adj.gradient(s)
    gmax = np.abs(g).max()
    print("adjoint: P2p %.6e, max|g| %.3e, %.1f s" % (P, gmax, time.time() - t0), flush=True)
    adj.population(s)                    # last launch: the final 50 forward steps
    est = adj.engine.diagnostics()
    print("Krylov error estimate of the last 50 forward steps: median %.2e, max %.2e"
          % (np.median(est[:, 0]), est[:, 0].max()), flush=True)
    assert est[:, 0].max() < 1e-9
    idx = np.sort(np.random.default_rng(7).choice(K, size=8, replace=False))
    noise = _population_noise_floor(adj, s)
    # a central difference subtracts two runs, each off by up to the floor,
    # and perturbing s_k changes the rounding of every later step: 3 floors
    for h, tol in [(1e-4, 1e-6 * gmax), (1e-6, 1e-6 * gmax + 3 * 2 * noise / (2 * 1e-6))]:
        t0 = time.time()
        fd = central_differences(adj, s, idx, h)
        err = np.abs(g[idx] - fd)
        print("GPU central differences h = %g at %s: max |g - fd| / max|g| = %.2e, tolerance %.2e "
              "(population noise floor %.2e; %.1f s)" % (h, list(idx), err.max() / gmax, tol / gmax, noise,
                                                          time.time() - t0), flush=True)
        assert err.max() <= tol
    # (b) the oracle's FP64 map, prefix of 120 samples
    K0 = 120
    P0, g0 = adj.gradient(s[:K0])
    h_or = O.Hydrogen(rmax=800.0, nbp=397, order=7, lmax=31, frequency=0.3, cycles=10, E0=0.05)
    ham = h_or.ham()
    psi0 = adj.psi0.cpu().numpy()
    S = h_or.overlap_dense()
    m = prob.Propagator.BasisSize

    def pop(sv):
        psi = ham.propagate_pamp(psi0, K0, dt, np.stack([np.ones(K0), sv], axis=1), m)
        return abs(np.sum(np.conj(target) * (psi @ S.T))) ** 2

    assert abs(pop(s[:K0]) - P0) <= 1e-8 * P0    # Pade scaling differs (1-norm on the GPU)
    for k in (17, 88):
        sp, sm = s[:K0].copy(), s[:K0].copy()
        sp[k] += 1e-6
        sm[k] -= 1e-6
        fdo = (pop(sp) - pop(sm)) / 2e-6
        print("k", k, "adjoint", g0[k], "oracle central difference", fdo, flush=True)
        assert abs(g0[k] - fdo) <= 1e-6 * np.abs(g0).max()


def _sensitivity_tol(run, psi0, floor=1e-12):
    """The oracle's own sensitivity: 10x the change a 1e-16 relative
    perturbation of psi0 causes (fixed-m Krylov steps on a stiff operator
    amplify rounding), at least `floor` of max|psi|."""
    ref = run(psi0)
    rng = np.random.default_rng(11)
    ref2 = run(psi0 * (1 + 1e-16 * rng.standard_normal(psi0.shape)))
    return ref, max(floor * np.abs(ref).max(), 10 * np.abs(ref2 - ref).max())


def test_hydrogen_large_matches_oracle():
    """The large hydrogen workload through the config API (order 7, N = 800,
    l <= 63, m = 24, dt = 0.005; the fiber-split step kernel) for 6 steps
    from the ground state at t = 50 (field on), against the oracle's pAMP:
    psi within the oracle's sensitivity, norm and bound populations to
    1e-10."""
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H
    print("hydrogen_large: 6 steps from t = 50 vs oracle", flush=True)
    conf = H.load("hydrogen_large")
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    prob.SetWavefunction(H.ground_state(prob))
    prob.psi.Normalize()
    psi0 = prob.psi.GetData().cpu().numpy()
    t0, n = 50.0, 6
    dt = prob.TimeStep.real
    prob.PropagatedTime = t0
    prob.AdvanceSteps(n)
    got = prob.psi.GetData().cpu().numpy()
    h = O.Hydrogen(rmax=400.0, nbp=797, order=7, lmax=63, frequency=0.5, cycles=10, E0=0.1)
    ham = h.ham()
    sc = np.array([[1.0, h.scaling(t)] for t in O.step_times(n, dt, t0)])
    ref, tol = _sensitivity_tol(lambda x: ham.propagate_pamp(x, n, dt, sc, 24), psi0)
    err = np.abs(got - ref).max()
    print("  max |psi - psi_ref| = %.3e (tol %.3e)" % (err, tol), flush=True)
    assert err <= tol
    assert abs(prob.psi.GetNorm() - h.norm(ref)) <= 1e-10
    tot, per = H.populations(prob, got)
    tref, pref = h.populations(ref)
    pmax = max(np.abs(a - b).max() for a, b in zip(per, pref) if len(a))
    print("  bound population %.15f (oracle %.15f), largest single-state difference %.2e"
          % (tot, tref, pmax), flush=True)
    assert abs(tot - tref) <= 1e-10
    assert pmax <= 1e-10


def test_batch_scan_of_512_instances():
    """The batched scan at its size: 512 seeded instances (N 300-600, lmax
    20-60, pulses drawn per instance, m = 16, dt = 0.01), autotuned chunk
    lengths, 20 steps from the ground state in the middle of every pulse
    through one HIP graph replay.  FP64 form: every instance keeps its norm to 1e-9 and uses the
    full basis; 4 instances spanning N and lmax match the oracle within its
    sensitivity.  Mixed form (the graph recorded again): bound-population
    error against the FP64 batch on 8 instances, within 1e-8 or 10x the
    FP64 form's own floor (its largest change under three independent 1e-16
    perturbations of psi0:
    at m = 16, dt = 0.01 and l up to 60, |H^| dt reaches ~6000, and the
    fixed-m steps amplify rounding to ~1e-4 in 20 steps)."""
    import pyprop_amd
    from pyprop_amd import batch as Bt
    from pyprop_amd.examples import hydrogen as H
    K, count = 20, 512
    draws = Bt.instance_draws(count)
    t0 = time.time()
    tuned = Bt.autotune_chunks([(d["N"], d["lmax"] + 1) for d in draws], kd=6)
    print("batch512: autotuned %d radial sizes in %.1f s" % (len(tuned), time.time() - t0), flush=True)
    probs, psi0s = [], []
    t0 = time.time()
    for i, d in enumerate(draws):
        prob = pyprop_amd.Problem(H.batch_instance_config(d, precision="fp64"))
        prob.SetupStep()
        prob.SetWavefunction(H.ground_state(prob))
        prob.psi.Normalize()
        prob.PropagatedTime = prob.StartTime = 0.5 * 2 * np.pi / d["omega"] * d["cycles"]
        probs.append(prob)
        psi0s.append(prob.psi.GetData().cpu().numpy())
        if (i + 1) % 64 == 0:
            print("  set up %d/%d instances (%.0f s)" % (i + 1, count, time.time() - t0), flush=True)
    batch = Bt.Batch(probs)
    batch.CaptureGraph(K)
    t0 = time.time()
    batch.AdvanceSteps(K)
    torch.cuda.synchronize()
    print("  FP64 graph replay of %d steps: %.2f s" % (K, time.time() - t0), flush=True)
    got = [p.psi.GetData().cpu().numpy() for p in probs]
    for i, p in enumerate(probs):
        d = p.Propagator.GetDiagnostics()
        assert np.all(np.isfinite(d)) and np.all(d[:, 1] == 16), i
        assert abs(p.psi.GetNorm() - 1.0) <= 1e-9, (i, p.psi.GetNorm())
    Ns = np.array([d["N"] for d in draws])
    Ls = np.array([d["lmax"] for d in draws])
    pick = sorted({int(np.argmin(Ns)), int(np.argmax(Ns)), int(np.argmin(Ls)), int(np.argmax(Ls))})
    for i in pick:
        d = draws[i]
        h = O.Hydrogen(rmax=0.5 * d["N"], nbp=d["N"] - 3, order=7, lmax=d["lmax"], frequency=d["omega"],
                       cycles=d["cycles"], E0=d["E0"])
        ham = h.ham()
        sc = np.array([[1.0, h.scaling(t)] for t in O.step_times(K, 0.01, probs[i].StartTime)])
        ref, tol = _sensitivity_tol(lambda x: ham.propagate_pamp(x, K, 0.01, sc, 16), psi0s[i])
        err = np.abs(got[i] - ref).max()
        print("  instance %d (N %d, lmax %d): max |psi - psi_ref| %.3e (tol %.3e)" % (i, d["N"], d["lmax"], err, tol),
              flush=True)
        assert err <= tol
    # the FP64 form's own floor: the same batch from psi0 (1 + 1e-16 eta), three
    # independent perturbations (one draw is a noisy estimate of the sensitivity)
    perts = []
    for seed in (13, 14, 15):
        rng = np.random.default_rng(seed)
        for p, x in zip(probs, psi0s):
            p.SetWavefunction(x * (1 + 1e-16 * rng.standard_normal(x.shape)))
            p.PropagatedTime = p.StartTime
        batch.AdvanceSteps(K)
        perts.append([p.psi.GetData().cpu().numpy() for p in probs])
    # mixed form: same instances, the graph recorded again
    for p, x in zip(probs, psi0s):
        p.SetWavefunction(x)
        p.PropagatedTime = p.StartTime
        p.Propagator.Engine.set_mixed([j for j, pot in enumerate(p.Propagator._pots) if pot.IsTimeDependent])
    batch.CaptureGraph(K)
    batch.AdvanceSteps(K)
    for i in range(0, count, count // 8):
        proj = H.bound_projectors(probs[i])
        a = H.populations(probs[i], got[i], projectors=proj)[1]
        b = H.populations(probs[i], probs[i].psi.GetData().cpu().numpy(), projectors=proj)[1]
        mixed = max(float(np.abs(x - y).max()) for x, y in zip(a, b) if len(x))
        floor = 0.0
        for pert in perts:
            c = H.populations(probs[i], pert[i], projectors=proj)[1]
            floor = max(floor, max(float(np.abs(x - y).max()) for x, y in zip(a, c) if len(x)))
        print("  instance %d (N %d, lmax %d): mixed vs FP64 bound-population error %.2e, FP64 floor %.2e"
              % (i, draws[i]["N"], draws[i]["lmax"], mixed, floor), flush=True)
        assert mixed <= max(1e-8, 10 * floor)


def test_helium_workload_matches_oracle():
    """The helium workload (66 channels, N = 120, 950,400 amplitudes,
    multipoles to l = 10, m = 20, dt = 0.01): 5 laser steps from the
    imaginary-time ground state at t = 10 against the oracle's factored
    propagation of the same terms (the reference layout of the e-e term
    would be ~48 GB): ground-state population and norm to 1e-10, psi within
    the oracle's sensitivity."""
    import pyprop_amd
    from pyprop_amd.examples import helium as HE
    from pyprop_amd.helium import KronTerm
    print("helium workload: ground state, then 5 laser steps vs the oracle", flush=True)
    psi0, E0 = HE.ground_state("helium", steps=150, dt=0.05)
    prob = pyprop_amd.Problem(HE.load("helium"))
    prob.SetupStep()
    prob.SetWavefunction(psi0)
    prob.psi.Normalize()
    ground = prob.psi.Copy()
    psi0 = prob.psi.GetData().cpu().numpy()
    C, N, _ = psi0.shape
    t0, n = 10.0, 5
    dt = prob.TimeStep.real
    prob.PropagatedTime = t0
    prob.AdvanceSteps(n)
    got = prob.psi.GetData().cpu().numpy()
    b = prob.psi.GetRepresentation().GetRepresentation(1).GetBSplineObject()
    Ub, info = O.zpbtrf_upper(b.overlap_upper(), N, 6)
    assert info == 0
    ham = O.Ham((C, N, N), [], {1: (Ub, 6), 2: (Ub, 6)})
    dep = []
    for pot in prob.Propagator.BasePropagator.PotentialList:
        for t in pot.Terms:
            if isinstance(t, KronTerm):
                ham.add_factored(t.pairs, t.coef, A=t.A, B=t.B)
            else:
                ham.add_factored(t.pairs, t.coef, R=t.R)
            dep.append(pot)
    sc = np.array([[p.GetTimeScaling(tt) for p in dep] for tt in O.step_times(n, dt, t0)])
    tc = time.time()
    ref, tol = _sensitivity_tol(lambda x: ham.propagate_pamp(x, n, dt, sc, 20), psi0)
    print("  oracle: 2 x %d steps in %.1f s" % (n, time.time() - tc), flush=True)
    err = np.abs(got - ref).max()
    print("  max |psi - psi_ref| = %.3e (tol %.3e)" % (err, tol), flush=True)
    assert err <= tol
    g0 = ground.GetData().cpu().numpy()
    pop_ref = abs(ham.inner(g0, ref)) ** 2
    pop = abs(ground.InnerProduct(prob.psi)) ** 2
    print("  ground-state population %.15f (oracle %.15f)" % (pop, pop_ref), flush=True)
    assert abs(pop - pop_ref) <= 1e-10
    assert abs(prob.psi.GetNorm() - np.sqrt(ham.inner(ref, ref).real)) <= 1e-10


def test_helium_workload_cayley_preconditioned():
    """The helium workload (66 channels, N = 120) under CayleyPropagator with
    the config's RadialPreconditioner (GMRES basis 20, tolerance
    eps^(2/3)): 2 laser steps from the 1s^2 product state at t = 10 against
    the oracle's cayley_step on the factored terms, whose preconditioner
    solves each channel's b S X S + a (H_l1 X S + S X H_l2) = Y as a
    Sylvester equation by Bartels-Stewart (complex Schur forms and LAPACK
    ztrsyl; H_l = -D2/2 + l(l+1)/(2r^2) - 2/r from the oracle's own B-spline
    projections).  GMRES converges in <= 12 operator calls, the same count
    as the oracle's, psi within 1e-9 of max|psi|."""
    import scipy.linalg
    from scipy.linalg import lapack
    import pyprop_amd
    from pyprop_amd.examples import helium as HE
    from pyprop_amd.helium import KronTerm
    print("helium workload: Cayley + RadialPreconditioner, 2 steps vs the oracle", flush=True)
    conf = HE.load("helium")
    conf.Propagation.propagator = pyprop_amd.CayleyPropagator
    conf.Propagation.preconditioner = "RadialPreconditioner"
    prob = pyprop_amd.Problem(conf)
    tc = time.time()
    prob.SetupStep()
    print("  setup %.1f s" % (time.time() - tc), flush=True)
    prob.SetWavefunction(HE.start_state(prob))
    prob.psi.Normalize()
    C, N, _ = prob.psi.GetData().shape
    t0 = 10.0
    dt = prob.TimeStep.real
    prob.PropagatedTime = t0
    bs = O.BSpline(O.linear_breakpoints(0.0, 60.0, 117), 7)
    assert bs.N == N
    s, bw = 6, 13

    def dense(f, d):
        nh = bs.project_bandnh(np.broadcast_to(f, bs.x.shape), d)
        M = np.zeros((N, N))
        for j in range(N):
            for i in range(max(0, j - s), min(N, j + s + 1)):
                M[i, j] = nh[j * bw + s + i - j].real
        return M
    x = bs.x
    S = dense(np.ones_like(x), 0)
    Hl = [-0.5 * dense(np.ones_like(x), 2) + l * (l + 1) / 2.0 * dense(1 / x ** 2, 0) - 2.0 * dense(1 / x, 0)
          for l in range(6)]
    Si = np.linalg.inv(S)
    aH = 0.5j * dt
    left = [scipy.linalg.schur(Si @ (0.5 * S + aH * H), output="complex") for H in Hl]
    right = [scipy.linalg.schur((0.5 * S + aH * H) @ Si, output="complex") for H in Hl]
    ch = prob.psi.GetRepresentation().GetRepresentation(0).Channels

    def msolve(y):
        out = np.empty_like(y)
        for c in range(C):
            (T, Z), (U, W) = left[ch[c].l1], right[ch[c].l2]
            q = Z.conj().T @ (Si @ y[c] @ Si) @ W
            z, scale, info = lapack.ztrsyl(T, U, q)
            assert info == 0 and scale == 1.0
            out[c] = Z @ z @ W.conj().T
        return out
    b = prob.psi.GetRepresentation().GetRepresentation(1).GetBSplineObject()
    Ub, info = O.zpbtrf_upper(b.overlap_upper(), N, 6)
    ham = O.Ham((C, N, N), [], {1: (Ub, 6), 2: (Ub, 6)})
    dep = []
    for pot in prob.Propagator.BasePropagator.PotentialList:
        for t in pot.Terms:
            if isinstance(t, KronTerm):
                ham.add_factored(t.pairs, t.coef, A=t.A, B=t.B)
            else:
                ham.add_factored(t.pairs, t.coef, R=t.R)
            dep.append(pot)

    def apply_s(v):
        return np.matmul(np.matmul(S, v), S.T)
    solver = prob.Propagator.Solver
    t = t0
    for k in range(2):
        prev = prob.psi.GetData().cpu().numpy()
        ham.set_scaling(np.array([p.GetTimeScaling(t + dt / 2) for p in dep]))
        tc = time.time()
        ref, err_ref, nops_ref = O.cayley_step(ham.potentials, apply_s, msolve, prev, dt, 20, 0.0)
        tref = time.time() - tc
        before = solver.GetOperatorCount()
        prob.AdvanceStep()
        t += dt
        nops = solver.GetOperatorCount() - before
        got = prob.psi.GetData().cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("  step %d: rel err %.3e, error estimate %.3e (oracle %.3e), operator calls %d (oracle %d), "
              "oracle %.1f s" % (k, err, solver.GetErrorEstimate(), err_ref, nops, nops_ref, tref), flush=True)
        assert nops == nops_ref and nops <= 12
        assert err < 1e-9
