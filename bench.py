"""Benchmark: pAMP Krylov steps/s of the large hydrogen workload on one MI355X,
with the helium workload beside it.

Default workload (config hydrogen_large): order-7 B-splines, N=800 on a
linear grid r in [0,400], l = 0..63 (n = 51,200 complex128 amplitudes),
velocity-gauge dipole coupling, Krylov dimension 24, dt = 0.005, FP64.  A
step is one PamPropagator.AdvanceStep (psi~ = U psi, 24 balanced-overlap
matvecs with classical Gram-Schmidt, Pade exponential, psi = U^-1 psi~);
inputs are resident on the GPU before timing starts.  Every run starts from
the 1s state at t = 50 a.u., inside the pulse.

The JSON line carries:
* roofline: algorithmic HBM bytes per step of the path's byte model over the
  step kernel's duration (HIP events on the launch stream), against 8 TB/s;
* parity: psi after the warm-up and timed steps against the oracle's pAMP
  from the same start (S-norm of the difference, bound populations, norm),
  next to the oracle's own change under a 1e-16 perturbation of psi0;
* mixed: the matvec_precision = "mixed" form (every term's radial bands in
  FP32 with FP32 band sums; Gram-Schmidt, overlap solves and Pade FP64): its
  speed over the timed steps, and its error against FP64 over the workload's
  whole window (20000 steps from t = 0, checkpoints every 2000) next to the
  FP64 run's own change under a 1e-16 perturbation of psi0 (the floor no
  precision change can go below);
* cpu_baseline: the C oracle (-O3 build for the host's instruction set) on
  the GPU box's host, 10 warm-up steps, median of 3 timed samples, on one
  thread and on all;
* helium: the two-electron workload (66 channels x 120 x 120) -- steps/s,
  the multipole stage's roofline, the mixed form, and an all-core CPU
  baseline of the oracle's factored restatement.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector peaks are equal)
METRIC = "Krylov steps/s + matvec HBM GB/s, FP64 vs mixed at <=1e-8 population error"
T_START = 50.0            # start time of the hydrogen runs (the pulse lasts 125.7 a.u.)

# HBM-side bytes per step of the pAMP step kernel (its fiber-split form, 4
# workgroups per fiber) on each workload, read by the FETCH_SIZE and
# WRITE_SIZE counter passes recorded in
# profiles/pp_pamp_kernel.hydrogen_large.pmc.md (FETCH_SIZE scaled x2 and
# WRITE_SIZE x1 by a 512 MiB device copy of known size).
PMC_TRAFFIC_PER_STEP = {"hydrogen_large": {"fetch": 53.3e6, "write": 77.9e6}}

# HBM-side bytes of one multipole stage (the three kernels' FETCH_SIZE +
# WRITE_SIZE per matvec, FETCH scaled x2 and WRITE x1 by a 512 MiB copy),
# from the counter passes recorded in profiles/multipole.helium.md.
HE_PMC_TRAFFIC = {"helium": 1303.0e6, "helium_mixed": 662.1e6}


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def step_bytes(n, m, N, s_mv=16, s_gs=16):
    """Algorithmic HBM bytes of one Krylov step:
    n [2 m s_mv + (m^2 + 5m + 1) s_gs] + m B_op,
    B_op = 8 N (13*4 + 7): four real radial bands + the Cholesky factor."""
    B_op = 8 * N * (13 * 4 + 7)
    return n * (2 * m * s_mv + (m * m + 5 * m + 1) * s_gs) + m * B_op


def oracle_for_baseline():
    """The oracle module with its fastest build for this host selected."""
    from oracle import oracle as O
    O.use_build(O.fastest_build())
    return O


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_rate(O, run, nsteps, threads, warm=10, reps=3):
    """Steps/s of run(n) (n oracle steps) on `threads` OpenMP threads: `warm`
    untimed steps, then the median of `reps` timed runs of nsteps."""
    import numpy as np
    prev = O.set_threads(threads)
    try:
        if warm:
            run(warm)
        ts = []
        for _ in range(reps):
            c0 = time.perf_counter()
            run(nsteps)
            ts.append(time.perf_counter() - c0)
        return nsteps / float(np.median(ts))
    finally:
        O.set_threads(prev)


def timed_advance(prob, nsteps):
    """prob.AdvanceSteps(nsteps) bracketed by synchronisations; returns (wall
    seconds, seconds between HIP events recorded on the launch stream)."""
    import torch
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    prob.AdvanceSteps(nsteps)
    ev1.record(stream)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, ev0.elapsed_time(ev1) / 1e3


def window_study(set_form, start, t0, total, chunk, measure):
    """The mixed form's error over a whole window: FP64, FP64 from a start
    perturbed by 1e-16 relative noise (the FP64 run's own floor) and mixed,
    each `total` steps from t0 in launches of `chunk` steps;
    measure(a, b) -> {name: value} compares two states at each checkpoint.
    set_form(mixed) switches the engine and returns the problem."""
    import numpy as np
    rng = np.random.default_rng(20261204)
    forms = [("fp64", False, start), ("fp64_perturbed", False, start * (1 + 1e-16 * rng.standard_normal(start.shape))),
             ("mixed", True, start)]
    snaps = {}
    for name, mixed, psi0 in forms:
        prob = set_form(mixed)
        prob.SetWavefunction(psi0)
        prob.PropagatedTime = t0
        out = []
        done = 0
        while done < total:
            n = min(chunk, total - done)
            prob.AdvanceSteps(n)
            done += n
            out.append(prob.psi.GetData().cpu().numpy())
        snaps[name] = out
        log("  %s: %d steps from t = %g" % (name, total, t0))
    set_form(False)
    rows = []
    for k in range(len(snaps["fp64"])):
        mx = measure(snaps["mixed"][k], snaps["fp64"][k])
        fl = measure(snaps["fp64_perturbed"][k], snaps["fp64"][k])
        rows.append({"step": min((k + 1) * chunk, total), "mixed": mx, "fp64_floor": fl})
    keys = list(rows[-1]["mixed"].keys())
    return {"steps": total, "start_time": t0, "checkpoint_every": chunk,
            "final": rows[-1], "max_over_checkpoints": {
                "mixed": {k: max(r["mixed"][k] for r in rows) for k in keys},
                "fp64_floor": {k: max(r["fp64_floor"][k] for r in rows) for k in keys}},
            "checkpoints": rows}


def run_hydrogen(args):
    """The default line: hydrogen_large FP64 (timed), parity against the
    oracle, the mixed form against the FP64 floor, the CPU baseline."""
    import numpy as np
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H

    t_setup = time.perf_counter()
    conf = H.load(args.workload)
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    psi_g = H.ground_state(prob)
    eng = prob.Propagator.Engine
    nterms = len(prob.Propagator._pots)
    L, N = prob.psi.GetRepresentation().GetFullShape()
    n = L * N
    m = prob.Propagator.BasisSize
    dt = prob.TimeStep
    W, K = args.warmup, args.steps
    log("hydrogen set-up %.1f s (L = %d, N = %d, m = %d)" % (time.perf_counter() - t_setup, L, N, m))

    def run(mask, psi0, timed):
        eng.set_mixed(mask)
        prob.SetWavefunction(psi0)
        prob.PropagatedTime = T_START
        if not timed:
            prob.AdvanceSteps(W + K)
            return prob.psi.GetData().cpu().numpy(), None, None
        prob.AdvanceSteps(W)
        wall, kern = timed_advance(prob, K)
        return prob.psi.GetData().cpu().numpy(), wall, kern

    psi64, wall, kernel_s = run([], psi_g, True)
    diag = prob.Propagator.GetDiagnostics()
    norm = prob.psi.GetNorm()
    steps_per_s = K / wall
    log("FP64: %d steps in %.3f s (%.1f steps/s)" % (K, wall, steps_per_s))

    proj = H.bound_projectors(prob)
    S = prob.psi.GetRepresentation().GetRepresentation(1).GetBSplineObject().overlap_dense()

    def snorm(d):
        return float(np.sqrt(sum(np.real(np.conj(d[l]) @ S @ d[l]) for l in range(d.shape[0]))))

    def pop_err(a, b):
        pa = H.populations(prob, a, projectors=proj)[1]
        pb = H.populations(prob, b, projectors=proj)[1]
        return max(float(np.abs(x - y).max()) for x, y in zip(pa, pb) if len(x))

    mixed = None
    if not args.no_mixed:
        _, wall32, kern32 = run(list(range(nterms)), psi_g, True)
        eng.set_mixed([])
        full = int(round(conf.Propagation.duration / dt.real))
        log("mixed form: %.1f steps/s; error over the workload's %d steps from t = 0" % (K / wall32, full))

        def set_form(m):
            eng.set_mixed(list(range(nterms)) if m else [])
            return prob
        window = window_study(set_form, psi_g, 0.0, full, 2000,
                              lambda a, b: {"population_error_max": pop_err(a, b), "psi_diff_snorm": snorm(a - b)})
        fin = window["final"]
        mixed = {"value": K / wall32, "unit": "steps/s", "ms_per_step": 1e3 * wall32 / K,
                 "speedup_vs_fp64": (K / wall32) / steps_per_s,
                 "population_error_max": fin["mixed"]["population_error_max"],
                 "psi_diff_snorm": fin["mixed"]["psi_diff_snorm"], "fp64_floor": fin["fp64_floor"],
                 "window": window,
                 "dtype": "c64 operator apply (every term: FP32 radial bands, FP32 band sums), "
                          "c128 exchange / Gram-Schmidt / overlap solves / Pade"}
        log("mixed: %.1f steps/s, population error %.2e after %d steps (FP64 floor %.2e)"
            % (mixed["value"], mixed["population_error_max"], full, fin["fp64_floor"]["population_error_max"]))

    O = None
    parity = None
    cpu = None
    rng = np.random.default_rng(20261204)
    if not args.no_cpu:
        O = oracle_for_baseline()
        h = O.Hydrogen(rmax=conf.RadialRepresentation.xmax, nbp=conf.RadialRepresentation.xsize, order=7,
                       lmax=conf.AngularRepresentation.maxl, frequency=conf.PulseParameters.frequency,
                       cycles=conf.PulseParameters.cycles, E0=conf.PulseParameters.field_strength)
        ham = h.ham()
        threads = host_threads()
        npar = W + K
        if npar > 200:     # long runs: parity on a separate 100-step GPU run
            npar = 100
            eng.set_mixed([])
            prob.SetWavefunction(psi_g)
            prob.PropagatedTime = T_START
            prob.AdvanceSteps(npar)
            got = prob.psi.GetData().cpu().numpy()
        else:
            got = psi64
        sc = np.array([[1.0, h.scaling(t)] for t in O.step_times(npar, dt.real, T_START)])
        O.set_threads(threads)
        ref = ham.propagate_pamp(psi_g, npar, dt.real, sc, m)
        ref2 = ham.propagate_pamp(psi_g * (1 + 1e-16 * rng.standard_normal(psi_g.shape)), npar, dt.real, sc, m)
        tot, _ = H.populations(prob, got, projectors=proj)
        tref, _ = H.populations(prob, ref, projectors=proj)
        parity = {"steps": npar, "start_time": T_START, "psi_diff_snorm": snorm(got - ref),
                  "population_error_max": pop_err(got, ref), "bound_population": float(tot),
                  "bound_population_oracle": float(tref), "norm": h.norm(got), "norm_oracle": h.norm(ref),
                  "oracle_sensitivity": {"psi_diff_snorm": snorm(ref2 - ref),
                                         "population_error_max": pop_err(ref2, ref)}}
        log("parity vs oracle after %d steps: S-norm %.2e (oracle sensitivity %.2e)"
            % (npar, parity["psi_diff_snorm"], parity["oracle_sensitivity"]["psi_diff_snorm"]))
        scb = np.array([[1.0, h.scaling(t)] for t in O.step_times(400, dt.real, T_START)])

        def orc_run(k):
            ham.propagate_pamp(psi_g, k, dt.real, scb[:k], m)

        one = cpu_rate(O, orc_run, 30, 1)
        allc = cpu_rate(O, orc_run, 100, threads) if threads > 1 else one
        cpu = {"value": allc, "unit": "steps/s", "cores": threads, "kind": "port", "single_core_value": one,
               "sample": "pAMP steps of the same workload from the same state (C oracle, complex128, -O3 %s build); "
                         "10 warm-up steps, median of 3 samples of 100 steps on %d OpenMP threads (angular rows, "
                         "fibers, whole dot products, elements; bitwise equal to 1 thread) and of 30 steps on one"
                         % (O.fastest_build(), threads)}
        log("CPU baseline: %.2f steps/s on %d threads, %.2f on one" % (allc, threads, one))

    Bstep = step_bytes(n, m, N)
    q = min(8, __import__("torch").cuda.get_device_properties(0).multi_processor_count // L, ((N + 31) // 32) // 2)
    kernel_name = ("pp_pamp_split_kernel (%d workgroups per fiber)" % q) if q >= 2 else "pp_pamp_kernel"
    pmc = PMC_TRAFFIC_PER_STEP.get(args.workload)
    achieved = Bstep * K / kernel_s / 1e9
    return {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (deterministic hydrogen operator, 1s start state by inverse iteration)",
        "config": {"workload": args.workload, "n": n, "L": L, "N": N, "order": 7, "krylov_dim": m,
                   "dt": dt.real, "start_time": T_START, "form": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if pmc is None else pmc["fetch"] + pmc["write"],
                     "traffic_unit": "bytes per step (counter passes)",
                     "kernel": kernel_name, "algorithmic_bytes_per_step": Bstep,
                     "kernel_ms_per_step": 1e3 * kernel_s / K,
                     "launch": "one launch of %d steps" % K},
        "cpu_baseline": cpu,
        "parity": parity,
        "checks": {"norm_after": norm, "krylov_error_estimate_last": float(diag[-1, 0]),
                   "basis_used_last": int(diag[-1, 1])},
        "mixed": mixed,
    }


def run_helium(args, steps, warmup, name="helium"):
    """Two-electron workload: 66 coupled channels x 120 x 120 B-splines
    (950,400 amplitudes), e-e multipoles to l = 10, velocity-gauge pulse,
    Krylov dimension 20, dt = 0.01, FP64, from the imaginary-time ground
    state at t = 10.  The roofline is the multipole stage's (its three
    kernels bracketed by HIP events on the launch stream), against HBM
    bandwidth; the mixed form runs that stage in FP32."""
    import numpy as np
    import pyprop_amd
    from pyprop_amd.examples import helium as HE
    from pyprop_amd.helium import KronTerm, MultipoleTerm, radial_supports

    t0 = time.perf_counter()
    psi_g, E0 = HE.ground_state(name, steps=args.imag_steps, dt=0.05)
    gs_s = time.perf_counter() - t0
    log("helium ground state E0 = %.10f after %d imaginary-time steps (%.1f s)" % (E0, args.imag_steps, gs_s))
    t_start = 10.0
    rng = np.random.default_rng(20261205)

    def run_form(mixed, psi0, timed=True):
        over = {("Propagation", "matvec_precision"): "mixed"} if mixed else None
        prob = pyprop_amd.Problem(HE.load(name, over))
        prob.SetupStep()
        prob.SetWavefunction(psi0)
        prob.psi.Normalize()
        prob.PropagatedTime = t_start
        eng = prob.Propagator.Engine
        if not timed:
            prob.AdvanceSteps(warmup + steps)
            return dict(prob=prob, psi=prob.psi.GetData().cpu().numpy())
        prob.AdvanceSteps(warmup)
        eng.set_timing(True)
        wall, kern = timed_advance(prob, steps)
        ee_ms, ee_launches = eng.kernel_time()
        eng.set_timing(False)
        return dict(prob=prob, eng=eng, wall=wall, step_s=kern, ee_ms=ee_ms, ee_launches=ee_launches,
                    psi=prob.psi.GetData().cpu().numpy())

    fp = run_form(False, psi_g)
    prob, eng, wall = fp["prob"], fp["eng"], fp["wall"]
    log("helium FP64: %d steps in %.2f s (%.2f steps/s)" % (steps, wall, steps / wall))
    C, N, _ = prob.psi.GetData().shape
    n = C * N * N
    m = prob.Propagator.BasisSize
    dt = prob.TimeStep
    ground = prob.psi.Copy()
    ground.GetData().copy_(pyprop_amd.device.to_device(psi_g))
    ground.Normalize()
    flops = eng.multipole_flops()
    ee_avg_s = fp["ee_ms"] / max(fp["ee_launches"], 1) / 1e3
    # algorithmic bytes of one multipole stage (mixing, near band, far field):
    # read the input state, read-modify-write the output, read the operator
    # (R on the near band, the P_l, Q_l bands; the mixing tables are tiny)
    lo, hi = radial_supports(prob.psi.GetRepresentation().GetRepresentation(1).GetBSplineObject())
    near_w = sum(int(np.sum(~((hi[i1] < lo) | (hi < lo[i1])))) for i1 in range(N))
    nterm = sum(1 for pot in prob.Propagator.BasePropagator.PotentialList for t in pot.Terms
                if isinstance(t, MultipoleTerm))
    ee_bytes = 3 * 16 * n + nterm * 8 * (near_w * 13 * 13 + 2 * N * 13)
    achieved = ee_bytes / ee_avg_s / 1e9
    pop_g = abs(ground.InnerProduct(prob.psi)) ** 2
    diag = prob.Propagator.GetDiagnostics()

    def pop_of(x):
        w = prob.psi.Copy()
        w.GetData().copy_(pyprop_amd.device.to_device(x))
        return abs(ground.InnerProduct(w)) ** 2

    def snorm_diff(a, b):
        w = prob.psi.Copy()
        w.GetData().copy_(pyprop_amd.device.to_device(a - b))
        return w.GetNorm()

    pert = run_form(False, psi_g * (1 + 1e-16 * rng.standard_normal(psi_g.shape)), timed=False)
    floor = {"population_error": abs(pop_of(pert["psi"]) - pop_g), "psi_diff_snorm": snorm_diff(pert["psi"], fp["psi"])}
    del pert
    mixed = None
    if not args.no_mixed:
        mx = run_form(True, psi_g)
        mixed = {"value": steps / mx["wall"], "unit": "steps/s", "ms_per_step": 1e3 * mx["wall"] / steps,
                 "speedup_vs_fp64": wall / mx["wall"],
                 "multipole_stage_ms_avg": mx["ee_ms"] / max(mx["ee_launches"], 1),
                 "multipole_stage_traffic": HE_PMC_TRAFFIC.get(name + "_mixed"),
                 "population_error": abs(pop_of(mx["psi"]) - pop_g),
                 "psi_diff_snorm": snorm_diff(mx["psi"], fp["psi"]),
                 "fp64_floor": floor, "steps_compared": warmup + steps,
                 "dtype": "multipole terms FP32 (complex64 intermediates); one-electron and laser terms, "
                          "Gram-Schmidt, overlap solves, Pade FP64"}
        log("helium mixed: %.2f steps/s, population error %.2e (FP64 floor %.2e)"
            % (mixed["value"], mixed["population_error"], floor["population_error"]))
        del mx
        if getattr(args, "mixed_study", False):
            full = int(round(prob.Duration / dt.real))
            log("helium mixed form over the workload's %d steps from t = 0" % full)

            def set_form(m):
                eng.set_mixed(m)
                return prob
            mixed["window"] = window_study(set_form, ground.GetData().cpu().numpy(), 0.0, full, 500,
                                           lambda a, b: {"population_error": abs(pop_of(a) - pop_of(b)),
                                                         "psi_diff_snorm": snorm_diff(a, b)})
            fin = mixed["window"]["final"]
            log("helium mixed after %d steps: population error %.2e (FP64 floor %.2e)"
                % (full, fin["mixed"]["population_error"], fin["fp64_floor"]["population_error"]))

    cpu = None
    if not args.no_cpu:
        # the oracle's factored restatement (the reference layout of the e-e
        # term would be ~48 GB): same terms, same start state
        O = oracle_for_baseline()
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
        ts = O.step_times(12, dt.real, t_start)
        sc = np.array([[p.GetTimeScaling(tt) for p in dep] for tt in ts])
        threads = host_threads()
        psi0 = ground.GetData().cpu().numpy()

        def orc_run(k):
            ham.propagate_pamp(psi0, k, dt.real, sc[:k], m)

        rate = cpu_rate(O, orc_run, 10, threads, warm=1, reps=1)
        cpu = {"value": rate, "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": "10 pAMP steps of the same workload after 1 warm-up step, one sample (C oracle, factored "
                         "operator, complex128, -O3 %s build; OpenMP over output channels, fibers, whole dot "
                         "products and elements on %d threads, bitwise equal to 1 thread; Pade serial)"
                         % (O.fastest_build(), threads)}
        log("helium CPU baseline: %.3f steps/s on %d threads" % (rate, threads))

    return {
        "value": steps / wall, "unit": "steps/s", "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * wall / steps, "dtype": "c128",
        "config": {"workload": name, "channels": C, "N": N, "n": n, "order": 7, "krylov_dim": m,
                   "dt": dt.real, "multipole_lmax": 10, "start_time": t_start, "form": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": HE_PMC_TRAFFIC.get(name),
                     "kernel": "multipole stage (p3_mixT + p3_near + p3_farL), one launch each per matvec",
                     "algorithmic_bytes_per_launch": ee_bytes, "kernel_ms_avg": 1e3 * ee_avg_s,
                     "launches": fp["ee_launches"], "kernel_share_of_step": fp["ee_ms"] / 1e3 / fp["step_s"],
                     "algorithmic_flops_per_launch": flops,
                     "fp64_tflops": flops / ee_avg_s / 1e12, "fp64_peak_tflops": FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": None if cpu is None else (steps / wall) / cpu["value"],
        "checks": {"ground_energy": E0, "ground_state_seconds": gs_s, "imag_steps": args.imag_steps,
                   "norm_after": prob.psi.GetNorm(), "ground_population_after": pop_g,
                   "krylov_error_estimate_last": float(diag[-1, 0]), "basis_used_last": int(diag[-1, 1])},
        "mixed": mixed,
    }


def run_helium_line(args):
    """--workload helium: the helium object as its own line."""
    he = run_helium(args, args.steps, args.warmup, args.workload)
    line = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "data": "synthetic (deterministic helium operator, ground state by imaginary-time pAMP)"}
    line.update(he)
    print(json.dumps(line))


def run_batch(args):
    """Batched scan (config 3): 512 seeded hydrogen instances (N 300-600,
    lmax 20-60, pulse drawn per instance), Krylov dimension 16, dt = 0.01,
    mixed form, advanced together by packed persistent launches recorded as
    one HIP graph (replayed for the timed steps); chunk lengths of the band
    solves picked by the autotuner."""
    import numpy as np
    import torch
    import pyprop_amd
    from pyprop_amd import batch as Bt
    from pyprop_amd.examples import hydrogen as H

    count = int(args.workload[len("batch"):] or 512)
    draws = Bt.instance_draws(count)
    t0 = time.perf_counter()
    tuned = Bt.autotune_chunks([(d["N"], d["lmax"] + 1) for d in draws], kd=6)
    tune_s = time.perf_counter() - t0
    print("autotuned %d radial sizes in %.1f s" % (len(tuned), tune_s), file=sys.stderr, flush=True)
    probs, psi0s = [], []
    t0 = time.perf_counter()
    for i, d in enumerate(draws):
        prob = pyprop_amd.Problem(H.batch_instance_config(d, precision="mixed"))
        prob.SetupStep()
        psi0 = H.ground_state(prob)
        prob.SetWavefunction(psi0)
        prob.PropagatedTime = prob.StartTime = 0.5 * 2 * np.pi / d["omega"] * d["cycles"]   # mid-pulse
        probs.append(prob)
        psi0s.append(psi0)
        if (i + 1) % 64 == 0:
            print("set up %d/%d instances (%.0f s)" % (i + 1, count, time.perf_counter() - t0), file=sys.stderr,
                  flush=True)
    batch = Bt.Batch(probs)
    K = args.steps

    def run(mixed, perturb=False):
        rng = np.random.default_rng(13)
        for p, psi0 in zip(probs, psi0s):
            p.SetWavefunction(psi0 * (1 + 1e-16 * rng.standard_normal(psi0.shape)) if perturb else psi0)
            p.PropagatedTime = p.StartTime
            eng = p.Propagator.Engine
            eng.set_mixed(list(range(len(p.Propagator._pots))) if mixed else [])
        batch.AdvanceSteps(args.warmup)
        # the timed advance replays one HIP graph of all packed launches
        batch.CaptureGraph(K)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record()
        batch.AdvanceSteps(K)
        ev1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, ev0.elapsed_time(ev1) / 1e3, [p.psi.GetData().cpu().numpy() for p in probs]

    wall_m, kern_m, psi_m = run(True)
    wall_d, kern_d, psi_d = run(False)
    _, _, psi_p = run(False, perturb=True)
    # per-instance error of the mixed form against the FP64 batch, next to
    # the FP64 batch's own change under a 1e-16 perturbation of psi0: S-norm
    # of the differences for every instance, bound populations on a sample
    def snorm(p, d):
        S = p.psi.GetRepresentation().GetRepresentation(1).GetBSplineObject().overlap_dense()
        return float(np.sqrt(sum(np.real(np.conj(d[l]) @ S @ d[l]) for l in range(d.shape[0]))))
    dn = [snorm(p, a - b) for p, a, b in zip(probs, psi_m, psi_d)]
    fl = [snorm(p, a - b) for p, a, b in zip(probs, psi_p, psi_d)]
    sample = list(range(0, count, max(1, count // 8)))[:8]
    perr, pfloor = [], []
    for i in sample:
        proj = H.bound_projectors(probs[i])
        pa = H.populations(probs[i], psi_m[i], projectors=proj)[1]
        pb = H.populations(probs[i], psi_d[i], projectors=proj)[1]
        pc = H.populations(probs[i], psi_p[i], projectors=proj)[1]
        perr.append(max(float(np.max(np.abs(x - y))) if len(x) else 0.0 for x, y in zip(pa, pb)))
        pfloor.append(max(float(np.max(np.abs(x - y))) if len(x) else 0.0 for x, y in zip(pc, pb)))
    n_tot = sum(p.psi.GetData().numel() for p in probs)
    Bsum = sum(step_bytes(p.psi.GetData().numel(), p.Propagator.BasisSize, p.psi.GetData().shape[1]) for p in probs)
    achieved = Bsum * K / kern_m / 1e9

    cpu = None
    if not args.no_cpu:
        O = oracle_for_baseline()
        d = draws[0]
        h = O.Hydrogen(rmax=0.5 * d["N"], nbp=d["N"] - 3, order=7, lmax=d["lmax"], frequency=d["omega"],
                       cycles=d["cycles"], E0=d["E0"])
        ham = h.ham()
        ts = O.step_times(60, 0.01, probs[0].StartTime)
        sc = np.array([[1.0, h.scaling(t)] for t in ts])
        g0 = h.ground_state()
        threads = host_threads()

        def orc_run(k):
            ham.propagate_pamp(g0, k, 0.01, sc[:k], 16)

        one = cpu_rate(O, orc_run, 10, 1)
        allc = cpu_rate(O, orc_run, 40, threads) if threads > 1 else one
        cpu = {"value": allc, "unit": "instance-steps/s", "cores": threads, "kind": "port", "single_core_value": one,
               "sample": "pAMP steps of instance 0 (N=%d, lmax=%d; C oracle, complex128, -O3 %s build), 10 warm-up "
                         "steps, median of 3 samples of 40 steps on %d threads and of 10 on one"
                         % (d["N"], d["lmax"], O.fastest_build(), threads)}
    line = {
        "metric": METRIC,
        "value": count * K / wall_m,
        "unit": "instance-steps/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_m / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128 (operator apply c64)",
        "data": "synthetic (seeded instance draws, 1s start states by inverse iteration)",
        "config": {"workload": args.workload, "instances": count, "amplitudes": n_tot, "krylov_dim": 16,
                   "dt": 0.01, "form": "mixed",
                   "launch": "one HIP graph replay per advance: packed cooperative kernel nodes, autotuned chunks"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "pp_pamp_batch_kernel",
                     "algorithmic_bytes_per_batch_step": Bsum},
        "cpu_baseline": cpu,
        "fp64": {"value": count * K / wall_d, "unit": "instance-steps/s", "ms_per_step": 1e3 * wall_d / K},
        "mixed_vs_fp64": {"psi_diff_snorm_max": max(dn), "psi_diff_snorm_median": float(np.median(dn)),
                          "population_error_per_sampled_instance": perr, "sample": sample,
                          "fp64_floor": {"psi_diff_snorm_max": max(fl), "psi_diff_snorm_median": float(np.median(fl)),
                                         "population_error_per_sampled_instance": pfloor},
                          "start_time": "middle of every pulse"},
        "autotune": {"sizes": len(tuned), "seconds": tune_s,
                     "chosen": {str(c): sum(1 for b, _, _ in tuned.values() if b == c) for c in (16, 32, 64)},
                     "two_workgroups_per_cu": sum(1 for _, _, o in tuned.values() if o >= 2)},
    }
    print(json.dumps(line))


def run_cayley(args):
    """CayleyPropagator: Crank-Nicolson steps solved by the single-cycle GMRES
    (tolerance eps^(2/3)) with the config's left preconditioner.
    hydrogen_cayley: the hydrogen_large operator (L = 64, N = 800), GMRES
    basis 20, BSplinePreconditioner, from the 1s state at t = 20.
    helium_cayley: the helium workload (66 channels x 120 x 120), GMRES basis
    20, RadialTwoElectronPreconditioner (fast diagonalisation; its solve
    timed with HIP events on the launch stream and priced against the FP64
    peak: 4 passes of 4 N^3 flops per channel), from the ground state at
    t = 10."""
    import torch
    import pyprop_amd
    helium = args.workload == "helium_cayley"
    if helium:
        from pyprop_amd.examples import helium as HE
        psi_g, E0 = HE.ground_state("helium", steps=args.imag_steps, dt=0.05)
        conf = HE.load("helium")
        t_start = 10.0
    else:
        from pyprop_amd.examples import hydrogen as H
        conf = H.load("hydrogen_large")
        t_start = 20.0
    conf.Propagation.propagator = pyprop_amd.CayleyPropagator
    conf.Propagation.preconditioner = "RadialPreconditioner"
    conf.Propagation.krylov_basis_size = 20
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    prob.SetWavefunction(psi_g if helium else H.ground_state(prob))
    prob.psi.Normalize()
    prob.PropagatedTime = t_start
    warm = min(args.warmup, 5)
    log("cayley: %d warmup steps" % warm)
    prob.AdvanceSteps(warm)
    solver = prob.Propagator.Solver
    ops0 = solver.GetOperatorCount()
    K = args.steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prob.AdvanceSteps(K)
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - t0
    shape = tuple(prob.psi.GetData().shape)
    line = {
        "metric": METRIC,
        "value": K / dt_s, "unit": "steps/s",
        "n_gpus": 1, "steps": K, "warmup": warm, "ms_per_step": 1e3 * dt_s / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": ("synthetic (deterministic helium operator, ground state by imaginary-time pAMP)" if helium else
                 "synthetic (deterministic hydrogen operator, 1s start state by inverse iteration)"),
        "config": {"workload": args.workload, "shape": shape, "gmres_basis": 20, "dt": prob.TimeStep.real,
                   "start_time": t_start,
                   "preconditioner": type(prob.Propagator.Preconditioner).__name__},
        "operator_calls_per_step": (solver.GetOperatorCount() - ops0) / K,
        "gmres_error_estimate_last": solver.GetErrorEstimate(),
        "norm_after": prob.psi.GetNorm(),
    }
    if helium:
        C, N, _ = shape
        prec = prob.Propagator.Preconditioner
        x = prob.psi.Copy()
        reps = 50
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            prec.Solve(x)
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3 / reps
        flops = 4 * 4.0 * C * N ** 3
        line["roofline"] = {"bound": "fp64", "achieved": flops / sec / 1e12, "peak": FP64_PEAK_TFLOPS,
                            "unit": "TFLOP/s", "frac": flops / sec / 1e12 / FP64_PEAK_TFLOPS, "traffic": None,
                            "kernel": "pr_mm x 4 (RadialTwoElectronPreconditioner.Solve)",
                            "algorithmic_flops_per_solve": flops, "solve_us": sec * 1e6,
                            "algorithmic_bytes_per_solve": 4 * 2 * 16 * C * N * N}
        line["checks"] = {"ground_energy": E0}
    print(json.dumps(line))


def run_adjoint(args):
    """Adjoint gradient (config 4): hydrogen N=400 on r <= 800, l=0..31, m=24,
    dt = 0.005, 4000 control samples, objective the final 2p population;
    dP/ds_k for every sample by the adjoint sweep (checkpoints every 50
    steps), checked against central differences on 32 seeded samples
    (h = 1e-4, and h = 1e-6 next to the population's own rounding floor), all
    forward runs of a check in one batched set of packed launches."""
    import numpy as np
    import torch
    import pyprop_amd
    from pyprop_amd.adjoint import PopulationAdjoint, central_differences
    from pyprop_amd.examples import hydrogen as H

    conf = H.load("hydrogen_adjoint")
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    prob.SetWavefunction(H.ground_state(prob))
    target, _ = H.bound_state(prob, 1, 0)
    dt = prob.TimeStep.real
    K = args.steps
    s = H.adjoint_control_samples(K, dt)
    adj = PopulationAdjoint(prob, target, checkpoint=50)
    adj.gradient(s[:60])   # warm up both engines
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P, g = adj.gradient(s)
    grad_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    adj.population(s)
    fwd_s = time.perf_counter() - t0
    krylov_est = adj.engine.diagnostics()
    log("adjoint: gradient of %d samples in %.1f s (forward %.1f s)" % (K, grad_s, fwd_s))
    idx = np.sort(np.random.default_rng(7).choice(K, size=min(32, K), replace=False))
    checks = {}
    for h in (1e-4, 1e-6):
        t0 = time.perf_counter()
        fd = central_differences(adj, s, idx, h)
        err = np.abs(g[idx] - fd)
        checks["h=%g" % h] = {"samples": len(idx), "max_abs_err": float(err.max()),
                              "max_rel_err_vs_max_grad": float(err.max() / np.abs(g).max()),
                              "batched_forward_runs": 2 * len(idx), "seconds": time.perf_counter() - t0}
    # the population's rounding floor: runs from psi0 (1 + 1e-16 eta)
    from pyprop_amd import device as D
    eng0, pots = adj.engine, prob.Propagator._pots
    pairs = pots[1].GeometryList[0].GetBasisPairs()
    rng = np.random.default_rng(3)
    x0 = adj.psi0.cpu().numpy()
    engines, psis = [], []
    for j in range(4):
        e = D.PampEngine(eng0.L, eng0.N, eng0.m, eng0.overlap)
        e.add_term("Diag", pots[0].DeviceData())
        e.add_term("Simp", pots[1].DeviceData(), pairs=pairs)
        e.set_pade_scaling(1)
        engines.append(e)
        psis.append(D.to_device(x0 if j == 0 else x0 * (1 + 1e-16 * rng.standard_normal(x0.shape))))
    D.pamp_batch_propagate(engines, psis, K, [dt] * 4, [np.stack([np.ones(K), s], axis=1)] * 4)
    pops = np.array([abs(adj._sdot(adj.target, p)) ** 2 for p in psis])
    noise = float(np.abs(pops[1:] - pops[0]).max())
    oracle_fd = None
    if not args.no_cpu:
        O = oracle_for_baseline()
        O.set_threads(host_threads())
        K0 = 120
        _, g0 = adj.gradient(s[:K0])
        h_or = O.Hydrogen(rmax=conf.RadialRepresentation.xmax, nbp=conf.RadialRepresentation.xsize, order=7,
                          lmax=conf.AngularRepresentation.maxl, frequency=0.3, cycles=10, E0=0.05)
        ham = h_or.ham()
        S = h_or.overlap_dense()
        m = prob.Propagator.BasisSize

        def pop(sv):
            psi = ham.propagate_pamp(x0, K0, dt, np.stack([np.ones(K0), sv], axis=1), m)
            return abs(np.sum(np.conj(target) * (psi @ S.T))) ** 2

        rows = []
        for k in (17, 88):
            sp, sm = s[:K0].copy(), s[:K0].copy()
            sp[k] += 1e-6
            sm[k] -= 1e-6
            rows.append(abs(g0[k] - (pop(sp) - pop(sm)) / 2e-6) / np.abs(g0).max())
        oracle_fd = {"prefix_samples": K0, "indices": [17, 88], "h": 1e-6, "max_rel_err_vs_max_grad": max(rows)}
    line = {
        "metric": METRIC,
        "value": K / grad_s, "unit": "adjoint-sweep steps/s",
        "n_gpus": 1, "steps": K, "warmup": 60, "ms_per_step": 1e3 * grad_s / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded control samples, 1s start state by inverse iteration)",
        "config": {"workload": "hydrogen_adjoint", "n": adj.L * adj.N, "L": adj.L, "N": adj.N, "rmax":
                   conf.RadialRepresentation.xmax, "krylov_dim": adj.m, "dt": dt, "checkpoint": 50,
                   "objective": "final 2p population"},
        "population_2p": P, "forward_s": fwd_s, "gradient_s": grad_s,
        "fd_checks": checks, "population_noise_floor": noise,
        "fd_noise_bound_h1e-6": 3 * 2 * noise / 2e-6 / float(np.abs(g).max()),
        "oracle_fd_check": oracle_fd,
        "krylov_error_estimate_median": float(np.median(krylov_est[:, 0])),
        "grad_norm": float(np.linalg.norm(g)),
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the workload's own count; 20000 for hydrogen_large)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="hydrogen_large",
                    help="hydrogen_large (default; with the helium object), hydrogen_small, helium, "
                         "batch512, hydrogen_adjoint, hydrogen_cayley, helium_cayley")
    ap.add_argument("--imag-steps", type=int, default=150, help="imaginary-time steps to the helium ground state")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision runs")
    ap.add_argument("--no-helium", action="store_true", help="default line without the helium object")
    ap.add_argument("--mixed-study", action="store_true",
                    help="helium: also compare the mixed form with FP64 and the FP64 floor over the workload's "
                         "whole duration (5000 steps, several minutes)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB before the workload, so that a "
                         "rocprofv3 --pmc run can scale FETCH_SIZE/WRITE_SIZE by a copy of known size")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"hydrogen_adjoint": 4000, "helium": 300, "helium_small": 50, "hydrogen_cayley": 400,
                      "helium_cayley": 300}.get(
            args.workload, 20 if args.workload.startswith("batch") else 20000)
    if args.gpus != 1:
        print("bench.py: this build runs one GPU holding the whole wavefunction; "
              "--gpus %d is not supported (sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration_mib > 0:
        import torch
        nb = args.copy_calibration_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload.startswith("batch"):
        return run_batch(args)
    if args.workload == "hydrogen_adjoint":
        return run_adjoint(args)
    if args.workload in ("hydrogen_cayley", "helium_cayley"):
        return run_cayley(args)
    if args.workload.startswith("helium"):
        return run_helium_line(args)
    line = run_hydrogen(args)
    if args.workload == "hydrogen_large" and not args.no_helium:
        line["helium"] = run_helium(args, min(args.steps, 300), min(args.warmup, 10))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
