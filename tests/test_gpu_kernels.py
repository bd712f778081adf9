"""HIP path vs the CPU oracle on identical inputs (hydrogen operator)."""
import numpy as np
import pytest

from oracle import oracle as O

try:
    import torch  # noqa: F401  (first import can take a minute on a fresh box)
except ImportError:  # pragma: no cover
    torch = None

pytestmark = pytest.mark.gpu


def _hydrogen(lmax=6, nbp=40, rmax=40.0, order=7):
    return O.Hydrogen(rmax=rmax, nbp=nbp, order=order, lmax=lmax, frequency=0.057, cycles=6, E0=0.05)


@pytest.mark.parametrize("order", [3, 4, 5, 6, 7, 8, 9])
def test_overlap_solves_match_oracle(order):
    """U^-1, U^-H and U x on the B-spline rank for every supported order
    (kd = order - 1 = 2..8) against the oracle's ztbsv / ztbmv restatement."""
    from pyprop_amd import device as D
    torch = D.torch()
    h = _hydrogen(order=order)
    ov = D.Overlap(h.S, h.N, h.kd)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(h.shape) + 1j * rng.standard_normal(h.shape)
    for op, oop in [(0, O.OP_SOLVE_N), (1, O.OP_SOLVE_C), (2, O.OP_MUL_N)]:
        xd = D.to_device(x)
        ov.apply(op, 1, xd)
        ref = O.band_apply(oop, 1, x, h.U, h.kd)
        got = xd.cpu().numpy()
        print("op", op, np.abs(got - ref).max())
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_tensor_multiply_matches_oracle():
    from pyprop_amd import device as D
    h = _hydrogen()
    rng = np.random.default_rng(2)
    x = rng.standard_normal(h.shape) + 1j * rng.standard_normal(h.shape)
    for st, pot, pairs in [((O.ST_DIAG, O.ST_BANDNH), h.H0, None), ((O.ST_SIMP, O.ST_BANDNH), h.Hint, [h.pairs, None])]:
        ref = np.zeros_like(x)
        O.tp_multiply(st, pot, 0.7, x, ref, pairs)
        xd = D.to_device(x)
        yd = D.to_device(np.zeros_like(x))
        D.tp_multiply(st, D.to_device(pot), 0.7, xd, yd, pairs)
        got = yd.cpu().numpy()
        print(st, np.abs(got - ref).max())
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def _storage_case(st, n, rng):
    """Potential extent and pairs of one rank for storage id st on extent n
    (the generator's layouts: Ident 1, Diag n, Simp/Herm one per pair,
    BandNH n*5 (two off-diagonals), Band n*3 (lower hermitian, k = 3))."""
    if st == O.ST_IDENT:
        return 1, None
    if st == O.ST_DIAG:
        return n, None
    if st in (O.ST_SIMP, O.ST_HERM):
        allp = [(a, b) for a in range(n) for b in range(n) if st == O.ST_SIMP or a <= b]
        pick = rng.choice(len(allp), size=min(len(allp), 2 * n), replace=False)
        pr = np.array([allp[i] for i in sorted(pick)], dtype=np.int32)
        return len(pr), pr
    if st == O.ST_BANDNH:
        return n * 5, None
    return n * 3, None


@pytest.mark.parametrize("storages", [
    (O.ST_IDENT,), (O.ST_DIAG,), (O.ST_SIMP,), (O.ST_HERM,), (O.ST_BANDNH,), (O.ST_BAND,),
    (O.ST_DIAG, O.ST_BANDNH), (O.ST_SIMP, O.ST_BANDNH), (O.ST_HERM, O.ST_BAND), (O.ST_IDENT, O.ST_BAND),
    (O.ST_DIAG, O.ST_DIAG), (O.ST_SIMP, O.ST_SIMP, O.ST_BANDNH), (O.ST_DIAG, O.ST_HERM, O.ST_BAND),
    (O.ST_IDENT, O.ST_SIMP, O.ST_BANDNH), (O.ST_BAND, O.ST_BANDNH), (O.ST_BAND, O.ST_DIAG), (O.ST_BAND, O.ST_BAND),
    (O.ST_BAND, O.ST_HERM, O.ST_BAND), (O.ST_HERM, O.ST_BAND, O.ST_BANDNH)])
def test_tensor_multiply_all_storages_match_oracle(storages):
    """pp_tp_multiply against the oracle's restatement of the generated
    TensorPotentialMultiply_<ids> loops for every storage id of the
    generator (Ident, Diag, Simp, Herm with its conjugate mirror, BandNH,
    Band: zhbmv's reading on the innermost rank, the upper-band loops on an
    outer rank) at ranks 1-3, random complex potentials, dst accumulated
    (+=): relative 1e-13."""
    from pyprop_amd import device as D
    rng = np.random.default_rng(sum(storages) * 7 + len(storages))
    shape = (9, 7, 11)[:len(storages)]
    ext, pairs = [], []
    for st, n in zip(storages, shape):
        e, pr = _storage_case(st, n, rng)
        ext.append(e)
        pairs.append(pr)
    pot = rng.standard_normal(ext) + 1j * rng.standard_normal(ext)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    y0 = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = y0.copy()
    O.tp_multiply(storages, pot, -0.45, x, ref, pairs)
    yd = D.to_device(y0)
    D.tp_multiply(storages, D.to_device(pot), -0.45, D.to_device(x), yd, pairs)
    err = np.abs(yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    print("storages", storages, "rel err", err, flush=True)
    assert err <= 1e-13


@pytest.mark.parametrize("lmax,nbp,m,nsteps,dt,reorth", [
    (6, 40, 12, 5, 0.05, False), (10, 100, 20, 20, 0.01, False), (10, 100, 20, 20, 0.05, False),
    (6, 40, 12, 5, 0.05, True), (63, 120, 24, 6, 0.005, False),
    (6, 40, 5, 5, 0.05, False), (10, 100, 13, 10, 0.01, False), (10, 100, 40, 5, 0.02, False)])
def test_pamp_persistent_matches_oracle(lmax, nbp, m, nsteps, dt, reorth):
    """psi after nsteps pAMP steps, GPU vs oracle.  The bar is the oracle's own
    sensitivity: the change a 1e-16 relative perturbation of psi0 causes (an
    unconverged Krylov space amplifies rounding), times 10, floor 1e-12.
    Krylov dimensions 5, 13 and 40 take the Pade's partial 16 x 16 product
    tiles, an odd pivot-column length and more than two tiles per side."""
    from pyprop_amd import device as D
    h = _hydrogen(lmax=lmax, nbp=nbp, rmax=float(nbp))
    psi0 = h.ground_state()
    # a little excited component so the dipole term matters from step 0
    psi0[1] = 0.1 * psi0[0]
    ts = O.step_times(nsteps, dt, t0=10.0)
    sc = np.array([[1.0, h.scaling(t)] for t in ts])
    ham = h.ham()
    ref = ham.propagate_pamp(psi0, nsteps, dt, sc, m, reorth)
    rng = np.random.default_rng(11)
    ref2 = ham.propagate_pamp(psi0 * (1 + 1e-16 * rng.standard_normal(psi0.shape)), nsteps, dt, sc, m, reorth)
    tol = max(1e-12, 10 * np.abs(ref2 - ref).max())
    ov = D.Overlap(h.S, h.N, h.kd)
    eng = D.PampEngine(h.L, h.N, m, ov, reorth=reorth)
    eng.add_term("Diag", D.to_device(h.H0))
    eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
    psi = D.to_device(psi0)
    eng.propagate(psi, nsteps, dt, sc)
    got = psi.cpu().numpy()
    err = np.abs(got - ref).max()
    print("pamp", lmax, nbp, m, nsteps, dt, reorth, "err", err, "tol", tol, eng.diagnostics()[-1])
    assert err <= tol


@pytest.mark.parametrize("order", [3, 4, 8])
def test_pamp_other_spline_orders_match_oracle(order):
    """pAMP steps (fiber-split and one-workgroup-per-fiber kernels) and an
    Expokit step on B-splines of order 3, 4 and 8 (kd = 2, 3, 7: the band
    width of the operator and of the overlap factor change with the order)
    against the oracle; same sensitivity bar as the order-7 tests."""
    from pyprop_amd import device as D
    h = _hydrogen(lmax=6, nbp=40, rmax=40.0, order=order)
    psi0 = h.ground_state()
    psi0[1] = 0.1 * psi0[0]
    nsteps, dt, m = 5, 0.05, 12
    ts = O.step_times(nsteps, dt, t0=10.0)
    sc = np.array([[1.0, h.scaling(t)] for t in ts])
    ham = h.ham()
    ref = ham.propagate_pamp(psi0, nsteps, dt, sc, m)
    rng = np.random.default_rng(13)
    ref2 = ham.propagate_pamp(psi0 * (1 + 1e-16 * rng.standard_normal(psi0.shape)), nsteps, dt, sc, m)
    tol = max(1e-12, 10 * np.abs(ref2 - ref).max())
    ov = D.Overlap(h.S, h.N, h.kd)
    for split in (0, 1):
        eng = D.PampEngine(h.L, h.N, m, ov)
        eng.add_term("Diag", D.to_device(h.H0))
        eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
        eng.set_fiber_split(split)
        psi = D.to_device(psi0)
        eng.propagate(psi, nsteps, dt, sc)
        err = np.abs(psi.cpu().numpy() - ref).max()
        print("order", order, "split", split, "pAMP err", err, "tol", tol, flush=True)
        assert err <= tol
    eref, _ = ham.propagate_expokit(psi0, 1, dt, sc[:1], m)
    eng = D.PampEngine(h.L, h.N, m, ov)
    eng.add_term("Diag", D.to_device(h.H0))
    eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
    psi = D.to_device(psi0)
    eng.propagate_expokit(psi, 1, dt, sc[:1])
    err = np.abs(psi.cpu().numpy() - eref).max() / np.abs(eref).max()
    print("order", order, "Expokit rel err", err, flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("lmax,nbp,m,nsteps,dt,split", [
    (6, 40, 12, 5, 0.05, 0), (6, 40, 12, 5, 0.05, 1), (10, 100, 20, 20, 0.01, 0), (10, 100, 21, 10, 0.01, 1)])
def test_pamp_eigenvector_exponential_matches_oracle(lmax, nbp, m, nsteps, dt, split):
    """krylov_exponentiation_method = 1 (pamp.h:567-607): the step kernels'
    Jacobi eigen-decomposition against the oracle's, which calls LAPACK's
    zgeev through numpy as the reference does; fiber-split and
    one-workgroup-per-fiber kernels, even and odd m.  Same bar as the Pade
    test (10x the oracle's 1e-16 sensitivity, floor 1e-12); and the
    eigenvector and Pade forms of the GPU step agree to 1e-11 (both are
    exp(-i dt H_m) e1 for the Hermitian Krylov matrix)."""
    from pyprop_amd import device as D
    h = _hydrogen(lmax=lmax, nbp=nbp, rmax=float(nbp))
    psi0 = h.ground_state()
    psi0[1] = 0.1 * psi0[0]
    ts = O.step_times(nsteps, dt, t0=10.0)
    sc = np.array([[1.0, h.scaling(t)] for t in ts])
    ham = h.ham()
    calls = O.EIG_EXPFN_CALLS[0]
    ref = ham.propagate_pamp(psi0, nsteps, dt, sc, m, False, exponentiation=1)
    assert O.EIG_EXPFN_CALLS[0] == calls + nsteps
    rng = np.random.default_rng(12)
    ref2 = ham.propagate_pamp(psi0 * (1 + 1e-16 * rng.standard_normal(psi0.shape)), nsteps, dt, sc, m, False,
                              exponentiation=1)
    tol = max(1e-12, 10 * np.abs(ref2 - ref).max())
    ov = D.Overlap(h.S, h.N, h.kd)
    out = {}
    for method in (1, 0):
        eng = D.PampEngine(h.L, h.N, m, ov)
        eng.add_term("Diag", D.to_device(h.H0))
        eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
        eng.set_fiber_split(split)
        eng.set_exponentiation(method)
        psi = D.to_device(psi0)
        eng.propagate(psi, nsteps, dt, sc)
        out[method] = psi.cpu().numpy()
    err = np.abs(out[1] - ref).max()
    print("eigenvector exponential", lmax, nbp, m, nsteps, dt, split, "err", err, "tol", tol,
          "vs GPU Pade", np.abs(out[1] - out[0]).max(), flush=True)
    assert err <= tol
    assert np.abs(out[1] - out[0]).max() <= 1e-11 * np.abs(out[0]).max()


@pytest.mark.parametrize("lmax,nbp,m,nsteps,dt", [(10, 100, 20, 20, 0.01), (63, 120, 24, 6, 0.005)])
@pytest.mark.parametrize("mask", [(1,), (0, 1)])
def test_pamp_mixed_close_to_oracle(lmax, nbp, m, nsteps, dt, mask):
    """Mixed form against the FP64 oracle, with the dipole term alone or
    every term (the matvec_precision = "mixed" form) applied in complex
    FP32: the difference must exceed 1e-13 of max|psi| (no silent FP64
    path) and stay within FP32 rounding amplified by the step (1e-6 for the
    dipole alone, 1e-4 for every term: the kinetic and centrifugal bands
    are large and nearly cancel) or 100x the oracle's own sensitivity to a
    1e-16 perturbation; the norm is conserved to 1e-8."""
    from pyprop_amd import device as D
    h = _hydrogen(lmax=lmax, nbp=nbp, rmax=float(nbp))
    psi0 = h.ground_state()
    psi0[1] = 0.1 * psi0[0]
    ts = O.step_times(nsteps, dt, t0=10.0)
    sc = np.array([[1.0, h.scaling(t)] for t in ts])
    ham = h.ham()
    ref = ham.propagate_pamp(psi0, nsteps, dt, sc, m, False)
    rng = np.random.default_rng(11)
    ref2 = ham.propagate_pamp(psi0 * (1 + 1e-16 * rng.standard_normal(psi0.shape)), nsteps, dt, sc, m, False)
    sens = np.abs(ref2 - ref).max() / np.abs(ref).max()
    ov = D.Overlap(h.S, h.N, h.kd)
    eng = D.PampEngine(h.L, h.N, m, ov)
    eng.add_term("Diag", D.to_device(h.H0))
    eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
    eng.set_mixed(list(mask))
    psi = D.to_device(psi0)
    eng.propagate(psi, nsteps, dt, sc)
    got = psi.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("mixed", mask, lmax, nbp, m, nsteps, dt, "rel err %.3e, oracle sensitivity %.3e" % (err, sens), flush=True)
    assert 1e-13 < err < max(1e-6 if mask == (1,) else 1e-4, 100 * sens)
    assert abs(h.norm(got) - h.norm(ref)) < 1e-8


@pytest.mark.parametrize("lmax,nbp,m,q,mixed", [
    (63, 120, 24, 0, False), (63, 120, 24, 2, False), (10, 100, 20, 3, False), (10, 100, 20, 7, False),
    (31, 300, 16, 8, False), (63, 120, 24, 0, True)])
def test_pamp_fiber_split_matches_one_workgroup_per_fiber(lmax, nbp, m, q, mixed):
    """Several workgroups per fiber (automatic and forced counts, including
    ones that do not divide the chunk count) against one workgroup per fiber
    on the same inputs: only the grouping of the partial sums differs, so the
    results agree to rounding (FP64: 1e-12; mixed: 1e-9, since a different
    last bit of an FP64 input can round its FP32 conversion the other way);
    the norm is conserved and the Krylov diagnostics match."""
    from pyprop_amd import device as D
    h = _hydrogen(lmax=lmax, nbp=nbp, rmax=float(nbp))
    psi0 = h.ground_state()
    psi0[1] = 0.1 * psi0[0]
    nsteps, dt = 6, 0.01
    ts = O.step_times(nsteps, dt, t0=10.0)
    sc = np.array([[1.0, h.scaling(t)] for t in ts])
    out, diags = [], []
    for split in (1, q):
        ov = D.Overlap(h.S, h.N, h.kd)
        eng = D.PampEngine(h.L, h.N, m, ov)
        eng.add_term("Diag", D.to_device(h.H0))
        eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
        if mixed:
            eng.set_mixed([1])
        eng.set_fiber_split(split)
        print("fiber split", lmax, nbp, m, "q", split, "mixed", mixed, flush=True)
        psi = D.to_device(psi0)
        eng.propagate(psi, nsteps, dt, sc)
        out.append(psi.cpu().numpy())
        diags.append(eng.diagnostics())
    err = np.abs(out[1] - out[0]).max() / np.abs(out[0]).max()
    print("split vs one workgroup per fiber, rel diff", err)
    assert err < (1e-9 if mixed else 1e-12)
    assert abs(h.norm(out[1]) - h.norm(out[0])) < (1e-10 if mixed else 1e-12)
    assert np.array_equal(diags[0][:, 1], diags[1][:, 1])
    assert np.allclose(diags[0][:, 2], diags[1][:, 2], rtol=1e-13, atol=0)


@pytest.mark.parametrize("lmax,nbp,m,nsteps,dt,anorm", [(6, 40, 12, 3, 0.05, 1e6), (10, 100, 20, 3, 0.01, 1e6),
                                                         (10, 100, 16, 4, 0.02, 50.0)])
def test_expokit_persistent_matches_oracle(lmax, nbp, m, nsteps, dt, anorm):
    """Expokit zgexpv steps on the GPU against the oracle's zgexpv: the same
    substep control (matvecs, substeps, rejections per step) and psi within
    1e-11 of max|psi| (the GPU orthogonalises with CGS2, the oracle with MGS,
    equal in exact arithmetic)."""
    from pyprop_amd import device as D
    h = _hydrogen(lmax=lmax, nbp=nbp, rmax=float(nbp))
    psi0 = h.ground_state()
    psi0[1] = 0.1 * psi0[0]
    ts = O.step_times(nsteps, dt, t0=10.0)
    sc = np.array([[1.0, h.scaling(t)] for t in ts])
    ham = h.ham()
    ref = psi0.copy()
    ref_stats = []
    for s in range(nsteps):   # one oracle call per step to collect its stats
        ref, st = ham.propagate_expokit(ref, 1, dt, sc[s:s + 1], m, 0.0, anorm)
        ref_stats.append(st.copy())
    ov = D.Overlap(h.S, h.N, h.kd)
    eng = D.PampEngine(h.L, h.N, m, ov)
    eng.add_term("Diag", D.to_device(h.H0))
    eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
    psi = D.to_device(psi0)
    eng.propagate_expokit(psi, nsteps, dt, sc, 0.0, anorm)
    got = psi.cpu().numpy()
    st = eng.expokit_stats()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("expokit", lmax, nbp, m, nsteps, dt, anorm, "rel err", err, "gpu stats", st[-1], "oracle", ref_stats[-1])
    for s in range(nsteps):
        # oracle stats: nmult, nexph, nstep, nreject, mbrkdwn
        assert list(st[s, :5].astype(int)) == list(ref_stats[s]), (s, st[s], ref_stats[s])
    assert err < 1e-11


@pytest.mark.parametrize("m", [24, 32])
def test_expokit_bench_shape_matches_oracle(m):
    """Expokit at the hydrogen_large shape (L = 64, N = 800, dt = 0.005):
    at m = 24 on the overlap's default solver plan, at m = 32 (where that
    plan leaves no LDS for zgexpv's workspace) on the longer chunk
    expokit_chunk picks: same substep control as the oracle's zgexpv and
    psi within 1e-11 of max|psi| after 3 steps."""
    from pyprop_amd import device as D
    from pyprop_amd.propagator import expokit_chunk
    h = O.Hydrogen(rmax=400.0, nbp=797, order=7, lmax=63, frequency=0.5, cycles=10, E0=0.1)
    nsteps, dt = 3, 0.005
    B = expokit_chunk(h.L, h.N, m, h.kd, 0)
    print("expokit hydrogen_large shape: L %d N %d m %d chunk %s" % (h.L, h.N, m, B), flush=True)
    assert (B is None) if m == 24 else (B == 64)
    psi0 = h.ground_state()
    psi0[1] = 0.1 * psi0[0]
    ts = O.step_times(nsteps, dt, t0=10.0)
    sc = np.array([[1.0, h.scaling(t)] for t in ts])
    ham = h.ham()
    ref = psi0.copy()
    ref_stats = []
    for s in range(nsteps):
        ref, st = ham.propagate_expokit(ref, 1, dt, sc[s:s + 1], m, 0.0, 1e6)
        ref_stats.append(st.copy())
    ov = D.Overlap(h.S, h.N, h.kd, chunk=B)
    eng = D.PampEngine(h.L, h.N, m, ov)
    eng.add_term("Diag", D.to_device(h.H0))
    eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
    psi = D.to_device(psi0)
    eng.propagate_expokit(psi, nsteps, dt, sc, 0.0, 1e6)
    got = psi.cpu().numpy()
    st = eng.expokit_stats()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("  rel err", err, "gpu stats", st[-1], "oracle", ref_stats[-1], flush=True)
    for s in range(nsteps):
        assert list(st[s, :5].astype(int)) == list(ref_stats[s]), (s, st[s], ref_stats[s])
    assert err < 1e-11


def test_expokit_imaginary_time_matches_oracle():
    """dt = -i tau: w = exp(-tau H^) psi~ (imaginary-time zgexpv) against the
    oracle's zgexpv with imag_time, wrapped in U / U^-1 by hand."""
    from pyprop_amd import device as D
    h = _hydrogen(lmax=4, nbp=40, rmax=40.0)
    psi0 = h.ground_state()
    psi0[1] = 0.3 * psi0[0]
    tau, m = 0.1, 12
    sc = np.array([[1.0, h.scaling(0.0)]])
    ham = h.ham()
    ham.set_scaling(sc[0])
    x = O.band_apply(O.OP_MUL_N, 1, psi0, h.U, h.kd)
    w, rc, st = ham.zgexpv(x, tau, m, 0.0, 1e6, imag_time=True)
    assert rc == 0
    ref = O.band_apply(O.OP_SOLVE_N, 1, w, h.U, h.kd)
    ov = D.Overlap(h.S, h.N, h.kd)
    eng = D.PampEngine(h.L, h.N, m, ov)
    eng.add_term("Diag", D.to_device(h.H0))
    eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
    psi = D.to_device(psi0)
    eng.propagate_expokit(psi, 1, -1j * tau, sc, 0.0, 1e6)
    got = psi.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("expokit imaginary time rel err", err, eng.expokit_stats()[0], st)
    assert err < 1e-11


def test_batch_matches_single_instances():
    """pp_pamp_batch_propagate over 30 independent instances of different
    sizes and pulses (more workgroups than one cooperative launch holds, so
    the batch is split) against each instance propagated alone: identical
    code per fiber, so the results agree to rounding (1e-14)."""
    from pyprop_amd import device as D
    rng = np.random.default_rng(5)
    engines, psis, refs, dts, scs = [], [], [], [], []
    nsteps = 4
    for i in range(30):
        lmax = int(rng.integers(6, 13))
        nbp = int(rng.integers(30, 60))
        h = O.Hydrogen(rmax=float(nbp), nbp=nbp, order=7, lmax=lmax, frequency=float(rng.uniform(0.1, 0.5)),
                       cycles=4, E0=float(rng.uniform(0.02, 0.1)))
        psi0 = h.ground_state()
        psi0[1] = 0.1 * psi0[0]
        dt = 0.02
        ts = O.step_times(nsteps, dt, t0=3.0)
        sc = np.array([[1.0, h.scaling(t)] for t in ts])
        ov = D.Overlap(h.S, h.N, h.kd)
        eng = D.PampEngine(h.L, h.N, 12, ov)
        eng.add_term("Diag", D.to_device(h.H0))
        eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
        ref = D.to_device(psi0)
        eng.propagate(ref, nsteps, dt, sc)
        refs.append(ref.cpu().numpy())
        engines.append(eng)
        psis.append(D.to_device(psi0))
        dts.append(dt)
        scs.append(sc)
    D.pamp_batch_propagate(engines, psis, nsteps, dts, scs)
    worst = max(np.abs(p.cpu().numpy() - r).max() / np.abs(r).max() for p, r in zip(psis, refs))
    print("batch vs single worst rel diff", worst, "total workgroups", sum(e.L for e in engines))
    assert worst < 1e-14
    for e in engines:
        d = e.diagnostics()
        assert np.all(d[:, 1] == 12)


def test_batch_graph_replays_match_direct_launches():
    """The batch recorded as one HIP graph (barrier-reset kernel nodes and
    cooperative batch-kernel nodes, two packs) replayed three times, with
    eager work and synchronisation between the replays: every replay equals
    the direct packed launches on the same inputs bitwise (same kernels,
    same arguments); a replay after an engine changed is refused."""
    from pyprop_amd import device as D
    from pyprop_amd import _native as nat
    t = D.torch()
    rng = np.random.default_rng(7)
    engines, psa, psb, dts, hs = [], [], [], [], []
    nsteps, dt = 3, 0.02
    for i in range(30):
        lmax = int(rng.integers(6, 13))
        nbp = int(rng.integers(30, 60))
        h = O.Hydrogen(rmax=float(nbp), nbp=nbp, order=7, lmax=lmax, frequency=float(rng.uniform(0.1, 0.5)),
                       cycles=4, E0=float(rng.uniform(0.02, 0.1)))
        psi0 = h.ground_state()
        psi0[1] = 0.1 * psi0[0]
        ov = D.Overlap(h.S, h.N, h.kd)
        eng = D.PampEngine(h.L, h.N, 12, ov)
        eng.add_term("Diag", D.to_device(h.H0))
        eng.add_term("Simp", D.to_device(h.Hint), pairs=h.pairs)
        engines.append(eng)
        psa.append(D.to_device(psi0))
        psb.append(D.to_device(psi0))
        dts.append(dt)
        hs.append(h)
    print("batch graph: 30 instances, %d workgroups" % sum(e.L for e in engines), flush=True)
    graph = D.BatchGraph(engines, psb, nsteps, dts)
    for r in range(3):
        ts = O.step_times(nsteps, dt, t0=3.0 + r * nsteps * dt)
        scs = [np.array([[1.0, h.scaling(tt)] for tt in ts]) for h in hs]
        D.pamp_batch_propagate(engines, psa, nsteps, dts, scs)
        t.cuda.synchronize()
        graph.launch(scs, psb)
        t.cuda.synchronize()
        norms = [float(t.linalg.vector_norm(p).item()) for p in psb]   # eager work between replays
        assert all(np.isfinite(norms))
        for a, b in zip(psa, psb):
            assert t.equal(a, b), "replay %d differs from the direct launches" % r
        print("  replay %d equals the direct launches" % r, flush=True)
    engines[0].set_mixed([1])
    with pytest.raises(nat.NativeError):
        graph.launch(scs, psb)


def _adjoint_problem(K, seed=3):
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H
    over = {("AngularRepresentation", "maxl"): 5, ("RadialRepresentation", "xsize"): 37,
            ("RadialRepresentation", "xmax"): 40.0, ("Propagation", "krylov_basis_size"): 12,
            ("Propagation", "timestep"): 0.05}
    prob = pyprop_amd.Problem(H.load("hydrogen_small", over))
    prob.SetupStep()
    prob.SetWavefunction(H.ground_state(prob))
    target, _ = H.bound_state(prob, 1, 0)   # 2p
    rng = np.random.default_rng(seed)
    t = 0.05 * np.arange(K)
    s = 0.3 * np.sin(np.pi * t / t[-1]) ** 2 * np.cos(0.3 * t) + 0.005 * rng.standard_normal(K)
    return prob, target, s


def test_adjoint_gradient_matches_finite_differences():
    """dP(2p)/ds_k by the adjoint sweep (checkpoints every 7 steps, so the
    last segment is partial) against central finite differences h = 1e-6 of
    the GPU forward map on 6 samples, and against central differences of the
    oracle's FP64 propagation of the same operator on 2 samples."""
    from pyprop_amd.adjoint import PopulationAdjoint, population_function
    torch = pytest.importorskip("torch")
    K = 40
    prob, target, s = _adjoint_problem(K)
    adj = PopulationAdjoint(prob, target, checkpoint=7)
    P, g = adj.gradient(s)
    assert P > 0
    h = 1e-6
    idx = [0, 5, 13, 20, 33, 39]
    for k in idx:
        sp, sm = s.copy(), s.copy()
        sp[k] += h
        sm[k] -= h
        fd = (adj.population(sp) - adj.population(sm)) / (2 * h)
        print("k", k, "adjoint", g[k], "fd", fd)
        assert abs(g[k] - fd) <= 1e-6 * np.abs(g).max() + 1e-10
    # oracle reference: the same operator assembled in the oracle, pAMP FP64
    h_or = O.Hydrogen(rmax=40.0, nbp=37, order=7, lmax=5, frequency=0.057, cycles=6, E0=0.05)
    ham = h_or.ham()
    psi0 = adj.psi0.cpu().numpy()
    S = h_or.overlap_dense()

    def pop(sv):
        psi = ham.propagate_pamp(psi0, K, 0.05, np.stack([np.ones(K), sv], axis=1), 12)
        return abs(np.sum(np.conj(target) * (psi @ S.T))) ** 2

    assert abs(pop(s) - P) < 1e-8 * max(P, 1e-30) + 1e-14   # Pade scaling differs (1-norm here)
    for k in (7, 29):
        sp, sm = s.copy(), s.copy()
        sp[k] += h
        sm[k] -= h
        fd = (pop(sp) - pop(sm)) / (2 * h)
        print("k", k, "adjoint", g[k], "oracle fd", fd)
        assert abs(g[k] - fd) <= 1e-6 * np.abs(g).max() + 1e-10
    # torch.autograd wiring
    x = torch.tensor(s, dtype=torch.float64, requires_grad=True)
    f = population_function(adj)
    y = f(x)
    y.backward()
    assert np.allclose(x.grad.numpy(), g, rtol=1e-12, atol=0)


def test_step_kernels_repeat_bitwise():
    """The same inputs give bitwise the same state on every run, for the
    fiber-split step kernel (the default on this workload: 4 workgroups per
    fiber) and the one-workgroup-per-fiber kernel.  Regression test for the
    barrier ordering: every wave now waits for its exchanged stores before
    the arrival word is published; before, a reader could see the arrival
    first and take a stale chunk-boundary value (differences of 1e-33 within
    100 steps on this workload, amplified to 1e-4 over 20,000)."""
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H
    print("hydrogen_large: 3 x 200 steps, split and unsplit", flush=True)
    prob = pyprop_amd.Problem(H.load("hydrogen_large"))
    prob.SetupStep()
    psi0 = H.ground_state(prob)
    eng = prob.Propagator.Engine
    for q in (0, 1):
        eng.set_fiber_split(q)
        out = []
        for rep in range(3):
            prob.SetWavefunction(psi0)
            prob.PropagatedTime = 0.0
            prob.AdvanceSteps(200)
            out.append(prob.psi.GetData().cpu().numpy())
        for rep in (1, 2):
            assert np.array_equal(out[0], out[rep]), "q=%d run %d differs by %.3e" % (
                q, rep, np.abs(out[0] - out[rep]).max())


def test_step_kernels_launch_split_bitwise():
    """Steps carry nothing from one to the next but psi: 6 steps in one
    launch, in two launches of 3 and in six launches of 1 (the same time
    factors) give bitwise the same state, for the fiber-split step kernel
    (one grid barrier per Arnoldi iteration, U^-1 u handed to the coupled
    fibers through per-workgroup production counts) and the
    one-workgroup-per-fiber kernel."""
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H
    print("hydrogen_large: 6 steps as 1, 2 and 6 launches, split and unsplit", flush=True)
    prob = pyprop_amd.Problem(H.load("hydrogen_large"))
    prob.SetupStep()
    psi0 = H.ground_state(prob)
    prop = prob.Propagator
    eng = prop.Engine
    dt = prob.TimeStep
    times = 50.0 + dt.real * np.arange(6)
    for q in (0, 1):
        eng.set_fiber_split(q)
        out = []
        for parts in ((6,), (3, 3), (1,) * 6):
            prob.SetWavefunction(psi0)
            k = 0
            for n in parts:
                prop.AdvanceSteps(times[k:k + n], dt)
                k += n
            out.append(prob.psi.GetData().cpu().numpy())
        for i in (1, 2):
            assert np.array_equal(out[0], out[i]), "q=%d: launch split %d differs by %.3e" % (
                q, i, np.abs(out[0] - out[i]).max())


@pytest.mark.parametrize("prop", ["PamPropagator", "ExpokitPropagator"])
def test_zero_wavefunction_raises(prop):
    """A zero wavefunction through Problem.AdvanceStep: the reference prints
    "Wavefunction is zero" (expokitpropagator.cpp:48-54) and then divides by
    the zero norm; here the step raises RuntimeError with that message, and
    the state stays zero (exp(-iH dt) 0 = 0)."""
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H
    print("zero wavefunction through", prop, flush=True)
    conf = H.load("hydrogen_small")
    conf.SetValue("AngularRepresentation", "maxl", 3)
    conf.SetValue("RadialRepresentation", "xsize", 30)
    conf.SetValue("RadialRepresentation", "xmax", 30.0)
    conf.SetValue("Propagation", "krylov_basis_size", 10)
    conf = pyprop_amd.Config(conf.cfgObj)
    conf.SetValue("Propagation", "propagator", getattr(pyprop_amd, prop), ignoreSerializationError=True)
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    psi0 = H.ground_state(prob)
    prob.SetWavefunction(psi0)
    prob.AdvanceStep()                        # a normal step first
    prob.psi.GetData().zero_()
    with pytest.raises(RuntimeError, match="Wavefunction is zero"):
        prob.AdvanceStep()
    assert not prob.psi.GetData().abs().max().item() > 0.0


def test_observables_match_oracle():
    """Wavefunction.GetNorm, InnerProduct (<phi|S|psi> by the overlap kernel
    and the device zdotc) and the bound populations (analysis.py:92-140,
    344-364) of a propagated state, against the oracle's S-norm, its
    orc_ham_inner and its populations: same psi (after 10 GPU steps of the
    small hydrogen config with the field on), 1e-13 relative."""
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H
    print("observables: norm / inner product / populations vs oracle", flush=True)
    conf = H.load("hydrogen_small")
    conf.SetValue("AngularRepresentation", "maxl", 4)
    conf.SetValue("RadialRepresentation", "xsize", 40)
    conf.SetValue("RadialRepresentation", "xmax", 40.0)
    conf.SetValue("PulseParameters", "frequency", 0.3)
    conf = pyprop_amd.Config(conf.cfgObj)
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    psi0 = H.ground_state(prob)
    psi0[1] = 0.3 * psi0[0]
    prob.SetWavefunction(psi0)
    prob.PropagatedTime = 30.0
    prob.AdvanceSteps(10)
    psi = prob.psi.GetData().cpu().numpy()
    h = O.Hydrogen(rmax=40.0, nbp=40, order=7, lmax=4, frequency=0.3, cycles=6, E0=0.05)
    ham = h.ham()
    # norm
    nref = h.norm(psi)
    assert abs(prob.psi.GetNorm() - nref) <= 1e-13 * nref
    # <phi|S|psi> against the oracle's inner product (U-factored S)
    phi_h, _ = H.bound_state(prob, 1, 0)          # 2p
    phi = prob.psi.Copy()
    phi.GetData().copy_(dev_tensor(phi_h))
    ip = phi.InnerProduct(prob.psi)
    iref = ham.inner(phi_h, psi)
    assert abs(ip - iref) <= 1e-13 * max(abs(iref), 1e-300) + 1e-15
    assert abs(prob.psi.InnerProduct(prob.psi) - nref ** 2) <= 1e-13
    # bound populations of the product's analysis against the oracle's
    tot, per = H.populations(prob, psi)
    tref, pref = h.populations(psi)
    assert abs(tot - tref) <= 1e-12
    for a, b in zip(per, pref):
        assert len(a) == len(b) and np.abs(np.asarray(a) - np.asarray(b)).max() <= 1e-12


def dev_tensor(a):
    from pyprop_amd import device as D
    return D.to_device(a)


def test_checkpoint_resume_is_bitwise(tmp_path):
    """Checkpoint and resume through the reference's interface: 12 steps,
    Problem.SaveWavefunctionHDF, 8 more steps; a second Problem built from
    the stored config (LoadConfigFromFile) with InitialCondition.type =
    File loads the checkpoint in SetupStep, RestartPropagation sets the
    time, and its 8 steps equal the uninterrupted run bitwise."""
    import pyprop_amd
    from pyprop_amd.examples import hydrogen as H
    print("checkpoint / resume, small hydrogen", flush=True)
    conf = H.load("hydrogen_small")
    conf.SetValue("AngularRepresentation", "maxl", 4)
    conf.SetValue("RadialRepresentation", "xsize", 40)
    conf.SetValue("RadialRepresentation", "xmax", 40.0)
    conf.SetValue("PulseParameters", "frequency", 0.3)
    conf = pyprop_amd.Config(conf.cfgObj)
    prob = pyprop_amd.Problem(conf)
    prob.SetupStep()
    prob.SetWavefunction(H.ground_state(prob))
    prob.PropagatedTime = 20.0
    prob.AdvanceSteps(12)
    t_ckpt = prob.PropagatedTime
    f = str(tmp_path / "ckpt.h5")
    prob.SaveWavefunctionHDF(f, "/wavefunction")
    prob.AdvanceSteps(8)
    ref = prob.psi.GetData().cpu().numpy()
    conf2 = pyprop_amd.LoadConfigFromFile(f, "/wavefunction")
    conf2.SetValue("InitialCondition", "type", pyprop_amd.InitialConditionType.File)
    conf2.SetValue("InitialCondition", "format", pyprop_amd.WavefunctionFileFormat.HDF)
    conf2.SetValue("InitialCondition", "filename", f)
    conf2.SetValue("InitialCondition", "dataset", "/wavefunction")
    prob2 = pyprop_amd.Problem(conf2)
    prob2.SetupStep()
    prob2.RestartPropagation(prob2.TimeStep, t_ckpt, prob2.Duration)
    prob2.AdvanceSteps(8)
    assert np.array_equal(prob2.psi.GetData().cpu().numpy(), ref)
    psi3 = pyprop_amd.CreateWavefunctionFromFile(f, "/wavefunction")
    assert tuple(psi3.GetData().shape) == ref.shape


def test_potential_generation_on_device_matches_host():
    """generate_on_device = True (the B-spline ranks projected by
    pp_projector_apply) against the host projection for every potential of
    the hydrogen_small config (kinetic d = 2, SAE, the velocity-gauge d = 1
    and 1/r parts): relative 1e-14, and the consolidated list the same."""
    import pyprop_amd as P
    from pyprop_amd.core import Wavefunction
    from pyprop_amd.examples import hydrogen as H
    from pyprop_amd.problem import CreateRepresentation
    conf = H.load("hydrogen_small")
    bp = P.BasisPropagator(Wavefunction(CreateRepresentation(conf)))
    conf.Apply(bp)
    confd = H.load("hydrogen_small")
    for n in confd.Propagation.grid_potential_list:
        confd.SetValue(n, "generate_on_device", True)
    bpd = P.BasisPropagator(Wavefunction(CreateRepresentation(confd)))
    for n in confd.Propagation.grid_potential_list:
        a = bp.GeneratePotential(conf.GetSection(n)).PotentialData
        b = bpd.GeneratePotential(confd.GetSection(n)).PotentialData
        err = np.abs(a - b).max() / max(np.abs(a).max(), 1e-300)
        print("potential", n, "rel err %.2e" % err, flush=True)
        assert err <= 1e-14
    confd.Apply(bpd)
    for a, b in zip(bp.PotentialList, bpd.PotentialList):
        assert a.Name == b.Name
        assert np.abs(a.PotentialData - b.PotentialData).max() <= 1e-14 * np.abs(a.PotentialData).max()
