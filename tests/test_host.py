"""Host-side logic on the CPU: config semantics, operator assembly against the
oracle, and the C-ABI library's exports (no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyprop_amd as P
from oracle import oracle as O
from pyprop_amd.core import Wavefunction
from pyprop_amd.examples import hydrogen as H
from pyprop_amd.problem import CreateRepresentation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_config_semantics(tmp_path):
    (tmp_path / "inc.ini").write_text("[A]\nx = 2\ny = x * 3\n[B]\nbase = 'A'\nz = y + 1\n")
    (tmp_path / "main.ini").write_text(
        "[Import]\nfiles = ['inc.ini']\n\n[C]\nfreq = 0.5\ncycles = 4\n"
        "dur = 2 * pi / %(freq)s * %(cycles)s\nCamelCase = 'abc'\n")
    conf = P.Load(str(tmp_path / "main.ini"))
    assert conf.A.y == 6
    assert conf.B.z == 7 and conf.B.x == 2          # base= copies A first
    assert abs(conf.C.dur - 2 * np.pi / 0.5 * 4) < 1e-15
    assert conf.C.name == "C" and conf.C.camelcase == "abc"   # option names are lower-cased
    conf.SetValue("C", "freq", 0.25)
    assert conf.C.freq == 0.25 and conf.cfgObj.get("C", "freq") == "0.25"


def test_workload_configs_parse():
    c = H.load("hydrogen_large")
    assert c.RadialRepresentation.xsize == 797 and c.AngularRepresentation.maxl == 63
    assert c.Propagation.krylov_basis_size == 24 and c.Propagation.timestep == 0.005
    assert c.LaserPotentialVelocity1.time_function is H.LaserFunctionVelocity
    assert abs(c.LaserPotentialVelocity2.amplitude - 0.1 / 0.5) < 1e-15


def _assemble(name):
    conf = H.load(name)
    repr = CreateRepresentation(conf)
    psi = Wavefunction(repr)
    bp = P.BasisPropagator(psi)
    conf.Apply(bp)
    return conf, repr, bp


def test_assembly_matches_oracle_small():
    conf, repr, bp = _assemble("hydrogen_small")
    assert repr.GetFullShape() == [11, 103]
    names = [p.Name for p in bp.PotentialList]
    assert names == ["RadialKineticEnergy+AngularKineticEnergy+SAEPotential",
                     "LaserPotentialVelocity1+LaserPotentialVelocity2+LaserPotentialVelocity3"]
    h = O.Hydrogen(rmax=100.0, nbp=100, order=7, lmax=10, frequency=0.057, cycles=6, E0=0.05)
    H0, Hi = bp.PotentialList[0].PotentialData, bp.PotentialList[1].PotentialData
    assert np.abs(H0 - h.H0).max() <= 1e-13 * np.abs(h.H0).max()
    assert np.abs(Hi - h.Hint).max() <= 1e-13 * np.abs(h.Hint).max()
    assert np.array_equal(bp.PotentialList[1].GeometryList[0].GetBasisPairs(), h.pairs)
    S = repr.GetRepresentation(1).GetBSplineObject().overlap_upper()
    assert np.abs(S - h.S).max() < 1e-14
    for t in [0.0, 3.7, 50.0, 120.0]:
        assert bp.PotentialList[1].GetTimeScaling(t) == pytest.approx(h.scaling(t), abs=1e-16)


def test_initial_state_matches_oracle():
    h = O.Hydrogen(rmax=100.0, nbp=100, order=7, lmax=10)
    conf, repr, bp = _assemble("hydrogen_small")

    class _Prop:
        BasePropagator = bp

    class _Problem:
        Propagator = _Prop
        psi = Wavefunction(repr)

    psi = H.ground_state(_Problem)
    ref = h.ground_state()
    assert np.abs(psi - ref).max() < 1e-10


def test_capi_exports_match_header():
    from pyprop_amd import _native as nat
    decl = open(os.path.join(ROOT, "include", "pyprop_amd.h")).read()
    names = set(re.findall(r"\b(?:int|long)\s+(pp_\w+)\s*\(", decl))
    assert names == set(nat.EXPORTS)
    lib = ctypes.CDLL(nat.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_tensor_multiply_name_dispatch():
    from pyprop_amd.core import core
    f = core.TensorPotentialMultiply_Simp_BandNH
    assert f.__name__ == "TensorPotentialMultiply_Simp_BandNH"
    with pytest.raises(AttributeError):
        core.TensorPotentialMultiply_Bogus_BandNH


def test_lowrank_factor_finds_exact_rank_of_batch_operators():
    """The library's host factorisation (pp_lowrank_factor) of assembled
    hydrogen operators: rank 2 for the static and for the dipole tensor of
    batch instances with 500-600 B-splines (regression: one-pass residual
    updates used to leave ~1e-13 of the pivot rows and inflate the rank to
    the cap), reconstruction within 1e-13 of max|V|."""
    import ctypes
    from pyprop_amd import _native as nat
    from pyprop_amd import batch as Bt
    lib = nat.lib()
    for d in Bt.instance_draws(3):
        conf = H.batch_instance_config(d)
        repr = CreateRepresentation(conf)
        bp = P.BasisPropagator(Wavefunction(repr))
        conf.Apply(bp)
        for pot in bp.PotentialList:
            V = np.ascontiguousarray(pot.PotentialData, dtype=np.complex128)
            P0, M = V.shape
            mr = min(2 * P0, 16)
            coef, bands = np.zeros(P0 * mr * 2), np.zeros(mr * M)
            rk, err = ctypes.c_int(), ctypes.c_double()
            rc = lib.pp_lowrank_factor(P0, M, nat.dptr(V.view(np.float64)), mr, ctypes.byref(rk), nat.dptr(coef),
                                       nat.dptr(bands), ctypes.byref(err))
            assert rc == 0 and rk.value == 2 and err.value < 1e-13, (d, pot.StorageIds(), rk.value, err.value)


def test_expokit_chunk_fits_lds():
    """The Expokit engine's solver chunk: the default plan where zgexpv's
    Hessenberg workspace still fits 160 KB of LDS (small shapes, and the
    hydrogen_large shape L = 64, N = 800, kd = 6 at m = 24), a longer chunk
    where it does not (the same shape at m = 32: the default B = 32 plan and
    the Pade's 5 m^2 buffers leave less than the 19,040 B of zgexpv
    workspace)."""
    from pyprop_amd import _native as nat
    from pyprop_amd.propagator import expokit_chunk
    lib = nat.lib()
    assert lib.pp_pamp_lds_estimate(64, 800, 24, 3, 6, 0) == lib.pp_pamp_lds_estimate(64, 800, 24, 3, 6, 32)
    assert expokit_chunk(11, 103, 20, 6, 0) is None
    assert expokit_chunk(64, 800, 24, 6, 0) is None
    assert lib.pp_pamp_lds_estimate(64, 800, 24, 3, 6, 32) + 16 * (26 * 26 + 26) <= 160 * 1024
    B = expokit_chunk(64, 800, 32, 6, 0)
    assert B == 64
    assert lib.pp_pamp_lds_estimate(64, 800, 32, 3, 6, B) + 16 * (34 * 34 + 34) <= 160 * 1024
    assert lib.pp_pamp_lds_estimate(64, 800, 32, 3, 6, 32) + 16 * (34 * 34 + 34) > 160 * 1024


def test_krylov_status_mapping_raises_like_the_reference():
    """Per-step diagnostics -> exceptions (no GPU): the reference throws
    runtime_error for a singular Pade denominator (pamp.h:544-550) and for
    zgexpv's iflag != 0 (expokitpropagator.cpp:91-100), and divides by the
    norm of a zero wavefunction (expokitpropagator.cpp:48-54); here each is a
    RuntimeError naming the step."""
    from pyprop_amd import device as D
    ok = np.array([[1e-12, 20, 1.0, 0], [3e-11, 20, 0.99, 0]])
    D.check_pamp_steps(ok)
    cases = [(1, 2, 0.0, "Wavefunction is zero"), (1, 2, np.nan, "Wavefunction is zero"),
             (0, 0, -1.0, "singular Pade"), (1, 0, np.inf, "non-finite")]
    for row, col, val, msg in cases:
        d = ok.copy()
        d[row, col] = val
        with pytest.raises(RuntimeError, match=msg):
            D.check_pamp_steps(d)
    st = np.zeros((3, 6))
    D.check_expokit_steps(st)
    for flag, msg in [(1, "Expokit returned nonzero value 1"), (-9, "nonzero value -9"), (-8, "nonzero value -8"),
                      (-10, "Wavefunction is zero")]:
        s = st.copy()
        s[2, 5] = flag
        with pytest.raises(RuntimeError, match=msg):
            D.check_expokit_steps(s)
    from pyprop_amd import _native as nat
    with pytest.raises(RuntimeError, match="grid barrier"):
        nat.check(nat.PP_ETIMEOUT, "pp_pamp_diagnostics")


def test_config_import_order_and_python2_rules(tmp_path):
    """Import priority (the importing file over its imports, an earlier entry
    of files over a later one, breadth first), Python 2 ConfigParser rules
    (repeated keys: the last counts; ';' inline comments only; '#' stays in
    the value and is an expression comment), Python 2 integer division, base
    chains, and a cycle of bases refused."""
    (tmp_path / "sub").mkdir()
    (tmp_path / "sub" / "deep.ini").write_text("[S]\na = 'deep'\nd = 'deep'\n")
    (tmp_path / "f1.ini").write_text("[Import]\nfiles = ['sub/deep.ini']\n[S]\na = 'f1'\nb = 'f1'\n")
    (tmp_path / "f2.ini").write_text("[S]\na = 'f2'\nb = 'f2'\nc = 'f2'\n")
    (tmp_path / "main.ini").write_text(
        "[Import]\nfiles = ['f1.ini', 'f2.ini']\n"
        "[S]\na = 'main'\n"
        "[T]\nx = 5\nx = 7\nh = 7 / 2\nf = 7.0 / 2\nneg = -7 / 2\ns = 'a#b' # trailing comment\n"
        "t = 3 ; an inline comment\n"
        "[U]\nbase = 'T'\nh = h * 10\n"
        "[V]\nbase = 'W'\n[W]\nbase = 'V'\n")
    files = P.config.config_files(str(tmp_path / "main.ini"))
    assert [os.path.basename(f) for f in files] == ["deep.ini", "f2.ini", "f1.ini", "main.ini"]
    parser = P.config._parser()
    for f in files:
        P.config._read_file(parser, f)
    sec = P.config.Section("S", parser)
    assert (sec.a, sec.b, sec.c, sec.d) == ("main", "f1", "f2", "deep")
    T = P.config.Section("T", parser)
    assert T.x == 7 and T.h == 3 and T.f == 3.5 and T.neg == -4 and T.s == "a#b" and T.t == 3
    U = P.config.Section("U", parser)
    assert U.h == 30 and U.x == 7
    with pytest.raises(ValueError, match="cycle"):
        P.config.Section("V", parser)


def test_exponential_linear_grid_of_the_sae_paper_config():
    """InitBSpline on an exponential-linear section with the SAE paper's
    grid (xmin 0, xpartition 10, xmax 160, 30 inner points, gamma 1): 310
    breakpoints, as pylab.frange's closed interval gives."""
    from pyprop_amd import bspline as B
    conf = P.LoadString("[R]\nbpstype = 'exponentiallinear'\nxmin = 0.0\nxmax = 160.0\nxsize = 30\n"
                        "gamma = 1.0\nxpartition = 10\norder = 7\ncontinuity = 'zero'\n").R
    basis = B.InitBSpline(conf)
    assert len(basis.Breakpoints) == 310
    assert basis.Breakpoints[-1] > 159.7


def test_wavefunction_archive_round_trip(tmp_path):
    """SaveWavefunctionHDF / LoadWavefunctionHDF / LoadConfigFromFile on
    host tensors: bitwise round trip, several datasets per file addressed
    by path, replacing a dataset, the stored config object evaluating to
    the same values, shape checks, and a refusal of real HDF5 files."""
    import torch
    from pyprop_amd import serialization as SZ

    class _Repr:
        def __init__(self, shape):
            self.shape = shape

        def GetFullShape(self):
            return self.shape

    class _Psi:
        def __init__(self, data):
            self.data = data

        def GetData(self):
            return self.data

        def GetRepresentation(self):
            return _Repr(tuple(self.data.shape))

    rng = np.random.default_rng(4)
    a = rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7))
    b = rng.standard_normal((3, 4, 4)) + 1j * rng.standard_normal((3, 4, 4))
    conf = P.LoadString("[Propagation]\ntimestep = 0.01\nduration = 2 * pi / 4\nname2 = 'x#y'\n")
    f = str(tmp_path / "wf.h5")
    SZ.SaveWavefunctionHDF(f, "/wavefunction", _Psi(torch.from_numpy(a.copy())), conf=conf)
    SZ.SaveWavefunctionHDF(f, "/run1/psi", _Psi(torch.from_numpy(b.copy())))
    assert SZ.ListDatasets(f) == ["/run1/psi", "/wavefunction"]
    out = _Psi(torch.zeros((5, 7), dtype=torch.complex128))
    SZ.LoadWavefunctionHDF(f, "/wavefunction", out)
    assert np.array_equal(out.data.numpy(), a)
    out3 = _Psi(torch.zeros((3, 4, 4), dtype=torch.complex128))
    SZ.LoadWavefunctionHDF(f, "run1/psi", out3)
    assert np.array_equal(out3.data.numpy(), b)
    # replace a dataset; the other one stays
    SZ.SaveWavefunctionHDF(f, "/wavefunction", _Psi(torch.from_numpy(2 * a)), conf=conf)
    SZ.LoadWavefunctionHDF(f, "/wavefunction", out)
    assert np.array_equal(out.data.numpy(), 2 * a)
    SZ.LoadWavefunctionHDF(f, "/run1/psi", out3)
    assert np.array_equal(out3.data.numpy(), b)
    c = SZ.LoadConfigFromFile(f, "/wavefunction")
    assert c.Propagation.timestep == 0.01 and c.Propagation.duration == 2 * np.pi / 4
    assert c.Propagation.name2 == "x#y"
    with pytest.raises(ValueError, match="Invalid shape"):
        SZ.LoadWavefunctionHDF(f, "/wavefunction", out3)
    with pytest.raises(KeyError):
        SZ.LoadWavefunctionHDF(f, "/missing", out)
    h5 = tmp_path / "ref.h5"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(IOError, match="HDF5"):
        SZ.LoadWavefunctionHDF(str(h5), "/wavefunction", out)


def test_radial_preconditioner_decomposition_matches_oracle():
    """RadialTwoElectronPreconditioner.Decompose on the small helium config:
    the eigenvectors are S-orthonormal, one decomposition per distinct l, and
    S V diag(lam) V^T S rebuilds each channel's H1 (x) S + S (x) H2, whose
    action equals the oracle's four channel-diagonal one-electron potentials
    (radial kinetic on both ranks, angular kinetic, nuclear attraction,
    projected on the product grid) to rounding; the channel-diagonal e-e
    multipoles are the omitted, non-separable terms."""
    from pyprop_amd.examples import helium as HE
    conf = HE.load("helium_small", field_free=True)
    repr = CreateRepresentation(conf)
    psi = Wavefunction(repr)
    bp = P.BasisPropagator(psi)
    conf.Apply(bp)
    prec = conf.RadialPreconditioner.type(psi)
    conf.RadialPreconditioner.Apply(prec)
    assert isinstance(prec, P.RadialTwoElectronPreconditionerIfpack) and prec.Cutoff == 0

    class _Prop:
        BasePropagator = bp
    idx1, idx2, V, lam = prec.Decompose(_Prop())
    h = O.Helium(lmax=1, Ls=(0, 1), nbp=6, rmax=8.0, laser=False)
    C, N = h.C, h.N
    assert (len(idx1), V.shape[1]) == (C, N)
    assert V.shape[0] == 2                                   # l = 0, 1
    ch = repr.GetRepresentation(0).Channels
    assert all(idx1[c] == idx1[d] for c in range(C) for d in range(C) if ch[c].l1 == ch[d].l1)
    assert all(idx1[c] == idx2[d] for c in range(C) for d in range(C) if ch[c].l1 == ch[d].l2)
    from pyprop_amd.helium import band_dense
    S = band_dense(repr.RadialMatrices()["S"])
    for m in range(V.shape[0]):
        assert np.abs(V[m].T @ S @ V[m] - np.eye(N)).max() < 1e-11
    Hm = [S @ V[m] @ np.diag(lam[m]) @ V[m].T @ S for m in range(V.shape[0])]
    statics = []
    for kind, pk, pot in h.terms[:4]:
        assert pk == "diag"
        statics.append((O.ST_SIMP, O.ST_SIMP, O.ST_SIMP, pot,
                        [h.pairs[pk].astype(np.int32), h.packed.astype(np.int32), h.packed.astype(np.int32)]))
    ham = O.Ham(h.shape, statics, {1: (h.U, h.kd), 2: (h.U, h.kd)})
    rng = np.random.default_rng(5)
    x = rng.standard_normal(h.shape) + 1j * rng.standard_normal(h.shape)
    ref = ham.potentials(x)
    got = np.stack([Hm[idx1[c]] @ x[c] @ S + S @ x[c] @ Hm[idx2[c]] for c in range(C)])
    assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()
    assert prec.OmittedTerms and all("ElectronicCouplingPotentialMonopoleTerm" in t for t in prec.OmittedTerms)


def test_tensor_potential_files(tmp_path):
    """SaveTensorPotential / LoadTensorPotential and a potential section's
    `filename` / `dataset` keys (TensorPotentialHDF.py:10-171,
    TensorPotential.py:56-71): every hydrogen potential saved and loaded back
    through the config equals the generated one bitwise, also after
    consolidation; a potential stored for another radial grid is refused with
    the reference's "regenerate" error; '/' is refused as a group; the
    factored two-electron terms of the helium config round-trip bitwise
    (their reference layout included)."""
    from pyprop_amd import serialization as SZ
    f = str(tmp_path / "pots.h5")
    conf, repr, bp = _assemble("hydrogen_small")
    names = list(conf.Propagation.grid_potential_list)
    gen = {}
    for n in names:
        pot = bp.GeneratePotential(conf.GetSection(n))
        gen[n] = pot.PotentialData.copy()
        SZ.SaveTensorPotential(f, "/potentials/" + n, pot, None, conf)
    with pytest.raises(Exception, match="bad idea"):
        SZ.SaveTensorPotential(f, "/", pot, None)
    conf2 = H.load("hydrogen_small")
    for n in names:
        conf2.SetValue(n, "filename", f)
        conf2.SetValue(n, "dataset", "/potentials/" + n)
    repr2 = CreateRepresentation(conf2)
    bp2 = P.BasisPropagator(Wavefunction(repr2))
    for n in names:
        pot = bp2.GeneratePotential(conf2.GetSection(n))
        assert np.array_equal(pot.PotentialData, gen[n])
    conf2.Apply(bp2)
    assert [p.Name for p in bp2.PotentialList] == [p.Name for p in bp.PotentialList]
    for a, b in zip(bp.PotentialList, bp2.PotentialList):
        assert np.array_equal(a.PotentialData, b.PotentialData)
    # another grid: the stored radial pairs no longer match
    conf3 = H.load("hydrogen_small")
    conf3.SetValue("RadialRepresentation", "xsize", 90)
    conf3.SetValue("RadialKineticEnergy", "filename", f)
    conf3.SetValue("RadialKineticEnergy", "dataset", "/potentials/RadialKineticEnergy")
    bp3 = P.BasisPropagator(Wavefunction(CreateRepresentation(conf3)))
    with pytest.raises(Exception, match="regenerate potential"):
        bp3.GeneratePotential(conf3.GetSection("RadialKineticEnergy"))
    # two-electron potentials: factored terms
    from pyprop_amd.examples import helium as HE
    hconf = HE.load("helium_small")
    hrepr = CreateRepresentation(hconf)
    hbp = P.BasisPropagator(Wavefunction(hrepr))
    sec = hconf.GetSection("ElectronicCouplingPotential")
    ee = hbp.GeneratePotential(sec)
    SZ.SaveTensorPotential(f, "/helium/ee", ee, None, hconf)
    hconf.SetValue("ElectronicCouplingPotential", "filename", f)
    hconf.SetValue("ElectronicCouplingPotential", "dataset", "/helium/ee")
    ee2 = hbp.GeneratePotential(hconf.GetSection("ElectronicCouplingPotential"))
    assert len(ee2.Terms) == len(ee.Terms) > 0
    for a, b in zip(ee.Terms, ee2.Terms):
        assert type(a) is type(b) and np.array_equal(a.pairs, b.pairs) and np.array_equal(a.coef, b.coef)
    assert np.array_equal(ee.ReferenceLayout(), ee2.ReferenceLayout())
    with pytest.raises(ValueError, match="factored"):
        SZ.LoadTensorPotential(f, "/potentials/RadialKineticEnergy", ee2)
