"""Device-side objects over the C-ABI: overlap factor, tensor multiply and the
persistent pAMP engine.  Device buffers are torch tensors on ``cuda:0``; torch
is used only for allocation and the current stream.
"""
import ctypes

import numpy as np

from . import _native as nat


def _release(obj):
    """Finalizer body of the native wrappers: destroy the handle unless the
    interpreter is shutting down (then the exit hook of _native did it)."""
    token = getattr(obj, "_token", None)
    if token is not None and nat is not None:
        nat.release(token)
        obj._token = None


def torch():
    import torch as _t
    return _t


def device():
    t = torch()
    if not t.cuda.is_available():
        raise nat.NativeError("no GPU visible: the pyprop_amd compute path runs on MI355X only")
    return t.device("cuda:0")


def stream_ptr():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def to_device(a):
    """numpy -> contiguous complex128/float64/int32 torch tensor on the GPU."""
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a)).to(device())


def ptr(tensor):
    if not tensor.is_contiguous():
        raise ValueError("device buffers must be C-contiguous (MapToRank3 requires it)")
    return ctypes.c_void_p(tensor.data_ptr())


# Solver chunk length per (N, kd) chosen by the autotuner (batch.autotune_chunks);
# shapes without an entry use the library default.
CHUNK_TABLE = {}


class Overlap:
    """Banded overlap S = U^H U of one B-spline rank, factorised on creation
    (OverlapMatrix::Setup, overlapmatrix.cpp:9-104).  chunk: solver chunk
    length (None: the tuned value for (N, kd) if any, else the default)."""

    def __init__(self, S_upper, N, kd, chunk=None):
        L = nat.lib()
        S = np.ascontiguousarray(S_upper, dtype=np.complex128)
        if S.size != N * (kd + 1):
            raise ValueError("overlap band has %d entries, expected %d" % (S.size, N * (kd + 1)))
        if chunk is None:
            chunk = CHUNK_TABLE.get((N, kd), 0)
        h = ctypes.c_void_p()
        info = ctypes.c_int(0)
        rc = L.pp_overlap_create_chunked(N, kd, nat.dptr(S.view(np.float64)), int(chunk), ctypes.byref(h),
                                         ctypes.byref(info))
        if rc > 0:
            raise RuntimeError("overlap matrix is not positive definite (zpbtrf info=%d)" % info.value)
        nat.check(rc, "pp_overlap_create")
        self.h = h
        self._token = nat.track("pp_overlap_destroy", h)
        self.N, self.kd, self.chunk = N, kd, chunk

    def factor(self):
        U = np.zeros(self.N * (self.kd + 1))
        nat.check(nat.lib().pp_overlap_factor(self.h, nat.dptr(U)), "pp_overlap_factor")
        return U

    def apply(self, op, rank, x):
        shape = np.array(x.shape, dtype=np.int32)
        nat.check(nat.lib().pp_overlap_apply(self.h, op, rank, x.dim(), nat.iptr(shape), ptr(x), stream_ptr()),
                  "pp_overlap_apply")
        return x

    def __del__(self, _release=_release):   # bound now: module globals are gone at teardown
        _release(self)


def tp_multiply(storage_ids, potential, scaling, src, dst, pairs=None):
    """core.TensorPotentialMultiply_<ids>(potential, scaling, src, dst, *pairs): dst += s * V src."""
    R = src.dim()
    if len(storage_ids) != R or potential.dim() != R or dst.shape != src.shape:
        raise ValueError("rank mismatch in tensor multiply")
    st = np.array([nat.STORAGE_IDS[s] if isinstance(s, str) else int(s) for s in storage_ids], dtype=np.int32)
    ps = np.array(potential.shape, dtype=np.int32)
    sh = np.array(src.shape, dtype=np.int32)
    keep = []
    arr = (ctypes.POINTER(ctypes.c_int) * R)()
    for r in range(R):
        p = None if pairs is None else pairs[r]
        if p is not None:
            p = np.ascontiguousarray(p, dtype=np.int32)
            keep.append(p)
            arr[r] = nat.iptr(p)
        else:
            arr[r] = ctypes.POINTER(ctypes.c_int)()
    rc = nat.lib().pp_tp_multiply(R, nat.iptr(st), nat.iptr(ps), ptr(potential), float(scaling), nat.iptr(sh),
                                  ptr(src), ptr(dst), arr, stream_ptr())
    nat.check(rc, "pp_tp_multiply")
    return dst


def check_pamp_steps(diag, what="pAMP"):
    """Raise the way the reference's Krylov step does when a step of the
    last launch failed.  diag: per step (|last Krylov coefficient| or -1
    for a singular Pade denominator, basis size used, |U psi|, ...).

    * |U psi| = 0 (or not finite): "Wavefunction is zero".  The reference
      prints that (expokitpropagator.cpp:48-54) and then divides by the zero
      norm (pamp.h:619-620, zgexpv.cpp:304,322); here the step raises.
    * Pade denominator singular: "Error in pAMP" (pamp.h:544-550).
    * A non-finite Krylov coefficient: the step produced NaN/Inf.
    """
    d = np.asarray(diag)
    for n in range(d.shape[0]):
        beta = d[n, 2]
        if not np.isfinite(beta) or beta == 0.0:
            raise RuntimeError("Wavefunction is zero: initial norm = %r (%s step %d of the launch)" % (beta, what, n))
        if d[n, 0] == -1.0:
            raise RuntimeError("Error in %s: singular Pade denominator (step %d of the launch)" % (what, n))
        if not np.isfinite(d[n, 0]):
            raise RuntimeError("Error in %s: non-finite Krylov coefficient (step %d of the launch)" % (what, n))


def check_expokit_steps(stats):
    """zgexpv's iflag of every step of the last launch (expokitpropagator.cpp:
    96-100 throws on iflag != 0); -10 marks a zero wavefunction."""
    st = np.asarray(stats)
    bad = np.nonzero(st[:, 5] != 0)[0]
    if len(bad):
        flag = int(st[bad[0], 5])
        if flag == -10:
            raise RuntimeError("Wavefunction is zero (Expokit step %d of the launch)" % bad[0])
        raise RuntimeError("Expokit returned nonzero value %d (step %d of the launch)" % (flag, bad[0]))


class PampEngine:
    """Persistent-kernel pAMP propagator for rank-2 (angular x B-spline) states.

    Terms: (angular storage id, angular pairs or None, device potential of shape
    (P0, N*bw)) with radial BandNH storage.  ``propagate`` runs nsteps whole
    PamPropagator.AdvanceStep's (PamPropagator.py:50-59) in one launch."""

    def __init__(self, L, N, m, overlap, reorth=False):
        h = ctypes.c_void_p()
        nat.check(nat.lib().pp_pamp_create(L, N, m, int(bool(reorth)), overlap.h, ctypes.byref(h)), "pp_pamp_create")
        self.h = h
        self._token = nat.track("pp_pamp_destroy", h)
        self.L, self.N, self.m = L, N, m
        self.overlap = overlap
        self.terms = []

    def add_term(self, angular_storage, potential, pairs=None, radial_storage="BandNH"):
        P0, width = potential.shape
        bw = width // self.N
        if bw * self.N != width:
            raise ValueError("radial band width does not divide the potential extent")
        aid = nat.STORAGE_IDS[angular_storage] if isinstance(angular_storage, str) else int(angular_storage)
        rid = nat.STORAGE_IDS[radial_storage] if isinstance(radial_storage, str) else int(radial_storage)
        pp = ctypes.POINTER(ctypes.c_int)()
        keep = None
        if pairs is not None:
            keep = np.ascontiguousarray(pairs, dtype=np.int32)
            pp = nat.iptr(keep)
        nat.check(nat.lib().pp_pamp_add_term(self.h, aid, P0, pp, rid, bw, ptr(potential)), "pp_pamp_add_term")
        self.terms.append((potential, keep))

    def propagate(self, psi, nsteps, dt, scalings, check=True):
        """nsteps pAMP steps in one launch.  check: read the launch's status
        (barrier time-out, zero wavefunction, singular Pade denominator)
        and raise as the reference does; this synchronises the stream."""
        dt = complex(dt)
        sc = np.ascontiguousarray(np.asarray(scalings, dtype=np.float64).reshape(nsteps, len(self.terms)))
        nat.check(nat.lib().pp_pamp_propagate(self.h, ptr(psi), int(nsteps), dt.real, dt.imag, nat.dptr(sc),
                                              stream_ptr()), "pp_pamp_propagate")
        self._last = nsteps
        if check and nsteps > 0:
            check_pamp_steps(self.diagnostics(nsteps))
        return psi

    def propagate_expokit(self, psi, nsteps, dt, scalings, tol=0.0, anorm=1e6, check=True):
        """nsteps ExpokitPropagator.AdvanceStep's in one launch (zgexpv);
        check: raise on a barrier time-out or iflag != 0 (synchronises)."""
        dt = complex(dt)
        sc = np.ascontiguousarray(np.asarray(scalings, dtype=np.float64).reshape(nsteps, len(self.terms)))
        nat.check(nat.lib().pp_expokit_propagate(self.h, ptr(psi), int(nsteps), dt.real, dt.imag, nat.dptr(sc),
                                                 float(tol), float(anorm), stream_ptr()), "pp_expokit_propagate")
        self._last = nsteps
        if check and nsteps > 0:
            self.expokit_stats(nsteps)
        return psi

    def expokit_stats(self, nsteps=None, check=True):
        """Per step: matvecs, exponentials, substeps, rejections, basis size, iflag."""
        n = self._last if nsteps is None else nsteps
        d = np.zeros((n, 6))
        nat.check(nat.lib().pp_expokit_diagnostics(self.h, n, nat.dptr(d), stream_ptr()), "pp_expokit_diagnostics")
        if check:
            check_expokit_steps(d)
        return d

    def diagnostics(self, nsteps=None):
        n = self._last if nsteps is None else nsteps
        d = np.zeros((n, 4))
        nat.check(nat.lib().pp_pamp_diagnostics(self.h, n, nat.dptr(d), stream_ptr()), "pp_pamp_diagnostics")
        return d

    def set_mixed(self, terms=()):
        """Mixed-precision form: the listed term indices (order of add_term)
        are applied in complex FP32, everything else stays FP64."""
        mask = 0
        for t in terms:
            if not 0 <= int(t) < len(self.terms):
                raise ValueError("no term %r" % (t,))
            mask |= 1 << int(t)
        nat.check(nat.lib().pp_pamp_set_mixed(self.h, mask), "pp_pamp_set_mixed")

    def set_pade_scaling(self, mode):
        """0: largest element (pAMP); 1: 1-norm, accurate and smooth."""
        nat.check(nat.lib().pp_pamp_set_pade_scaling(self.h, int(mode)), "pp_pamp_set_pade_scaling")

    def set_exponentiation(self, method):
        """Exponential of the Krylov matrix: 0 scaling + Pade(6), 1 the
        eigenvector exponential (krylov_exponentiation_method)."""
        nat.check(nat.lib().pp_pamp_set_exponentiation(self.h, int(method)), "pp_pamp_set_exponentiation")

    def set_fiber_split(self, q=0):
        """Workgroups per angular fiber in propagate: 0 automatic (fill the
        CUs), 1 one per fiber, q >= 2 fixed."""
        nat.check(nat.lib().pp_pamp_set_fiber_split(self.h, int(q)), "pp_pamp_set_fiber_split")

    def set_timing(self, enable=True):
        nat.check(nat.lib().pp_pamp_set_timing(self.h, int(bool(enable))), "pp_pamp_set_timing")

    def timings(self):
        """Wall-clock stamps (seconds, 100 MHz counter) of workgroup 0 during
        step 0 of the last timed launch."""
        buf = (ctypes.c_longlong * 512)()
        cnt = ctypes.c_int(0)
        nat.check(nat.lib().pp_pamp_timings(self.h, buf, 512, ctypes.byref(cnt)), "pp_pamp_timings")
        return np.array(buf[:cnt.value], dtype=np.float64) / 1e8

    def __del__(self, _release=_release):   # bound now: module globals are gone at teardown
        _release(self)


class BatchGraph:
    """The packed launches of a batch recorded as one HIP graph
    (pp_pamp_batch_graph_*): launch(scalings) advances every psi by the
    recorded nsteps.  The graph keeps the psi tensors and engines alive and
    refuses a replay if a psi tensor was replaced."""

    def __init__(self, engines, psis, nsteps, dts):
        n = len(engines)
        if not (len(psis) == len(dts) == n) or n == 0:
            raise ValueError("engines, psis and dts must have the same non-zero length")
        self.engines, self.psis, self.nsteps = list(engines), list(psis), int(nsteps)
        self._ptrs = [ptr(p).value for p in psis]
        H = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
        Pp = (ctypes.c_void_p * n)(*self._ptrs)
        dt = np.ascontiguousarray([[complex(d).real, complex(d).imag] for d in dts], dtype=np.float64)
        h = ctypes.c_void_p()
        nat.check(nat.lib().pp_pamp_batch_graph_create(n, H, Pp, self.nsteps, nat.dptr(dt), ctypes.byref(h)),
                  "pp_pamp_batch_graph_create")
        self.h = h
        self._token = nat.track("pp_pamp_batch_graph_destroy", h)

    def launch(self, scalings, psis=None, check=True):
        if psis is not None and [ptr(p).value for p in psis] != self._ptrs:
            raise RuntimeError("a wavefunction buffer changed since the graph was recorded")
        n = len(self.engines)
        keep = [np.ascontiguousarray(np.asarray(sc, dtype=np.float64).reshape(self.nsteps, len(e.terms)))
                for e, sc in zip(self.engines, scalings)]
        S = (ctypes.POINTER(ctypes.c_double) * n)(*[nat.dptr(k) for k in keep])
        nat.check(nat.lib().pp_pamp_batch_graph_launch(self.h, S, stream_ptr()), "pp_pamp_batch_graph_launch")
        for e in self.engines:
            e._last = self.nsteps
        if check:
            for i, e in enumerate(self.engines):
                check_pamp_steps(e.diagnostics(self.nsteps), "pAMP (batch instance %d)" % i)

    def __del__(self, _release=_release):   # bound now: module globals are gone at teardown
        _release(self)


def pamp_batch_propagate(engines, psis, nsteps, dts, scalings, check=True):
    """nsteps steps of many independent engines in packed cooperative
    launches (pp_pamp_batch_propagate).  scalings[i]: (nsteps, nterms_i).
    check: raise if any instance's step failed (synchronises)."""
    n = len(engines)
    if not (len(psis) == len(dts) == len(scalings) == n):
        raise ValueError("engines, psis, dts and scalings must have the same length")
    H = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
    Pp = (ctypes.c_void_p * n)(*[ptr(p).value for p in psis])
    dt = np.ascontiguousarray([[complex(d).real, complex(d).imag] for d in dts], dtype=np.float64)
    keep = [np.ascontiguousarray(np.asarray(sc, dtype=np.float64).reshape(nsteps, len(e.terms)))
            for e, sc in zip(engines, scalings)]
    S = (ctypes.POINTER(ctypes.c_double) * n)(*[nat.dptr(k) for k in keep])
    nat.check(nat.lib().pp_pamp_batch_propagate(n, H, Pp, int(nsteps), nat.dptr(dt), S, stream_ptr()),
              "pp_pamp_batch_propagate")
    for e in engines:
        e._last = nsteps
    if check:
        for i, e in enumerate(engines):
            check_pamp_steps(e.diagnostics(nsteps), "pAMP (batch instance %d)" % i)


class Projector:
    """pp_projector_*: the sparse map from weighted quadrature samples to the
    band pairs of one B-spline rank, applied on the device."""

    def __init__(self, MT):
        MT = MT.tocsr()
        MT.sort_indices()
        ptr = np.ascontiguousarray(MT.indptr, dtype=np.int32)
        idx = np.ascontiguousarray(MT.indices, dtype=np.int32)
        val = np.ascontiguousarray(MT.data, dtype=np.float64)
        if np.iscomplexobj(MT.data):
            raise ValueError("the projection map is real")
        self.nout, self.G = MT.shape
        h = ctypes.c_void_p()
        nat.check(nat.lib().pp_projector_create(self.nout, self.G, nat.iptr(ptr), nat.iptr(idx), nat.dptr(val),
                                                ctypes.byref(h)), "pp_projector_create")
        self.h = h
        self._token = nat.track("pp_projector_destroy", h)

    def __del__(self, _release=_release):
        _release(self)

    def apply(self, fw):
        """fw (rows x G complex128, host): the weighted samples; returns
        (rows x nout) complex128 on the host."""
        t = torch()
        fw = np.ascontiguousarray(fw, dtype=np.complex128).reshape(-1, self.G)
        src = to_device(fw)
        out = t.empty((fw.shape[0], self.nout), dtype=t.complex128, device=device())
        nat.check(nat.lib().pp_projector_apply(self.h, int(fw.shape[0]), ptr(src), ptr(out), stream_ptr()),
                  "pp_projector_apply")
        return out.cpu().numpy()
