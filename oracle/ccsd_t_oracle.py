"""Host-side driver of the CPU oracle (test infrastructure only).

Restates the reference's preparation and back-transformation around the C
kernel (pyscfad/cc/ccsd_t.py:82-192) in NumPy, and calls the C restatement
``ccsd_t_oracle.c`` through ctypes, the same way the reference calls its
``ccsd_t_energy_vjp`` (ccsd_t.py:107-133).

Inputs use the reference's tensor conventions (pyscfad/cc/ccsd.py:470-498):
    t1 (o,v), t2 (o,o,v,v), mo_energy (n,), fock (n,n), ovoo (o,v,o,o),
    ovov (o,v,o,v), ovvv packed (o,v,v(v+1)/2) or unpacked (o,v,v,v).
"""
import ctypes
import os

import numpy as np

from . import build as _build

_lib = None


def _load():
    global _lib
    if _lib is None:
        path = _build.LIB
        if not os.path.exists(path):
            _build.build()
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        i, lg = ctypes.c_int, ctypes.c_long
        lib.oracle_ccsd_t_energy_range.restype = ctypes.c_double
        lib.oracle_ccsd_t_energy_range.argtypes = [i, i, i] + [dp] * 6 + [lg, lg]
        lib.oracle_ccsd_t_energy_vjp_range.restype = None
        lib.oracle_ccsd_t_energy_vjp_range.argtypes = (
            [i, i, i] + [dp] * 6 + [ctypes.c_double] + [dp] * 6 + [lg, lg, dp])
        lib.oracle_njobs.restype = lg
        lib.oracle_njobs.argtypes = [i]
        lib.oracle_lno_ccsd_t_energy_range.restype = ctypes.c_double
        lib.oracle_lno_ccsd_t_energy_range.argtypes = [i, i] + [dp] * 7 + [lg, lg]
        lib.oracle_lno_ccsd_t_energy_vjp_range.restype = None
        lib.oracle_lno_ccsd_t_energy_vjp_range.argtypes = (
            [i, i] + [dp] * 7 + [ctypes.c_double] + [dp] * 7 + [lg, lg, dp])
        lib.oracle_lno_njobs.restype = lg
        lib.oracle_lno_njobs.argtypes = [i]
        lib.oracle_num_threads.restype = i
        lib.oracle_set_dgemm.restype = None
        lib.oracle_set_dgemm.argtypes = [ctypes.c_void_p]
        _lib = lib
    return _lib


def _scipy_dgemm():
    """Address of the Fortran-ABI dgemm of the BLAS that scipy ships (its
    cython_blas capsule), or None when scipy is absent."""
    try:
        from scipy.linalg import cython_blas
    except ImportError:
        return None
    cap = cython_blas.__pyx_capi__["dgemm"]
    api = ctypes.pythonapi
    api.PyCapsule_GetName.restype = ctypes.c_char_p
    api.PyCapsule_GetName.argtypes = [ctypes.py_object]
    api.PyCapsule_GetPointer.restype = ctypes.c_void_p
    api.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return api.PyCapsule_GetPointer(cap, api.PyCapsule_GetName(cap))


class blas:
    """Context manager: inside it the oracle's products run on scipy's BLAS,
    limited to one BLAS thread per OpenMP thread (the oracle's own threads
    call it concurrently); outside, on the oracle's plain loops.  The
    reference makes the same products with dgemm_ (ccsd_t.c:137-142,
    ccsd_t_vjp.c:139-226)."""

    def __enter__(self):
        import threadpoolctl
        ptr = _scipy_dgemm()
        if ptr is None:
            raise RuntimeError("scipy's BLAS is not available")
        self._limit = threadpoolctl.threadpool_limits(limits=1, user_api="blas")
        _load().oracle_set_dgemm(ptr)
        return self

    def __exit__(self, *exc):
        _load().oracle_set_dgemm(None)
        self._limit.restore_original_limits()
        return False


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def unpack_ovvv(ovvv, nvir):
    """(o,v,v(v+1)/2) lower-triangle packed -> (o,v,v,v), symmetric in the last
    two indices (pyscf lib.unpack_tril, used by eris.get_ovvv at ccsd_t.py:103).
    The pair (f,b) with f>=b is stored at f*(f+1)/2 + b."""
    ovvv = np.asarray(ovvv)
    if ovvv.ndim == 4:
        return ovvv
    idx_f, idx_b = np.tril_indices(nvir)
    full = np.empty(ovvv.shape[:2] + (nvir, nvir), dtype=np.float64)
    full[:, :, idx_f, idx_b] = ovvv
    full[:, :, idx_b, idx_f] = ovvv
    return full


def pack_ovvv_bar(ovvv_bar_full):
    """Adjoint of unpack_ovvv: off-diagonal pairs summed, diagonal kept
    (ccsd_t.py:186-191; equals the unpack_tril VJP, _numpy_helper_opt.py:32-55)."""
    nvir = ovvv_bar_full.shape[-1]
    x = ovvv_bar_full + ovvv_bar_full.transpose(0, 1, 3, 2)
    di = np.arange(nvir)
    x[:, :, di, di] *= 0.5
    idx_f, idx_b = np.tril_indices(nvir)
    return np.ascontiguousarray(x[:, :, idx_f, idx_b])


def host_arrays(t1, t2, mo_energy, fock, ovoo, ovov, ovvv, vrow=None):
    """The reference's host arrays (ccsd_t.py:88-106).  With ``vrow`` < nvir,
    only the rows the first njobs(vrow) jobs read (see Prepared)."""
    nocc, nvir = np.shape(t1)
    nmo = nocc + nvir
    w = nvir if vrow is None else int(vrow)
    t1T = np.ascontiguousarray(np.asarray(t1, dtype=np.float64).T[:w])
    t2T = np.ascontiguousarray(np.asarray(t2, dtype=np.float64).transpose(2, 3, 1, 0)[:w])
    mo_energy = np.ascontiguousarray(mo_energy, dtype=np.float64)
    fvo = np.ascontiguousarray(np.asarray(fock, dtype=np.float64)[nocc:nocc + w, :nocc])
    vooo = np.ascontiguousarray(np.asarray(ovoo, dtype=np.float64).transpose(1, 0, 3, 2)[:w])
    vvop = np.empty((w, w, nocc, nmo))
    vvop[:, :, :, :nocc] = np.asarray(ovov, dtype=np.float64)[:, :w, :, :w].transpose(1, 3, 0, 2)
    if np.ndim(ovvv) == 3 and w < nvir:
        # vvop[a,b,i,o+f] = ovvv[i,a,f,b] for a, b < w: unpack the (f, b) pairs with b < w only
        vv = np.empty((nocc, w, nvir, w))
        for f in range(nvir):
            lo = f * (f + 1) // 2
            nb = min(f + 1, w)
            vv[:, :, f, :nb] = ovvv[:, :w, lo:lo + nb]                 # b <= f: stored at (f, b)
            for b in range(f + 1, w):                                  # b > f: stored at (b, f)
                vv[:, :, f, b] = ovvv[:, :w, b * (b + 1) // 2 + f]
        vvop[:, :, :, nocc:] = vv.transpose(1, 3, 0, 2)
    else:
        vvop[:, :, :, nocc:] = unpack_ovvv(ovvv, nvir)[:, :w, :, :w].transpose(1, 3, 0, 2)
    return t1T, t2T, mo_energy, fvo, vooo, vvop


class Prepared:
    """Host arrays of one instance (built once) for repeated timed calls on
    slices of the job list -- the CPU baseline times only the kernel.

    ``vrow`` < nvir keeps only what the first njobs(vrow) jobs read (their
    job slots a, b, c lie below vrow; the contracted virtual index still runs
    over all nvir): the first vrow rows of t1T, t2T, vooo, fvo and the
    vrow x vrow leading block of vvop.  That makes a bounded sample of a large
    instance affordable in host memory."""

    def __init__(self, t1, t2, mo_energy, fock, ovoo, ovov, ovvv, vrow=None):
        self.nocc, self.nvir = np.shape(t1)
        self.vrow = self.nvir if vrow is None else int(vrow)
        (self.t1T, self.t2T, self.me, self.fvo, self.vooo,
         self.vvop) = host_arrays(t1, t2, mo_energy, fock, ovoo, ovov, ovvv, self.vrow)
        self.bars = None

    def njobs(self):
        return njobs(self.vrow)

    def energy(self, j0, j1):
        return _load().oracle_ccsd_t_energy_range(
            self.nocc, self.nvir, self.vrow, _ptr(self.me), _ptr(self.t1T), _ptr(self.t2T),
            _ptr(self.vooo), _ptr(self.fvo), _ptr(self.vvop), j0, j1)

    def energy_vjp(self, j0, j1, et_bar=1.0):
        if self.bars is None:
            self.bars = [np.zeros_like(x) for x in (self.me, self.t1T, self.t2T, self.vooo, self.fvo,
                                                    self.vvop)]
        e = ctypes.c_double(0.0)
        _load().oracle_ccsd_t_energy_vjp_range(
            self.nocc, self.nvir, self.vrow, _ptr(self.me), _ptr(self.t1T), _ptr(self.t2T),
            _ptr(self.vooo), _ptr(self.fvo), _ptr(self.vvop), float(et_bar),
            *[_ptr(b) for b in self.bars], j0, j1, ctypes.byref(e))
        return e.value


def njobs(nvir):
    return _load().oracle_njobs(nvir)


def num_threads():
    return _load().oracle_num_threads()


def energy(t1, t2, mo_energy, fock, ovoo, ovov, ovvv, job_range=None):
    """E(T).  ``job_range=(j0, j1)`` restricts the sum to a slice of the
    a>=b>=c job list (used for bounded CPU timing samples)."""
    nocc, nvir = np.shape(t1)
    t1T, t2T, me, fvo, vooo, vvop = host_arrays(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
    j0, j1 = job_range if job_range is not None else (0, njobs(nvir))
    return _load().oracle_ccsd_t_energy_range(nocc, nvir, nvir, _ptr(me), _ptr(t1T), _ptr(t2T),
                                              _ptr(vooo), _ptr(fvo), _ptr(vvop), j0, j1)


def energy_vjp(t1, t2, mo_energy, fock, ovoo, ovov, ovvv, et_bar=1.0, job_range=None,
               return_energy=False):
    """Cotangents of E(T), in the reference's output conventions
    (ccsd_t.py:179-192): dict with t1, t2, fock (full (n,n), only the vo block
    nonzero), mo_energy, ovoo, ovov, ovvv (packed if the input was packed).
    With ``return_energy`` also E(T) of the same jobs: (dict, E)."""
    nocc, nvir = np.shape(t1)
    nmo = nocc + nvir
    t1T, t2T, me, fvo, vooo, vvop = host_arrays(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
    me_b = np.zeros_like(me)
    t1T_b = np.zeros_like(t1T)
    t2T_b = np.zeros_like(t2T)
    vooo_b = np.zeros_like(vooo)
    fvo_b = np.zeros_like(fvo)
    vvop_b = np.zeros_like(vvop)
    j0, j1 = job_range if job_range is not None else (0, njobs(nvir))
    e = ctypes.c_double(0.0)
    _load().oracle_ccsd_t_energy_vjp_range(
        nocc, nvir, nvir, _ptr(me), _ptr(t1T), _ptr(t2T), _ptr(vooo), _ptr(fvo), _ptr(vvop),
        float(et_bar), _ptr(me_b), _ptr(t1T_b), _ptr(t2T_b), _ptr(vooo_b), _ptr(fvo_b),
        _ptr(vvop_b), j0, j1, ctypes.byref(e))
    del vvop
    fock_b = np.zeros((nmo, nmo))
    fock_b[nocc:, :nocc] = fvo_b
    ovvv_full_b = vvop_b[:, :, :, nocc:].transpose(2, 0, 3, 1)
    packed = np.ndim(ovvv) == 3
    out = {
        "t1": np.ascontiguousarray(t1T_b.T),
        "t2": np.ascontiguousarray(t2T_b.transpose(3, 2, 0, 1)),
        "fock": fock_b,
        "mo_energy": me_b,
        "ovoo": np.ascontiguousarray(vooo_b.transpose(1, 0, 3, 2)),
        "ovov": np.ascontiguousarray(vvop_b[:, :, :, :nocc].transpose(2, 0, 3, 1)),
        "ovvv": pack_ovvv_bar(ovvv_full_b) if packed else np.ascontiguousarray(ovvv_full_b),
    }
    return (out, e.value) if return_energy else out


# ---------------------------------------------------------------------------
# LNO (T) with an occupied projector (pyscfad/lno/ccsd_t.py:70-237)

def lno_njobs(nvir):
    """Jobs (a, b >= c) of the LNO job list (lno_ccsd_t.c:73-94)."""
    return _load().oracle_lno_njobs(nvir)


def lno_host_arrays(t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv):
    """The host arrays of pyscfad/lno/ccsd_t.py:78-93 (inputs already in the
    t1T (v,o), t2T (v,v,o,o), fvo (v,o) forms that function takes)."""
    nvir, nocc = np.shape(t1T)
    nmo = nocc + nvir
    t1T = np.ascontiguousarray(t1T, dtype=np.float64)
    t2T = np.ascontiguousarray(t2T, dtype=np.float64)
    mo_energy = np.ascontiguousarray(mo_energy, dtype=np.float64)
    fvo = np.ascontiguousarray(fvo, dtype=np.float64)
    vooo = np.ascontiguousarray(np.asarray(ovoo, dtype=np.float64).transpose(1, 0, 3, 2))
    vvop = np.empty((nvir, nvir, nocc, nmo))
    vvop[:, :, :, :nocc] = np.asarray(ovov, dtype=np.float64).transpose(1, 3, 0, 2)
    vvop[:, :, :, nocc:] = unpack_ovvv(ovvv, nvir).transpose(1, 3, 0, 2)
    return t1T, t2T, mo_energy, fvo, vooo, vvop


def lno_energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, job_range=None):
    """E of the LNO (T) (the value of _ccsd_t_energy, lno/ccsd_t.py:70-113,
    including its factor 2/3).  ``mat`` is any (o, o) matrix."""
    nvir, nocc = np.shape(t1T)
    t1T, t2T, me, fvo, vooo, vvop = lno_host_arrays(t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv)
    mat = np.ascontiguousarray(mat, dtype=np.float64)
    j0, j1 = job_range if job_range is not None else (0, lno_njobs(nvir))
    return _load().oracle_lno_ccsd_t_energy_range(nocc, nvir, _ptr(mat), _ptr(me), _ptr(t1T), _ptr(t2T),
                                                  _ptr(vooo), _ptr(fvo), _ptr(vvop), j0, j1)


def lno_energy_vjp(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, et_bar=1.0, job_range=None,
                   return_energy=False):
    """Cotangents of the LNO E, in the order and layouts of
    _ccsd_t_energy_bwd (lno/ccsd_t.py:121-235): (mat_bar, t1T_bar, t2T_bar,
    mo_energy_bar, fvo_bar, ovoo_bar, ovov_bar, ovvv_bar), ovvv_bar packed if
    the input was packed.  With ``return_energy`` also E: (tuple, E)."""
    nvir, nocc = np.shape(t1T)
    t1T, t2T, me, fvo, vooo, vvop = lno_host_arrays(t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv)
    mat = np.ascontiguousarray(mat, dtype=np.float64)
    bars = [np.zeros_like(x) for x in (mat, me, t1T, t2T, vooo, fvo, vvop)]
    j0, j1 = job_range if job_range is not None else (0, lno_njobs(nvir))
    e = ctypes.c_double(0.0)
    _load().oracle_lno_ccsd_t_energy_vjp_range(
        nocc, nvir, _ptr(mat), _ptr(me), _ptr(t1T), _ptr(t2T), _ptr(vooo), _ptr(fvo), _ptr(vvop),
        float(et_bar), *[_ptr(b) for b in bars], j0, j1, ctypes.byref(e))
    mat_b, me_b, t1T_b, t2T_b, vooo_b, fvo_b, vvop_b = bars
    ovvv_full_b = vvop_b[:, :, :, nocc:].transpose(2, 0, 3, 1)
    out = (mat_b, t1T_b, t2T_b, me_b, fvo_b,
           np.ascontiguousarray(vooo_b.transpose(1, 0, 3, 2)),
           np.ascontiguousarray(vvop_b[:, :, :, :nocc].transpose(2, 0, 3, 1)),
           pack_ovvv_bar(ovvv_full_b) if np.ndim(ovvv) == 3 else np.ascontiguousarray(ovvv_full_b))
    return (out, e.value) if return_energy else out


class LnoPrepared:
    """Host arrays of one LNO instance (built once) for timed calls on slices
    of the LNO job list -- the CPU baseline times only the kernel."""

    def __init__(self, mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv):
        self.nvir, self.nocc = np.shape(t1T)
        self.mat = np.ascontiguousarray(mat, dtype=np.float64)
        (self.t1T, self.t2T, self.me, self.fvo, self.vooo,
         self.vvop) = lno_host_arrays(t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv)
        self.bars = None

    def njobs(self):
        return lno_njobs(self.nvir)

    def energy(self, j0, j1):
        return _load().oracle_lno_ccsd_t_energy_range(
            self.nocc, self.nvir, _ptr(self.mat), _ptr(self.me), _ptr(self.t1T), _ptr(self.t2T),
            _ptr(self.vooo), _ptr(self.fvo), _ptr(self.vvop), j0, j1)

    def energy_vjp(self, j0, j1, et_bar=1.0):
        if self.bars is None:
            self.bars = [np.zeros_like(x) for x in (self.mat, self.me, self.t1T, self.t2T, self.vooo,
                                                    self.fvo, self.vvop)]
        e = ctypes.c_double(0.0)
        _load().oracle_lno_ccsd_t_energy_vjp_range(
            self.nocc, self.nvir, _ptr(self.mat), _ptr(self.me), _ptr(self.t1T), _ptr(self.t2T),
            _ptr(self.vooo), _ptr(self.fvo), _ptr(self.vvop), float(et_bar), *[_ptr(b) for b in self.bars],
            j0, j1, ctypes.byref(e))
        return e.value
