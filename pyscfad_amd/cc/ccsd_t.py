"""Drop-in for ``pyscfad.cc.ccsd_t`` (pyscfad/cc/ccsd_t.py) on MI355X.

Same entry points, argument meaning and cotangent conventions as the
reference:

* ``kernel(mycc, eris, t1=None, t2=None, verbose=None)`` -- E(T) with a
  ``jax.custom_vjp`` over ``(eris, t1, t2)`` whose backward maps the
  cotangents onto the eris leaves by name and zero-fills the others
  (ccsd_t.py:28-80).  JAX is optional: without it ``kernel`` returns E(T)
  as a float.
* ``_ccsd_t_energy_vjp(eris, t1, t2, et_bar, max_memory)`` -- returns
  ``(t1_bar, t2_bar, fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, ovvv_bar)``
  in the layouts of ccsd_t.py:179-192 (ovvv_bar lower-triangle packed).
* ``CCSDTEnergy`` -- the same pair as a ``torch.autograd.Function`` over
  ``(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)``.
* ``energy_df`` / ``energy_vjp_df`` / ``CCSDTEnergyDF`` -- the same with the
  DF factors of pyscfad's DF-CCSD (pyscfad/cc/dfccsd.py:81-91) in place of
  ovvv: the o*v^3 slabs are built on the device as they are needed, and the
  cotangent of ovvv comes back as the cotangents of Lov and Lvv.

Both passes run on the GPU through the C ABI of ``libpyscfad_amd.so``;
arrays may be NumPy (copied to the device), torch tensors (used in place when
already contiguous FP64 on the device) or any DLPack array (a JAX device
array is taken without a copy).  ``stream=`` runs a call on another torch
stream: it waits for the caller's current stream first, and the current
stream waits for it afterwards, so results can be used right away on either.  ``precision`` selects the
arithmetic of the W/V contractions: 'f64' (default, the drop-in), 'f32',
'bf16x3', or 'bf16x6' (FP32-level accuracy from three-part bf16 splits);
denominators, the E(T) sum and cotangent accumulation stay FP64.
"""
import numpy
import torch

from .. import _lib

DEFAULT_PRECISION = "f64"


def _device():
    if not torch.cuda.is_available():
        raise _lib.T3Error("pyscfad_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _as_torch(x):
    """x as a torch tensor: itself, a zero-copy DLPack view (JAX and NumPy
    arrays), or a NumPy copy for anything else array-like."""
    if isinstance(x, torch.Tensor):
        return x
    if hasattr(x, "__dlpack__"):
        try:
            return torch.from_dlpack(x)
        except (RuntimeError, TypeError, BufferError):
            pass
    return torch.from_numpy(numpy.ascontiguousarray(x, dtype=numpy.float64))


def _dev(x):
    """Contiguous FP64 device tensor view/copy of x."""
    return _as_torch(x).to(device=_device(), dtype=torch.float64).contiguous()


class _OnStream:
    """Runs a call's allocations, copies and kernels on ``stream`` (default:
    the current stream).  On another stream the call first waits for the
    current one; caller tensors read in place are recorded on it (the caching
    allocator keeps them until it is done); afterwards the current stream
    waits for it and the outputs are recorded on the current stream."""

    def __init__(self, stream):
        _device()
        self.cur = torch.cuda.current_stream()
        self.s = self.cur if stream is None else stream
        self.side = self.s != self.cur
        self._ctx = None

    def __enter__(self):
        if self.side:
            self.s.wait_stream(self.cur)
            self._ctx = torch.cuda.stream(self.s)
            self._ctx.__enter__()
        return self

    def uses(self, *tensors):
        if self.side:
            for t in tensors:
                if isinstance(t, torch.Tensor) and t.is_cuda:
                    t.record_stream(self.s)

    def outputs(self, *tensors):
        if self.side:
            for t in tensors:
                if isinstance(t, torch.Tensor) and t.is_cuda:
                    t.record_stream(self.cur)

    def __exit__(self, *exc):
        if self.side:
            self._ctx.__exit__(*exc)
            self.cur.wait_stream(self.s)
        return False


def _ptr(t):
    return t.data_ptr()


def _prec(precision):
    if isinstance(precision, int):
        return precision
    try:
        return _lib.PRECISIONS[precision.lower()]
    except KeyError:
        raise _lib.T3Error(f"unknown precision {precision!r}; use one of {sorted(_lib.PRECISIONS)}")


def _inputs(t1, t2, mo_energy, fock, ovoo, ovov, ovvv):
    t1d, t2d = _dev(t1), _dev(t2)
    nocc, nvir = t1d.shape
    me, fk, oo, ov, vv = (_dev(mo_energy), _dev(fock), _dev(ovoo), _dev(ovov), _dev(ovvv))
    n = nocc + nvir
    packed = vv.dim() == 3
    expect = {
        "t2": (t2d, (nocc, nocc, nvir, nvir)), "mo_energy": (me, (n,)), "fock": (fk, (n, n)),
        "ovoo": (oo, (nocc, nvir, nocc, nocc)), "ovov": (ov, (nocc, nvir, nocc, nvir)),
        "ovvv": (vv, (nocc, nvir, nvir * (nvir + 1) // 2) if packed else (nocc, nvir, nvir, nvir)),
    }
    for name, (arr, shape) in expect.items():
        if tuple(arr.shape) != shape:
            raise ValueError(f"{name} has shape {tuple(arr.shape)}, expected {shape}")
    return nocc, nvir, t1d, t2d, me, fk, oo, ov, vv, packed


def num_triples(nocc):
    """Number of unique occupied triples i >= j >= k (the unit of work sharding)."""
    return _lib.load().pyscfad_amd_ccsd_t_num_triples(nocc)


def _ws(workspace):
    """(pointer, bytes) of a caller-provided device workspace (a torch tensor), or (None, 0)."""
    if workspace is None:
        return None, 0
    return workspace.data_ptr(), workspace.numel() * workspace.element_size()


def energy(t1, t2, mo_energy, fock, ovoo, ovov, ovvv, precision=DEFAULT_PRECISION, stream=None,
           triples=None, workspace=None):
    """E(T) as a 0-d FP64 device tensor (the forward of the custom_vjp).
    ``triples=(begin, end)`` restricts the sum to that slice of the unique
    occupied triples (a shard of the work).  ``workspace``: an optional device
    tensor of at least ``workspace_bytes(..., with_vjp=False)`` bytes, 256-byte
    aligned; without it the library allocates and frees its own."""
    lib = _lib.load()
    t0, t1_ = triples if triples is not None else (0, num_triples(numpy.shape(t1)[0]))
    with _OnStream(stream) as on:
        nocc, nvir, t1d, t2d, me, fk, oo, ov, vv, packed = _inputs(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
        on.uses(t1d, t2d, me, fk, oo, ov, vv, workspace)
        out = torch.zeros((), dtype=torch.float64, device=t1d.device)
        _lib.check(lib.pyscfad_amd_ccsd_t_energy_slice(
            nocc, nvir, _ptr(me), _ptr(t1d), _ptr(t2d), _ptr(fk), _ptr(oo), _ptr(ov), _ptr(vv),
            int(packed), _prec(precision), int(t0), int(t1_), _ptr(out), *_ws(workspace), on.s.cuda_stream))
        on.outputs(out)
    return out


def energy_vjp(t1, t2, mo_energy, fock, ovoo, ovov, ovvv, et_bar=1.0, precision=DEFAULT_PRECISION,
               stream=None, return_energy=False, triples=None, workspace=None):
    """Cotangents of E(T) as FP64 device tensors, in the order and layouts of
    the reference's _ccsd_t_energy_vjp: (t1, t2, fock, mo_energy, ovoo, ovov, ovvv)."""
    lib = _lib.load()
    if hasattr(et_bar, "item"):
        et_bar = et_bar.item()
    t0, t1_ = triples if triples is not None else (0, num_triples(numpy.shape(t1)[0]))
    with _OnStream(stream) as on:
        nocc, nvir, t1d, t2d, me, fk, oo, ov, vv, packed = _inputs(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
        on.uses(t1d, t2d, me, fk, oo, ov, vv, workspace)
        bars = [torch.empty_like(a) for a in (t1d, t2d, fk, me, oo, ov, vv)]
        out = torch.zeros((), dtype=torch.float64, device=t1d.device)
        _lib.check(lib.pyscfad_amd_ccsd_t_energy_vjp_slice(
            nocc, nvir, _ptr(me), _ptr(t1d), _ptr(t2d), _ptr(fk), _ptr(oo), _ptr(ov), _ptr(vv),
            int(packed), float(et_bar), _prec(precision), int(t0), int(t1_), *[_ptr(b) for b in bars],
            _ptr(out), *_ws(workspace), on.s.cuda_stream))
        on.outputs(out, *bars)
    if return_energy:
        return tuple(bars), out
    return tuple(bars)


def _instances(problems, precision, with_vjp, outputs=None):
    """ctypes array of PyscfadAmdInstance for a list of problems (dicts with the
    seven inputs and optionally et_bar), plus the device tensors it points at.
    ``outputs``: the result list of an earlier call, whose tensors are reused."""
    import ctypes
    arr = (_lib.Instance * len(problems))()
    keep = []
    for x, pr in enumerate(problems):
        nocc, nvir, t1d, t2d, me, fk, oo, ov, vv, packed = _inputs(
            pr["t1"], pr["t2"], pr["mo_energy"], pr["fock"], pr["ovoo"], pr["ovov"], pr["ovvv"])
        if outputs is not None:
            out = outputs[x][0]
            bars = list(outputs[x][1]) if with_vjp else [None] * 7
        else:
            out = torch.zeros((), dtype=torch.float64, device=t1d.device)
            bars = [torch.empty_like(a) for a in (t1d, t2d, fk, me, oo, ov, vv)] if with_vjp else [None] * 7
        et_bar = pr.get("et_bar", 1.0)
        if hasattr(et_bar, "item"):
            et_bar = et_bar.item()
        I = arr[x]
        I.nocc, I.nvir, I.ovvv_packed, I.et_bar = nocc, nvir, int(packed), float(et_bar)
        for name, t in zip(("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv"), (t1d, t2d, me, fk, oo, ov, vv)):
            setattr(I, name, t.data_ptr())
        for name, t in zip(("t1_bar", "t2_bar", "fock_bar", "mo_energy_bar", "ovoo_bar", "ovov_bar", "ovvv_bar"),
                           bars):
            setattr(I, name, t.data_ptr() if t is not None else None)
        I.e_out = out.data_ptr()
        keep.append((t1d, t2d, me, fk, oo, ov, vv, out, bars))
    return arr, keep


def batched_workspace_bytes(problems, precision=DEFAULT_PRECISION, with_vjp=True):
    arr, keep = _instances(problems, precision, with_vjp)
    return _lib.load().pyscfad_amd_ccsd_t_batched_workspace_bytes(len(problems), arr, _prec(precision),
                                                                  int(with_vjp))


def batched(problems, precision=DEFAULT_PRECISION, with_vjp=True, stream=None, workspace=None, outputs=None):
    """E(T) (and, with with_vjp, the cotangents) of many independent instances
    of uneven size in one call: the instances, dealt to up to four stream
    groups (largest first), run as captured hipGraphs on the groups' streams,
    forked from and joined back into ``stream``, and the graphs are replayed
    when the same call comes again.  ``problems`` is a list
    of dicts with t1, t2, mo_energy, fock, ovoo, ovov, ovvv (and optionally
    et_bar).  Returns a list of (E, bars) per instance, bars in the order of
    energy_vjp (None without with_vjp).  The arrays are used in place when they
    are contiguous FP64 device tensors; a repeated call with the same input
    tensors (new values allowed), the same ``workspace`` and ``outputs`` = the
    previous result replays the cached graph."""
    lib = _lib.load()
    with _OnStream(stream) as on:
        arr, keep = _instances(problems, precision, with_vjp, outputs)
        need = lib.pyscfad_amd_ccsd_t_batched_workspace_bytes(len(problems), arr, _prec(precision), int(with_vjp))
        if need == 0:
            raise _lib.T3Error(lib.pyscfad_amd_last_error().decode())
        if workspace is None:
            workspace = torch.empty(need, dtype=torch.uint8, device=keep[0][0].device)
        for k in keep:
            on.uses(*k[:7])
        on.uses(workspace)
        # the workspace of this call stays allocated until the stream is past it
        workspace.record_stream(on.s)
        _lib.check(lib.pyscfad_amd_ccsd_t_batched(len(problems), arr, _prec(precision), int(with_vjp),
                                                  *_ws(workspace), on.s.cuda_stream))
        for k in keep:
            on.outputs(k[7], *(k[8] if with_vjp else ()))
    return [(k[7], tuple(k[8]) if with_vjp else None) for k in keep]


TUNED_FAMILIES = ("w_build", "ovvv_adjoint", "t2_particle_adjoint", "ovoo_adjoint", "t2_hole_adjoint")
TILE_CONFIGS = {0: (64, 64), 1: (128, 128), 3: (64, 32), 4: (32, 64), 5: (128, 64), 6: (64, 128)}


def autotune(nocc, nvir, precision=DEFAULT_PRECISION, stream=None):
    """Pick the batch size and time every tile configuration of each
    contraction family for this (nocc, nvir) size; later calls in this
    process use the fastest (query workspace sizes after tuning).
    Returns {family: (tile_m, tile_n) or None if left at the default (the
    backward families when only the forward workspace fits), "batch": triples
    per batch}."""
    import ctypes
    out = (ctypes.c_int * 6)()
    s = stream if stream is not None else torch.cuda.current_stream()
    _lib.check(_lib.load().pyscfad_amd_ccsd_t_autotune(nocc, nvir, _prec(precision), out, s.cuda_stream))
    res = {f: TILE_CONFIGS.get(out[x]) for x, f in enumerate(TUNED_FAMILIES)}
    res["batch"] = out[5]
    return res


def autotune_problems(problems, precision=DEFAULT_PRECISION, stream=None):
    """Autotune every padded size class of a list of problems (the largest
    nocc of each class stands for it); returns {(nocc, nvir): choice}."""
    bk = 16 if _prec(precision) == _lib.PREC_F64 else 32
    cls = {}
    for pr in problems:
        o, v = pr["t1"].shape
        key = ((v + 31) // 32, (o + bk - 1) // bk)
        if key not in cls or cls[key][0] < o:
            cls[key] = (o, v)
    return {ov: autotune(ov[0], ov[1], precision, stream) for ov in cls.values()}


def workspace_bytes(nocc, nvir, precision=DEFAULT_PRECISION, with_vjp=True):
    return _lib.load().pyscfad_amd_ccsd_t_workspace_bytes(nocc, nvir, _prec(precision), int(with_vjp))


# ---- density-fitted variant: ovvv built on the device from the DF factors ----

def _df_inputs(t1, t2, mo_energy, fock, ovoo, ovov, Lov, Lvv):
    """Device copies of the inputs of the density-fitted calls.  Lov is
    (naux, nocc, nvir) or, as pyscfad/cc/dfccsd.py:82 holds it, (naux,
    nocc*nvir); Lvv is (naux, nvir(nvir+1)/2) lower-triangle packed
    (dfccsd.py:83)."""
    t1d, t2d = _dev(t1), _dev(t2)
    nocc, nvir = t1d.shape
    me, fk, oo, ov, lo, lv = (_dev(x) for x in (mo_energy, fock, ovoo, ovov, Lov, Lvv))
    naux = lo.shape[0]
    if lo.dim() == 2 and lo.shape[1] == nocc * nvir:
        lo = lo.reshape(naux, nocc, nvir)
    n = nocc + nvir
    expect = {
        "t2": (t2d, (nocc, nocc, nvir, nvir)), "mo_energy": (me, (n,)), "fock": (fk, (n, n)),
        "ovoo": (oo, (nocc, nvir, nocc, nocc)), "ovov": (ov, (nocc, nvir, nocc, nvir)),
        "Lov": (lo, (naux, nocc, nvir)), "Lvv": (lv, (naux, nvir * (nvir + 1) // 2)),
    }
    for name, (arr, shape) in expect.items():
        if tuple(arr.shape) != shape:
            raise ValueError(f"{name} has shape {tuple(arr.shape)}, expected {shape}")
    return nocc, nvir, naux, t1d, t2d, me, fk, oo, ov, lo, lv


def df_workspace_bytes(nocc, nvir, naux, precision=DEFAULT_PRECISION, with_vjp=True):
    return _lib.load().pyscfad_amd_ccsd_t_df_workspace_bytes(nocc, nvir, naux, _prec(precision), int(with_vjp))


def energy_df(t1, t2, mo_energy, fock, ovoo, ovov, Lov, Lvv, precision=DEFAULT_PRECISION, stream=None,
              triples=None, workspace=None):
    """E(T) with ovvv = Lov^T Lvv (pyscfad/cc/dfccsd.py:91) built slab by slab
    on the device, never stored whole.  ``triples`` selects a slice of the
    unique triples in this variant's blocked order."""
    t0, t1_ = triples if triples is not None else (0, num_triples(numpy.shape(t1)[0]))
    with _OnStream(stream) as on:
        nocc, nvir, naux, t1d, t2d, me, fk, oo, ov, lo, lv = _df_inputs(t1, t2, mo_energy, fock, ovoo, ovov,
                                                                         Lov, Lvv)
        on.uses(t1d, t2d, me, fk, oo, ov, lo, lv, workspace)
        out = torch.zeros((), dtype=torch.float64, device=t1d.device)
        _lib.check(_lib.load().pyscfad_amd_ccsd_t_df_energy(
            nocc, nvir, naux, _ptr(me), _ptr(t1d), _ptr(t2d), _ptr(fk), _ptr(oo), _ptr(ov), _ptr(lo), _ptr(lv),
            _prec(precision), int(t0), int(t1_), _ptr(out), *_ws(workspace), on.s.cuda_stream))
        on.outputs(out)
    return out


def energy_vjp_df(t1, t2, mo_energy, fock, ovoo, ovov, Lov, Lvv, et_bar=1.0, precision=DEFAULT_PRECISION,
                  stream=None, return_energy=False, triples=None, workspace=None):
    """Cotangents (t1, t2, fock, mo_energy, ovoo, ovov, Lov, Lvv) of the
    density-fitted E(T): those of energy_vjp for the first six, and for the DF
    factors the cotangent of ovvv carried through ovvv = Lov^T Lvv
    (Lov_bar in the shape Lov was given in)."""
    lov_shape = tuple(Lov.shape)
    if hasattr(et_bar, "item"):
        et_bar = et_bar.item()
    t0, t1_ = triples if triples is not None else (0, num_triples(numpy.shape(t1)[0]))
    with _OnStream(stream) as on:
        nocc, nvir, naux, t1d, t2d, me, fk, oo, ov, lo, lv = _df_inputs(t1, t2, mo_energy, fock, ovoo, ovov,
                                                                         Lov, Lvv)
        on.uses(t1d, t2d, me, fk, oo, ov, lo, lv, workspace)
        bars = [torch.empty_like(a) for a in (t1d, t2d, fk, me, oo, ov, lo, lv)]
        out = torch.zeros((), dtype=torch.float64, device=t1d.device)
        _lib.check(_lib.load().pyscfad_amd_ccsd_t_df_energy_vjp(
            nocc, nvir, naux, _ptr(me), _ptr(t1d), _ptr(t2d), _ptr(fk), _ptr(oo), _ptr(ov), _ptr(lo), _ptr(lv),
            float(et_bar), _prec(precision), int(t0), int(t1_), *[_ptr(b) for b in bars], _ptr(out),
            *_ws(workspace), on.s.cuda_stream))
        on.outputs(out, *bars)
    bars[6] = bars[6].reshape(lov_shape)
    if return_energy:
        return tuple(bars), out
    return tuple(bars)


def _like(t, ref):
    """t on the device and in the dtype of the autograd input ref (autograd
    requires a gradient to match its input; inputs may live on the host)."""
    return t.to(device=ref.device, dtype=ref.dtype)


class CCSDTEnergyDF(torch.autograd.Function):
    """The density-fitted E(T) as a torch autograd function of
    (t1, t2, mo_energy, fock, ovoo, ovov, Lov, Lvv)."""

    @staticmethod
    def forward(ctx, t1, t2, mo_energy, fock, ovoo, ovov, Lov, Lvv, precision=DEFAULT_PRECISION):
        ctx.save_for_backward(t1, t2, mo_energy, fock, ovoo, ovov, Lov, Lvv)
        ctx.precision = precision
        return _like(energy_df(t1, t2, mo_energy, fock, ovoo, ovov, Lov, Lvv, precision=precision), t1)

    @staticmethod
    def backward(ctx, gbar):
        t1, t2, me, fk, oo, ov, lo, lv = ctx.saved_tensors
        t1b, t2b, fkb, meb, oob, ovb, lob, lvb = energy_vjp_df(t1, t2, me, fk, oo, ov, lo, lv, et_bar=float(gbar),
                                                              precision=ctx.precision)
        grads = (t1b, t2b, meb, fkb, oob, ovb, lob, lvb)
        return tuple(_like(g, x) for g, x in zip(grads, (t1, t2, me, fk, oo, ov, lo, lv))) + (None,)


def _ccsd_t_energy_vjp(eris, t1, t2, et_bar, max_memory=None, precision=DEFAULT_PRECISION):
    """Same signature and return value as pyscfad/cc/ccsd_t.py:82 (max_memory is
    accepted for compatibility; device memory is sized by the library).
    Returns NumPy arrays, like the reference."""
    bars = energy_vjp(t1, t2, eris.mo_energy, eris.fock, eris.ovoo, eris.ovov, eris.ovvv,
                      et_bar=et_bar, precision=precision)
    return tuple(b.cpu().numpy() for b in bars)


class CCSDTEnergy(torch.autograd.Function):
    """E(T) with the hand-written VJP, as a torch autograd function of
    (t1, t2, mo_energy, fock, ovoo, ovov, ovvv)."""

    @staticmethod
    def forward(ctx, t1, t2, mo_energy, fock, ovoo, ovov, ovvv, precision=DEFAULT_PRECISION):
        ctx.save_for_backward(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
        ctx.precision = precision
        return _like(energy(t1, t2, mo_energy, fock, ovoo, ovov, ovvv, precision=precision), t1)

    @staticmethod
    def backward(ctx, gbar):
        t1, t2, me, fk, oo, ov, vv = ctx.saved_tensors
        t1b, t2b, fkb, meb, oob, ovb, vvb = energy_vjp(t1, t2, me, fk, oo, ov, vv,
                                                        et_bar=float(gbar), precision=ctx.precision)
        grads = (t1b, t2b, meb, fkb, oob, ovb, vvb)
        return tuple(_like(g, x) for g, x in zip(grads, (t1, t2, me, fk, oo, ov, vv))) + (None,)


class ERIs:
    """Minimal container with the eris leaves the (T) path reads
    (the dynamic attributes of pyscfad's _ChemistsERIs, ccsd.py:30-31)."""

    def __init__(self, fock, mo_energy, ovoo, ovov, ovvv, **others):
        self.fock = fock
        self.mo_energy = mo_energy
        self.ovoo = ovoo
        self.ovov = ovov
        self.ovvv = ovvv
        for k, val in others.items():
            setattr(self, k, val)


COTANGENT_LEAVES = ("fock", "mo_energy", "ovoo", "ovov", "ovvv")


def eris_cotangent_leaves(paths_and_leaves, named):
    """Cotangent leaves for a flattened eris pytree (the list of (path, leaf)
    pairs of jax.tree_util.tree_flatten_with_path): a leaf whose path starts
    at an attribute named in ``named`` gets that cotangent, every other leaf
    zeros of its own shape.  This is the pytree contract of the reference's
    backward (pyscfad/cc/ccsd_t.py:52-77): leaves are matched by their
    GetAttrKey name, and oooo, vvvv, Lvv or any other leaf gets zeros."""
    return [named[path[0].name] if path and getattr(path[0], "name", None) in named
            else numpy.zeros(numpy.shape(leaf)) for path, leaf in paths_and_leaves]


def kernel(mycc, eris, t1=None, t2=None, verbose=None, precision=DEFAULT_PRECISION):
    """E(T) of ``mycc``'s amplitudes (pyscfad/cc/ccsd_t.py:28).

    With JAX importable this is a ``custom_vjp`` over (eris, t1, t2) whose
    backward returns cotangents for the leaves fock, mo_energy, ovoo, ovov and
    ovvv by name and zeros for every other leaf (ccsd_t.py:48-77).  Without
    JAX it returns E(T) as a float.  Device arrays reach the library through
    DLPack without a host copy.
    """
    if t1 is None:
        t1 = mycc.t1
    if t2 is None:
        t2 = mycc.t2

    def _e(eris, t1, t2):
        return float(energy(t1, t2, eris.mo_energy, eris.fock, eris.ovoo, eris.ovov, eris.ovvv,
                            precision=precision))

    try:
        from jax import custom_vjp
        from jax.tree_util import tree_flatten_with_path, tree_unflatten
    except ImportError:
        return _e(eris, t1, t2)

    @custom_vjp
    def _ccsd_t_kernel(eris, t1, t2):
        return _e(eris, t1, t2)

    def _fwd(eris, t1, t2):
        return _ccsd_t_kernel(eris, t1, t2), (eris, t1, t2)

    def _bwd(res, ybar):
        eris, t1, t2 = res
        t1_bar, t2_bar, fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, ovvv_bar = _ccsd_t_energy_vjp(
            eris, t1, t2, float(ybar), precision=precision)
        flat, treedef = tree_flatten_with_path(eris)
        named = dict(zip(COTANGENT_LEAVES, (fock_bar, mo_energy_bar, ovoo_bar, ovov_bar, ovvv_bar)))
        return tree_unflatten(treedef, eris_cotangent_leaves(flat, named)), t1_bar, t2_bar

    _ccsd_t_kernel.defvjp(_fwd, _bwd)
    return _ccsd_t_kernel(eris, t1, t2)
