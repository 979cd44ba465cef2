"""Drop-in for ``pyscfad.lno.ccsd_t`` (pyscfad/lno/ccsd_t.py) on MI355X: the
impurity (T) correction of the LNO method, with an occupied projector.

Same entry points, argument meaning and cotangent conventions as the
reference:

* ``kernel(mycc, eris, ulo, t1=None, t2=None, verbose=None)`` -- forms
  ``mat = ulo^T ulo``, ``t1T``, ``t2T = t2.transpose(2,3,1,0)`` and
  ``fvo = fock[o:, :o]`` (lno/ccsd_t.py:35-60) and returns the energy.
* ``_ccsd_t_energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv,
  max_memory)`` -- the energy, including the reference's factor 2/3
  (lno/ccsd_t.py:70-113); a ``jax.custom_vjp`` when JAX is importable.
* ``_ccsd_t_energy_fwd`` / ``_ccsd_t_energy_bwd(max_memory, res, et_bar)`` --
  its two passes; the backward returns ``(mat_bar, t1T_bar, t2T_bar,
  mo_energy_bar, fvo_bar, ovoo_bar, ovov_bar, ovvv_tril_bar)`` as NumPy
  arrays (lno/ccsd_t.py:121-235).
* ``LNOCCSDTEnergy`` -- the same pair as a ``torch.autograd.Function``.

``mat`` may be any (o, o) matrix (the gradient is the unsymmetrised
mat_bar of lno_ccsd_t_vjp.c:104-106).  Both passes run on the GPU through the
C ABI of ``libpyscfad_amd.so`` (``pyscfad_amd_lno_ccsd_t_*``); arrays follow
the conventions of ``pyscfad_amd.cc.ccsd_t`` (NumPy, torch or DLPack; device
tensors used in place; ``stream=``, ``workspace=`` and ``precision`` as
there).  ovvv is packed (o, v, v(v+1)/2) as the reference reads it
(lno/ccsd_t.py:62-68), or dense (o, v, v, v).
"""
import numpy
import torch

from .. import _lib
from ..cc.ccsd_t import DEFAULT_PRECISION, _OnStream, _dev, _like, _prec, _ptr, _ws


def _inputs(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv):
    t1d = _dev(t1T)
    nvir, nocc = t1d.shape
    md, t2d, me, fv, oo, ov, vv = (_dev(x) for x in (mat, t2T, mo_energy, fvo, ovoo, ovov, ovvv))
    n = nocc + nvir
    packed = vv.dim() == 3
    expect = {
        "mat": (md, (nocc, nocc)), "t2T": (t2d, (nvir, nvir, nocc, nocc)), "mo_energy": (me, (n,)),
        "fvo": (fv, (nvir, nocc)), "ovoo": (oo, (nocc, nvir, nocc, nocc)), "ovov": (ov, (nocc, nvir, nocc, nvir)),
        "ovvv": (vv, (nocc, nvir, nvir * (nvir + 1) // 2) if packed else (nocc, nvir, nvir, nvir)),
    }
    for name, (arr, shape) in expect.items():
        if tuple(arr.shape) != shape:
            raise ValueError(f"{name} has shape {tuple(arr.shape)}, expected {shape}")
    return nocc, nvir, md, t1d, t2d, me, fv, oo, ov, vv, packed


def workspace_bytes(nocc, nvir, precision=DEFAULT_PRECISION, with_vjp=True):
    return _lib.load().pyscfad_amd_lno_ccsd_t_workspace_bytes(nocc, nvir, _prec(precision), int(with_vjp))


def energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, precision=DEFAULT_PRECISION, stream=None,
           workspace=None):
    """The LNO (T) energy (x 2/3 included) as a 0-d FP64 device tensor."""
    lib = _lib.load()
    with _OnStream(stream) as on:
        nocc, nvir, md, t1d, t2d, me, fv, oo, ov, vv, packed = _inputs(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov,
                                                                        ovvv)
        on.uses(md, t1d, t2d, me, fv, oo, ov, vv, workspace)
        out = torch.zeros((), dtype=torch.float64, device=t1d.device)
        _lib.check(lib.pyscfad_amd_lno_ccsd_t_energy(
            nocc, nvir, _ptr(md), _ptr(t1d), _ptr(t2d), _ptr(me), _ptr(fv), _ptr(oo), _ptr(ov), _ptr(vv),
            int(packed), _prec(precision), _ptr(out), *_ws(workspace), on.s.cuda_stream))
        on.outputs(out)
    return out


def energy_vjp(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, et_bar=1.0, precision=DEFAULT_PRECISION,
               stream=None, return_energy=False, workspace=None):
    """Cotangents as FP64 device tensors, in the order and layouts of the
    reference's backward: (mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv)."""
    lib = _lib.load()
    if hasattr(et_bar, "item"):
        et_bar = et_bar.item()
    with _OnStream(stream) as on:
        nocc, nvir, md, t1d, t2d, me, fv, oo, ov, vv, packed = _inputs(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov,
                                                                        ovvv)
        on.uses(md, t1d, t2d, me, fv, oo, ov, vv, workspace)
        bars = [torch.empty_like(a) for a in (md, t1d, t2d, me, fv, oo, ov, vv)]
        out = torch.zeros((), dtype=torch.float64, device=t1d.device)
        _lib.check(lib.pyscfad_amd_lno_ccsd_t_energy_vjp(
            nocc, nvir, _ptr(md), _ptr(t1d), _ptr(t2d), _ptr(me), _ptr(fv), _ptr(oo), _ptr(ov), _ptr(vv),
            int(packed), float(et_bar), _prec(precision), *[_ptr(b) for b in bars], _ptr(out), *_ws(workspace),
            on.s.cuda_stream))
        on.outputs(out, *bars)
    if return_energy:
        return tuple(bars), out
    return tuple(bars)


def _ccsd_t_energy_fwd(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, max_memory=None,
                       precision=DEFAULT_PRECISION):
    """(energy, residuals) as the reference's forward rule (lno/ccsd_t.py:115-119)."""
    et = float(energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, precision=precision))
    return et, (mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv)


def _ccsd_t_energy_bwd(max_memory, res, et_bar, precision=DEFAULT_PRECISION):
    """The reference's backward rule (lno/ccsd_t.py:121-235): NumPy cotangents
    (mat_bar, t1T_bar, t2T_bar, mo_energy_bar, fvo_bar, ovoo_bar, ovov_bar,
    ovvv_tril_bar).  max_memory is accepted for compatibility; device memory
    is sized by the library."""
    bars = energy_vjp(*res, et_bar=float(et_bar), precision=precision)
    return tuple(b.cpu().numpy() for b in bars)


def _make_ccsd_t_energy():
    """_ccsd_t_energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv,
    max_memory=None, precision='f64'): a jax.custom_vjp whose non-differentiable
    arguments are max_memory (as in lno/ccsd_t.py:70) and precision, or, without
    JAX, the energy as a float."""
    try:
        from functools import partial
        from jax import custom_vjp
    except ImportError:
        def _ccsd_t_energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, max_memory=None,
                           precision=DEFAULT_PRECISION):
            return _ccsd_t_energy_fwd(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, max_memory, precision)[0]
        return _ccsd_t_energy

    @partial(custom_vjp, nondiff_argnums=(8, 9))
    def _ccsd_t_energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, max_memory=None,
                       precision=DEFAULT_PRECISION):
        return _ccsd_t_energy_fwd(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, max_memory, precision)[0]

    def _fwd(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, max_memory, precision):
        return _ccsd_t_energy_fwd(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, max_memory, precision)

    def _bwd(max_memory, precision, res, et_bar):
        return _ccsd_t_energy_bwd(max_memory, res, et_bar, precision)

    _ccsd_t_energy.defvjp(_fwd, _bwd)
    return _ccsd_t_energy


_ccsd_t_energy = _make_ccsd_t_energy()


def kernel(mycc, eris, ulo, t1=None, t2=None, verbose=None, precision=DEFAULT_PRECISION):
    """The LNO (T) energy of ``mycc``'s amplitudes with the fragment's local
    orbital coefficients ``ulo`` (lno/ccsd_t.py:35-60)."""
    if t1 is None:
        t1 = mycc.t1
    if t2 is None:
        t2 = mycc.t2
    ulo = numpy.asarray(ulo) if not isinstance(ulo, torch.Tensor) else ulo
    mat = ulo.T @ ulo
    nocc = numpy.shape(t1)[0]
    if mat.shape[0] != nocc:
        raise ValueError(f"ulo^T ulo is {tuple(mat.shape)}, expected ({nocc}, {nocc})")
    t1T = t1.T
    t2T = t2.transpose(2, 3, 1, 0) if not isinstance(t2, torch.Tensor) else t2.permute(2, 3, 1, 0)
    fvo = eris.fock[nocc:, :nocc]
    return _ccsd_t_energy(mat, t1T, t2T, eris.mo_energy, fvo, eris.ovoo, eris.ovov, eris.ovvv,
                          getattr(mycc, "max_memory", None), precision)


class LNOCCSDTEnergy(torch.autograd.Function):
    """The LNO (T) energy with the hand-written VJP, as a torch autograd
    function of (mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv)."""

    @staticmethod
    def forward(ctx, mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, precision=DEFAULT_PRECISION):
        ctx.save_for_backward(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv)
        ctx.precision = precision
        return _like(energy(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv, precision=precision), t1T)

    @staticmethod
    def backward(ctx, gbar):
        xs = ctx.saved_tensors
        bars = energy_vjp(*xs, et_bar=float(gbar), precision=ctx.precision)
        return tuple(_like(g, x) for g, x in zip(bars, xs)) + (None,)
