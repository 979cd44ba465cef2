"""The drop-in boundary under the reference's own names
(pyscfad/cc/ccsd_t.py:28-80, 82-192): kernel(mycc, eris), the numpy 7-tuple
of _ccsd_t_energy_vjp, the torch autograd Functions on host and device
inputs, calls on a side stream, and the batched graph cache.

Bar: the FP64 drop-in bar of the parity tests, |dE| <= 1e-10 Eh and
max|got - oracle| <= 1e-9 max|oracle| per cotangent.
"""
import ctypes
import types

import numpy as np
import pytest
import torch

from oracle import ccsd_t_oracle as orc
from oracle import df as odf
from pyscfad_amd.synthetic import make_instance

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
REL_TOL = 1e-9
KEYS = ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")
NAMES = ("t1", "t2", "fock", "mo_energy", "ovoo", "ovov", "ovvv")


def _inst(o=4, v=23, seed=31, fs=0.03):
    d = make_instance(o, v, seed, fvo_scale=fs)
    rng = np.random.default_rng(seed)
    d["t2"] = d["t2"] + 1e-3 * rng.normal(size=d["t2"].shape)
    return d


def _close(got, ref, name):
    got = np.asarray(got)
    assert got.shape == ref.shape, name
    assert np.abs(got - ref).max() <= REL_TOL * max(np.abs(ref).max(), 1e-300), name


def _eris(d, **others):
    from pyscfad_amd.cc import ccsd_t
    return ccsd_t.ERIs(d["fock"], d["mo_energy"], d["ovoo"], d["ovov"], d["ovvv"], **others)


def test_kernel_reference_signature(gpu_lib):
    """kernel(mycc, eris) reads t1/t2 from mycc when they are not passed."""
    from pyscfad_amd.cc import ccsd_t
    d = _inst()
    mycc = types.SimpleNamespace(t1=d["t1"], t2=d["t2"])
    eris = _eris(d, oooo=np.zeros((4, 4, 4, 4)))
    ref = orc.energy(**d)
    e = ccsd_t.kernel(mycc, eris)
    assert isinstance(e, float)
    assert abs(e - ref) <= E_TOL
    assert abs(ccsd_t.kernel(None, eris, t1=d["t1"], t2=d["t2"], verbose=4) - ref) <= E_TOL


def test_ccsd_t_energy_vjp_reference_tuple(gpu_lib):
    """_ccsd_t_energy_vjp(eris, t1, t2, et_bar, max_memory) returns the
    reference's 7-tuple of NumPy arrays, ovvv_bar lower-triangle packed."""
    from pyscfad_amd.cc import ccsd_t
    d = _inst()
    ref = orc.energy_vjp(**d, et_bar=0.35)
    out = ccsd_t._ccsd_t_energy_vjp(_eris(d), d["t1"], d["t2"], 0.35, 2000)
    assert isinstance(out, tuple) and len(out) == 7
    for name, got in zip(NAMES, out):
        assert isinstance(got, np.ndarray) and got.dtype == np.float64, name
        _close(got, ref[name], name)
    v = d["t1"].shape[1]
    assert out[6].shape == (4, v, v * (v + 1) // 2)


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_autograd_function(gpu_lib, device):
    """CCSDTEnergy under torch.autograd.grad, with host or device inputs:
    the energy and every gradient come back on the inputs' device."""
    from pyscfad_amd.cc import ccsd_t
    d = _inst()
    ref = orc.energy_vjp(**d, et_bar=1.0)
    xs = [torch.tensor(d[k], device=device, requires_grad=True) for k in KEYS]
    e = ccsd_t.CCSDTEnergy.apply(*xs)
    assert e.device.type == device
    assert abs(float(e) - orc.energy(**d)) <= E_TOL
    # a downstream factor: the cotangent reaching the Function is 2.5
    grads = torch.autograd.grad(2.5 * e, xs)
    for k, g, x in zip(KEYS, grads, xs):
        assert g.device == x.device and g.dtype == x.dtype, k
        _close(g.detach().cpu().numpy(), 2.5 * ref[k], k)


def test_autograd_function_df(gpu_lib):
    """CCSDTEnergyDF: the Lov/Lvv gradients are the chain rule of the
    oracle's ovvv cotangent through ovvv = Lov^T Lvv (pyscfad/cc/dfccsd.py:91)."""
    from pyscfad_amd.cc import ccsd_t
    o, v = 4, 21
    d = make_instance(o, v, 33, fvo_scale=0.02, df=True)
    Lov, Lvv = d.pop("Lov"), d.pop("Lvv")
    g = orc.energy_vjp(**d, ovvv=odf.ovvv_from_df(Lov, Lvv), et_bar=1.0)
    lov_bar, lvv_bar = odf.df_vjp(Lov, Lvv, g["ovvv"])
    keys = ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov")
    xs = [torch.tensor(d[k], device="cuda", requires_grad=True) for k in keys]
    xs += [torch.tensor(Lov, device="cuda", requires_grad=True), torch.tensor(Lvv, device="cuda", requires_grad=True)]
    e = ccsd_t.CCSDTEnergyDF.apply(*xs)
    assert abs(float(e) - orc.energy(**d, ovvv=odf.ovvv_from_df(Lov, Lvv))) <= E_TOL
    e.backward()
    for k, x in zip(keys, xs):
        _close(x.grad.cpu().numpy(), g[k], k)
    _close(xs[6].grad.cpu().numpy(), lov_bar, "Lov")
    _close(xs[7].grad.cpu().numpy(), lvv_bar, "Lvv")


def test_side_stream_matches_current_stream(gpu_lib):
    """A call on another stream waits for the caller's pending work (here the
    in-place change of t1 just enqueued on the current stream) and its
    results are usable on the current stream right after it returns."""
    from pyscfad_amd.cc import ccsd_t
    d = _inst(5, 31, 41)
    a = [torch.as_tensor(d[k]).cuda() for k in KEYS]
    ref_bars, ref_e = ccsd_t.energy_vjp(*[x.clone() for x in a], et_bar=0.7, return_energy=True)
    side = torch.cuda.Stream()
    for rep in range(3):
        t1 = a[0].clone()
        t1.mul_(1.0)   # pending work on the current stream that the side-stream call must see
        bars, e = ccsd_t.energy_vjp(t1, *a[1:], et_bar=0.7, return_energy=True, stream=side)
        e2 = ccsd_t.energy(t1, *a[1:], stream=side)
        del t1    # the caching allocator must keep it until the side stream is done
        assert torch.equal(e, ref_e) and torch.equal(e2, ref_e)
        for x, y in zip(bars, ref_bars):
            assert torch.equal(x, y)
    torch.cuda.synchronize()


def test_batched_graph_cache_ignores_padding(gpu_lib):
    """Two instance arrays with the same field values but different bytes in
    the struct's padding select the same cached graph (a replay, no new
    capture)."""
    from pyscfad_amd import _lib
    from pyscfad_amd.cc import ccsd_t
    probs = []
    for o, v, seed in ((3, 17, 51), (4, 24, 52)):
        p = make_instance(o, v, seed, fvo_scale=0.02, device="cuda")
        p["et_bar"] = 1.0
        probs.append(p)
    arr, keep = ccsd_t._instances(probs, "f64", True)
    need = gpu_lib.pyscfad_amd_ccsd_t_batched_workspace_bytes(len(probs), arr, _lib.PREC_F64, 1)
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    pad_off = _lib.Instance.nocc.offset + 3 * ctypes.sizeof(ctypes.c_int)
    pad_len = _lib.Instance.et_bar.offset - pad_off
    assert pad_len > 0
    counts = []
    for junk in (b"\x00", b"\xa5", b"\x3c"):
        fresh = (_lib.Instance * len(probs))()
        ctypes.memmove(fresh, arr, ctypes.sizeof(arr))
        for x in range(len(probs)):
            ctypes.memmove(ctypes.addressof(fresh[x]) + pad_off, junk * pad_len, pad_len)
        _lib.check(gpu_lib.pyscfad_amd_ccsd_t_batched(len(probs), fresh, _lib.PREC_F64, 1,
                                                      ws.data_ptr(), need, stream))
        counts.append(_lib.graph_cache_stats())
    torch.cuda.synchronize()
    (r0, c0), (r1, c1), (r2, c2) = counts
    assert (c1, c2) == (c0, c0), counts
    assert (r1, r2) == (r0 + 1, r0 + 2), counts
