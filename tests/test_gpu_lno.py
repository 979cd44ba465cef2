"""GPU parity of the LNO (T) engine (pyscfad_amd.lno.ccsd_t, the
pyscfad_amd_lno_ccsd_t_* entry points) against the LNO oracle
(oracle/ccsd_t_oracle.c, pinned in tests/test_oracle_lno.py).

Bars:
  * FP64 against the oracle: |dE| <= 1e-10 Eh and, per cotangent,
    max|gpu - oracle| <= 1e-9 max|oracle| (the canonical drop-in's bar), on a
    non-symmetric random mat and a Gram mat, packed and dense ovvv, t2 with
    and without pair symmetry;
  * mat = I: the LNO energy and cotangents equal the canonical GPU path's
    (both FP64) to 1e-11 relative;
  * FP32 against FP64 at molecular scale (integral_scale = 1/sqrt(n)):
    max|d|/max|ref| <= 1e-6 per cotangent, |dE| <= 1e-6 Eh;
  * the reference-named wrappers: kernel(mycc, eris, ulo), the custom_vjp's
    backward rule and the torch autograd function on host tensors.
"""
import numpy as np
import pytest
import torch

from oracle import ccsd_t_oracle as orc
from pyscfad_amd.synthetic import make_lno_instance

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
REL_TOL = 1e-9
KEYS = ("mat", "t1T", "t2T", "mo_energy", "fvo", "ovoo", "ovov", "ovvv")
INPUTS = KEYS


def _inputs(x):
    return {k: x[k] for k in INPUTS}


def _rel(g, r):
    g = g.cpu().numpy() if isinstance(g, torch.Tensor) else np.asarray(g)
    return float(np.abs(g - r).max() / max(np.abs(r).max(), 1e-300))


CASES = [
    # nocc, nvir, seed, mat kind, packed, t2 noise
    (2, 3, 1, "random", True, 0.0),
    (5, 19, 2, "random", True, 1e-3),
    (4, 24, 3, "gram", False, 1e-3),
    (6, 33, 4, "gram", True, 0.0),
    (3, 40, 5, "random", False, 1e-3),
    (1, 9, 6, "random", True, 0.0),
]


def _case(o, v, seed, kind, packed, noise):
    x = make_lno_instance(o, v, seed, mat=kind, packed=packed)
    if noise:
        rng = np.random.default_rng(seed + 300)
        x["t2T"] = x["t2T"] + noise * rng.normal(size=x["t2T"].shape)
    return x


@pytest.mark.parametrize("o,v,seed,kind,packed,noise", CASES)
def test_lno_f64_against_oracle(gpu_lib, o, v, seed, kind, packed, noise):
    from pyscfad_amd.lno import ccsd_t as lno
    x = _case(o, v, seed, kind, packed, noise)
    ins = _inputs(x)
    e_ref = orc.lno_energy(**ins)
    assert abs(float(lno.energy(**ins)) - e_ref) <= E_TOL
    ref, e2 = orc.lno_energy_vjp(**ins, et_bar=0.7, return_energy=True)
    bars, e = lno.energy_vjp(**ins, et_bar=0.7, return_energy=True)
    assert abs(float(e) - e_ref) <= E_TOL
    for name, g, r in zip(KEYS, bars, ref):
        assert tuple(g.shape) == r.shape, name
        err = float(np.abs(g.cpu().numpy() - r).max())
        assert err <= max(REL_TOL * np.abs(r).max(), 1e-17), (name, err, np.abs(r).max())


@pytest.mark.parametrize("o,v,seed", [(5, 19, 7), (8, 64, 8)])
def test_identity_projector_equals_canonical_gpu(gpu_lib, o, v, seed):
    from pyscfad_amd.cc import ccsd_t
    from pyscfad_amd.lno import ccsd_t as lno
    x = make_lno_instance(o, v, seed, mat="identity")
    d = x["canonical"]
    cb, ce = ccsd_t.energy_vjp(d["t1"], d["t2"], d["mo_energy"], d["fock"], d["ovoo"], d["ovov"], d["ovvv"],
                               et_bar=1.3, return_energy=True)
    lb, le = lno.energy_vjp(**_inputs(x), et_bar=1.3, return_energy=True)
    assert abs(float(le) - float(ce)) <= 1e-11 * abs(float(ce))
    t1b, t2b, fkb, meb, oob, ovb, vvb = (b.cpu().numpy() for b in cb)
    ref = {"t1T": t1b.T, "t2T": t2b.transpose(2, 3, 1, 0), "mo_energy": meb, "fvo": fkb[o:, :o],
           "ovoo": oob, "ovov": ovb, "ovvv": vvb}
    got = dict(zip(KEYS, lb))
    for k, r in ref.items():
        assert _rel(got[k], r) <= 1e-11, k
    # E is linear in mat: <I, mat_bar> = et_bar E
    assert abs(float(torch.trace(got["mat"])) - 1.3 * float(ce)) <= 1e-11 * abs(float(ce))


def test_lno_larger_f64_against_oracle(gpu_lib):
    """o=8, v=80: several triple batches and merged adjoint problems."""
    from pyscfad_amd.lno import ccsd_t as lno
    x = make_lno_instance(8, 80, 9, mat="gram", integral_scale=1.0 / np.sqrt(88))
    ins = _inputs(x)
    with orc.blas():
        ref, e_ref = orc.lno_energy_vjp(**ins, et_bar=1.0, return_energy=True)
    bars, e = lno.energy_vjp(**ins, return_energy=True)
    assert abs(float(e) - e_ref) <= E_TOL
    for name, g, r in zip(KEYS, bars, ref):
        assert _rel(g, r) <= REL_TOL, name


@pytest.mark.parametrize("prec", ["f32", "bf16x6"])
def test_lno_reduced_precision_against_f64(gpu_lib, prec):
    from pyscfad_amd.lno import ccsd_t as lno
    o, v = 12, 100
    x = make_lno_instance(o, v, 10, mat="gram", integral_scale=1.0 / np.sqrt(o + v))
    ins = _inputs(x)
    b64, e64 = lno.energy_vjp(**ins, return_energy=True)
    b32, e32 = lno.energy_vjp(**ins, precision=prec, return_energy=True)
    assert abs(float(e32) - float(e64)) <= 1e-6
    for name, g, r in zip(KEYS, b32, b64):
        assert _rel(g, r.cpu().numpy()) <= 1e-6, name


def test_lno_kernel_bwd_rule_and_autograd(gpu_lib):
    """kernel(mycc, eris, ulo) forms mat = ulo^T ulo, t1T, t2T, fvo as
    lno/ccsd_t.py:35-60 does; the backward rule returns the reference's
    8-tuple of NumPy arrays; the autograd function works on host tensors."""
    from pyscfad_amd.cc.ccsd_t import ERIs
    from pyscfad_amd.lno import ccsd_t as lno
    o, v = 4, 21
    x = make_lno_instance(o, v, 11, mat="identity")
    d = x["canonical"]
    rng = np.random.default_rng(5)
    ulo = rng.normal(size=(3, o))
    mat = ulo.T @ ulo

    class CC:
        pass
    mycc = CC()
    mycc.t1, mycc.t2, mycc.max_memory = d["t1"], d["t2"], 4000
    eris = ERIs(d["fock"], d["mo_energy"], d["ovoo"], d["ovov"], d["ovvv"])
    ins = dict(_inputs(x), mat=mat)
    e_ref = orc.lno_energy(**ins)
    assert abs(float(lno.kernel(mycc, eris, ulo)) - e_ref) <= E_TOL
    et, res = lno._ccsd_t_energy_fwd(*[ins[k] for k in INPUTS], 4000)
    assert abs(et - e_ref) <= E_TOL
    bars = lno._ccsd_t_energy_bwd(4000, res, 0.5)
    ref = orc.lno_energy_vjp(**ins, et_bar=0.5)
    assert len(bars) == 8
    for name, g, r in zip(KEYS, bars, ref):
        assert isinstance(g, np.ndarray) and g.shape == r.shape, name
        assert _rel(g, r) <= REL_TOL, name
    ts = [torch.tensor(ins[k], dtype=torch.float64, requires_grad=True) for k in INPUTS]
    e = lno.LNOCCSDTEnergy.apply(*ts)
    grads = torch.autograd.grad(2.0 * e, ts)
    assert abs(float(e.detach()) - e_ref) <= E_TOL
    ref2 = orc.lno_energy_vjp(**ins, et_bar=2.0)
    for name, g, r in zip(KEYS, grads, ref2):
        assert g.device.type == "cpu", name
        assert _rel(g.numpy(), r) <= REL_TOL, name


def test_lno_rejects_bad_arguments(gpu_lib):
    from pyscfad_amd import _lib
    from pyscfad_amd.lno import ccsd_t as lno
    x = make_lno_instance(3, 5, 12)
    ins = _inputs(x)
    with pytest.raises(ValueError):
        lno.energy(**dict(ins, mat=np.eye(4)))
    assert lno.workspace_bytes(129, 10) == 0   # nocc above the LNO form's limit
    lib = _lib.load()
    assert lib.pyscfad_amd_lno_ccsd_t_energy(3, 5, None, None, None, None, None, None, None, None, 1, 0, None,
                                             None, 0, None) == 1
