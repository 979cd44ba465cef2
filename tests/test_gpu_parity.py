"""GPU parity of the FP64 path against the CPU oracle (same seeded inputs).

Bar (FP64 drop-in): |dE(T)| <= 1e-10 Eh and, for every cotangent,
max|gpu - oracle| <= 1e-9 * max|oracle|.
"""
import numpy as np
import pytest

from oracle import ccsd_t_oracle as orc
from pyscfad_amd.synthetic import make_instance

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
REL_TOL = 1e-9
ABS_FLOOR = 1e-17   # cotangents that vanish exactly (e.g. nocc = 1) are compared absolutely
NAMES = ("t1", "t2", "fock", "mo_energy", "ovoo", "ovov", "ovvv")

CASES = [
    # (nocc, nvir, seed, fvo_scale, packed)
    (2, 3, 1, 0.0, True),
    (3, 5, 2, 0.05, True),
    (5, 19, 3, 0.02, True),
    (4, 40, 4, 0.0, True),
    (7, 33, 5, 0.03, False),
    (1, 9, 6, 0.0, True),
    (9, 1, 7, 0.05, True),
]


def _inst(o, v, seed, fs, packed):
    d = make_instance(o, v, seed, fvo_scale=fs, packed=packed)
    return d


@pytest.mark.parametrize("o,v,seed,fs,packed", CASES)
def test_energy_f64(gpu_lib, o, v, seed, fs, packed):
    from pyscfad_amd.cc import ccsd_t
    d = _inst(o, v, seed, fs, packed)
    ref = orc.energy(**d)
    got = float(ccsd_t.energy(d["t1"], d["t2"], d["mo_energy"], d["fock"], d["ovoo"], d["ovov"], d["ovvv"]))
    assert abs(got - ref) <= E_TOL, (got, ref)


@pytest.mark.parametrize("o,v,seed,fs,packed", CASES)
def test_vjp_f64(gpu_lib, o, v, seed, fs, packed):
    from pyscfad_amd.cc import ccsd_t
    d = _inst(o, v, seed, fs, packed)
    # an unsymmetric t2 pins the element-wise cotangent convention
    rng = np.random.default_rng(seed + 100)
    d["t2"] = d["t2"] + 1e-3 * rng.normal(size=d["t2"].shape)
    ref = orc.energy_vjp(**d, et_bar=0.7)
    bars, e = ccsd_t.energy_vjp(d["t1"], d["t2"], d["mo_energy"], d["fock"], d["ovoo"], d["ovov"],
                                d["ovvv"], et_bar=0.7, return_energy=True)
    assert abs(float(e) - orc.energy(**d)) <= E_TOL
    for name, b in zip(NAMES, bars):
        r = ref[name]
        g = b.cpu().numpy()
        assert g.shape == r.shape, name
        scale = max(np.abs(r).max(), 1e-300)
        err = np.abs(g - r).max()
        assert err <= max(REL_TOL * scale, ABS_FLOOR), (name, err, scale)


@pytest.mark.parametrize("with_ws", [False, True])
def test_xla_legacy_handlers(gpu_lib, with_ws):
    """The XLA custom-call entry points give the results of the direct ABI:
    et_bar arrives as a device scalar operand (a traced cotangent), and with
    a caller workspace result the call allocates nothing."""
    import ctypes
    import torch
    from pyscfad_amd import _lib
    from pyscfad_amd.cc import ccsd_t
    o, v = 4, 21
    d = _inst(o, v, 8, 0.02, True)
    keys = ("mo_energy", "t1", "t2", "fock", "ovoo", "ovov", "ovvv")
    ins = [torch.as_tensor(d[k]).cuda().contiguous() for k in keys]
    stream = torch.cuda.current_stream().cuda_stream
    nws = ccsd_t.workspace_bytes(o, v, "f64", with_vjp=True) if with_ws else 0
    ws = [torch.empty(nws, dtype=torch.uint8, device="cuda")] if with_ws else []
    op = _lib.XlaOpaque(o, v, _lib.PREC_F64, 1, nws)
    e_out = torch.zeros((), dtype=torch.float64, device="cuda")
    bufs = (ctypes.c_void_p * (8 + len(ws)))(*[t.data_ptr() for t in ins + [e_out] + ws])
    gpu_lib.pyscfad_amd_xla_ccsd_t_energy(stream, bufs, bytes(op), ctypes.sizeof(op))
    torch.cuda.synchronize()
    assert abs(float(e_out) - orc.energy(**d)) <= E_TOL
    args = [d[k] for k in ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")]
    for et_bar in (0.5, 0.3):
        outs = [torch.empty_like(ins[i]) for i in (1, 2, 3, 0, 4, 5, 6)]   # t1,t2,fock,mo_energy,ovoo,ovov,ovvv
        eb = torch.tensor(et_bar, dtype=torch.float64, device="cuda")
        bufs = (ctypes.c_void_p * (15 + len(ws)))(*[t.data_ptr() for t in ins + [eb] + outs + ws])
        gpu_lib.pyscfad_amd_xla_ccsd_t_energy_vjp(stream, bufs, bytes(op), ctypes.sizeof(op))
        torch.cuda.synchronize()
        ref = ccsd_t.energy_vjp(*args, et_bar=et_bar)
        for a, b in zip(outs, ref):
            if et_bar == 0.5:
                # a power-of-two factor scales every step exactly
                assert torch.equal(a, b)
            else:
                assert torch.allclose(a, b, rtol=1e-13, atol=1e-15 * float(torch.abs(b).max()))
