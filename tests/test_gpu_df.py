"""Density-fitted mode on the GPU: the ovvv slabs are built on the device from
the DF factors Lov, Lvv (pyscfad/cc/dfccsd.py:81-91) and the ovvv cotangent
comes back as Lov_bar, Lvv_bar.

Bar (FP64): |dE(T)| <= 1e-10 Eh and every cotangent within 1e-9 of max|ref|,
against the CPU oracle on ovvv = Lov^T Lvv with the chain rule of
oracle/df.py.  Small occupied blocks (PYSCFAD_AMD_DF_BLOCK) make the slab
cache evict, rebuild and fold blocks.
"""
import numpy as np
import pytest

from oracle import ccsd_t_oracle as orc
from oracle import df as odf
from pyscfad_amd.synthetic import make_instance

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
REL_TOL = 1e-9
NAMES = ("t1", "t2", "fock", "mo_energy", "ovoo", "ovov", "Lov", "Lvv")
BASE = ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov")


def _args(d):
    return [d[k] for k in ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "Lov", "Lvv")]


def _oracle(d, et_bar):
    ovvv = odf.ovvv_from_df(d["Lov"], d["Lvv"])
    base = {k: d[k] for k in BASE}
    g = orc.energy_vjp(**base, ovvv=ovvv, et_bar=et_bar)
    out = {k: g[k] for k in NAMES[:6]}
    out["Lov"], out["Lvv"] = odf.df_vjp(d["Lov"], d["Lvv"], g["ovvv"])
    return out, orc.energy(**base, ovvv=ovvv)


def _close(name, got, ref, rel):
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max()
    assert got.shape == ref.shape, name
    assert err <= max(rel * scale, 1e-17), (name, err, scale)


# (nocc, nvir, seed, fvo_scale, block): blocks of 1 and 2 with 5-7 occupied
# indices give 4-7 blocks, more than the three cache slots
CASES = [(5, 19, 3, 0.02, "1"), (6, 24, 4, 0.0, "1"), (7, 33, 6, 0.03, "2"), (4, 40, 5, 0.0, None)]


@pytest.mark.parametrize("o,v,seed,fs,blk", CASES)
def test_df_f64_vs_oracle(gpu_lib, monkeypatch, o, v, seed, fs, blk):
    from pyscfad_amd.cc import ccsd_t
    if blk:
        monkeypatch.setenv("PYSCFAD_AMD_DF_BLOCK", blk)
    print(f"df f64 o={o} v={v} seed={seed} block={blk}", flush=True)
    d = make_instance(o, v, seed, fvo_scale=fs, df=True)
    rng = np.random.default_rng(seed + 100)
    d["t2"] = d["t2"] + 1e-3 * rng.normal(size=d["t2"].shape)   # element-wise t2 convention
    ref, e_ref = _oracle(d, 0.7)
    bars, e = ccsd_t.energy_vjp_df(*_args(d), et_bar=0.7, return_energy=True)
    assert abs(float(e) - e_ref) <= E_TOL
    assert abs(float(ccsd_t.energy_df(*_args(d))) - e_ref) <= E_TOL
    for name, b in zip(NAMES, bars):
        _close(name, b.cpu().numpy(), ref[name], REL_TOL)


def test_df_slices_partition(gpu_lib, monkeypatch):
    """Slices of the blocked triple order sum to the full energy and cotangents."""
    from pyscfad_amd.cc import ccsd_t
    monkeypatch.setenv("PYSCFAD_AMD_DF_BLOCK", "2")
    o, v = 6, 20
    d = make_instance(o, v, 9, fvo_scale=0.02, df=True)
    full, e_full = ccsd_t.energy_vjp_df(*_args(d), return_energy=True)
    n = ccsd_t.num_triples(o)
    cuts = [0, 5, 17, 40, n]
    acc = [np.zeros(b.shape) for b in full]
    e_sum = 0.0
    for a, b in zip(cuts[:-1], cuts[1:]):
        bars, e = ccsd_t.energy_vjp_df(*_args(d), triples=(a, b), return_energy=True)
        e_sum += float(e)
        assert abs(float(ccsd_t.energy_df(*_args(d), triples=(a, b))) - float(e)) <= 1e-13
        for x, t in zip(acc, bars):
            x += t.cpu().numpy()
    assert abs(e_sum - float(e_full)) <= 1e-12
    for name, x, t in zip(NAMES, acc, full):
        _close(name, x, t.cpu().numpy(), 1e-12)


@pytest.mark.parametrize("prec,rel", [("f64", 1e-12), ("f32", 1e-6)])
def test_df_matches_dense(gpu_lib, monkeypatch, prec, rel):
    """The DF path against the dense path on the same ovvv (default block
    size and blocks of 3): same energy and cotangents, and Lov_bar, Lvv_bar
    equal to the dense ovvv_bar carried through ovvv = Lov^T Lvv."""
    import torch
    from pyscfad_amd.cc import ccsd_t
    o, v = 9, 64
    d = make_instance(o, v, 12, fvo_scale=0.01, df=True)
    ovvv = odf.ovvv_from_df(d["Lov"], d["Lvv"])
    dense, e_dense = ccsd_t.energy_vjp(*[d[k] for k in BASE[:2]], d["mo_energy"], d["fock"], d["ovoo"],
                                       d["ovov"], ovvv, precision=prec, return_energy=True)
    dense = [b.cpu().numpy() for b in dense]
    lb, vb = odf.df_vjp(d["Lov"], d["Lvv"], dense[6])
    want = dense[:6] + [lb, vb]
    for blk in (None, "3"):
        if blk:
            monkeypatch.setenv("PYSCFAD_AMD_DF_BLOCK", blk)
        print(f"df vs dense {prec} block={blk}", flush=True)
        bars, e = ccsd_t.energy_vjp_df(*_args(d), precision=prec, return_energy=True)
        assert abs(float(e) - float(e_dense)) <= rel * abs(float(e_dense))
        for name, b, w in zip(NAMES, bars, want):
            _close(name, b.cpu().numpy(), w, rel)
    torch.cuda.synchronize()
