"""Size-independent properties of the GPU path at sizes the CPU oracle would
take too long for: Euler homogeneity of E(T) in the amplitudes, in the
integrals and in the orbital energies, shift invariance, determinism of
repeated calls, and agreement of the FP32 path with the FP64 path."""
import numpy as np
import pytest
import torch

from pyscfad_amd.synthetic import make_instance

pytestmark = pytest.mark.gpu


def _run(d, prec="f64"):
    from pyscfad_amd.cc import ccsd_t
    args = [d[k] for k in ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")]
    bars, e = ccsd_t.energy_vjp(*args, et_bar=1.0, precision=prec, return_energy=True)
    return float(e), [b.cpu().numpy() for b in bars]


@pytest.mark.parametrize("o,v", [(12, 96), (16, 130)])
def test_homogeneity_large(gpu_lib, o, v):
    d = make_instance(o, v, 9, fvo_scale=0.01)
    e, (t1b, t2b, fkb, meb, oob, ovb, vvb) = _run(d)
    tol = 1e-10 * abs(e)
    assert abs(np.vdot(d["t1"], t1b) + np.vdot(d["t2"], t2b) - 2 * e) < tol
    fock_vo = np.zeros_like(d["fock"])
    fock_vo[o:, :o] = d["fock"][o:, :o]
    s = np.vdot(d["ovvv"], vvb) + np.vdot(d["ovoo"], oob) + np.vdot(d["ovov"], ovb) + np.vdot(fock_vo, fkb)
    assert abs(s - 2 * e) < tol
    assert abs(np.vdot(d["mo_energy"], meb) + e) < tol
    assert abs(meb.sum()) < 1e-10 * np.abs(meb).max()


def test_deterministic(gpu_lib):
    d = make_instance(10, 70, 5, fvo_scale=0.02)
    e1, b1 = _run(d)
    e2, b2 = _run(d)
    assert e1 == e2
    for x, y in zip(b1, b2):
        assert np.array_equal(x, y)


def test_f32_agrees_with_f64(gpu_lib):
    d = make_instance(12, 100, 6)
    e64, b64 = _run(d, "f64")
    e32, b32 = _run(d, "f32")
    assert abs(e32 - e64) <= 1e-6 * abs(e64)
    for x, y in zip(b32, b64):
        assert np.abs(x - y).max() <= 1e-6 * np.abs(y).max()


def test_triple_slices_partition(gpu_lib):
    """The shares of a partition of the unique-triple list sum to the full
    E(T) and cotangents (the unit of multi-GPU sharding)."""
    from pyscfad_amd.cc import ccsd_t
    d = make_instance(7, 45, 12, fvo_scale=0.02)
    args = [d[k] for k in ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")]
    n = ccsd_t.num_triples(7)
    assert n == 7 * 8 * 9 // 6
    full, e_full = ccsd_t.energy_vjp(*args, return_energy=True)
    cuts = [0, 5, 37, n]
    acc = None
    e_sum = 0.0
    for a, b in zip(cuts[:-1], cuts[1:]):
        part, e = ccsd_t.energy_vjp(*args, return_energy=True, triples=(a, b))
        e_sum += float(e)
        assert abs(float(ccsd_t.energy(*args, triples=(a, b))) - float(e)) <= 1e-12 * abs(float(e_full))
        acc = [p.clone() for p in part] if acc is None else [x + p for x, p in zip(acc, part)]
    assert abs(e_sum - float(e_full)) <= 1e-12 * abs(float(e_full))
    for x, y in zip(acc, full):
        assert torch.allclose(x, y, rtol=1e-11, atol=1e-14 * float(torch.max(torch.abs(y))))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_ovvv_adjoint_grouping_changes_only_summation_order(gpu_lib, monkeypatch, prec, tol):
    """The ovvv adjoint of several consecutive batches runs as one launch that
    merges the triples sharing an output slab (PYSCFAD_AMD_B1_GROUP sets how
    many): every grouping gives the same E(T) and cotangents up to summation
    order."""
    d = make_instance(12, 100, 21, fvo_scale=0.02)
    monkeypatch.setenv("PYSCFAD_AMD_BATCH", "2")
    res = {}
    for g in ("1", "4", "8"):
        monkeypatch.setenv("PYSCFAD_AMD_B1_GROUP", g)
        res[g] = _run(d, prec)
    e1, b1 = res["1"]
    for g in ("4", "8"):
        e, b = res[g]
        assert abs(e - e1) <= 1e-12 * abs(e1), (g, e, e1)
        for x, y in zip(b, b1):
            assert np.abs(x - y).max() <= tol * np.abs(y).max(), (g, prec)
