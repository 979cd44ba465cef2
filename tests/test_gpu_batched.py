"""Batched (T): many instances of uneven size in one captured hipGraph.

Each instance of a batched call runs the same plan and kernels as a
single-instance call, so results must agree bit for bit; a replay of the
cached graph after the inputs change in place must give the single-call
results of the new inputs.
"""
import pytest
import torch

from pyscfad_amd.synthetic import make_instance

pytestmark = pytest.mark.gpu

KEYS = ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")
SHAPES = [(3, 17, 11), (5, 24, 12), (4, 40, 13), (2, 9, 14)]


def _problems(shapes, fvo_scale=0.02):
    out = []
    for o, v, seed in shapes:
        d = make_instance(o, v, seed, fvo_scale=fvo_scale, device="cuda")
        d["et_bar"] = 0.5 + 0.1 * seed
        out.append(d)
    return out


def _single(ccsd_t, d, precision, with_vjp):
    args = [d[k] for k in KEYS]
    if not with_vjp:
        return ccsd_t.energy(*args, precision=precision), None
    bars, e = ccsd_t.energy_vjp(*args, et_bar=d["et_bar"], precision=precision, return_energy=True)
    return e, bars


def _same(got, ref, with_vjp):
    for (ge, gb), (re_, rb) in zip(got, ref):
        assert torch.equal(ge, re_), (float(ge), float(re_))
        if with_vjp:
            for x, y in zip(gb, rb):
                assert torch.equal(x, y)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("with_vjp", [True, False])
def test_batched_matches_single(gpu_lib, precision, with_vjp):
    from pyscfad_amd.cc import ccsd_t
    probs = _problems(SHAPES)
    got = ccsd_t.batched(probs, precision=precision, with_vjp=with_vjp)
    ref = [_single(ccsd_t, d, precision, with_vjp) for d in probs]
    _same(got, ref, with_vjp)


def test_batched_graph_replay(gpu_lib):
    from pyscfad_amd.cc import ccsd_t
    probs = _problems(SHAPES)
    ws = torch.empty(ccsd_t.batched_workspace_bytes(probs, "f64"), dtype=torch.uint8, device="cuda")
    first = ccsd_t.batched(probs, precision="f64", workspace=ws)
    for rep in range(3):
        # new values in the same buffers; the replay must see them
        for d in probs:
            d["t1"].mul_(0.9)
            d["t2"].add_(1e-4)
            torch.cuda.synchronize()
        got = ccsd_t.batched(probs, precision="f64", workspace=ws, outputs=first)
        ref = [_single(ccsd_t, d, "f64", True) for d in probs]
        _same(got, ref, True)


def test_batched_many_graphs(gpu_lib, monkeypatch):
    """One triple per batch makes the call long enough (thousands of kernels)
    to be captured as several bounded graphs launched back to back."""
    from pyscfad_amd.cc import ccsd_t
    monkeypatch.setenv("PYSCFAD_AMD_BATCH", "1")
    probs = _problems([(12, 40, 21), (3, 17, 22)])
    got = ccsd_t.batched(probs, precision="f32")
    ref = [_single(ccsd_t, d, "f32", True) for d in probs]
    _same(got, ref, True)


def test_autotune_keeps_results(gpu_lib):
    """The autotuner only changes tile shapes: results stay within the FP32
    reduction-order tolerance of the default configuration, and a tuned
    single call and a tuned batched call agree bit for bit."""
    from pyscfad_amd.cc import ccsd_t
    probs = _problems([(6, 72, 31)])
    d = probs[0]
    e0, b0 = _single(ccsd_t, d, "f32", True)
    chosen = ccsd_t.autotune(6, 72, "f32")
    assert all(v is not None for v in chosen.values()), chosen   # the backward fits here
    assert chosen["batch"] >= 1
    e1, b1 = _single(ccsd_t, d, "f32", True)
    assert abs(float(e1) - float(e0)) <= 1e-6 * abs(float(e0))
    for x, y in zip(b0, b1):
        assert float((x - y).abs().max()) <= 1e-5 * float(y.abs().max())
    got = ccsd_t.batched(probs, precision="f32")
    _same(got, [(e1, b1)], True)
