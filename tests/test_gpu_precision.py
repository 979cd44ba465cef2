"""FP32-MFMA, split-bf16x3 and split-bf16x6 paths against the FP64 oracle.

These paths contract in reduced precision (FP32 accumulation) while the
denominators, the E(T) sum and all cotangent accumulation stay FP64.
Tolerances written here: FP32 and bf16x6 |dE| <= 1e-6 |E| and max relative
cotangent error (||gpu - oracle||_inf / ||oracle||_inf) <= 1e-6 (the target bound for FP32 work); bf16x3
|dE| <= 1e-4 |E| and <= 1e-3 per cotangent.  The bench reports the errors
actually reached at the benchmark size.
"""
import numpy as np
import pytest

from oracle import ccsd_t_oracle as orc
from pyscfad_amd.synthetic import make_instance

pytestmark = pytest.mark.gpu
NAMES = ("t1", "t2", "fock", "mo_energy", "ovoo", "ovov", "ovvv")
TOL = {"f32": (1e-6, 1e-6), "bf16x3": (1e-4, 1e-3), "bf16x6": (1e-6, 1e-6)}


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16x6"])
@pytest.mark.parametrize("o,v,seed,fs", [(3, 7, 1, 0.02), (5, 40, 2, 0.0), (6, 33, 3, 0.03)])
def test_reduced_precision(gpu_lib, prec, o, v, seed, fs):
    from pyscfad_amd.cc import ccsd_t
    d = make_instance(o, v, seed, fvo_scale=fs)
    ref_e = orc.energy(**d)
    ref = orc.energy_vjp(**d, et_bar=1.0)
    e = float(ccsd_t.energy(d["t1"], d["t2"], d["mo_energy"], d["fock"], d["ovoo"], d["ovov"], d["ovvv"],
                            precision=prec))
    etol, ctol = TOL[prec]
    assert abs(e - ref_e) <= etol * abs(ref_e), (prec, e, ref_e)
    bars = ccsd_t.energy_vjp(d["t1"], d["t2"], d["mo_energy"], d["fock"], d["ovoo"], d["ovov"], d["ovvv"],
                             precision=prec)
    for name, b in zip(NAMES, bars):
        r = ref[name]
        rel = np.abs(b.cpu().numpy() - r).max() / max(np.abs(r).max(), 1e-300)
        assert rel <= ctol, (prec, name, rel)


def test_bf16x6_resolves_what_bf16x3_cannot():
    """The three-part split and its six products carry 24 significant bits:
    bf16x6 must reach FP32-level errors (<= 1e-6 here) and land at least 8x
    closer to the oracle than bf16x3 on every cotangent (a dropped middle part
    or product would not; both share the FP32 storage of the G and H cubes,
    which bounds how far apart they can be)."""
    from pyscfad_amd.cc import ccsd_t
    d = make_instance(6, 40, 7, fvo_scale=0.02)
    ref = orc.energy_vjp(**d, et_bar=1.0)
    args = [d[k] for k in ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")]
    err = {}
    for prec in ("bf16x3", "bf16x6"):
        bars = ccsd_t.energy_vjp(*args, precision=prec)
        err[prec] = [np.abs(b.cpu().numpy() - ref[n]).max() / np.abs(ref[n]).max() for n, b in zip(NAMES, bars)]
    for n, e3, e6 in zip(NAMES, err["bf16x3"], err["bf16x6"]):
        assert e6 <= 1e-6 and e6 * 8 <= e3, (n, e3, e6)
