"""Every GEMM tile configuration at every precision against the FP64 path.

The grouped GEMM has six tile shapes (64x64, 128x128, 64x32, 32x64, 128x64,
64x128; no 128x128 in FP64), each with its own LDS layout and unit
permutations per precision.  PYSCFAD_AMD_TILE forces one shape on every
contraction of a call.  The instance (nocc=6, nvir=130: padded V = 160, so
the 128-row tiles run both full and clamped) is small enough that the
default-tile FP64 result, itself checked against the oracle at smaller
sizes, serves as the reference.  Tolerances: FP64 1e-10 relative (energy and
every cotangent, reduction order only); FP32 and bf16x6 1e-6 on the energy
and 1e-5 per cotangent; bf16x3 1e-4 and 1e-3.
"""
import pytest

from pyscfad_amd.synthetic import make_instance

pytestmark = pytest.mark.gpu
TILES = {0: "64x64", 1: "128x128", 3: "64x32", 4: "32x64", 5: "128x64", 6: "64x128"}
TOL = {"f64": (1e-10, 1e-10), "f32": (1e-6, 1e-5), "bf16x3": (1e-4, 1e-3), "bf16x6": (1e-6, 1e-5)}


@pytest.fixture(scope="module")
def reference(gpu_lib):
    from pyscfad_amd.cc import ccsd_t
    d = make_instance(6, 130, 11, fvo_scale=0.02)
    args = [d[k] for k in ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")]
    e = float(ccsd_t.energy(*args, precision="f64"))
    bars = [b.clone() for b in ccsd_t.energy_vjp(*args, precision="f64")]
    return args, e, bars


@pytest.mark.parametrize("cfg", sorted(TILES))
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3", "bf16x6"])
def test_tile_configuration(reference, monkeypatch, prec, cfg):
    if prec == "f64" and cfg == 1:
        pytest.skip("no 128x128 FP64 tile (it would spill)")
    from pyscfad_amd.cc import ccsd_t
    args, e_ref, bars_ref = reference
    monkeypatch.setenv("PYSCFAD_AMD_TILE", str(cfg))
    etol, ctol = TOL[prec]
    e = float(ccsd_t.energy(*args, precision=prec))
    assert abs(e - e_ref) <= etol * abs(e_ref), (prec, TILES[cfg], e, e_ref)
    bars = ccsd_t.energy_vjp(*args, precision=prec)
    for i, (b, r) in enumerate(zip(bars, bars_ref)):
        rel = float((b - r).abs().max()) / max(float(r.abs().max()), 1e-300)
        assert rel <= ctol, (prec, TILES[cfg], i, rel)
