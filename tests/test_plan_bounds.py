"""Host-side bounds check of the launch plans (no GPU).

tests/plan_check.cpp builds the plan of a backward call with the library's
own planning code and checks every GEMM and split-K reduction access against
the workspace and the caller's in-place outputs (see its header), also for
the LNO form's plan.  It is
compiled here with hipcc against the current sources and the kernel objects
of the last build.  Covers the dense and the density-fitted plans, every
precision, uneven sizes, batch merging (many triples per output slab),
small occupied blocks (cache evictions) and the workspace layout of the
batched call (region parts per stream group, descriptor blobs after them)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "pyscfad_amd", "_build")
OBJS = [os.path.join(BUILD, f) for f in ("t3_gemm.hip.o", "t3_epilogue.hip.o", "t3_prep.hip.o", "t3_noalloc.hip.o",
                                         "t3_lno.hip.o")]
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC) or not all(os.path.exists(o) for o in OBJS):
        pytest.skip("needs hipcc and a built library (python -m pyscfad_amd.build)")
    d = tmp_path_factory.mktemp("plan_check")
    obj, exe = str(d / "plan_check.o"), str(d / "plan_check")
    drv = os.path.join(ROOT, "pyscfad_amd", "csrc", "t3_driver.cpp")
    cc = [HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(ROOT, "pyscfad_amd", "csrc"), f'-DDRIVER="{drv}"', "-c",
          os.path.join(ROOT, "tests", "plan_check.cpp"), "-o", obj]
    subprocess.run(cc, check=True, capture_output=True)
    subprocess.run([HIPCC, "--offload-arch=gfx950", obj] + OBJS + ["-o", exe], check=True, capture_output=True)
    yield exe
    shutil.rmtree(d, ignore_errors=True)


def _run(exe, *args, block=None, tile=None, batch=None, b1_group=None):
    env = dict(os.environ)
    env.pop("PYSCFAD_AMD_BATCH", None)
    env.pop("PYSCFAD_AMD_TILE", None)
    env.pop("PYSCFAD_AMD_B1_GROUP", None)
    if batch is not None:
        env["PYSCFAD_AMD_BATCH"] = str(batch)
    if b1_group is not None:
        env["PYSCFAD_AMD_B1_GROUP"] = str(b1_group)
    if block is not None:
        env["PYSCFAD_AMD_DF_BLOCK"] = str(block)
    if tile is not None:
        env["PYSCFAD_AMD_TILE"] = str(tile)
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, env=env, timeout=120)


CASES = [
    # nocc, nvir, precision (0 f64, 1 f32, 2 bf16x3), naux, DF block
    (12, 100, 1, 0, None), (12, 100, 2, 0, None), (12, 100, 0, 0, None),
    (16, 130, 1, 0, None), (6, 72, 1, 0, None), (2, 3, 1, 0, None), (9, 1, 0, 0, None), (1, 9, 1, 0, None),
    (32, 256, 1, 0, None), (32, 256, 0, 0, None), (64, 512, 1, 0, None),
    (7, 33, 1, 120, 2), (6, 24, 0, 90, 1), (9, 64, 1, 216, None), (80, 640, 1, 2160, 4),
    (80, 640, 1, 32, 1),   # the autotuner's density-fitted trial plan where the dense one does not fit
]


@pytest.mark.parametrize("o,v,prec,naux,block", CASES)
def test_plan_in_bounds(checker, o, v, prec, naux, block):
    r = _run(checker, o, v, prec, naux, block=block)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout and "plan_error=''" in r.stdout, r.stdout


def test_plan_interleave_and_split_counts(checker):
    """A two-triple batch plan interleaves its hole-side adjoints and gives
    merged problems (twice the K) their own split count."""
    r = _run(checker, 12, 100, 1, 0, batch=2)
    assert r.returncode == 0 and "bad=0" in r.stdout, r.stdout + r.stderr
    m = re.search(r"interleaved=(\d+), split problems=(\d+) \((\d+) with", r.stdout)
    assert m, r.stdout
    assert int(m.group(1)) > 0 and int(m.group(2)) > 0 and int(m.group(3)) > 0, r.stdout


@pytest.mark.parametrize("o,v,prec", [(5, 19, 0), (6, 40, 1), (12, 100, 0), (12, 100, 1), (7, 33, 2), (9, 64, 3),
                                      (32, 256, 1), (2, 3, 0)])
def test_lno_plan_in_bounds(checker, o, v, prec):
    """The LNO form's plan: the S build from the projected operands and the
    adjoint families' extra terms and outputs."""
    r = _run(checker, o, v, prec, 0, "lno")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout and "plan_error=''" in r.stdout, r.stdout


def test_checker_reports_planted_overflow(checker):
    r = _run(checker, 12, 100, 1, 0, "plant")
    assert r.returncode != 0 and "OOB" in r.stdout, r.stdout


@pytest.mark.parametrize("tile", [0, 1, 3, 4, 5, 6])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_plan_in_bounds_forced_tile(checker, prec, tile):
    """Every contraction on one forced tile shape (PYSCFAD_AMD_TILE), as the
    per-tile GPU test runs them (FP64 ignores the 128x128 shape it lacks)."""
    r = _run(checker, 6, 130, prec, 0, tile=tile)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout and "plan_error=''" in r.stdout, r.stdout
    if not (prec == 0 and tile == 1):
        assert f"cfgs={tile},\n" in r.stdout, r.stdout   # every GEMM launch on the forced tile


@pytest.mark.parametrize("o,v,prec,batch,group", [(12, 100, 1, 2, 8), (12, 100, 0, 3, 5), (7, 33, 3, 4, 8),
                                                  (64, 512, 1, 2, 4)])
def test_plan_in_bounds_ovvv_adjoint_groups(checker, o, v, prec, batch, group):
    """Several batches per ovvv-adjoint launch: the G cubes of the whole
    group stay resident, and the merged problems' accesses stay in bounds."""
    r = _run(checker, o, v, prec, 0, batch=batch, b1_group=group)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout and "plan_error=''" in r.stdout, r.stdout


def _batched_args():
    from pyscfad_amd.synthetic import batched_shapes
    return [x for o, v, _ in batched_shapes(64, 4) for x in (o, v)]


@pytest.mark.parametrize("prec", [0, 1, 3])
def test_batched_plan_in_bounds(checker, prec):
    """The workspace layout of a batched backward call over the 64 uneven
    instances (nocc 4-24, nvir 40-200, seed 4): every instance's plan stays in
    the region part of its stream group and in its own outputs, and each
    descriptor blob lies after the parts, where its launches read it, apart
    from the others and inside the size the entry point reports."""
    r = _run(checker, "batched", prec, *_batched_args())
    assert r.returncode == 0, r.stdout + r.stderr
    assert "groups=4" in r.stdout and "bad=0" in r.stdout, r.stdout


@pytest.mark.parametrize("batch", [None, 3, 64])
def test_batched_plan_in_bounds_small_instances(checker, batch):
    """Tiny instances (as the GPU batched tests use) with the default, a
    small and the largest batch of triples."""
    r = _run(checker, "batched", 1, 3, 17, 5, 24, 4, 40, 2, 9, 1, 5, batch=batch)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout, r.stdout


def test_batched_checker_reports_overlapping_parts(checker):
    """Region parts planted 256 bytes too small: the largest instances no
    longer fit their part, and the checker must say so."""
    r = _run(checker, "batched_plant", 1, *_batched_args())
    assert r.returncode != 0 and "exceed the part" in r.stdout, r.stdout


@pytest.mark.parametrize("o,v,prec,naux,batch,expect", [
    (64, 512, 1, 0, 2, 4),      # FP32, tuned two triples per batch: four batches
    (64, 512, 0, 0, 2, 2),      # FP64: two, which keeps the whole solve clear of 288 GB
    (64, 512, 0, 0, 1, 4),
    (80, 640, 1, 2160, 1, 4),   # density-fitted o=80/v=640
    (32, 256, 1, 0, None, None),
    (12, 100, 0, 0, None, None),
])
def test_ovvv_adjoint_group_memory_bound(checker, o, v, prec, naux, batch, expect):
    """The G cubes that one ovvv-adjoint group keeps resident stay within the
    cache budget (512 MB) or, for batches whose cubes alone exceed it, within
    about 4.3 GB and four batches: the bound that keeps an FP64 o=64/v=512
    backward below the device memory."""
    r = _run(checker, o, v, prec, naux, batch=batch)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"B=(\d+) S=(\d+) cube=(\d+)", r.stdout)
    assert m, r.stdout
    B, S, cube = (int(x) for x in m.groups())
    held = S * B * cube
    assert 1 <= S <= 4, r.stdout
    if cube * B > 512e6:
        assert held <= 4.3e9, r.stdout
    else:
        assert S == 1 or held <= 512e6, r.stdout
    if expect is not None:
        assert S == expect, r.stdout
