"""The C ABI library loads, exports every symbol include/pyscfad_amd.h declares,
sizes its workspace and rejects bad arguments -- all without a GPU (no call
here reaches the HIP runtime)."""
import ctypes
import os
import re

import pytest

from pyscfad_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pyscfad_amd.h")).read()
    return sorted(set(re.findall(r"\b(pyscfad_amd_\w+)\s*\(", text)))


def test_header_matches_binding():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_all_symbols():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.pyscfad_amd_abi_version() == 2


def test_workspace_bytes():
    lib = _lib.load()
    small = lib.pyscfad_amd_ccsd_t_workspace_bytes(5, 19, 0, 0)
    assert small > 0
    assert lib.pyscfad_amd_ccsd_t_workspace_bytes(5, 19, 0, 1) > small
    # FP32 operands need less than FP64
    assert lib.pyscfad_amd_ccsd_t_workspace_bytes(32, 256, 1, 1) < lib.pyscfad_amd_ccsd_t_workspace_bytes(32, 256, 0, 1)
    assert lib.pyscfad_amd_ccsd_t_workspace_bytes(0, 19, 0, 0) == 0
    assert lib.pyscfad_amd_ccsd_t_workspace_bytes(5, 19, 7, 0) == 0


@pytest.mark.parametrize("nocc,nvir,prec,code", [(0, 5, 0, 1), (3, -1, 0, 1), (3, 5, 9, 2)])
def test_bad_arguments(nocc, nvir, prec, code):
    lib = _lib.load()
    dummy = ctypes.c_void_p(16)
    rc = lib.pyscfad_amd_ccsd_t_energy(nocc, nvir, *([dummy] * 7), 1, prec, dummy, None, 0, None)
    assert rc == code
    assert lib.pyscfad_amd_last_error()


def test_null_pointers_rejected():
    lib = _lib.load()
    rc = lib.pyscfad_amd_ccsd_t_energy(3, 5, *([None] * 7), 1, 0, None, None, 0, None)
    assert rc == 1
    assert b"null" in lib.pyscfad_amd_last_error()


def test_workspace_too_small_rejected():
    lib = _lib.load()
    dummy = ctypes.c_void_p(256)
    rc = lib.pyscfad_amd_ccsd_t_energy(3, 5, *([dummy] * 7), 1, 0, dummy, dummy, 16, None)
    assert rc == 4
    assert b"workspace" in lib.pyscfad_amd_last_error()


def _fake_instances(shapes):
    """Instance array with non-null placeholder pointers: sizing and argument
    checks read only the shapes and never dereference a pointer."""
    arr = (_lib.Instance * len(shapes))()
    for x, (o, v) in enumerate(shapes):
        I = arr[x]
        I.nocc, I.nvir, I.ovvv_packed, I.et_bar = o, v, 1, 1.0
        for name, _ in _lib.Instance._fields_[4:]:
            setattr(I, name, 4096)
    return arr


def test_batched_workspace_bytes(monkeypatch):
    lib = _lib.load()
    shapes = [(4, 40), (9, 73), (6, 120), (5, 33), (7, 41)]
    arr = _fake_instances(shapes)
    for prec in (0, 1):
        for vjp in (0, 1):
            single = max(lib.pyscfad_amd_ccsd_t_workspace_bytes(o, v, prec, vjp) for o, v in shapes)
            total = sum(lib.pyscfad_amd_ccsd_t_workspace_bytes(o, v, prec, vjp) for o, v in shapes)
            # the instances of one stream group share a region part; each keeps its own descriptors
            monkeypatch.setenv("PYSCFAD_AMD_BATCH_STREAMS", "1")
            one = lib.pyscfad_amd_ccsd_t_batched_workspace_bytes(len(shapes), arr, prec, vjp)
            assert single <= one < total
            # four groups (the default): four region parts of the largest instance's size
            monkeypatch.delenv("PYSCFAD_AMD_BATCH_STREAMS")
            need = lib.pyscfad_amd_ccsd_t_batched_workspace_bytes(len(shapes), arr, prec, vjp)
            assert one < need <= 4 * single + one
    assert lib.pyscfad_amd_ccsd_t_batched_workspace_bytes(0, arr, 1, 1) == 0
    bad = _fake_instances([(4, 40), (0, 10)])
    assert lib.pyscfad_amd_ccsd_t_batched_workspace_bytes(2, bad, 1, 1) == 0


def test_batched_argument_errors():
    lib = _lib.load()
    arr = _fake_instances([(4, 40), (5, 19)])
    # no workspace: rejected before any device work
    assert lib.pyscfad_amd_ccsd_t_batched(2, arr, 1, 1, None, 0, None) == 1
    arr[1].ovvv = None
    assert lib.pyscfad_amd_ccsd_t_batched(2, arr, 1, 1, None, 0, None) == 1
    assert b"instance 1" in lib.pyscfad_amd_last_error()


def test_df_workspace_bytes(monkeypatch):
    lib = _lib.load()
    assert lib.pyscfad_amd_ccsd_t_df_workspace_bytes(5, 19, 72, 0, 0) > 0
    assert lib.pyscfad_amd_ccsd_t_df_workspace_bytes(5, 19, 0, 0, 0) == 0
    assert lib.pyscfad_amd_ccsd_t_df_workspace_bytes(0, 19, 72, 0, 0) == 0
    # with several occupied blocks only three blocks of slabs (and, backward,
    # their FP64 accumulators) are held, instead of all nocc
    monkeypatch.setenv("PYSCFAD_AMD_DF_BLOCK", "2")
    o, v, naux = 24, 96, 3 * 120
    for prec in (0, 1):
        dense = lib.pyscfad_amd_ccsd_t_workspace_bytes(o, v, prec, 1)
        df = lib.pyscfad_amd_ccsd_t_df_workspace_bytes(o, v, naux, prec, 1)
        elt = 8 if prec == 0 else 4
        saved = (o - 6) * v ** 3 * (elt + 8)
        factors = 2 * (o * v + v * v) * naux * 8   # re-laid-out factors and their accumulators (unpadded)
        assert df <= dense - saved + 2 * factors, (prec, dense, df)
        assert df < dense


def test_df_argument_errors():
    lib = _lib.load()
    dummy = ctypes.c_void_p(256)
    # naux < 1
    rc = lib.pyscfad_amd_ccsd_t_df_energy(3, 5, 0, *([dummy] * 8), 0, 0, -1, dummy, None, 0, None)
    assert rc == 1
    # null DF factor
    rc = lib.pyscfad_amd_ccsd_t_df_energy(3, 5, 24, *([dummy] * 6), None, dummy, 0, 0, -1, dummy, None, 0, None)
    assert rc == 1 and b"null" in lib.pyscfad_amd_last_error()
    # null Lvv_bar in the VJP
    rc = lib.pyscfad_amd_ccsd_t_df_energy_vjp(3, 5, 24, *([dummy] * 8), 1.0, 0, 0, -1, *([dummy] * 7), None,
                                              dummy, None, 0, None)
    assert rc == 1 and b"null" in lib.pyscfad_amd_last_error()
    # workspace too small
    rc = lib.pyscfad_amd_ccsd_t_df_energy(3, 5, 24, *([dummy] * 8), 0, 0, -1, dummy, dummy, 16, None)
    assert rc == 4
