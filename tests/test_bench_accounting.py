"""The benchmark's credited work (CPU only).

The metric credits 3 F_E per unit of work, F_E = 2 o^3 v^3 (v + o) for a whole
instance (the E(T) contraction over all o^3 occupied orderings, each a
v^3 (v + o) contraction of 2 flops per multiply-add).  bench.py credits a
slice of unique occupied triples i >= j >= k by the number of orderings each
stands for (6, 3 or 1).  Summed over every unique triple that must give o^3
orderings, i.e. exactly F_E; slices must add up; and a wrong weight must break
the closed form.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _ntrip(o):
    return o * (o + 1) * (o + 2) // 6


@pytest.mark.parametrize("o,v", [(1, 1), (2, 3), (5, 19), (7, 33), (13, 40), (32, 256)])
def test_whole_slice_is_f_e(o, v):
    from pyscfad_amd.synthetic import flops_energy
    b = _bench()
    assert b.slice_flops(o, v, 0, _ntrip(o)) == pytest.approx(flops_energy(o, v), rel=1e-15)


def test_slices_add_up():
    b = _bench()
    o, v = 12, 50
    cuts = [0, 1, 17, 100, 101, 250, _ntrip(o)]
    parts = sum(b.slice_flops(o, v, a, c) for a, c in zip(cuts, cuts[1:]))
    assert parts == pytest.approx(b.slice_flops(o, v, 0, _ntrip(o)), rel=1e-15)
    assert b.slice_flops(o, v, 5, 5) == 0.0


def test_benchmark_slice_weights():
    """The default o=64 step [20000, 20064): 64 unique triples, each counted
    with the number of distinct orderings of its (i, j, k)."""
    b = _bench()
    o, v = 64, 512
    idx, orderings = 0, 0
    for i in range(o):
        for j in range(i + 1):
            for k in range(j + 1):
                if 20000 <= idx < 20064:
                    orderings += len({(i, j, k), (i, k, j), (j, i, k), (j, k, i), (k, i, j), (k, j, i)})
                idx += 1
    assert b.slice_flops(o, v, 20000, 20064) == pytest.approx(orderings * 2.0 * v ** 3 * (v + o), rel=1e-15)


def test_wrong_weight_breaks_closed_form():
    """Counting every unique triple as six orderings (a plausible slip)
    over-credits, and the closed form catches it."""
    from pyscfad_amd.synthetic import flops_energy
    o, v = 9, 20
    wrong = 6 * _ntrip(o) * 2.0 * v ** 3 * (v + o)
    assert wrong > 1.1 * flops_energy(o, v)
