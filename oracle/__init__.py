"""CPU oracle for the CCSD(T) (T) energy and its VJP.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or, for the bench, as the timed CPU restatement).  The product package
``pyscfad_amd`` never imports it.

Contents:
  * ``ccsd_t_oracle.c``  -- FP64 C/OpenMP restatement of the reference's
    virtual-triple algorithm (forward + reverse), built to
    ``liboracle_ccsd_t.so`` by ``build.py``.
  * ``ccsd_t_oracle.py`` -- the reference's host-side preparation and
    back-transformation around that kernel (pyscfad/cc/ccsd_t.py:82-192).
  * ``slow_forms.py``    -- NumPy closed forms of the energy used to pin the
    C restatement (pyscfad/cc/ccsd_t_slow.py:68-125 and 138-220).

Every function here is pinned by tests/test_oracle.py; the checks used for
each are listed in that file's docstring.
"""
