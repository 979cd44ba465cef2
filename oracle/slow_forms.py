"""NumPy closed forms of E(T) (test infrastructure only).

Two formulations that are independent of the C restatement's loop structure,
used to pin its energy:

* ``energy_vectorised``: the reference's pure-JAX kernel ``_compute_et``
  (pyscfad/cc/ccsd_t_slow.py:68-125) -- a loop over a>=b>=c with the 6-fold
  sum written as explicit transposes, ``ZZ = r3(WW + VV/2)/d3``,
  ``E = 2 sum WW*ZZ``.
* ``energy_dense6``: the dense six-index form (ccsd_t_slow.py:138-220) --
  ``t3 = P6(W+V)/D``, ``E = sum t3 * r3(P6(W)/D) * D / 3``.

Both use ``t2T = t2.transpose(2,3,0,1)`` exactly as ccsd_t_slow.py does.  For
a t2 with the CCSD symmetry t2[i,j,a,b] = t2[j,i,b,a] this reads the same values
as the fast path's ``t2.transpose(2,3,1,0)`` (ccsd_t.py:90).  For any other t2
the two paths differ: the particle term reads the same t2 element in both,
while the hole and V terms read its pair partner.  Only use these two with such t2.
Their maps of ovvv, ovoo and ovov (ccsd_t_slow.py:40-42) equal the fast path's
(ccsd_t.py:98-106) element for element, so they pin those maps on any input.

* ``energy_dense6_fastmap``: the dense six-index form written with the fast
  path's own map ``t2T = t2.transpose(2,3,1,0)`` (ccsd_t.py:90) in each of the
  terms of get_wv (ccsd_t.c:137-150).  It holds for any t2, and it pins the
  element of t2 that each term reads.
"""
import numpy as np

from .ccsd_t_oracle import unpack_ovvv


def _prep(t1, t2, mo_energy, fock, ovoo, ovov, ovvv):
    nocc, nvir = t1.shape
    vvov = unpack_ovvv(ovvv, nvir).transpose(1, 3, 0, 2)   # ccsd_t_slow.py:40
    vooo = np.asarray(ovoo).transpose(1, 0, 3, 2)          # :41
    vvoo = np.asarray(ovov).transpose(1, 3, 0, 2)          # :42
    fvo = np.asarray(fock)[nocc:, :nocc]                   # :43
    return vvov, vooo, vvoo, fvo


def _r3(w):
    return (4 * w + w.transpose(1, 2, 0) + w.transpose(2, 0, 1)
            - 2 * w.transpose(2, 1, 0) - 2 * w.transpose(0, 2, 1)
            - 2 * w.transpose(1, 0, 2))


def energy_vectorised(t1, t2, mo_energy, fock, ovoo, ovov, ovvv):
    """ccsd_t_slow.py:29-125 restated with NumPy loops."""
    nocc, nvir = t1.shape
    vvov, vooo, vvoo, fvo = _prep(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
    e_occ, e_vir = mo_energy[:nocc], mo_energy[nocc:]
    eijk = e_occ[:, None, None] + e_occ[None, :, None] + e_occ[None, None, :]
    t1T = t1.T
    t2T = t2.transpose(2, 3, 0, 1)

    def get_w(a, b, c):
        w = np.einsum('if,fkj->ijk', vvov[a, b], t2T[c, :])
        w -= np.einsum('ijm,mk->ijk', vooo[a, :], t2T[b, c])
        return w

    def get_v(a, b, c):
        v = np.einsum('ij,k->ijk', vvoo[a, b], t1T[c])
        v += np.einsum('ij,k->ijk', t2T[a, b], fvo[c])
        return v

    et = 0.0
    for a in range(nvir):
        for b in range(a + 1):
            for c in range(b + 1):
                fac = 6 if a == c else (2 if (a == b or b == c) else 1)
                d3 = (eijk - e_vir[a] - e_vir[b] - e_vir[c]) * fac
                WW = (get_w(a, b, c) + get_w(a, c, b).transpose(0, 2, 1)
                      + get_w(b, a, c).transpose(1, 0, 2) + get_w(b, c, a).transpose(2, 0, 1)
                      + get_w(c, a, b).transpose(1, 2, 0) + get_w(c, b, a).transpose(2, 1, 0))
                VV = (get_v(a, b, c) + get_v(a, c, b).transpose(0, 2, 1)
                      + get_v(b, a, c).transpose(1, 0, 2) + get_v(b, c, a).transpose(2, 0, 1)
                      + get_v(c, a, b).transpose(1, 2, 0) + get_v(c, b, a).transpose(2, 1, 0))
                ZZ = _r3(WW + 0.5 * VV) / d3
                et += np.einsum('ijk,ijk', WW, ZZ)
    return 2.0 * et


def _p6(w):
    return (w + w.transpose(1, 0, 2, 4, 3, 5) + w.transpose(2, 1, 0, 5, 4, 3)
            + w.transpose(0, 2, 1, 3, 5, 4) + w.transpose(1, 2, 0, 4, 5, 3)
            + w.transpose(2, 0, 1, 5, 3, 4))


def _r3_6d(w):
    return (4 * w + w.transpose(2, 0, 1, 3, 4, 5) + w.transpose(1, 2, 0, 3, 4, 5)
            - 2 * w.transpose(2, 1, 0, 3, 4, 5) - 2 * w.transpose(0, 2, 1, 3, 4, 5)
            - 2 * w.transpose(1, 0, 2, 3, 4, 5))


def energy_dense6(t1, t2, mo_energy, fock, ovoo, ovov, ovvv):
    """ccsd_t_slow.py:138-220: E = sum t3 * r3(P6(W)/D) * D / 3, t3 = P6(W+V)/D."""
    nocc, nvir = t1.shape
    vvov, vooo, vvoo, fvo = _prep(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
    t2T = t2.transpose(2, 3, 0, 1)
    w = np.einsum('abif,cfkj->ijkabc', vvov, t2T)
    w -= np.einsum('aijm,bcmk->ijkabc', vooo, t2T)
    v = 0.5 * (np.einsum('abij,ck->ijkabc', vvoo, t1.T) + np.einsum('abij,ck->ijkabc', t2T, fvo))
    eia = mo_energy[:nocc, None] - mo_energy[None, nocc:]
    d3 = (eia[:, None, None, :, None, None] + eia[None, :, None, None, :, None]
          + eia[None, None, :, None, None, :])
    t3 = _p6(w + v) / d3
    wd = _p6(w) / d3
    return float(np.einsum('ijkabc,ijkabc', t3, _r3_6d(wd) * d3) / 3.0)


def energy_dense6_fastmap(t1, t2, mo_energy, fock, ovoo, ovov, ovvv):
    """E = (1/3) sum P6(W+V) r3(P6(W)) / D, with the terms of get_wv
    (ccsd_t.c:137-150) over all (a,b,c) and (i,j,k):
        W = sum_f vvop[a,b,i,o+f] t2T[c,f,j,k] - sum_m vooo[a,i,j,m] t2T[c,b,m,k]
        V = 1/2 vvop[a,b,i,j] t1T[c,k] + 1/2 t2T[b,a,i,j] fvo[c,k]
    and t2T = t2.transpose(2,3,1,0) (ccsd_t.py:90)."""
    nocc, nvir = t1.shape
    vvov, vooo, vvoo, fvo = _prep(t1, t2, mo_energy, fock, ovoo, ovov, ovvv)
    t2T = np.asarray(t2).transpose(2, 3, 1, 0)
    w = np.einsum('abif,cfjk->ijkabc', vvov, t2T)
    w -= np.einsum('aijm,cbmk->ijkabc', vooo, t2T)
    v = 0.5 * (np.einsum('abij,ck->ijkabc', vvoo, t1.T) + np.einsum('baij,ck->ijkabc', t2T, fvo))
    eia = mo_energy[:nocc, None] - mo_energy[None, nocc:]
    d3 = (eia[:, None, None, :, None, None] + eia[None, :, None, None, :, None]
          + eia[None, None, :, None, None, :])
    return float(np.einsum('ijkabc,ijkabc', _p6(w + v), _r3_6d(_p6(w)) / d3) / 3.0)


def lno_energy_dense6(mat, t1T, t2T, mo_energy, fvo, ovoo, ovov, ovvv):
    """The LNO (T) energy (lno_ccsd_t.c:6-29, 96-138, x 2/3 at
    lno/ccsd_t.py:111) as a dense six-index sum over all (a,b,c) and (i,j,k):
        E = (1/3) sum_{abc,ijk} (sum_l mat[l,i] w0[l,j,k]) z0[i,j,k] / D
    with w0 = P6(W), z0 = r3(P6(W + V)) and W, V the terms of get_wv
    (ccsd_t.c:137-150) on the t2T the LNO caller passes (lno/ccsd_t.py:46)."""
    nvir, nocc = np.shape(t1T)
    vvov = unpack_ovvv(ovvv, nvir).transpose(1, 3, 0, 2)
    vooo = np.asarray(ovoo).transpose(1, 0, 3, 2)
    vvoo = np.asarray(ovov).transpose(1, 3, 0, 2)
    w = np.einsum('abif,cfjk->ijkabc', vvov, t2T)
    w -= np.einsum('aijm,cbmk->ijkabc', vooo, t2T)
    v = 0.5 * (np.einsum('abij,ck->ijkabc', vvoo, t1T) + np.einsum('baij,ck->ijkabc', t2T, fvo))
    eia = mo_energy[:nocc, None] - mo_energy[None, nocc:]
    d3 = (eia[:, None, None, :, None, None] + eia[None, :, None, None, :, None]
          + eia[None, None, :, None, None, :])
    w0 = _p6(w)
    z0 = _r3_6d(_p6(w + v))
    wm = np.einsum('li,ljkabc->ijkabc', mat, w0)
    return float(np.einsum('ijkabc,ijkabc', wm, z0 / d3) / 3.0)
