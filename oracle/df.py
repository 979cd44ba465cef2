"""Density-fitted ovvv and the chain rule of its cotangent (test infrastructure only).

The reference's DF-CCSD builds the packed ovvv from the DF factors
(pyscfad/cc/dfccsd.py:81-91):

    Lov (naux, o*v) = Lpq[:, :o, o:],   Lvv (naux, v(v+1)/2) = pack_tril(Lpq[:, o:, o:]),
    ovvv = dot(Lov.T, Lvv).reshape(o, v, v(v+1)/2),

and passes it to the (T) kernel; differentiating through that product maps
the packed cotangent ovvv_bar (pyscfad/cc/ccsd_t.py:179-191) onto the factors:

    Lov_bar = dot(Lvv, ovvv_bar.reshape(o*v, -1).T)      (naux, o*v)
    Lvv_bar = dot(Lov, ovvv_bar.reshape(o*v, -1))        (naux, v(v+1)/2)

Checked against central finite differences of the oracle energy in
tests/test_oracle.py.
"""
import numpy as np


def ovvv_from_df(Lov, Lvv):
    """Packed ovvv (o, v, v(v+1)/2) from Lov (naux, o, v) and packed Lvv."""
    naux, o, v = Lov.shape
    return (Lov.reshape(naux, o * v).T @ Lvv).reshape(o, v, Lvv.shape[1])


def df_vjp(Lov, Lvv, ovvv_bar):
    """(Lov_bar (naux, o, v), Lvv_bar (naux, v(v+1)/2)) from the packed ovvv cotangent."""
    naux, o, v = Lov.shape
    ob = ovvv_bar.reshape(o * v, -1)
    lov_bar = (Lvv @ ob.T).reshape(naux, o, v)
    lvv_bar = Lov.reshape(naux, o * v) @ ob
    return lov_bar, lvv_bar
