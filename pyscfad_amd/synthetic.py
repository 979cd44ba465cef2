"""Seeded synthetic (T) inputs with the structure of canonical RHF-CCSD data.

One generator for every synthetic workload (single instances and the batched
set); all randomness comes from ``numpy.random.default_rng(seed)`` (PCG64) on
the host, so a seed fixes the instance bit for bit wherever it is built.

For an instance with ``nocc = o``, ``nvir = v``, ``n = o + v``:
  * mo_energy: occupied ~ U[-1.5, -0.3] Eh, virtual ~ U[0.1, 3.0] Eh, each
    sorted ascending (so every denominator satisfies |D| >= 1.2 Eh).
  * a density-fitting factor L of shape (naux = 3n, n, n): the upper triangle
    (diagonal included) is drawn from N(0, 0.05^2) -- 0.05 is the standard
    deviation -- and mirrored, so L[P,p,q] = L[P,q,p].
  * integrals in chemists' notation built from L (never the full n^4 tensor):
        ovov[i,a,j,b] = sum_P L[P,i,o+a] L[P,j,o+b]
        ovoo[i,a,k,j] = sum_P L[P,i,o+a] L[P,k,j]
        ovvv[i,a,b,c] = sum_P L[P,i,o+a] L[P,o+b,o+c], stored lower-triangle
                        packed over (b,c) as (o, v, v(v+1)/2), b >= c.
  * t1 ~ N(0, 1e-2^2).
  * t2 = ovov.transpose(0,2,1,3) / D_ijab + N(0, 1e-3^2), then symmetrised
    t2 <- (t2 + t2.transpose(1,0,3,2)) / 2, so t2[i,j,a,b] = t2[j,i,b,a];
    D_ijab = e_i + e_j - e_a - e_b.
  * fock = diag(mo_energy) (canonical, so the vo block is zero); with
    ``fvo_scale > 0`` the vo and ov blocks get N(0, fvo_scale^2) entries to
    exercise the t2*fvo term.
Draw order: occupied energies, virtual energies, L, t1, t2 noise, fvo.

``integral_scale = s`` multiplies L by s and the t2 noise by s^2 (the same
draws): the integrals and t2 shrink by s^2.  With the generator as stated
(s = 1) the integrals grow with n -- |E(T)| is about 3e8 Eh at nocc=32,
nvir=256 -- and s = 1/sqrt(n) gives amplitudes and an E(T) of molecular size
(|t2| ~ 1e-2, |E(T)| ~ 1e-2 Eh), the scale at which absolute accuracy in Eh
is meaningful.

Large instances can build the integral slices with torch on a GPU
(``device='cuda'``); the random draws are identical either way.
"""
import numpy as np


def _draws(nocc, nvir, seed, fvo_scale):
    rng = np.random.default_rng(seed)
    n = nocc + nvir
    e_occ = np.sort(rng.uniform(-1.5, -0.3, nocc))
    e_vir = np.sort(rng.uniform(0.1, 3.0, nvir))
    mo_energy = np.concatenate([e_occ, e_vir])
    naux = 3 * n
    iu = np.triu_indices(n)
    L = np.zeros((naux, n, n))
    upper = rng.normal(0.0, 0.05, size=(naux, iu[0].size))
    L[:, iu[0], iu[1]] = upper
    L[:, iu[1], iu[0]] = upper
    del upper
    t1 = rng.normal(0.0, 1e-2, size=(nocc, nvir))
    t2_noise = rng.normal(0.0, 1e-3, size=(nocc, nocc, nvir, nvir))
    fvo = rng.normal(0.0, fvo_scale, size=(nvir, nocc)) if fvo_scale > 0 else None
    return mo_energy, L, t1, t2_noise, fvo


def make_instance(nocc, nvir, seed, fvo_scale=0.0, packed=True, device=None, df=False, integral_scale=1.0):
    """Return a dict with t1, t2, mo_energy, fock, ovoo, ovov, ovvv (FP64).

    With ``device=None`` everything is NumPy on the host.  With a torch device
    string (e.g. 'cuda') the arrays are torch FP64 tensors on that device and
    the O(o v^3) slice is built there.  With ``df=True`` the dict holds the DF
    factors in the layouts of pyscfad/cc/dfccsd.py:82-83 in place of ovvv:
    Lov (naux, o, v) and Lvv (naux, v(v+1)/2), lower-triangle packed, with
    ovvv = Lov^T Lvv exactly the packed ovvv above.
    """
    o, v = nocc, nvir
    mo_energy, L, t1, t2_noise, fvo = _draws(o, v, seed, fvo_scale)
    if integral_scale != 1.0:
        L *= integral_scale
        t2_noise *= integral_scale ** 2
    n = o + v
    fock = np.diag(mo_energy)
    if fvo is not None:
        fock[o:, :o] = fvo
        fock[:o, o:] = fvo.T
    e_occ, e_vir = mo_energy[:o], mo_energy[o:]
    d_ijab = (e_occ[:, None, None, None] + e_occ[None, :, None, None]
              - e_vir[None, None, :, None] - e_vir[None, None, None, :])
    idx_b, idx_c = np.tril_indices(v)[0], np.tril_indices(v)[1]
    if device is None:
        Lov = L[:, :o, o:].reshape(L.shape[0], o * v)
        ovov = (Lov.T @ Lov).reshape(o, v, o, v)
        ovoo = (Lov.T @ L[:, :o, :o].reshape(L.shape[0], o * o)).reshape(o, v, o, o)
        Lvv = L[:, o:, o:]
        if df:
            ovvv = None
        elif packed:
            ovvv = (Lov.T @ Lvv[:, idx_b, idx_c]).reshape(o, v, v * (v + 1) // 2)
        else:
            ovvv = (Lov.T @ Lvv.reshape(L.shape[0], v * v)).reshape(o, v, v, v)
        t2 = ovov.transpose(0, 2, 1, 3) / d_ijab + t2_noise
        t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
        out = dict(t1=t1, t2=np.ascontiguousarray(t2), mo_energy=mo_energy, fock=fock, ovoo=ovoo, ovov=ovov)
        if df:
            out["Lov"] = np.ascontiguousarray(Lov.reshape(L.shape[0], o, v))
            out["Lvv"] = np.ascontiguousarray(Lvv[:, idx_b, idx_c])
        else:
            out["ovvv"] = ovvv
        return out
    import torch
    dev = torch.device(device)
    Lt = torch.from_numpy(L).to(dev)
    del L
    naux = Lt.shape[0]
    Lov = Lt[:, :o, o:].reshape(naux, o * v)
    ovov = (Lov.T @ Lov).reshape(o, v, o, v)
    ovoo = (Lov.T @ Lt[:, :o, :o].reshape(naux, o * o)).reshape(o, v, o, o)
    Lvv = Lt[:, o:, o:]
    if df:
        ib = torch.from_numpy(idx_b).to(dev)
        ic = torch.from_numpy(idx_c).to(dev)
        Lvv_p = Lvv[:, ib, ic].contiguous()
        Lov_3 = Lov.reshape(naux, o, v).contiguous()
        ovvv = None
    elif packed:
        ib = torch.from_numpy(idx_b).to(dev)
        ic = torch.from_numpy(idx_c).to(dev)
        Lvv_p = Lvv[:, ib, ic]
        ovvv = torch.empty((o, v, v * (v + 1) // 2), dtype=torch.float64, device=dev)
        # build one occupied slab at a time to bound the temporary
        for i in range(o):
            ovvv[i] = Lov[:, i * v:(i + 1) * v].T @ Lvv_p
        del Lvv_p
    else:
        ovvv = (Lov.T @ Lvv.reshape(naux, v * v)).reshape(o, v, v, v)
    d = torch.from_numpy(d_ijab).to(dev)
    t2 = ovov.permute(0, 2, 1, 3) / d + torch.from_numpy(t2_noise).to(dev)
    t2 = (0.5 * (t2 + t2.permute(1, 0, 3, 2))).contiguous()
    out = dict(t1=torch.from_numpy(t1).to(dev), t2=t2,
               mo_energy=torch.from_numpy(mo_energy).to(dev),
               fock=torch.from_numpy(fock).to(dev), ovoo=ovoo.contiguous(),
               ovov=ovov.contiguous())
    if df:
        out["Lov"] = Lov_3
        out["Lvv"] = Lvv_p
    else:
        out["ovvv"] = ovvv
    return out


def batched_shapes(count=64, seed=4, nocc_range=(4, 24), nvir_range=(40, 200)):
    """(nocc, nvir, seed) of each instance of the batched workload: sizes drawn
    uniformly from the given inclusive ranges with ``seed``, instance k built
    with seed 4000 + k."""
    rng = np.random.default_rng(seed)
    occ = rng.integers(nocc_range[0], nocc_range[1] + 1, size=count)
    vir = rng.integers(nvir_range[0], nvir_range[1] + 1, size=count)
    return [(int(o), int(v), 4000 + k) for k, (o, v) in enumerate(zip(occ, vir))]


def flops_energy(nocc, nvir):
    """Algorithmic flops of the E(T) contraction, F_E = 2 o^3 v^3 (v + o):
    v^3 W evaluations of a (o^2 x v)(v x o) and an (o x o)(o x o^2) product."""
    o, v = float(nocc), float(nvir)
    return 2.0 * o ** 3 * v ** 3 * (v + o)


def make_lno_instance(nocc, nvir, seed, mat="gram", fvo_scale=0.02, packed=True, integral_scale=1.0,
                      nlo=None):
    """The inputs of the LNO (T) (pyscfad/lno/ccsd_t.py:70) in the forms that
    function takes, from the instance ``make_instance`` builds: a dict with
    mat (o,o), t1T (v,o), t2T = t2.transpose(2,3,1,0) (v,v,o,o), mo_energy,
    fvo = fock[o:, :o], ovoo, ovov, ovvv, plus "canonical": the make_instance
    dict.  ``mat``: "gram" (ulo^T ulo with ulo of shape (nlo, o), entries
    N(0, 1/o), nlo = o//2 + 1 by default, as lno/ccsd_t.py:42 forms it),
    "identity", or "random" (I + N(0, 0.3^2), not symmetric).  The projector
    draws come from ``seed + 2000``, after the instance's own."""
    d = make_instance(nocc, nvir, seed, fvo_scale=fvo_scale, packed=packed, integral_scale=integral_scale)
    rng = np.random.default_rng(seed + 2000)
    if mat == "identity":
        m = np.eye(nocc)
    elif mat == "random":
        m = np.eye(nocc) + 0.3 * rng.normal(size=(nocc, nocc))
    elif mat == "gram":
        k = nlo if nlo is not None else nocc // 2 + 1
        ulo = rng.normal(0.0, 1.0 / np.sqrt(nocc), size=(k, nocc))
        m = ulo.T @ ulo
    else:
        raise ValueError(f"unknown mat kind {mat!r}")
    return {
        "mat": m,
        "t1T": np.ascontiguousarray(d["t1"].T),
        "t2T": np.ascontiguousarray(np.asarray(d["t2"]).transpose(2, 3, 1, 0)),
        "mo_energy": d["mo_energy"],
        "fvo": np.ascontiguousarray(np.asarray(d["fock"])[nocc:, :nocc]),
        "ovoo": d["ovoo"], "ovov": d["ovov"], "ovvv": d["ovvv"],
        "canonical": d,
    }
