#!/usr/bin/env python3
"""(T) energy+VJP throughput on one MI355X, FP32-MFMA against FP64.

Workload (default): the synthetic single instance nocc=64, nvir=512 (seed 3;
generator in pyscfad_amd/synthetic.py), inputs resident in HBM before the
timed region.  One step = the forward call (E(T), the custom_vjp forward)
plus the backward call (every cotangent; the backward recomputes W/V
instead of storing them) over one batch of --batch-triples unique occupied
triples; step s takes the s-th consecutive batch from --first-triple on.  A
whole solve (45760 triples) takes about 25 min in FP32, so it is not timed
here; the whole o=32/v=256 solve is (--nocc 32 --nvir 256 --batch-triples 0).

metric   : TFLOP/s of energy+VJP credited as 3 * F_E of the triples processed,
           F_E = 2 v^3 (v + o) per distinct occupied ordering (2 o^3 v^3 (v+o)
           for a whole instance)
value    : that rate for the --precision path (default f32 MFMA)
f64      : the FP64 path timed on the same batches in the same process, and
           the ratio of the two rates
bf16x6   : (f32 runs) the bf16x6 path on the same batches: FP32 operands
           split into three bf16 parts, six products per K-step on the bf16
           matrix cores, FP32 accumulation; its rate, GEMM roofline against
           the bf16 peak over six products, speedup over FP64 and accuracy
roofline : the grouped MFMA GEMM kernel, against the MI355X dense matrix peak
           of the precision used; achieved = the credited 3 F_E over the
           kernel's device time (hipEvents around every launch on the call's
           stream, in the first warm-up step); the kernel also executes the
           backward's W recompute (a fourth F_E), which is not credited
cpu_baseline: the repository's FP64 C/OpenMP restatement of the reference
           algorithm (oracle/, products on scipy's BLAS), energy + VJP of a
           bounded run of the a>=b>=c virtual-triple jobs of the same
           instance, on the host cores of the same machine
accuracy : the last timed batch in --precision against FP64: |dE| in Eh and
           relative, max|d|/max|ref| per cotangent
"""
import argparse
import json
import os
import sys
import time

MI355X_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16x3": 2516.6 / 3.0, "bf16x6": 2516.6 / 6.0}
METRIC = "(T) energy+VJP TFLOP/s & s/solve, FP32-MFMA vs FP64 on MI355X at <=1e-6 Eh"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nocc", type=int, default=64)
    ap.add_argument("--nvir", type=int, default=512)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--batch-triples", type=int, default=64,
                    help="unique occupied triples per step (0: the whole instance per step)")
    ap.add_argument("--first-triple", type=int, default=20000,
                    help="first triple of the timed batches (the warm-up batches precede it)")
    ap.add_argument("--precision", default="f32", choices=["f64", "f32", "bf16x3", "bf16x6"])
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 accuracy reference run")
    ap.add_argument("--no-bf16x6", action="store_true",
                    help="skip the bf16x6 leg (FP32-accurate products on the bf16 matrix cores) that an "
                         "f32 run times on the same batches")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--triples", default=None,
                    help="BEGIN:END slice of the unique occupied triples (a bounded sample of a large "
                         "instance, or a profiling run); default: all")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--progress", action="store_true",
                    help="wait for each timed step and log it (long steps, e.g. a whole solve in slices); "
                         "the host's per-call planning is then no longer hidden behind the previous step")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="skip the extra untimed step that carries the per-kernel timing when there is no "
                         "warm-up step (the roofline fields are then left empty)")
    ap.add_argument("--energy-only", action="store_true",
                    help="time the forward E(T) alone (F_E credited), for instances whose VJP "
                         "workspace does not fit")
    ap.add_argument("--batched", action="store_true",
                    help="the batched workload instead: 64 instances, nocc in [4,24], nvir in [40,200] "
                         "(seed 4), as one captured hipGraph, against 64 sequential FP64 calls")
    ap.add_argument("--df", action="store_true",
                    help="density-fitted inputs (Lov, Lvv of the same generator) in place of ovvv: the "
                         "o*v^3 slabs are built on the device block by block, which fits the o=80/v=640 "
                         "backward in HBM; --triples then counts in the blocked order ('auto': the first "
                         "block triple of three distinct blocks)")
    ap.add_argument("--df-block", type=int, default=4, help="occupied indices per block in --df mode")
    ap.add_argument("--lno", action="store_true",
                    help="the LNO (T) with an occupied projector (a Gram mat of nocc//2+1 local orbitals), "
                         "whole solves of --nocc x --nvir at molecular integral scale")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of this many MiB (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(inst, nocc, nvir, target_s, vrow=16):
    """Oracle energy + VJP on a bounded run of the instance's a>=b>=c jobs,
    timed on the host cores.  Only the host arrays those jobs read are
    built: the jobs [0, njobs(vrow)) have a, b, c < vrow, so their operands
    are the first vrow virtual rows of t1T, t2T, vooo and fvo, the vrow x vrow
    block of vvop, and ovvv[:, :vrow] (the contracted indices still run over
    all nocc and nvir).  Each job's work is the same for every job
    (12 o^3 (v + o) flops forward), so the rate stands for the whole list."""
    from oracle import ccsd_t_oracle as orc
    w = min(vrow, nvir)

    def host(x):
        return x.cpu().numpy() if hasattr(x, "cpu") else x

    prep = orc.Prepared(t1=host(inst["t1"]), t2=host(inst["t2"][:, :, :w, :]), mo_energy=host(inst["mo_energy"]),
                        fock=host(inst["fock"]), ovoo=host(inst["ovoo"][:, :w]),
                        ovov=host(inst["ovov"][:, :w, :, :w]), ovvv=host(inst["ovvv"][:, :w]), vrow=w)
    flops_per_job = 12.0 * nocc ** 3 * (nvir + nocc)          # forward, one a>=b>=c job
    nwin = prep.njobs()

    def run(j0, j1):
        t0 = time.perf_counter()
        prep.energy(j0, j1)
        prep.energy_vjp(j0, j1)
        return time.perf_counter() - t0

    with orc.blas():
        j0 = nwin // 4
        probe = max(2, min(16, nwin // 8))
        t_probe = run(j0, j0 + probe)
        n = int(max(probe, min(nwin - j0, probe * target_s / max(t_probe, 1e-3))))
        t = run(j0, j0 + n)
    rate = 3.0 * flops_per_job * n / t / 1e12
    return {"value": rate, "unit": "TFLOP/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"energy+VJP of a>=b>=c jobs [{j0}, {j0 + n}) of the {orc.njobs(nvir)} of the nocc={nocc}, "
                      f"nvir={nvir} instance (jobs with a, b, c < {w}, whose host arrays alone are built) in "
                      f"{t:.1f} s, FP64 C/OpenMP with scipy's BLAS (one BLAS thread per OpenMP thread), "
                      f"credited 3*F_E per job"}


def measured_traffic(precision, nocc, nvir, batch_triples):
    """HBM bytes per gemm_grouped launch from the committed FETCH_SIZE / WRITE_SIZE
    profile of this workload (rocprofv3 --pmc runs, calibrated on a copy kernel),
    or None when none was recorded for it."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "gemm_grouped_traffic.json")
    try:
        recs = json.load(open(path))
    except (OSError, ValueError):
        return None, None
    for rec in recs.get("records", []):
        if (rec.get("precision"), rec.get("nocc"), rec.get("nvir")) == (precision, nocc, nvir):
            return rec["bytes_per_launch"], rec.get("source")
    return None, None


def slice_flops(nocc, nvir, t0, t1):
    """Algorithmic F_E of the unique triples [t0, t1): each unique triple stands
    for its distinct orderings (6, 3 or 1), each costing 2 v^3 (v + o)."""
    per = 2.0 * nvir ** 3 * (nvir + nocc)
    total, idx = 0.0, 0
    for i in range(nocc):
        for j in range(i + 1):
            for k in range(j + 1):
                if t0 <= idx < t1:
                    total += per * (1 if i == k else (3 if (i == j or j == k) else 6))
                idx += 1
    return total


def df_schedule(nocc, blk):
    """The density-fitted plan's block triples in order, each with its unique
    triples and the number of blocks the slab cache builds before it (the
    library's policy: three block slots, a missing block takes the first slot
    holding no block the block triple needs)."""
    nb = (nocc + blk - 1) // blk
    slots = [-1] * min(3, nb)
    out = []
    for I in range(nb):
        for J in range(I + 1):
            for K in range(J + 1):
                need, builds = (I, J, K), 0
                for b in need:
                    if b in slots:
                        continue
                    u = next(u for u in range(len(slots)) if slots[u] not in need)
                    slots[u] = b
                    builds += 1
                rng = lambda B: range(B * blk, min(nocc, (B + 1) * blk))
                trip = [(i, j, k) for i in rng(I) for j in rng(J) if j <= i for k in rng(K) if k <= j]
                out.append(((I, J, K), trip, builds))
    return out


def df_slice_info(nocc, nvir, blk, t0, t1):
    """(F_E of the triples [t0, t1) of the blocked order, slab blocks built for
    that slice from a cold cache, slab blocks built by the full plan)."""
    per = 2.0 * nvir ** 3 * (nvir + nocc)
    fe, idx, full_builds, slots = 0.0, 0, 0, None
    nb = (nocc + blk - 1) // blk
    cache = [-1] * min(3, nb)
    slice_builds = 0
    for (I, J, K), trip, builds in df_schedule(nocc, blk):
        full_builds += builds
        lo, hi = max(idx, t0), min(idx + len(trip), t1)
        if lo < hi:
            need = (I, J, K)
            for b in need:
                if b in cache:
                    continue
                u = next(u for u in range(len(cache)) if cache[u] not in need)
                cache[u] = b
                slice_builds += 1
            for (i, j, k) in trip[lo - idx:hi - idx]:
                fe += per * (1 if i == k else (3 if (i == j or j == k) else 6))
        idx += len(trip)
    return fe, slice_builds, full_builds


def df_auto_slice(nocc, blk):
    """[t0, t1) of the first block triple with three distinct blocks."""
    idx = 0
    for (I, J, K), trip, _ in df_schedule(nocc, blk):
        if I > J > K:
            return idx, idx + len(trip)
        idx += len(trip)
    return 0, idx


def max_rel_chunked(a, b):
    """max_rel of a (host or device) against device b, moving a to the device
    a slice along the first axis at a time."""
    import torch
    step = max(1, (1 << 27) // max(1, b[0].numel()))
    scale = 0.0
    for x in range(0, b.shape[0], step):
        scale = max(scale, torch.max(torch.abs(b[x:x + step])).item())
    if scale == 0:
        return (0.0, 0.0)
    dmax, elem = 0.0, 0.0
    for x in range(0, b.shape[0], step):
        ac = a[x:x + step].to(b.device)
        bc = b[x:x + step]
        d = torch.abs(ac - bc)
        dmax = max(dmax, torch.max(d).item())
        mask = torch.abs(bc) > 1e-12 * scale
        if bool(mask.any()):
            elem = max(elem, torch.max(d[mask] / torch.abs(bc[mask])).item())
    return (dmax / scale, elem)


def max_rel(a, b):
    """(max|a-b| / max|b|, max over entries with |b| > 1e-12 max|b| of |a-b|/|b|)."""
    import torch
    scale = torch.max(torch.abs(b)).item()
    if scale == 0:
        return (0.0, 0.0)
    d = torch.abs(a - b)
    mask = torch.abs(b) > 1e-12 * scale
    elem = torch.max(d[mask] / torch.abs(b[mask])).item() if bool(mask.any()) else 0.0
    return (torch.max(d).item() / scale, elem)


def triple_flops(nocc, nvir):
    """Cumulative F_E over the unique triples in the library's order (i >= j >= k,
    i outermost): entry t is the F_E of triples [0, t).  Each unique triple
    stands for its distinct orderings (6, 3 or 1), each 2 v^3 (v + o)."""
    import numpy as np
    per = 2.0 * nvir ** 3 * (nvir + nocc)
    w = [per * (1 if i == k else (3 if (i == j or j == k) else 6))
         for i in range(nocc) for j in range(i + 1) for k in range(j + 1)]
    return np.concatenate([[0.0], np.cumsum(w)])


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(a GPU's share of the work is a slice of the occupied triples, "
                                   "see --triples)"}))
        sys.exit(2)
    if args.batched:
        return main_batched(args)
    if args.lno:
        return main_lno(args)
    import torch
    from pyscfad_amd import _lib
    from pyscfad_amd.cc import ccsd_t
    from pyscfad_amd.synthetic import make_instance, flops_energy

    o, v = args.nocc, args.nvir
    if args.df:
        # the block size fixes the blocked triple order: every run of this
        # process (FP32 and the FP64 reference) uses the same one
        os.environ["PYSCFAD_AMD_DF_BLOCK"] = str(args.df_block)
    log(f"[bench] building synthetic instance nocc={o} nvir={v} seed={args.seed}" + (" (DF factors)" if args.df else ""))
    inst = make_instance(o, v, args.seed, device="cuda", df=args.df)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()   # return the generator's temporaries before the workspace is allocated
    keys = ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov") + (("Lov", "Lvv") if args.df else ("ovvv",))
    a = [inst[k] for k in keys]
    naux = inst["Lov"].shape[0] if args.df else 0
    ntrip = ccsd_t.num_triples(o)
    W, K = max(args.warmup, 0), args.steps

    # the batch of each step
    whole = False
    if args.triples:
        t0, t1 = (df_auto_slice(o, args.df_block) if (args.triples == "auto" and args.df)
                  else (int(x) for x in args.triples.split(":")))
        batches = [(t0, t1)] * (W + K)
    elif args.batch_triples > 0 and not args.df:
        B = min(args.batch_triples, ntrip)
        nb = W + K
        start = max(0, min(args.first_triple - W * B, ntrip - nb * B))
        batches = [(start + x * B, min(ntrip, start + (x + 1) * B)) for x in range(nb)]
        if start + nb * B > ntrip:
            # more steps than batches in the list: cycle through it
            batches = [((x * B) % (ntrip - B + 1), (x * B) % (ntrip - B + 1) + B) for x in range(nb)]
    else:
        whole = True
        batches = [(0, ntrip)] * (W + K)
    warm, timed = batches[:W], batches[W:]
    if args.df:
        def fe_of(sl):
            return df_slice_info(o, v, args.df_block, sl[0], sl[1])[0]
    else:
        cum = triple_flops(o, v)

        def fe_of(sl):
            return float(cum[sl[1]] - cum[sl[0]])

    if args.calibrate_copy:
        # two plain copies of the given size for the FETCH_SIZE / WRITE_SIZE
        # scale: the runtime's blit kernel, then an elementwise kernel with
        # ordinary vector loads and stores (the kind the GEMM issues)
        src = torch.ones(args.calibrate_copy * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    # one caller-provided workspace shared by the forward and backward calls
    # (the library's own allocation per call costs seconds at these sizes)
    wsp = {}

    def workspace(prec):
        if prec not in wsp:
            wsp.clear()
            torch.cuda.empty_cache()
            nbytes = (ccsd_t.df_workspace_bytes(o, v, naux, prec, with_vjp=not args.energy_only) if args.df
                      else ccsd_t.workspace_bytes(o, v, prec, with_vjp=not args.energy_only))
            wsp[prec] = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        return wsp[prec]

    f_energy = ccsd_t.energy_df if args.df else ccsd_t.energy
    f_vjp = ccsd_t.energy_vjp_df if args.df else ccsd_t.energy_vjp

    def step(prec, sl, timing=False):
        _lib.set_timing(timing)
        e = f_energy(*a, precision=prec, triples=sl, workspace=workspace(prec))
        t_fwd = _lib.last_timing() if timing else None
        if args.energy_only:
            _lib.set_timing(False)
            return e, None, t_fwd, ({k: (0.0, 0) for k in t_fwd} if timing else None)
        bars = f_vjp(*a, et_bar=1.0, precision=prec, triples=sl, workspace=workspace(prec))
        t_bwd = _lib.last_timing() if timing else None
        _lib.set_timing(False)
        return e, bars, t_fwd, t_bwd

    e_sums = {}   # E(T) summed over the timed steps of each leg (disjoint batches: their total)

    def timed_run(prec, sls, label):
        """Time the steps over the batches sls: (seconds per step, per-step
        seconds, last energy, last cotangents)."""
        log(f"[bench] timing {len(sls)} {label} step(s) ({prec})")
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(sls) + 1)]
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        evs[0].record()
        e = bars = None
        es = []
        for x, sl in enumerate(sls):
            bars = None
            e, bars, _, _ = step(prec, sl)
            es.append(e)
            evs[x + 1].record()
            if args.progress:
                evs[x + 1].synchronize()
                log(f"[bench] {label} step {x + 1}/{len(sls)} (triples [{sl[0]},{sl[1]})) done at "
                    f"{time.perf_counter() - c0:.1f} s")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - c0) / len(sls)
        per = [evs[x].elapsed_time(evs[x + 1]) / 1e3 for x in range(len(sls))]
        e_sums[label] = sum(float(x) for x in es)
        return dt, per, e, bars

    # tile shapes from the autotuner (outside the timed region)
    tuned = None
    compare = not args.no_compare and args.precision != "f64"
    if not args.no_autotune:
        log(f"[bench] autotuning tile shapes ({args.precision})")
        tuned = ccsd_t.autotune(o, v, args.precision)
        if compare:
            ccsd_t.autotune(o, v, "f64")

    torch.cuda.reset_peak_memory_stats()   # peak_hbm_GB: resident inputs + workspace + outputs
    fam, fe_w = {}, None
    for w, sl in enumerate(warm):
        log(f"[bench] warm-up step {w} ({args.precision}, triples [{sl[0]},{sl[1]})" + (", kernel timing on)" if w == 0 else ")"))
        _, bars, tf, tb = step(args.precision, sl, timing=(w == 0))
        bars = None
        if w == 0:
            fam = {k: (tf[k][0] + tb[k][0], tf[k][1] + tb[k][1]) for k in tf}
            fe_w = fe_of(sl)
    dt, per32, e, bars = timed_run(args.precision, timed, "timed")
    fe = sum(fe_of(sl) for sl in timed)
    peak_hbm = torch.cuda.max_memory_allocated() / 1e9
    credit = 1.0 if args.energy_only else 3.0    # forward only: F_E; energy + VJP: 3 F_E
    tflops = credit * fe / (dt * K) / 1e12
    if fe_w is None and not args.no_kernel_timing:
        # no warm-up step carried the kernel timing: one more (untimed) step
        # does, after the last timed step's cotangents leave the device
        if bars is not None:
            bars = [b.cpu() for b in bars] if compare else None
        torch.cuda.empty_cache()
        _, bars_t, tf, tb = step(args.precision, timed[0], timing=True)
        del bars_t
        fam = {k: (tf[k][0] + tb[k][0], tf[k][1] + tb[k][1]) for k in tf}
        fe_w = fe_of(timed[0])
    kernel_timed = fe_w is not None
    if not kernel_timed:   # --no-kernel-timing without a warm-up step: no roofline fields
        fam, fe_w = {"gemm": (0.0, 0)}, 0.0
    gemm_ms = fam["gemm"][0]
    peak = MI355X_PEAK_TFLOPS[args.precision]
    # density-fitted mode: the FP64 slab builds and folds run through the same
    # GEMM kernel but are not part of the credited contractions; their family
    # time (which also holds their clears and reductions) is taken out
    df_gemm_ms = fam.get("df", (0.0, 0))[0]
    achieved = credit * fe_w / ((gemm_ms - df_gemm_ms) / 1e3) / 1e12 if kernel_timed else None
    # algorithmic flops of each contraction of the timed batch: W build
    # (forward + recompute) 2 F_E, of which F_E is credited; particle-term
    # adjoints (ovvv, t2) F_E v/(v+o) each; hole-term adjoints (ovoo, t2)
    # F_E o/(v+o) each
    part, hole = fe_w * v / (v + o), fe_w * o / (v + o)
    sub_flops = {"gemm_w": 2.0 * fe_w, "gemm_b1": part, "gemm_b2": part, "gemm_b3": hole, "gemm_b4": hole}
    if args.energy_only:
        sub_flops = {"gemm_w": fe_w}
    contraction = {k: {"ms": fam[k][0], "launches": fam[k][1],
                       "tflops_executed": (f / (fam[k][0] / 1e3) / 1e12) if fam[k][0] > 0 else None}
                   for k, f in sub_flops.items() if k in fam}
    traffic, traffic_src = measured_traffic(args.precision, o, v, None if whole else timed[0][1] - timed[0][0])
    full = flops_energy(o, v)

    # bf16x6 on the same batches: three bf16 parts per FP32 operand, six
    # products, FP32 accumulation (FP32-level accuracy on the bf16 matrix cores)
    b6 = None
    bars6 = None
    if args.precision == "f32" and not args.no_bf16x6 and not args.df:
        if bars is not None and o * v ** 3 * 8 > 8e9:
            bars = [b.cpu() for b in bars]
        wsp.clear()
        torch.cuda.empty_cache()
        n6 = len(timed) if not whole else min(len(timed), 2)
        if not args.no_autotune:
            log("[bench] autotuning tile shapes (bf16x6)")
            ccsd_t.autotune(o, v, "bf16x6")
        log("[bench] bf16x6 warm-up step")
        step("bf16x6", warm[0] if warm else timed[0])
        _, bars_w, tf6, tb6 = step("bf16x6", warm[-1] if warm else timed[0], timing=True)
        del bars_w
        dt6, per6, e6, bars6 = timed_run("bf16x6", timed[-n6:], "bf16x6")
        if bars6 is not None and o * v ** 3 * 8 > 8e9:
            bars6 = [b.cpu() for b in bars6]
        fe6_w = fe_of(warm[-1] if warm else timed[0])
        gemm6 = tf6["gemm"][0] + tb6["gemm"][0]
        ach6 = credit * fe6_w / (gemm6 / 1e3) / 1e12
        b6 = {"tflops": credit * sum(fe_of(sl) for sl in timed[-n6:]) / sum(per6) / 1e12,
              "ms_per_step": dt6 * 1e3, "steps": n6, "batches": f"the last {n6} of the timed batches",
              "gemm_ms_per_step": gemm6,
              "roofline": {"achieved": ach6, "peak": MI355X_PEAK_TFLOPS["bf16x6"], "unit": "TFLOP/s",
                           "frac": ach6 / MI355X_PEAK_TFLOPS["bf16x6"],
                           "peak_note": "the dense bf16 matrix peak over the six products"}}
        wsp.clear()
        torch.cuda.empty_cache()

    f64 = None
    accuracy = None
    if compare:
        # the FP64 path on the same batches; large cotangents wait in host
        # memory while the FP64 workspace is live
        last = timed[-1] if args.energy_only else None
        if bars is not None and o * v ** 3 * 8 > 8e9:
            bars = [b.cpu() for b in bars]
        wsp.clear()
        torch.cuda.empty_cache()
        n64 = len(timed) if not whole else min(len(timed), 2)
        log("[bench] FP64 warm-up step")
        step("f64", warm[0] if warm else timed[0])
        dt64, per64, e64, bars64 = timed_run("f64", timed[-n64:], "FP64")
        t32_same = sum(per32[-n64:])
        t64_same = sum(per64)
        if b6 is not None:
            b6["speedup_over_f64"] = sum(per64[-b6["steps"]:]) / sum(per6)
            b6["accuracy_vs_f64"] = {"batch": list(timed[-1]), "dE_Eh": abs(float(e6) - float(e64)),
                                     "dE_rel": abs(float(e6) - float(e64)) / max(abs(float(e64)), 1e-300)}
        f64 = {"tflops": credit * sum(fe_of(sl) for sl in timed[-n64:]) / t64_same / 1e12,
               "ms_per_step": dt64 * 1e3, "steps": n64,
               "batches": f"the last {n64} of the timed batches",
               "f32_over_f64": t64_same / t32_same if args.precision == "f32" else None,
               "speedup_over_f64": t64_same / t32_same}
        accuracy = {"batch": list(timed[-1]), "dE_Eh": abs(float(e) - float(e64)), "E_f64": float(e64),
                    "dE_rel": abs(float(e) - float(e64)) / max(abs(float(e64)), 1e-300)}
        if not args.energy_only:
            names = ("t1", "t2", "fock", "mo_energy", "ovoo", "ovov") + (("Lov", "Lvv") if args.df else ("ovvv",))
            rel, elem = {}, {}
            for n, b, r in zip(names, bars, bars64):
                rel[n], elem[n] = max_rel_chunked(b, r)
            accuracy["max_rel"] = rel
            accuracy["max_rel_elementwise"] = elem
            if b6 is not None and bars6 is not None:
                b6["accuracy_vs_f64"]["max_rel"] = {n: max_rel_chunked(b, r)[0]
                                                    for n, b, r in zip(names, bars6, bars64)}
                bars6 = None
        del last

    cpu = None
    if not args.no_cpu and not args.df:
        log(f"[bench] CPU baseline sample (~{args.cpu_seconds:.0f} s)")
        cpu = cpu_baseline(inst, o, v, args.cpu_seconds)

    wl = (f"single (T) instance nocc={o} nvir={v} seed={args.seed}"
          + (f", density-fitted (naux={naux}, occupied blocks of {args.df_block})" if args.df else ""))
    if whole:
        wl += f", all {ntrip} unique occupied triples per step"
    elif args.triples:
        wl += f", unique triples [{timed[0][0]},{timed[0][1]}) of {ntrip} per step" + (" (blocked order)" if args.df else "")
    else:
        wl += (f", {timed[0][1] - timed[0][0]} unique occupied triples per step: consecutive batches "
               f"[{timed[0][0]},{timed[-1][1]}) of the {ntrip}")
    out = {
        "metric": METRIC if not args.energy_only else "(T) energy TFLOP/s (F_E credited)",
        "value": tflops,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dt * 1e3,
        "s_per_solve": dt if whole else None,
        "projected_full_solve_s": None if whole else credit * full / (tflops * 1e12),
        "projection_note": (None if whole else
                            "not measured: the whole triple list at the measured rate (each batch here "
                            "also pays the per-call setup once)"),
        "peak_hbm_GB": peak_hbm,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32 (FP64 denominators and sums)",
                  "bf16x3": "bf16x3 (FP32 accumulate, FP64 denominators and sums)",
                  "bf16x6": "bf16x6 (three-part bf16 split, FP32 accumulate, FP64 denominators and sums)"}[args.precision],
        "data": "synthetic",
        "config": {"workload": wl, "precision": args.precision, "E_T_last_step": float(e),
                   "E_T_sum_timed": e_sums.get("timed")},
        "f64": f64,
        "bf16x6": b6,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak if achieved is not None else None,
                     "traffic": traffic, "traffic_unit": "bytes per launch",
                     "traffic_source": traffic_src,
                     "kernel": "gemm_grouped",
                     "credited": f"{credit:.0f} F_E of the first warm-up batch over its gemm_grouped device time "
                                 "(the backward's W recompute runs in the same kernel and is not credited)",
                     "gemm_ms_per_step": gemm_ms,
                     "gemm_launches_per_step": fam.get("gemm", (0, 0))[1],
                     "tiles": tuned,
                     "family_ms": {k: val[0] for k, val in fam.items() if not k.startswith("gemm_")},
                     "contractions": contraction},
        "cpu_baseline": cpu,
        "accuracy_vs_f64": accuracy,
    }
    print(json.dumps(out))


def main_batched(args):
    """Batched uneven instances: one graph launch for all 64 energy+VJP problems
    (the graph is captured in the warm-up and replayed in the timed steps),
    against 64 sequential single-instance FP64 calls."""
    import torch
    from pyscfad_amd import _lib
    from pyscfad_amd.cc import ccsd_t
    from pyscfad_amd.synthetic import make_instance, flops_energy, batched_shapes

    shapes = batched_shapes(64, seed=4)
    log(f"[bench] building {len(shapes)} instances")
    probs = []
    for o, v, seed in shapes:
        d = make_instance(o, v, seed, device="cuda")
        d["et_bar"] = 1.0
        probs.append(d)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    fe = sum(flops_energy(o, v) for o, v, _ in shapes)
    tuned = None
    if not args.no_autotune:
        # tile shapes and batch sizes of every size class, before the workspace is sized
        log("[bench] autotuning the size classes")
        tuned = {f"{o}x{v}": t for (o, v), t in ccsd_t.autotune_problems(probs, args.precision).items()}
    ws = torch.empty(ccsd_t.batched_workspace_bytes(probs, args.precision), dtype=torch.uint8, device="cuda")

    log("[bench] warm-up (graph capture)")
    out = ccsd_t.batched(probs, precision=args.precision, workspace=ws)
    for _ in range(max(args.warmup - 1, 0)):
        out = ccsd_t.batched(probs, precision=args.precision, workspace=ws, outputs=out)
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} graph replay(s)")
    c0 = time.perf_counter()
    for _ in range(args.steps):
        out = ccsd_t.batched(probs, precision=args.precision, workspace=ws, outputs=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - c0) / args.steps

    # per-kernel family times of one eager (timed) pass
    _lib.set_timing(True)
    ccsd_t.batched(probs, precision=args.precision, workspace=ws, outputs=out)
    fam = _lib.last_timing()
    _lib.set_timing(False)
    gemm_ms = fam["gemm"][0]
    peak = MI355X_PEAK_TFLOPS[args.precision]
    # credited 3 F_E over the GEMM device time (the W recompute is executed, not credited)
    achieved = 3.0 * fe / (gemm_ms / 1e3) / 1e12

    log("[bench] 64 sequential FP64 single-instance calls (default tiles)")
    ws64 = torch.empty(max(ccsd_t.workspace_bytes(o, v, "f64") for o, v, _ in shapes), dtype=torch.uint8,
                       device="cuda")
    keys = ("t1", "t2", "mo_energy", "fock", "ovoo", "ovov", "ovvv")
    ref = []
    for d in probs:   # warm-up of the sequential path on the first instance
        ccsd_t.energy_vjp(*[d[k] for k in keys], et_bar=1.0, precision="f64", workspace=ws64)
        break
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    for d in probs:
        args64 = [d[k] for k in keys]
        e64 = ccsd_t.energy(*args64, precision="f64", workspace=ws64)
        bars64 = ccsd_t.energy_vjp(*args64, et_bar=1.0, precision="f64", workspace=ws64)
        ref.append((e64, bars64))
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - c0

    names = ("t1", "t2", "fock", "mo_energy", "ovoo", "ovov", "ovvv")
    accuracy = None
    if args.precision != "f64":
        de = max(abs(float(g[0]) - float(r[0])) for g, r in zip(out, ref))
        rel = {n: max(max_rel(g[1][x], r[1][x])[0] for g, r in zip(out, ref)) for x, n in enumerate(names)}
        accuracy = {"max_dE_Eh": de, "max_abs_E_f64": max(abs(float(r[0])) for r in ref), "max_rel": rel}

    line = {
        "metric": "(T) energy+VJP TFLOP/s (3*F_E credited), batched uneven instances",
        "value": 3.0 * fe / dt / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": "64 independent instances, nocc in [4,24], nvir in [40,200] drawn with seed 4, "
                               "instance k with seed 4000+k; one captured hipGraph per call",
                   "precision": args.precision, "instances": len(shapes),
                   "sum_F_E": fe},
        "sequential_f64": {"s_per_batch": dt64, "tflops": 3.0 * fe / dt64 / 1e12, "speedup": dt64 / dt},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "kernel": "gemm_grouped",
                     "gemm_ms_per_step": gemm_ms, "tuned_classes": tuned,
                     "family_ms": {k: val[0] for k, val in fam.items() if not k.startswith("gemm_")}},
        "cpu_baseline": None,
        "accuracy_vs_f64": accuracy,
    }
    print(json.dumps(line))


def cpu_baseline_lno(x, nocc, nvir, target_s):
    """LNO oracle energy + VJP on a bounded run of the LNO job list (every a,
    then b >= c), timed on the host cores.  Every job costs the same
    12 o^3 (v + o) flops of W forward (plus the o^4 projector products), so
    the rate of a run stands for the list.  A job is credited three times its
    forward (forward and adjoint): the list has about 3 F_E / (12 o^3 (v+o))
    jobs, so this is the 9 F_E per solve the GPU path is credited with."""
    from oracle import ccsd_t_oracle as orc
    prep = orc.LnoPrepared(*[x[k] for k in ("mat", "t1T", "t2T", "mo_energy", "fvo", "ovoo", "ovov", "ovvv")])
    flops_per_job = 12.0 * nocc ** 3 * (nvir + nocc)
    nj = prep.njobs()

    def run(j0, j1):
        t0 = time.perf_counter()
        prep.energy(j0, j1)
        prep.energy_vjp(j0, j1)
        return time.perf_counter() - t0

    with orc.blas():
        j0 = nj // 4
        probe = max(2, min(64, nj // 8))
        t_probe = run(j0, j0 + probe)
        n = int(max(probe, min(nj - j0, probe * target_s / max(t_probe, 1e-3))))
        t = run(j0, j0 + n)
    rate = 3.0 * flops_per_job * n / t / 1e12
    return {"value": rate, "unit": "TFLOP/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"LNO energy+VJP of jobs [{j0}, {j0 + n}) of the {nj} (a, b >= c) of the nocc={nocc}, "
                      f"nvir={nvir} instance in {t:.1f} s, FP64 C/OpenMP with scipy's BLAS (one BLAS thread per "
                      f"OpenMP thread), credited 3 x 12 o^3 (v+o) per job"}


def main_lno(args):
    """LNO (T) with an occupied projector: whole energy+VJP solves.

    Credited flops: 9 F_E per energy+VJP, F_E = 2 o^3 v^3 (v + o).  The
    reference's LNO job list (all a, b >= c) is about 3x the canonical
    a >= b >= c list, so its forward W costs about 3 F_E and the adjoint of W
    twice that; the GPU engine forms the same 3 F_E forward as W (F_E) plus
    the slot-projected S (2 F_E), and 6 F_E of adjoint contractions (2 over
    G, 4 over the adjoint of S).  Its backward also recomputes W and S
    (3 F_E executed, not credited)."""
    import math
    import torch
    from pyscfad_amd import _lib
    from pyscfad_amd.lno import ccsd_t as lno
    from pyscfad_amd.synthetic import make_lno_instance, flops_energy

    o, v = args.nocc, args.nvir
    log(f"[bench] building the LNO instance nocc={o} nvir={v} seed={args.seed}")
    x = make_lno_instance(o, v, args.seed, mat="gram", integral_scale=1.0 / math.sqrt(o + v))
    keys = ("mat", "t1T", "t2T", "mo_energy", "fvo", "ovoo", "ovov", "ovvv")
    a = [torch.as_tensor(x[k]).cuda().contiguous() for k in keys]
    torch.cuda.synchronize()
    W, K = max(args.warmup, 1), args.steps
    fe = flops_energy(o, v)
    credit = 9.0
    wsp = {}

    def workspace(prec):
        if prec not in wsp:
            wsp.clear()
            torch.cuda.empty_cache()
            wsp[prec] = torch.empty(lno.workspace_bytes(o, v, prec, with_vjp=True), dtype=torch.uint8, device="cuda")
        return wsp[prec]

    def step(prec, timing=False):
        _lib.set_timing(timing)
        e = lno.energy(*a, precision=prec, workspace=workspace(prec))
        tf = _lib.last_timing() if timing else None
        bars = lno.energy_vjp(*a, et_bar=1.0, precision=prec, workspace=workspace(prec))
        tb = _lib.last_timing() if timing else None
        _lib.set_timing(False)
        fam = {k: (tf[k][0] + tb[k][0], tf[k][1] + tb[k][1]) for k in tf} if timing else None
        return e, bars, fam

    def timed(prec, n):
        log(f"[bench] timing {n} LNO step(s) ({prec})")
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        e = bars = None
        for _ in range(n):
            bars = None
            e, bars, _ = step(prec)
        torch.cuda.synchronize()
        return (time.perf_counter() - c0) / n, e, bars

    tuned = None
    if not args.no_autotune:
        # the tuning table is keyed by size and shared with the canonical
        # path, whose contraction families the LNO form runs
        log(f"[bench] autotuning tile shapes ({args.precision})")
        from pyscfad_amd.cc import ccsd_t
        tuned = ccsd_t.autotune(o, v, args.precision)
        if args.precision != "f64" and not args.no_compare:
            ccsd_t.autotune(o, v, "f64")
    torch.cuda.reset_peak_memory_stats()
    fam = None
    for w in range(W):
        log(f"[bench] LNO warm-up step {w} ({args.precision})" + (", kernel timing on" if w == 0 else ""))
        _, _, f = step(args.precision, timing=(w == 0))
        fam = fam or f
    dt, e, bars = timed(args.precision, K)
    peak_hbm = torch.cuda.max_memory_allocated() / 1e9
    gemm_ms = fam["gemm"][0]
    achieved = credit * fe / (gemm_ms / 1e3) / 1e12
    peak = MI355X_PEAK_TFLOPS[args.precision]
    f64 = accuracy = None
    if args.precision != "f64" and not args.no_compare:
        step("f64")
        dt64, e64, bars64 = timed("f64", K)
        f64 = {"tflops": credit * fe / dt64 / 1e12, "s_per_solve": dt64, "speedup_over_f64": dt64 / dt}
        rel = {}
        for n, b, r in zip(keys, bars, bars64):
            rel[n] = max_rel(b, r)[0]
        accuracy = {"dE_Eh": abs(float(e) - float(e64)), "E_f64": float(e64), "max_rel": rel}
    cpu = None
    if not args.no_cpu:
        log(f"[bench] LNO CPU baseline sample (~{args.cpu_seconds:.0f} s)")
        cpu = cpu_baseline_lno(x, o, v, args.cpu_seconds)
    line = {
        "metric": "LNO (T) energy+VJP TFLOP/s (9*F_E credited) & s/solve",
        "value": credit * fe / dt / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dt * 1e3,
        "s_per_solve": dt,
        "peak_hbm_GB": peak_hbm,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": f"LNO (T) instance nocc={o} nvir={v} seed={args.seed}, integral_scale "
                               f"1/sqrt(n), Gram projector of {o // 2 + 1} local orbitals; whole energy+VJP "
                               f"solve per step", "precision": args.precision, "E_last_step": float(e)},
        "f64": f64,
        "bf16x6": b6,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "kernel": "gemm_grouped",
                     "credited": "9 F_E of the first warm-up solve over its gemm_grouped device time (the "
                                 "backward's W and S recomputes run in the same kernel and are not credited)",
                     "gemm_ms_per_step": gemm_ms,
                     "tiles": tuned,
                     "family_ms": {k: val[0] for k, val in fam.items()}},
        "cpu_baseline": cpu,
        "accuracy_vs_f64": accuracy,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
