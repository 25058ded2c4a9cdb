#!/usr/bin/env python3
"""Throughput of the per-move wavefunction path on one MI355X.

A step is one metrop1 sweep of the whole walker population: for every
electron a drift-diffusion proposal, psie (distances, jastrow4e, orbitalse,
Sherman-Morrison ratio), the new drift, accept/reject and commit; then one
hpsi (full Jastrow, orbitals, LU inverse, ECP quadrature, local energy) per
walker.  Workload: 40 nuclei (20 Z_eff=4 with s/p non-local ECP + 20 H),
160 electrons (80/80), 600 AOs, 12-point quadrature, 4096 walkers.

Each path first runs the autotuner (champ_gpu_autotune: GEMM kinds and the
quadrature form timed on the live walkers).  The same walkers and RNG streams
are run on the plain FP64 path and on the mixed-precision path (FP32 Jastrow, per-move AOs and quadrature Jastrow terms;
FP64 determinants, inverse refresh every 80 moves and local energy).  `value`
is the mixed path's rate; the FP64 rate, their ratio and the per-walker
local-energy and log-psi differences on identical configurations are in the
same line.  The timed region launches every kernel with HIP events around
each kernel class (the per-kernel times and the roofline come from it); the
same sweeps replayed from the captured HIP graph are timed after it.

A second workload, reported beside the headline (`heterogeneous_batch`):
eight molecules of 20-300 electrons with 4096 down to 512 walkers swept
together (champ_gpu_metrop1_batch, FP64 and mixed) against the same
populations swept one molecule after the other on the FP64 path.

Prints one JSON line (walker-moves/s = walkers * electrons * sweeps / s).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Work per launch of the timed kernel classes.
#
# Reference count (the per-term operation count of CHAMP's loops, the basis of
# the work-equivalent rates):
#   move (jastrow4e + AO):  (Ne-1)*Nc*(25T+60) + AO
#   hpsi_nonloc (per quadrature point): (Ne-1)*(Nc*(4T+40)+60) + Nc*40 + Nao*42
#   hpsi_jastrow: Ne(Ne-1)/2 * Nc * (54T+80)
#   hpsi_nonloc_psi (the psi ratio at each quadrature point, its own kernel):
#       Nao*42 per point
# Executed matrix-core work of the quadrature kernel (k_nonloc_mma): every
# group of up to four active pairs of a walker runs, for each 16-partner tile
# and each step of four centres, 3 row tiles x naz table entries of one
# 16x16x4 MFMA (2048 flop).  T = e-e-n terms (13 at nordc = 4), naz = the
# non-zero A_kn table entries (12 at nordc = 4), Nc centres, Nao AOs.


def flops_move_jastrow(ne, nc, nt, nao, nrad_total):
    return (ne - 1) * nc * (25 * nt + 60) + nao * 20 + nrad_total * 40


def flops_nonloc_point(ne, nc, nt, nao):
    return (ne - 1) * (nc * (4 * nt + 40) + 60) + nc * 40 + nao * 42


def een_naz(nordc):
    """Non-zero entries of the e-e-n tables A_kn (een_poly.hpp): term (k, l, m)
    fills A[k][m] and A[k][l + m]."""
    from champ_amd import system as S
    return len({(k, m) for k, l, m in S.een_terms(nordc)} | {(k, l + m) for k, l, m in S.een_terms(nordc)})


def mfma_flops_nonloc(sy, nxq):
    """Executed MFMA flop of one k_nonloc_mma launch for the walkers' active
    pair counts (nxquad / nquad per walker)."""
    nq = max(1, len(sy.wq))
    pairs = np.asarray(nxq, dtype=np.int64) // nq
    groups = int(np.sum((pairs + 3) // 4))
    per_group = -(-sy.nelec // 16) * -(-sy.ncent // 4) * 3 * een_naz(sy.nordc) * 2048
    return float(groups) * per_group


def run_path(sy, W, precision, args, label):
    from champ_amd import system as S
    from champ_amd.engine import Engine

    eng = Engine(sy, W, precision=precision)
    eng.set_walkers(sy.initial_walkers(W, seed=1))
    eng.seed_rng(S.seed_from_string())
    g0 = eng.hpsi()
    tuned = eng.autotune() if args.autotune else None
    print(f"[bench] {label}: {sy.name}: Ne={sy.nelec} Nc={sy.ncent} Nao={sy.nbasis} W={W}; warmup {args.warmup}; "
          f"autotuned {tuned}", flush=True)
    if args.warmup:
        eng.metrop1(args.warmup, args.tau)
    eng.enable_timing(True)
    t0 = time.perf_counter()
    # blocks of one sweep (champ_gpu_vmc): the blocked mean and error of E_L
    vr = eng.vmc(args.steps, 1, nblkeq=0, tau=args.tau)
    t1 = time.perf_counter()
    kt = eng.kernel_times()
    eng.enable_timing(False)
    nxq = eng.array("nxquad")
    # the same number of sweeps replayed from the captured sweep graph, no events
    gwall = None
    if args.graph_check:
        eng.set_graphs(True)
        eng.metrop1(1, args.tau)  # captures the graph
        tg0 = time.perf_counter()
        eng.metrop1(args.steps, args.tau)
        gwall = time.perf_counter() - tg0
        eng.set_graphs(False)
    wall = t1 - t0
    print(f"[bench] {label}: {args.steps} sweeps in {wall:.3f}s; acceptance {vr['accept']:.3f}; "
          f"<E_L> {vr['efin']:.6f} +- {vr['eerr']:.6f}; flagged {vr['nflagged']}", flush=True)
    for k, (ms, n) in kt.items():
        if n:
            print(f"[bench]   {label} {k:20s} {ms:10.2f} ms  {n:6d} launches  {ms / n:9.4f} ms/launch", flush=True)
    if gwall is not None:
        print(f"[bench] {label}: graph replay {args.steps} sweeps in {gwall:.3f}s", flush=True)
    return {"wall": wall, "value": W * sy.nelec * args.steps / wall, "kt": kt, "vmc": vr, "nxq": nxq, "hpsi0": g0,
            "graph_value": None if gwall is None else W * sy.nelec * args.steps / gwall, "tuned": tuned, "eng": eng}


def accuracy_at_sampled(e64, e32, W):
    """The mixed path's per-walker errors at the configurations its timed
    sweeps reached: both paths take the mixed population's positions and
    streams, draw the same quadrature rotation and evaluate hpsi."""
    x, st = e32.get_walkers(), e32.get_rng()
    out = []
    for e in (e64, e32):
        e.set_walkers(x)
        e.set_rng(st)
        e.rotqua()
        out.append(e.hpsi())
    g64, gmx = out
    de = np.abs(gmx["energy"] - g64["energy"])
    dl = np.abs(np.log(np.abs(gmx["psid"])) + gmx["psij"] - np.log(np.abs(g64["psid"])) - g64["psij"])
    return {"sample": f"all {W} walkers at the configurations after the mixed path's timed sweeps, identical "
                      f"positions and quadrature on both paths",
            "dE_L_max_Ha": float(de.max()), "dE_L_p99_Ha": float(np.percentile(de, 99)),
            "dE_L_mean_Ha": float(de.mean()),
            "dE_L_rel_max": float((de / np.maximum(1.0, np.abs(g64["energy"]))).max()),
            "dlnpsi_max": float(dl.max()), "dlnpsi_mean": float(dl.mean()),
            "E_L_abs_median_Ha": float(np.median(np.abs(g64["energy"])))}


def run_hetero(args):
    """Eight molecules (20-300 electrons, Nao ~ 3.75 Ne, ECP, 4096 down to 512
    walkers) swept together with champ_gpu_metrop1_batch, each population on
    its own stream, beside the same populations swept one molecule after the
    other on the FP64 path.  Shape rules, no autotuning."""
    from champ_amd import system as S
    from champ_amd.engine import PREC_FP64, PREC_MIXED, Engine, metrop1_batch

    mols = S.hetero_batch(scale=args.hetero_scale)
    moves = sum(W * sy.nelec for sy, W in mols)
    steps = args.hetero_steps

    def make(prec):
        engs = []
        for sy, W in mols:
            e = Engine(sy, W, precision=prec)
            e.set_walkers(sy.initial_walkers(W, seed=1))
            e.seed_rng(S.seed_from_string())
            e.hpsi()
            engs.append(e)
        return engs

    out = {"molecules": [{"electrons": sy.nelec, "aos": sy.nbasis, "walkers": W} for sy, W in mols],
           "walker_moves_per_sweep": moves, "steps": steps, "unit": "walker-moves/s"}
    engs = make(PREC_FP64)
    for e in engs:
        e.metrop1(1, args.tau)
    t0 = time.perf_counter()
    for e in engs:
        e.metrop1(steps, args.tau)
    t_seq = time.perf_counter() - t0
    for e in engs:
        e.close()
    out["sequential_fp64"] = round(moves * steps / t_seq, 1)
    print(f"[bench] hetero: {len(mols)} molecules, {moves} walker-moves per sweep; sequential fp64 {steps} sweeps in "
          f"{t_seq:.3f}s", flush=True)
    for prec, key in ((PREC_FP64, "batched_fp64"), (PREC_MIXED, "batched_mixed")):
        engs = make(prec)
        metrop1_batch(engs, 1, args.tau)
        t0 = time.perf_counter()
        metrop1_batch(engs, steps, args.tau)
        tb = time.perf_counter() - t0
        for e in engs:
            e.close()
        out[key] = round(moves * steps / tb, 1)
        print(f"[bench] hetero: {key} {steps} sweeps in {tb:.3f}s", flush=True)
    out["batched_mixed_over_sequential_fp64"] = round(out["batched_mixed"] / out["sequential_fp64"], 3)
    out["batched_fp64_over_sequential_fp64"] = round(out["batched_fp64"] / out["sequential_fp64"], 3)
    return out


# HBM bytes per launch of the matrix-core quadrature kernel (k_nonloc_mma,
# the psi ratio excluded: it has its own kernel) at the default workload
# (c160, 4096 walkers, the timed sweep after ten warm-up sweeps): FETCH_SIZE
# x 2 + WRITE_SIZE x 1 from two rocprofv3 --pmc passes over the default bench,
# scaled by the same counters on a device-to-device copy of known size
# (profiles/k_nonloc_c160_traffic.md)
NONLOC_TRAFFIC = {("c160", 4096, "fp64"): 1.29e9, ("c160", 4096, "mixed"): 1.20e9}

PEAK = {"fp64-mfma": 78.6, "fp32-mfma": 157.3, "fp64-valu": 78.6,
        # FP32 without packed math (this build never emits v_pk_* FP32): one
        # FMA per lane per cycle, the FP64 rate
        "fp32-valu-unpacked": 78.6, "hbm": 8000.0}


def per_kernel(sy, W, kt, nxq, path, acc):
    """Each timed kernel class against its own roof: work per launch (stated
    per class) over its mean launch time."""
    from champ_amd import system as S
    nt = len(S.een_terms(sy.nordc))
    nrad = int(sum(sy.nrbas[t] for t in sy.iwctype))
    ns, ne, nc, nb, no = max(sy.nup, sy.ndn), sy.nelec, sy.ncent, sy.nbasis, sy.norb
    chunk = int(max(1, min(W, 3.0e9 / (ne * 5 * nb * 8))))
    wchunk = W / -(-W // chunk)  # walkers per hpsi chunk launch, on average
    valu = "fp32-valu-unpacked" if path == "mixed" else "fp64-valu"
    mfma = "fp32-mfma" if path == "mixed" else "fp64-mfma"
    work = {
        "move_jastrow_ao": (W * flops_move_jastrow(ne, nc, nt, nb, nrad), valu, "TFLOP/s",
                            "reference count (Ne-1) Nc (25T+60) + AO per walker", False),
        "move_mo_gemm": (2.0 * 4 * W * no * nb, "fp64-mfma", "TFLOP/s", "2 (4W) norb Nao", True),
        "move_ratio_commit": (W * 8.0 * (ns + 4 * no + 3 * ne) + W * acc * (16.0 * ns * ns + 8.0 * (3 * no + 20 * ne + 2 * nc)),
                              "hbm", "GB/s", "inverse row + orbitals per walker; inverse read+write and the "
                              "Jastrow/orbital commits per accepted walker", True),
        "hpsi_jastrow": (W * 0.5 * ne * (ne - 1) * nc * (54 * nt + 80), valu, "TFLOP/s",
                         "reference count Ne(Ne-1)/2 Nc (54T+80) per walker", False),
        "hpsi_mo_gemm": (2.0 * wchunk * ne * 5 * no * nb, "fp64-mfma", "TFLOP/s", "2 (5 Ne W_chunk) norb Nao", True),
        "hpsi_slater_lu": (wchunk * 2 * 2.0 * ns ** 3, "fp64-valu", "TFLOP/s", "2 Nsigma^3 per spin matrix", True),
        "hpsi_nonloc": (mfma_flops_nonloc(sy, nxq), mfma, "TFLOP/s", "executed 16x16x4 MFMA flop (groups x "
                        "partner tiles x centre quads x 3 row tiles x naz)", True),
        "hpsi_nonloc_psi": (float(np.sum(nxq)) * nb * 42, "fp64-valu", "TFLOP/s",
                            "reference count Nao*42 per quadrature point (AO values at the point, dot with W)", False),
    }
    out = {}
    for k, (ms, n) in kt.items():
        if not n or k not in work:
            continue
        w, pipe, unit, basis, util = work[k]
        t = ms / n / 1e3
        ach = w / t / (1e9 if unit == "GB/s" else 1e12)
        out[k] = {"pipe": pipe, "achieved": round(ach, 3), "peak": PEAK[pipe], "unit": unit,
                  "frac": round(ach / PEAK[pipe], 4), "avg_launch_ms": round(ms / n, 4), "work_per_launch": w,
                  "work_basis": basis,
                  # False: a reference per-term count the kernel does not execute
                  # (its factorised polynomial does less), a work-equivalent rate
                  "utilisation": util}
    return out


def roofline(sy, W, kt, nxq, path, acc):
    """The dominant kernel class (most time in the timed sweeps) against its roof."""
    from champ_amd import system as S
    pk = per_kernel(sy, W, kt, nxq, path, acc)
    dname = max((k for k in pk), key=lambda k: kt[k][0])
    d = pk[dname]
    bound = "hbm" if d["pipe"] == "hbm" else "mfma"
    r = {"bound": bound, "kernel": dname, "pipe": d["pipe"], "achieved": d["achieved"], "peak": d["peak"],
         "unit": d["unit"], "frac": d["frac"], "traffic": None, "avg_launch_ms": d["avg_launch_ms"],
         "work_per_launch": d["work_per_launch"], "work_basis": d["work_basis"]}
    if dname == "hpsi_nonloc":
        npts = float(np.sum(nxq))
        nt = len(S.een_terms(sy.nordc))
        ref = npts * flops_nonloc_point(sy.nelec, sy.ncent, nt, sy.nbasis)
        t = d["avg_launch_ms"] / 1e3
        r["traffic"] = NONLOC_TRAFFIC.get((sy.name, W, path))
        r["traffic_source"] = "profiles/k_nonloc_c160_traffic.md" if r["traffic"] else None
        r["quadrature_points_per_launch"] = int(npts)
        r["work_equivalent"] = {
            "achieved": round(ref / t / 1e12, 3), "unit": "TFLOP/s",
            "basis": "reference per-term count per point x points; the factorised kernel executes ~7x fewer "
                     "operations, so this rate is not a utilisation"}
    return r


def cpu_baseline(sy, args):
    """The oracle (one walker chain per thread, -O2 at the host's x86-64
    level) on the host cores this process may use, at most 16 (the GPU
    host's share per GPU); two walkers per thread for a fixed number of sweeps."""
    import ctypes as C
    from champ_amd import system as S
    from oracle import ffi
    model, ncpu, navail, _ = ffi.host_cpu()
    nth = args.cpu_threads or min(16, navail)
    L, level = ffi.baseline_lib()
    nwc = 2 * nth
    x0 = sy.initial_walkers(nwc, seed=1)
    streams = ffi.seeded_streams(S.seed_from_string(), nwc)
    es, ac = np.zeros(1), np.zeros(1)
    cs = sy.as_ctypes()
    tc = L.orc_vmc_run(C.byref(cs), nwc, ffi.P(x0), ffi.U(streams), args.cpu_sweeps, args.tau, nth, ffi.P(es),
                       ffi.P(ac))
    return {"value": round(nwc * sy.nelec * args.cpu_sweeps / tc, 2), "unit": "walker-moves/s", "cores": nth,
            "kind": "port", "cpu_model": model, "nproc": ncpu, "cpus_available": navail,
            "build": f"gcc -O2 -march={level} (the highest level this host supports)",
            "sample": f"{nwc} walkers x {args.cpu_sweeps} metrop1 sweeps of the same system, two walker chains per "
                      f"thread on {nth} threads, {tc:.1f} s"}


def _kernel_table(kt):
    tot = sum(ms for ms, n in kt.values() if n)
    return {k: {"ms": round(ms, 3), "launches": n, "ms_per_launch": round(ms / n, 4), "share": round(ms / tot, 4)}
            for k, (ms, n) in sorted(kt.items(), key=lambda kv: -kv[1][0]) if n}


def _cpu_threads(args):
    from oracle import ffi
    model, ncpu, navail, _ = ffi.host_cpu()
    return model, ncpu, navail, args.cpu_threads or min(16, navail)


def bench_metrop6(args):
    """metrop6 (CHAMP's default VMC mover) on the headline system: the same
    population as the metrop1 headline, blocks of one sweep, FP64 and mixed;
    the oracle's metrop6 sweeps on the host cores beside it."""
    from concurrent.futures import ThreadPoolExecutor

    from champ_amd import system as S
    from champ_amd.engine import PREC_FP64, PREC_MIXED, Engine

    sy = S.config(args.system)
    W = args.walkers
    res = {}
    for prec, label in ((PREC_FP64, "fp64"), (PREC_MIXED, "mixed")):
        eng = Engine(sy, W, precision=prec)
        eng.set_walkers(sy.initial_walkers(W, seed=1))
        eng.seed_rng(S.seed_from_string())
        eng.hpsi()
        if args.autotune:
            eng.autotune()
        eng.metrop6(args.warmup)
        eng.enable_timing(True)
        t0 = time.perf_counter()
        vr = eng.vmc(args.steps, 1, nblkeq=0, imetro=6)
        wall = time.perf_counter() - t0
        kt = eng.kernel_times()
        eng.close()
        res[label] = {"value": W * sy.nelec * args.steps / wall, "wall": wall, "vmc": vr, "kernels": _kernel_table(kt)}
        print(f"[bench] metrop6 {label}: {args.steps} sweeps in {wall:.3f}s, acceptance {vr['accept']:.3f}", flush=True)
    cpu = None
    if not args.no_cpu:
        from oracle import ffi  # the CPU baseline leg only
        model, ncpu, navail, nth = _cpu_threads(args)
        cs = sy.as_ctypes()
        x0 = sy.initial_walkers(nth, seed=1)
        st = ffi.seeded_streams(S.seed_from_string(), nth)
        walkers = []
        for k in range(nth):
            w = ffi.walker(sy, cs)
            w.set_x(x0[k])
            w.set_rng(st[k])
            w.set_quad(sy.xq0, sy.yq0, sy.zq0)
            w.hpsi()
            walkers.append(w)
        nsw = max(1, args.cpu_sweeps // 2)

        def run(w):
            for _ in range(nsw):
                w.sweep6()
                w.hpsi()

        t0 = time.perf_counter()
        with ThreadPoolExecutor(nth) as ex:
            list(ex.map(run, walkers))
        tc = time.perf_counter() - t0
        cpu = {"value": round(nth * sy.nelec * nsw / tc, 2), "unit": "walker-moves/s", "cores": nth, "kind": "port",
               "cpu_model": model, "nproc": ncpu, "cpus_available": navail,
               "sample": f"{nth} walker chains x {nsw} metrop6 sweeps (each followed by hpsi), one chain per thread, "
                         f"{tc:.1f} s"}
    out = {"metric": "walker-moves/sec, metrop6 (CHAMP's default VMC mover)", "value": round(res["mixed"]["value"], 1),
           "unit": "walker-moves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(1e3 * res["mixed"]["wall"] / args.steps, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": f"metrop6 sweep, {sy.nelec} e, {sy.ncent} nuclei, {sy.nbasis} AOs, {W} walkers",
                      "walkers": W, "electrons": sy.nelec},
           "fp64_value": round(res["fp64"]["value"], 1),
           "mixed_over_fp64": round(res["mixed"]["value"] / res["fp64"]["value"], 3),
           "acceptance": {k: round(v["vmc"]["accept"], 4) for k, v in res.items()},
           "per_kernel": {k: v["kernels"] for k, v in res.items()}, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_dmc(args):
    """DMC on the reference's own DMC CI case (butadiene, 500 determinants,
    BFD, tau 0.05, T-moves as its input asks): walker-steps per second of the
    population after equilibration, FP64 and mixed; the oracle's DMC on a few
    walkers beside it."""
    from champ_amd import champ_io as io
    from champ_amd import system as S
    from champ_amd.engine import PREC_FP64, PREC_MIXED, Engine

    sy = io.load_input(os.path.join(ROOT, "tests", "golden", "champ_inputs", "C4H6_dmc", "dmc.inp"), name="C4H6_dmc")
    cap = args.walkers
    nconf = cap // 2
    res = {}
    for prec, label in ((PREC_FP64, "fp64"), (PREC_MIXED, "mixed")):
        eng = Engine(sy, cap, precision=prec)
        eng.set_walkers(sy.initial_walkers(cap, seed=3))
        eng.seed_rng(S.seed_from_string())
        eng.hpsi()
        eng.vmc(2, 10, nblkeq=0, tau=0.2)
        # the branching stream: any non-zero xoshiro state (a fixed one, from numpy)
        pop = np.random.default_rng(2024).integers(1, 2 ** 63 - 1, size=4, dtype=np.uint64)
        eng.dmc_init(nconf, 0.05, -26.3, pop)
        eng.dmc_set_icasula(-1)
        for _ in range(args.warmup):
            eng.dmc_step()
        eng.enable_timing(True)
        nws = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            nws += eng.dmc_step()["nwalk"]
        wall = time.perf_counter() - t0
        kt = eng.kernel_times()
        eng.close()
        res[label] = {"value": nws / wall, "wall": wall, "kernels": _kernel_table(kt), "walkers": nws / args.steps}
        print(f"[bench] dmc {label}: {args.steps} steps in {wall:.3f}s, {nws / args.steps:.0f} walkers", flush=True)
    cpu = None
    if not args.no_cpu:
        from oracle import ffi  # the CPU baseline leg only
        model, ncpu, navail, _ = _cpu_threads(args)
        ncw, nst = 16, 30
        x0 = sy.initial_walkers(ncw, seed=3)
        st = ffi.seeded_streams(S.seed_from_string(), 2 * ncw + 1)
        orc = ffi.Dmc(sy, np.concatenate([x0, x0]), st[:2 * ncw], ncw, 0.05, -26.3, st[2 * ncw], icasula=-1)
        t0 = time.perf_counter()
        nws = 0
        for _ in range(nst):
            nws += orc.step()["nwalk"]
        tc = time.perf_counter() - t0
        cpu = {"value": round(nws / tc, 3), "unit": "walker-steps/s", "cores": 1, "kind": "port",
               "cpu_model": model, "nproc": ncpu, "cpus_available": navail,
               "sample": f"the oracle's DMC (one thread), {ncw} walkers x {nst} steps of the same case, {tc:.1f} s"}
    ne = sy.nelec
    out = {"metric": "DMC walker-steps/sec (dmc_ps with T-moves), reference DMC CI case",
           "value": round(res["mixed"]["value"], 1), "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * res["mixed"]["wall"] / args.steps, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
           "data": "the reference's DMC-C4H6-ci1010_pVTZ-500-dets input; walkers from a short VMC",
           "config": {"workload": f"DMC step, butadiene {ne} e, 500 determinants, BFD, tau 0.05, icasula -1, "
                                  f"{nconf} target walkers", "capacity": cap, "electrons": ne},
           "walker_moves_per_s": round(res["mixed"]["value"] * ne, 1),
           "fp64_value": round(res["fp64"]["value"], 1),
           "mixed_over_fp64": round(res["mixed"]["value"] / res["fp64"]["value"], 3),
           "per_kernel": {k: v["kernels"] for k, v in res.items()}, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_sr(args):
    """SR sampling on the reference's 5000-determinant optimisation case: the
    derivatives of all 12033 parameters (13 Jastrow, 4999 determinant, 7021
    orbital excitations), their map and the S, <O>, <E O> sums accumulated on
    the GPU per sample, then the SR step (S z on the GPU)."""
    from champ_amd import champ_io as io
    from champ_amd import optwf
    from champ_amd import system as S
    from champ_amd.engine import DERIV_MAPPED, Engine

    d = os.path.join(ROOT, "tests", "golden", "champ_inputs", "C4H6_5k")
    sy = io.load_input(os.path.join(d, "vmc_optimization_5000.inp"), name="C4H6_5k", nadorb=280)
    der = io.read_jastrow_der(os.path.join(d, "jastrow.der"), sy.nctype, sy.nspin2b)
    irrep = io.read_symmetry(os.path.join(d, "ci1010_pVTZ_1_symmetry.sym"), sy.coef.shape[0])
    terms = optwf.optorb_define(sy.iworbd, sy.nup, irrep, nadorb=280)
    p = optwf.ChampParams.build(sy, der, ci=True, orb_terms=terms)
    W = args.walkers
    eng = Engine(sy, W)
    eng.set_walkers(sy.initial_walkers(W, seed=3))
    eng.seed_rng(S.seed_from_string())
    eng.hpsi()
    eng.set_orb_terms(terms)
    eng.set_deriv_map(p.kinds, p.m)
    eng.metrop6(args.warmup)
    eng.sr_accumulate(DERIV_MAPPED, reset=True)
    eng.enable_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.sr_accumulate(DERIV_MAPPED)
    ts = time.perf_counter() - t0
    kt = eng.kernel_times()
    t0 = time.perf_counter()
    dparm, info = eng.sr_solve(sr_tau=0.025, sr_eps=1e-3, sr_adiag=0.01)
    tsol = time.perf_counter() - t0
    eng.close()
    print(f"[bench] sr: {args.steps} samples in {ts:.3f}s; solve {tsol:.3f}s ({info['cg_iters']} CG iterations)",
          flush=True)
    out = {"metric": "SR samples/sec (derivatives of every parameter + S, <O>, <E O> sums), "
                     "5000-determinant optimisation case",
           "value": round(W * args.steps / ts, 1), "unit": "walker-samples/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * ts / args.steps, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "the reference's VMC-C4H6-ci1010_pVTZ-05000-dets input",
           "config": {"workload": f"SR sample of {W} walkers, {p.nparm} parameters", "walkers": W,
                      "parameters": p.nparm},
           "solve_s": round(tsol, 3), "cg_iters": info["cg_iters"], "per_kernel": _kernel_table(kt),
           "cpu_baseline": None,
           "cpu_baseline_note": "the oracle has no derivative path (its checks are finite differences of ln psi, "
                                "2 nparm evaluations per walker)"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="metrop1", choices=["metrop1", "metrop6", "dmc", "sr"],
                    help="metrop1: the headline; the others measure the VMC default mover, DMC and SR sampling")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed sweeps first: the walkers spread from their initial Gaussian clouds and the ECP pair count settles")
    ap.add_argument("--walkers", type=int, default=4096)
    ap.add_argument("--system", default="c160")
    ap.add_argument("--tau", type=float, default=0.3)
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: the usable cores, at most 16)")
    ap.add_argument("--cpu-sweeps", type=int, default=15)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true", help="time the plain FP64 path only")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the shape rules instead of timing the candidate kernel forms first")
    ap.add_argument("--no-graph-check", dest="graph_check", action="store_false",
                    help="skip the second timing of the sweeps replayed from the captured HIP graph")
    ap.add_argument("--no-hetero", dest="hetero", action="store_false",
                    help="skip the heterogeneous batch (eight molecules of 20-300 electrons swept together)")
    ap.add_argument("--hetero-steps", type=int, default=2)
    ap.add_argument("--hetero-scale", type=float, default=1.0, help="scale of the batch's walker counts")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs the walker population on one GPU; "
              "multi-GPU runs (walker sharding, see DESIGN.md) are not part of this build.", file=sys.stderr)
        sys.exit(2)

    if args.workload != "metrop1":
        {"metrop6": bench_metrop6, "dmc": bench_dmc, "sr": bench_sr}[args.workload](args)
        return

    from champ_amd import system as S
    from champ_amd.engine import PREC_FP64, PREC_MIXED

    sy = S.config(args.system)
    W = args.walkers
    r64 = run_path(sy, W, PREC_FP64, args, "fp64")
    rmx = None if args.fp64_only else run_path(sy, W, PREC_MIXED, args, "mixed")
    head = rmx if rmx is not None else r64

    acc = None
    if rmx is not None:
        acc = accuracy_at_sampled(r64["eng"], rmx["eng"], W)
        e64, emx = r64["vmc"], rmx["vmc"]
        acc["vmc_mean_E_fp64"] = {"E": e64["efin"], "err": e64["eerr"]}
        acc["vmc_mean_E_mixed"] = {"E": emx["efin"], "err": emx["eerr"]}
        acc["vmc_mean_diff_in_fp64_errors"] = round(abs(emx["efin"] - e64["efin"]) / max(e64["eerr"], 1e-300), 3)
        acc["vmc_mean_note"] = (f"blocked mean of E_L over the {args.steps} timed sweeps of each run (blocks of one "
                                f"sweep of all walkers, same seeds)")
        print(f"[bench] accuracy at sampled configurations: |dE_L| max {acc['dE_L_max_Ha']:.3e} Ha p99 "
              f"{acc['dE_L_p99_Ha']:.3e} mean {acc['dE_L_mean_Ha']:.3e}; |dln psi| max {acc['dlnpsi_max']:.3e}; "
              f"VMC means {e64['efin']:.4f} +- {e64['eerr']:.4f} (fp64) {emx['efin']:.4f} +- {emx['eerr']:.4f} "
              f"(mixed)", flush=True)
    for r in (r64, rmx):
        if r is not None:
            r["eng"].close()

    het = run_hetero(args) if args.hetero else None
    cpu = None if args.no_cpu else cpu_baseline(sy, args)

    acc_rate = head["vmc"]["accept"]
    out = {
        "metric": "walker-moves/sec at 160 e\u207b on 1 MI355X, mixed vs FP64 GPU; \u0394E_L vs FP64",
        "value": round(head["value"], 1),
        "unit": "walker-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * head["wall"] / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if rmx is None else "f32+f64",
        "data": "synthetic (seeded generator: geometry, basis, MOs, Jastrow, ECP; walkers Gaussian around nuclei)",
        "config": {"workload": f"metrop1 sweep, {sy.nelec} e, {sy.ncent} nuclei, {sy.nbasis} AOs, ECP + 12-pt "
                               f"quadrature, {W} walkers", "walkers": W, "electrons": sy.nelec,
                   "nuclei": sy.ncent, "aos": sy.nbasis, "precision": "fp64" if rmx is None else "mixed",
                   "path": "plain FP64 GPU path" if rmx is None else "mixed-precision GPU path",
                   "parallelism": "walkers on 1 GPU"},
        "fp64_value": round(r64["value"], 1),
        "mixed_over_fp64": None if rmx is None else round(rmx["value"] / r64["value"], 3),
        "walker_sweeps_per_s": round(W * args.steps / head["wall"], 2),
        "autotuned": head["tuned"],
        "graph_replay": None if head["graph_value"] is None else {
            "value": round(head["graph_value"], 1), "unit": "walker-moves/s",
            "note": "the same number of sweeps replayed from the captured sweep graph, without per-kernel events "
                    "(value and roofline come from the event-timed eager launches)"},
        "acceptance": round(acc_rate, 4),
        "accuracy_mixed_vs_fp64": acc,
        "roofline": roofline(sy, W, head["kt"], head["nxq"], "fp64" if rmx is None else "mixed", acc_rate),
        "roofline_fp64_path": roofline(sy, W, r64["kt"], r64["nxq"], "fp64", r64["vmc"]["accept"]),
        "per_kernel": per_kernel(sy, W, head["kt"], head["nxq"], "fp64" if rmx is None else "mixed", acc_rate),
        "per_kernel_fp64_path": per_kernel(sy, W, r64["kt"], r64["nxq"], "fp64", r64["vmc"]["accept"]),
        "heterogeneous_batch": het,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
