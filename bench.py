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


def main():
    ap = argparse.ArgumentParser()
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
