#!/usr/bin/env python3
"""Benchmark: SCFT iterations/s of the PmmmDCT path at 128^3, Ns = 200.

One "step" is one full SCFT iteration (crysSCFT_sample.cpp:476-508): 400
contour steps of the two propagator chains, the densities, Q, the free
energy and the mixing update, all on the GPU with the fields resident in HBM.
Workload: f_A = 0.30, chiN = 25, cubic L = 4.2 Rg, 128^3 logical grid
(64^3 reduced), ds = 1/200 (NsA = 60, NsB = 140), BCC-seeded Pmmm fields,
simple mixing 0.1.

Prints ONE JSON line.  The roofline block is for the dominant kernel of the
contour step (the one-launch plane phase at this size, both chains per
launch).  Its launch duration is the mean per step of the iteration's own
propagation sequence, captured in a hipGraph and timed with HIP events on the
solver's stream (crys_scft_time_propagation); the same kernel launched alone
and repeated is reported beside it.  "contour_step" gives the whole step
against the step's minimum bytes.  The CPU baseline is the oracle (a C
restatement of the reference) timed on this host on three SCFT iterations of
the same workload on the process's CPU share, one iteration on one thread,
and the reference's triplet-only clock for both; the first iterations of
both GPU legs are compared with the oracle's.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_LOG = 128
FA, XN, L, DS, ACC = 0.30, 25.0, 4.2, 1.0 / 200, 0.1
HBM_PEAK_GBS = 8000.0
# MI355X dense matrix-core peaks (FP64 and FP32 MFMA)
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}

# Kernels of one contour step in issue order, by the number of launches: the
# one-launch plane phase (all axes in one length class <= 64), the fused z+y
# slab plus the x pass, or three axis passes.
STEP_KERNELS = {
    1: ["crys::plane_phase_half_kernel<double, true>" if os.environ.get("CRYS_PHASE_HALF") != "0"
        else "crys::plane_phase_kernel<double,64,true,1024>"],
    2: ["crys::slab_zy_kernel<double,64,true,true>",
        "crys::axis_pass_mfma_kernel<double,0,64,true,true>"],
    3: ["crys::axis_pass_mfma_kernel<double,2,64,true,true>",
        "crys::axis_pass_mfma_kernel<double,1,64,true,true>",
        "crys::axis_pass_mfma_kernel<double,0,64,true,true>"],
}
# HBM bytes per dispatch of both chains come from the rocprofv3 FETCH_SIZE
# and WRITE_SIZE passes of this workload (FETCH_SIZE scaled by its reading on
# a copy of known size), kept in this profile summary: the row of a kernel
# gives its calibrated MB per dispatch.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_plane_phase_pmmm128_fp64.md")


def measured_traffic(kernel):
    """(bytes per dispatch, profile row name) of `kernel` from TRAFFIC_PROFILE,
    or (None, None) when the profile has no row for it."""
    try:
        lines = open(TRAFFIC_PROFILE).read().splitlines()
    except OSError:
        return None, None
    want = kernel.replace(" ", "")
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) < 6 or not cells[0].startswith("`"):
            continue
        name = cells[0].split("`")[1]
        if name.replace(" ", "") == want:
            try:
                return float(cells[4]) * 1e6, name
            except ValueError:
                return None, None
    return None, None


def bytes_per_iteration(n, b, NsA, NsB):
    """Algorithmic HBM bytes of one SCFT iteration: 2 Ns' contour steps of
    3 n b each, the phi pass over both histories, 22 n doubles of field work."""
    Ns = NsA + NsB
    return 2 * Ns * 3 * n * b + 2 * (Ns + 2) * n * b + 22 * n * 8


def run_gpu(precision, steps, warmup, wA, wB):
    import crysscft_sample_amd as C

    solver = C.SolverPmmmDCT((N_LOG,) * 3, precision=precision)
    run = C.SCFT(solver, FA, DS, XN, L, ACC)
    run.set_fields(wA, wB)
    warm = run.run(max(warmup, 1), trace=True)
    C.lib().crys_device_synchronize()
    t0 = time.perf_counter()
    tr = run.run(steps, trace=True)
    C.lib().crys_device_synchronize()
    dt = time.perf_counter() - t0
    ms_step, ms_launch = run.time_step(reps=200)
    ms_seq, launches_seq = run.time_propagation(reps=5)
    phiA, phiB = run.fields()[:2]
    out = dict(iter_s=steps / dt, ms_per_iter=dt / steps * 1e3, trace=tr, warm=warm, NsA=run.NsA,
               NsB=run.NsB, ms_per_contour_step=ms_step, ms_per_launch=ms_launch,
               ms_per_step_in_sequence=ms_seq, launches_per_step=launches_seq,
               phiA=phiA, phiB=phiB)
    run.close()
    solver.close()
    return out


def checkpoint_timing(steps):
    """One JSON line: the headline workload with full histories and with
    checkpoints every ceil(sqrt(Ns)) slots (crys_scft_set_checkpointing):
    iterations/s, contour storage and the free-energy difference, FP64 and
    FP32; plus one gradient-and-stress evaluation of each."""
    import crysscft_sample_amd as C
    from crysscft_sample_amd import fields as F

    wfA, wfB = F.bcc_fields((N_LOG,) * 3, L, FA, XN)
    wA, wB = F.extract(wfA, "PmmmDCT"), F.extract(wfB, "PmmmDCT")
    out = {"metric": "SCFT iter/s at 128^3 Ns=200 PmmmDCT, full histories vs sqrt(Ns) "
                     "checkpoints", "unit": "iter/s", "higher_is_better": True, "steps": steps}
    for prec in ("fp64", "fp32"):
        res = {}
        for every in (0, None):
            s = C.SolverPmmmDCT((N_LOG,) * 3, precision=prec)
            run = C.SCFT(s, FA, DS, XN, L, ACC)
            Ns = run.NsA + run.NsB
            run.set_checkpointing(every if every is not None else math.ceil(math.sqrt(Ns)), 0)
            used, nbytes = run.checkpointing()
            run.set_fields(wA, wB)
            run.run(2)
            C.lib().crys_device_synchronize()
            t0 = time.perf_counter()
            tr = run.run(steps)
            C.lib().crys_device_synchronize()
            dt = time.perf_counter() - t0
            t1 = time.perf_counter()
            g = run.evaluate(wA, wB, stress=True)
            te = time.perf_counter() - t1
            res[used] = dict(iter_s=steps / dt, bytes=nbytes, trace=tr, eval_ms=te * 1e3, H=g["H"],
                             dL=g["dH_dL"])
            run.close()
            s.close()
        full, (every, ck) = res[0], [(k, v) for k, v in res.items() if k][0]
        out[prec] = {"every": every,
                     "full_iter_s": round(full["iter_s"], 3),
                     "checkpointed_iter_s": round(ck["iter_s"], 3),
                     "full_storage_mib": round(full["bytes"] / 2**20, 1),
                     "checkpointed_storage_mib": round(ck["bytes"] / 2**20, 1),
                     "max_rel_dF": float(np.abs(ck["trace"][:, 0] / full["trace"][:, 0] - 1).max()),
                     "evaluate_ms_full": round(full["eval_ms"], 2),
                     "evaluate_ms_checkpointed": round(ck["eval_ms"], 2),
                     "evaluate_rel_dH": abs(ck["H"] / full["H"] - 1),
                     "evaluate_rel_dstress": float(np.abs(ck["dL"] - full["dL"]).max()
                                                   / np.abs(full["dL"]).max())}
        print(f"bench: checkpoint {prec} done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return 0


def cpu_threads():
    """Threads of the CPU baseline: the process's CPU share (OMP_NUM_THREADS
    when set, else the CPUs this process may run on)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(wA, wB, threads, iters):
    """The oracle on `iters` SCFT iterations of the same workload from the
    same fields (a bounded sample), on `threads` OpenMP threads: iterations/s
    of the whole iteration, and the reference's own clock, the time spent in
    diffusion() alone (std::clock around solver.diffusion(),
    crysSCFT_sample.cpp:389-391), from timed triplets on the same solver."""
    from oracle import oracle as O

    O.set_num_threads(threads)
    s = O.Solver("PmmmDCT", (N_LOG,) * 3)
    t0 = time.perf_counter()
    tr, *_ = s.scft(FA, DS, XN, L, ACC, iters, wA, wB)
    dt = time.perf_counter() - t0
    io = s.iomat()
    io[...] = 1.0 + 0.01 * np.asarray(wA).reshape(io.shape)
    reps = 20
    s.diffusion()
    t1 = time.perf_counter()
    for _ in range(reps):
        s.diffusion()
    triplet_ms = (time.perf_counter() - t1) / reps * 1e3
    return iters / dt, tr, triplet_ms


def pmc_workload():
    """Fixed workload for the rocprofv3 --pmc traffic passes: 2 SCFT iterations
    of the FP64 workload, then one FP32 diffusion() whose conversion kernel
    (n doubles read, n floats written) calibrates FETCH_SIZE / WRITE_SIZE."""
    import crysscft_sample_amd as C
    from crysscft_sample_amd import fields as F

    wfA, wfB = F.bcc_fields((N_LOG,) * 3, L, FA, XN)
    wA, wB = F.extract(wfA, "PmmmDCT"), F.extract(wfB, "PmmmDCT")
    s = C.SolverPmmmDCT((N_LOG,) * 3)
    run = C.SCFT(s, FA, DS, XN, L, ACC)
    run.set_fields(wA, wB)
    run.run(2)
    s32 = C.SolverPmmmDCT((N_LOG,) * 3, precision="fp32")
    s32.set_contour_step(DS)
    s32.upload_iomat(wA)
    s32.diffusion()
    s32.synchronize()
    print("pmc workload done", flush=True)
    return 0


# The reference's triplet clock (Readme.md:41-76): total std::clock seconds
# spent in diffusion() over 100 iterations of its 64^3 sample with
# NsA + NsB = 100, i.e. 2 (NsA + NsB) = 200 triplets per iteration and 20000
# in all, on one thread of a Xeon E5-2690 v4.
README_TRIPLET_MS = {"P1": 67.15e3 / 20000, "3m": 11.43e3 / 20000, "PmmmDCT": 17.53e3 / 20000}


def triplet_timing(reps):
    """One JSON line: device milliseconds per diffusion() (one triplet) for
    the three solvers on the reference's 64^3 sample grid, beside the
    Readme's per-triplet CPU times."""
    import crysscft_sample_amd as C

    out = {}
    rng = np.random.default_rng(5)
    for name in ("P1", "3m", "PmmmDCT"):
        for prec in ("fp64", "fp32"):
            s = C.SOLVERS[name]((64, 64, 64), precision=prec)
            s.set_cell_para((4.0, 4.0, 4.0))
            s.set_contour_step(0.01)
            s.upload_iomat(rng.random(s.get_grid_size_phy()))
            ms = s.time_diffusion(reps)
            out[f"{name}_{prec}"] = round(ms, 6)
            s.close()
    line = {"metric": "ms per diffusion() triplet, 64^3 logical grid (reference sample)",
            "unit": "ms", "higher_is_better": False, "reps": reps,
            "gpu": out,
            "reference_cpu_readme": {k: round(v, 4) for k, v in README_TRIPLET_MS.items()},
            "speedup_fp64_vs_readme": {k: round(README_TRIPLET_MS[k] / out[f"{k}_fp64"], 1)
                                       for k in README_TRIPLET_MS}}
    print(json.dumps(line), flush=True)
    return 0


def batch_timing(count, steps):
    """One JSON line: the batched-instances workload: `count`
    PmmmDCT instances with reduced grids drawn from 16^3..64^3 (mean 84,685
    points) and Ns from {100, 150, 200}, advanced together by crys_batch_run
    (one hipGraph launch per iteration), on the path crys_batch_autotune
    picks.  Aggregate instance-iterations/s."""
    from crysscft_sample_amd.batch import Batch, random_instances

    insts = random_instances(count, seed=2024)
    rng = np.random.default_rng(2024)
    out = {"metric": f"batched SCFT: {count} PmmmDCT instances, reduced 16^3..64^3, "
                     "Ns in {100,150,200}",
           "unit": "instance-iterations/s", "higher_is_better": True, "instances": count,
           "steps": steps, "points_total": 0}
    fields = None
    traces, phis = {}, {}
    for prec in ("fp64", "fp32"):
        b = Batch(insts, precision=prec)
        out["points_total"] = int(sum(i[0] for i in b.info))
        if fields is None:
            fields = [inst.xN * 0.1 * rng.standard_normal(b.shape(k)) for k, inst in enumerate(insts)]
        for k, wA in enumerate(fields):
            b.set_fields(k, wA, -wA)
        path, ms_ph, ms_pa = b.autotune(reps=3)
        warm = b.run(1)  # warm: graph capture
        t0 = time.perf_counter()
        tr = b.run(steps)
        dt = time.perf_counter() - t0
        if not np.isfinite(tr[:, :, 0]).all():
            raise RuntimeError("non-finite free energy in the batch")
        traces[prec] = np.concatenate([warm, tr], axis=1)
        phis[prec] = [b.fields(k)[0] for k in range(count)]
        out[prec] = {"instance_iter_s": round(count * steps / dt, 2),
                     "ms_per_batch_iteration": round(dt / steps * 1e3, 3),
                     "path": path,
                     "phase_grid": b.phase_grid(),
                     "autotune_ms_steps": {"phases": None if ms_ph is None else round(ms_ph, 3),
                                           "passes": round(ms_pa, 3)}}
        b.close()
        print(f"bench: batch {prec} done", file=sys.stderr, flush=True)
    # per-instance error of the FP32 propagators against the FP64 batch, over
    # every iteration from the same starting fields
    dF = np.abs(traces["fp32"][:, :, 0] / traces["fp64"][:, :, 0] - 1).max(axis=1)
    dphi = np.array([np.abs(a - b).max() for a, b in zip(phis["fp32"], phis["fp64"])])
    out["fp32_vs_fp64"] = {"max_rel_dF": float(dF.max()), "median_rel_dF": float(np.median(dF)),
                           "max_abs_dphiA": float(dphi.max())}
    print(json.dumps(out), flush=True)
    return 0


SYNC_VARIANTS = ["launch", "empty_launch", "fenced", "two_level", "bypass", "barrier_only",
                 "xcd_local"]


def sync_timing(rounds):
    """One JSON line: microseconds per round of the exchange between contour
    steps (crys_time_grid_sync) for each synchronisation variant at 64, 128
    and 256 workgroups of 1024 threads."""
    import ctypes

    from crysscft_sample_amd._lib import check, lib

    out = {"metric": "us per round: 64 KB per workgroup written, then read by a workgroup on "
                     "another XCD (xcd_local: on the same XCD), synchronised by a launch or a "
                     "grid barrier",
           "unit": "us", "higher_is_better": False, "rounds": rounds, "workgroups": {}}
    for wg in (64, 128, 256):
        row = {}
        for v, name in enumerate(SYNC_VARIANTS):
            us = ctypes.c_double()
            bad, tmo = ctypes.c_uint(), ctypes.c_uint()
            print(f"bench: grid sync {name} {wg} workgroups", file=sys.stderr, flush=True)
            check(lib().crys_time_grid_sync(v, wg, rounds, ctypes.byref(us), ctypes.byref(bad),
                                            ctypes.byref(tmo)))
            row[name] = {"us": round(us.value, 3), "not_visible": bad.value,
                         "timeouts": tmo.value}
        out["workgroups"][str(wg)] = row
    print(json.dumps(out), flush=True)
    return 0


PASSES_CHILD = r"""
import json, sys, time
sys.path.insert(0, {root!r})
import crysscft_sample_amd as C
from crysscft_sample_amd import fields as F
name, N, L, T, prec, iters = json.loads(sys.argv[1])
N = tuple(N)
wA, wB = F.bcc_fields(N, L, 0.30, 25.0)
s = (C.SOLVERS[name](N, precision=prec) if T is None
     else C.SOLVERS[name](N, translational_part=T, precision=prec))
run = C.SCFT(s, 0.30, 1.0 / 200, 25.0, L, 0.1)
run.set_fields(F.extract(wA, name), F.extract(wB, name))
run.run(2)
C.lib().crys_device_synchronize()
t0 = time.perf_counter()
tr = run.run(iters)
C.lib().crys_device_synchronize()
dt = time.perf_counter() - t0
ms_step, per = run.time_step(reps=50)
ms_seq, nl = run.time_propagation(reps=3)
print(json.dumps({{"iter_s": iters / dt, "ms_step_repeated": ms_step, "ms_per_launch": per,
                  "ms_step_in_sequence": ms_seq, "launches": nl, "F_last": float(tr[-1, 0])}}))
"""

# (solver, logical grid, cell length): the headline grid (length class 64)
# and two grids of length class 128, where the contour step is three passes
PASS_GRIDS = [("PmmmDCT", 128, 4.2), ("PmmmDCT", 256, 4.2), ("P1", 128, 4.2)]
PASS_PATHS = {"phase": {}, "slab+x": {"CRYS_PHASE": "0"},
              "mfma_passes": {"CRYS_PHASE": "0", "CRYS_SLAB": "0", "CRYS_AXIS_PASS": "mfma"},
              "radix_passes": {"CRYS_PHASE": "0", "CRYS_SLAB": "0", "CRYS_AXIS_PASS": "radix"}}


def passes_timing(iters, only=None, paths=None):
    """One JSON line: ms per contour step (both chains) of each kernel path
    of the SCFT iteration, on the headline grid (axis length 64) and on two
    grids of axis length 128: the one-launch plane phase (matrix-core at 64,
    radix at 128), the fused z+y slab plus the x pass (likewise), three
    matrix-core passes (dense operators as GEMMs) and three radix FFT/DCT
    passes.  Each path runs in a child process because the path switches are
    read once per process."""
    import subprocess

    out = {"metric": "ms per contour step (both chains) by kernel path, SCFT iteration, Ns=200",
           "unit": "ms", "higher_is_better": False, "grids": {}}
    grids = PASS_GRIDS
    if only:
        # SOLVER:N, SOLVER:N:L (cell length, default 4.2) or SOLVER:N:L:T (3m with
        # the nine translational parts T, comma-separated)
        grids = []
        for g in only:
            f = g.split(":")
            T = [float(x) for x in f[3].split(",")] if len(f) > 3 else None
            if T is not None and len(T) != 9:
                raise ValueError(f"--pass-grid {g}: T needs 9 values")
            grids.append((f[0], int(f[1]), float(f[2]) if len(f) > 2 else 4.2, T))
    for name, N, L, *rest in grids:
        T = rest[0] if rest else None
        n = (N // 2) ** 3 if name != "P1" else N ** 3
        g = {}
        for prec in ("fp64", "fp32"):
            b = 8 if prec == "fp64" else 4
            for path, env_add in PASS_PATHS.items():
                if paths and path not in paths:
                    continue
                env = dict(os.environ, **env_add)
                arg = json.dumps([name, [N] * 3, L, T, prec, iters])
                print(f"bench: passes {name} {N}^3 {prec} {path}", file=sys.stderr, flush=True)
                r = subprocess.run([sys.executable, "-c", PASSES_CHILD.format(root=ROOT), arg],
                                   env=env, capture_output=True, text=True, timeout=300)
                if r.returncode != 0:
                    raise RuntimeError(f"passes child failed ({name} {N} {prec} {path}): "
                                       + r.stderr[-1500:])
                d = json.loads(r.stdout.strip().splitlines()[-1])
                d["step_gbs"] = round(2 * 3 * n * b / (d["ms_step_in_sequence"] * 1e-3) / 1e9, 1)
                g[f"{prec}/{path}"] = {k: (round(v, 6) if isinstance(v, float) else v)
                                       for k, v in d.items()}
        out["grids"][f"{name} {N}^3" + ("" if T is None else f" T={T}")] = g
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-propagator leg")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="run the fixed workload of the rocprofv3 --pmc passes and exit")
    ap.add_argument("--triplet", action="store_true",
                    help="time single diffusion() triplets (the reference's clock metric) and exit")
    ap.add_argument("--checkpoint", action="store_true",
                    help="time full histories against sqrt(Ns) checkpoints and exit")
    ap.add_argument("--passes", action="store_true",
                    help="time the contour step of every kernel path (phase, slab, matrix-core "
                         "passes, radix passes) and exit")
    ap.add_argument("--pass-grid", action="append", metavar="SOLVER:N[:L[:T]]",
                    help="with --passes: time these grids instead (e.g. PmmmDCT:256, 3m:64:4.0; "
                         "T: 3m's nine translational parts, comma-separated)")
    ap.add_argument("--pass-path", action="append", choices=list(PASS_PATHS),
                    help="with --passes: time only these kernel paths")
    ap.add_argument("--sync", action="store_true",
                    help="time the exchange between contour steps under a launch and under "
                         "grid barriers (crys_time_grid_sync) and exit")
    ap.add_argument("--batch", type=int, default=0, metavar="COUNT",
                    help="time COUNT batched instances (the batching workload) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not run "
              "(the path would shard by independent SCFT instances, one batch per GPU)", file=sys.stderr)
        return 2
    if args.sync:
        return sync_timing(400)
    if args.batch:
        return batch_timing(args.batch, max(args.steps // 10, 3))
    if args.pmc_workload:
        return pmc_workload()
    if args.checkpoint:
        return checkpoint_timing(max(args.steps // 2, 5))
    if args.triplet:
        return triplet_timing(200)
    if args.passes:
        return passes_timing(5, args.pass_grid, args.pass_path)

    from crysscft_sample_amd import fields as F

    wfA, wfB = F.bcc_fields((N_LOG,) * 3, L, FA, XN)
    wA, wB = F.extract(wfA, "PmmmDCT"), F.extract(wfB, "PmmmDCT")
    n = wA.size

    print(f"bench: FP64 PmmmDCT 128^3 Ns=200, {args.warmup} warmup + {args.steps} timed",
          file=sys.stderr, flush=True)
    f64 = run_gpu("fp64", args.steps, args.warmup, wA, wB)
    mixed = None
    if not args.no_mixed:
        print("bench: FP32 propagators + FP64 reductions", file=sys.stderr, flush=True)
        mixed = run_gpu("fp32", args.steps, args.warmup, wA, wB)

    NsA, NsB = f64["NsA"], f64["NsB"]
    b_iter64 = bytes_per_iteration(n, 8, NsA, NsB)
    # The unit of work is the contour step of one chain, 3 n b algorithmic
    # bytes (q_s, E, q_s+1).  The one-launch phase kernel does one step of
    # both chains per launch (6 n b); on the two/three-launch paths every
    # launch also reads a field and E and writes a field per chain (6 n b).
    names = STEP_KERNELS.get(len(f64["ms_per_launch"]), [])
    k_dom = int(np.argmax(f64["ms_per_launch"]))
    launch_bytes = 2 * 3 * n * 8
    # launch duration of the dominant kernel inside the iteration's own
    # sequence (a graph of the Ns-step propagation on its own buffers) when a
    # step is that one launch; else the kernel launched alone, repeated
    t_alone = f64["ms_per_launch"][k_dom]
    t_dom = f64["ms_per_step_in_sequence"] if f64["launches_per_step"] == 1 else t_alone
    achieved = launch_bytes / (t_dom * 1e-3) / 1e9
    kernel = names[k_dom] if k_dom < len(names) else f"launch {k_dom}"
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(kernel)[0],
                "traffic_source": (os.path.relpath(TRAFFIC_PROFILE, ROOT)
                                   if measured_traffic(kernel)[0] else None),
                "kernel": kernel + " (both chains)",
                "ms_per_launch": round(t_dom, 6),
                "ms_per_launch_repeated_alone": round(t_alone, 6),
                "bytes_per_launch": launch_bytes}
    # Matrix-core use of the same kernel: every plane of both chains takes
    # three half-size products per launch (M x M values, M / 2 multiply-adds
    # each, M = 64); the split kernel also repeats the finishing product in
    # the second workgroup of each plane, which is not counted.
    # (per contour step, timed in the iteration's own sequence: one launch on
    # the default path)
    M = N_LOG // 2
    flops_launch = 2 * M * 3 * (M * M * (M // 2) * 2)
    t_step = f64["ms_per_step_in_sequence"]
    mfma = {"achieved": round(flops_launch / (t_step * 1e-3) / 1e12, 3),
            "peak": MFMA_PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
            "frac": round(flops_launch / (t_step * 1e-3) / 1e12 / MFMA_PEAK_TFLOPS["fp64"], 5),
            "flops_per_step": flops_launch, "executed_flops_per_step": flops_launch * 4 // 3,
            "kernel": roofline["kernel"]}
    step_bytes = 2 * 3 * n * 8  # minimum of a contour step: q_s, E, q_s+1 per chain
    step = {"ms": round(f64["ms_per_step_in_sequence"], 6),
            "ms_repeated_first_step": round(f64["ms_per_contour_step"], 6),
            "ms_per_launch": {nm: round(t, 6) for nm, t in zip(names, f64["ms_per_launch"])},
            "algorithmic_bytes": step_bytes,
            "achieved_gbs": round(step_bytes / (f64["ms_per_step_in_sequence"] * 1e-3) / 1e9, 2)}

    cpu = None
    tr_cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        iters = 3
        print(f"bench: CPU baseline (oracle, {threads} threads, {iters} iterations; then 1 thread)",
              file=sys.stderr, flush=True)
        v, tr_cpu, trip = cpu_baseline(wA, wB, threads, iters)
        v1, tr1, trip1 = cpu_baseline(wA, wB, 1, 1)
        O_threads = threads
        dF = float(np.abs(f64["warm"][:iters, 0] / tr_cpu[:, 0] - 1).max())
        # the reference's own figure: triplet milliseconds, 2 (NsA + NsB) = 400
        # triplets per iteration (crysSCFT_sample.cpp:389-391, :512)
        cpu = {"value": round(v, 4), "unit": "iter/s", "cores": O_threads, "kind": "port",
               "sample": f"{iters} SCFT iterations (400 PmmmDCT diffusions + driver each) of the "
                         "same 128^3 Ns=200 workload from the same fields, own radix DCT, OpenMP "
                         f"over lines; 1 iteration on 1 thread",
               "host_cpus": os.cpu_count(),
               "value_1thread": round(v1, 5),
               "triplet_ms": round(trip, 3),
               "triplet_ms_1thread": round(trip1, 3),
               "triplet_s_per_iter": round(trip * 2 * (NsA + NsB) / 1e3, 4),
               "triplet_s_per_iter_1thread": round(trip1 * 2 * (NsA + NsB) / 1e3, 4),
               "gpu_fp64_rel_dF_vs_oracle": dF,
               "gpu_fp64_rel_dMaxInc_vs_oracle": float(np.abs(
                   f64["warm"][:iters, 5] / tr_cpu[:, 5] - 1).max()),
               "oracle_1thread_rel_dF": float(abs(tr1[0, 0] / tr_cpu[0, 0] - 1))}

    line = {
        "metric": "SCFT iter/s at 128^3 Ns=200 PmmmDCT: mixed vs FP64 GPU; HBM GB/s",
        "value": round(f64["iter_s"], 3),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms_per_iter"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (BCC-seeded Pmmm fields, splitmix64 seed 20241203)",
        "config": {"workload": "AB diblock SCFT, PmmmDCT, 128^3 logical / 64^3 reduced, "
                               "fA=0.30 chiN=25 L=4.2 ds=1/200 (NsA=60, NsB=140), mixing 0.1",
                   "solver": "PmmmDCT", "grid": [N_LOG] * 3, "NsA": NsA, "NsB": NsB},
        "hbm_gbs": round(b_iter64 * f64["iter_s"] / 1e9, 2),
        "bytes_per_iter": b_iter64,
        "roofline": roofline,
        "mfma": mfma,
        "contour_step": step,
        "cpu_baseline": cpu,
    }
    if mixed is not None:
        b_iter32 = bytes_per_iteration(n, 4, NsA, NsB) + 0  # FP64 field work unchanged
        dF = float(np.abs(mixed["trace"][:, 0] / f64["trace"][:, 0] - 1).max())
        line["mixed"] = {"iter_s": round(mixed["iter_s"], 3),
                         "speedup_vs_fp64": round(mixed["iter_s"] / f64["iter_s"], 3),
                         "hbm_gbs": round(b_iter32 * mixed["iter_s"] / 1e9, 2),
                         "max_rel_dF_vs_fp64_gpu": dF,
                         "max_rel_dMaxInc_vs_fp64_gpu": float(np.abs(
                             mixed["trace"][:, 5] / f64["trace"][:, 5] - 1).max()),
                         "max_abs_dphiA_vs_fp64_gpu": float(np.abs(mixed["phiA"] - f64["phiA"]).max()),
                         "max_abs_dphiB_vs_fp64_gpu": float(np.abs(mixed["phiB"] - f64["phiB"]).max()),
                         "max_rel_dF_vs_oracle": (None if tr_cpu is None else float(np.abs(
                             mixed["warm"][:len(tr_cpu), 0] / tr_cpu[:, 0] - 1).max())),
                         "max_rel_dMaxInc_vs_oracle": (None if tr_cpu is None else float(np.abs(
                             mixed["warm"][:len(tr_cpu), 5] / tr_cpu[:, 5] - 1).max())),
                         "oracle_iterations_compared": None if tr_cpu is None else len(tr_cpu),
                         "ms_per_contour_step": round(mixed["ms_per_step_in_sequence"], 6),
                         "mfma_fp32_frac": round(flops_launch / (mixed["ms_per_step_in_sequence"] * 1e-3)
                                                 / 1e12 / MFMA_PEAK_TFLOPS["fp32"], 5),
                         "ms_per_launch": {nm.replace("double", "float"): round(t, 6)
                                           for nm, t in zip(names, mixed["ms_per_launch"])}}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
