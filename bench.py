#!/usr/bin/env python3
"""Benchmark of the MI355X SPIKE banded solve (one JSON line on stdout).

Metric: solved RHS/sec at ||Ax-b||/||b|| <= 1e-10, mixed precision vs plain FP64 SPIKE,
with p50/p99 step latency.  A step is one full solve of the workload with its inputs already
resident in HBM: CSR -> blocks, block LU, spike tips, reduced system, recovery and the FP64
residual / refinement loop, ending when every column's relative residual meets the tolerance.

Workloads (synthetic, seeded):
  c4 (default) 4096 independent systems in one batched solve: n log-uniform [10^3, 2*10^5],
               half-bandwidth uniform in {4, 8, 16, 32, 64}, 30% fill, dominance 1.1, nrhs 1..8
  c2           n = 10^6, half-bandwidth 64, 15% in-band fill U[-1,1], diagonal 1.05*sum|offdiag|
               (floor 1), 16 RHS U[-1,1]; plain FP64 is also run at the configured 32 partitions
  c3           n = 2*10^6, half-bandwidth 128, 20% fill, per-row dominance U[1.01, 2.0], 64 RHS,
               refinement to 1e-12
  c1           n = 10^4 pentadiagonal, a_ii = sum|a_ij| + 1, 4 partitions, 1 RHS
  c5           request serving: Poisson arrivals (rate --rate), 80% real / 20% complex128, n
               log-uniform [10^4, 5*10^6], half-bandwidth 2..256, 30% fill,
               dominance 1.2, nrhs 1..64, 30% of requests repeat a resident matrix with a new RHS
The CPU baseline is the oracle (oracle/, a C restatement of the reference's SPIKE flow with a
banded partial-pivot LU in place of SuperLU_MT), timed on the host cores on a bounded sample of
the same workload: on every core at once (independent solves spread over the cores, each solve
on one core -- the reference's driver loops over its partitions serially,
OpenMP/spike_interfaces.c:105-237), and one solve after another on one core.

The timed steps run the factorization and then the solve replayed from a captured HIP graph (the
library's mode for repeated solves of one shape, bit-identical to its eager solve); the eager
throughput is reported beside it.  Every stage is reported against both peaks (`stage_roofline`,
from eager steps of the same workload); `roofline` is the stage with the largest time per step.  Per-launch times are the library's HIP events on the stream it launches
on; `traffic` (HBM bytes per launch) and MFMA busy cycles come from the committed rocprofv3
counter summaries in profiles/ for this workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4|c2|c3|c1|c5] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solved RHS/sec at ||Ax-b||/||b||<=1e-10, mixed vs FP64 SPIKE; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}

CONFIGS = {
    "c2": dict(n=1_000_000, k=64, fill=0.15, dom=(1.05, 1.05), diag_add=0.0, nrhs=16, P_cfg=32, tol=1e-10,
               desc="banded n=1e6, half-bandwidth 64, 15% in-band fill, diagonal 1.05*sum|offdiag|, 16 RHS"),
    "c3": dict(n=2_000_000, k=128, fill=0.20, dom=(1.01, 2.0), diag_add=0.0, nrhs=64, P_cfg=0, tol=1e-12,
               desc="banded n=2e6, half-bandwidth 128, 20% fill, per-row dominance U[1.01,2.0], 64 RHS, refine to 1e-12"),
    "c1": dict(n=10_000, k=2, fill=1.0, dom=(1.0, 1.0), diag_add=1.0, nrhs=1, P_cfg=4, tol=1e-10,
               desc="pentadiagonal n=1e4, a_ii = sum|a_ij| + 1, 1 RHS"),
}
SEED_A, SEED_B = 20261202, 7
C4 = dict(nsys=4096, n_lo=1e3, n_hi=2e5, ks=(4, 8, 16, 32, 64), fill=0.30, dom=1.1, nrhs_hi=8, tol=1e-10,
          seed=20261204)
C5 = dict(nreq=120, rate=40.0, frac_complex=0.2, n_lo=1e4, n_hi=5e6, k_lo=2, k_hi=256, k_hi_complex=256,
          nrhs_hi=64, repeat=0.3, fill=0.30, dom=1.2, tol=1e-10, seed=20261205)

# stage -> stats key of its time (the library's HIP events around the stage's launches)
STAGES = {"factor": "ms_factor", "spikes": "ms_spikes", "ysweep": "ms_ysweep", "recover": "ms_recover",
          "residual": "ms_residual"}
PER_PASS = ("ysweep", "recover", "residual")          # launched once per solve / refinement pass


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def c5_stream(nreq, seed, n_hi=None):
    """The C5 request stream: (arrival offset s, matrix id, nrhs, RHS seed) per request and
    (dtype complex?, n, k, matrix seed) per distinct matrix, drawn from `seed`; n_hi caps the
    row-count range (the tests draw the same stream shape at sizes the oracle solves quickly)."""
    rng = np.random.RandomState(seed)
    t = np.cumsum(rng.exponential(1.0, nreq))
    mats, reqs = [], []
    for i in range(nreq):
        if mats and rng.rand() < C5["repeat"]:
            mid = int(rng.randint(len(mats)))
        else:
            cplx = rng.rand() < C5["frac_complex"]
            n = int(np.exp(rng.uniform(np.log(C5["n_lo"]), np.log(n_hi or C5["n_hi"]))))
            k = int(rng.randint(C5["k_lo"], (C5["k_hi_complex"] if cplx else C5["k_hi"]) + 1))
            mats.append((cplx, n, k, seed * 131 + len(mats)))
            mid = len(mats) - 1
        reqs.append((float(t[i]), mid, int(rng.randint(1, C5["nrhs_hi"] + 1)), seed * 977 + i))
    return reqs, mats


def c4_systems(nsys, seed):
    """The C4 batch: (n, k, nrhs, matrix seed, RHS seed) per system, drawn from `seed`."""
    rng = np.random.RandomState(seed)
    ns = np.exp(rng.uniform(np.log(C4["n_lo"]), np.log(C4["n_hi"]), nsys)).astype(np.int64)
    ks = rng.choice(np.array(C4["ks"]), nsys)
    nr = rng.randint(1, C4["nrhs_hi"] + 1, nsys)
    return [(int(ns[i]), int(ks[i]), int(nr[i]), seed * 7919 + i, seed + 100003 * i) for i in range(nsys)]


def algorithmic_counts(name, n, nnz, ku, kl, nrhs, P, s):
    """Algorithmic flops and HBM bytes of one system's share of one launch of a stage: the
    counting formulas of the benchmark definition (band read/written once, RHS and solution
    panels once per pass); implementation overheads never enter."""
    if name == "factor":      # banded LU of all partitions
        return 2.0 * n * kl * ku, 2.0 * n * (kl + ku + 1) * s
    if name == "spikes":      # V/W spike tips over whole partitions: m_i = ku [i < P-1] + kl [i > 0]
        # columns per partition (the chains stop early; spike_counts charges the rows they swept)
        m_avg = (ku * (P - 1) + kl * (P - 1)) / max(P, 1)
        return 2.0 * n * (kl + ku) * m_avg, n * (kl + ku + 1) * s
    if name == "recover":     # forward + backward sweep over nrhs columns, f read, x written
        return 2.0 * n * (kl + ku) * nrhs, n * (kl + ku + 1) * s + 2.0 * n * nrhs * 8
    if name == "ysweep":      # RHS tips: forward sweep, f read
        return 2.0 * n * kl * nrhs + 2.0 * n * ku * nrhs, n * (kl + ku + 1) * s + n * nrhs * 8
    if name == "residual":    # FP64 CSR residual of every column
        return 2.0 * nnz * nrhs, nnz * 12.0 + 4.0 * (n + 1) + 3.0 * n * nrhs * 8
    raise KeyError(name)


def spike_counts(work, s):
    """Algorithmic flops and bytes of the V / W spike stage from the rows its chains swept in the
    last factorization (spike_get_spike_work): a W chain stops once its spike has decayed below
    2^-70 of its first block and a V chain covers only the blocks its right-hand side reaches, so
    the stage reads the band of those rows, not of the whole partition."""
    rows, band, flops = work
    return flops, band * s


def profile_table(config, kind):
    """Per-kernel figures from the committed counter summaries of `config`:
    kind 'traffic' -> HBM bytes per dispatch (profiles/spike_<config>_mixed_hbm_traffic.md),
    kind 'mfma' -> MFMA busy fraction (profiles/spike_<config>_mixed_mfma.md)."""
    name = {"traffic": "spike_%s_mixed_hbm_traffic.md", "mfma": "spike_%s_mixed_mfma.md"}[kind] % config
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if not os.path.exists(path):
        return out
    key = "traffic_bytes" if kind == "traffic" else "mfma_busy"
    for line in open(path):
        parts = line.split()
        if len(parts) == 5 and parts[0] == key and parts[1] == config:
            out[parts[2]] = float(parts[4])
    return out


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": usable, "cpu_model": model}


def oracle_baseline(jobs, unit_per_job, budget_s, serial_budget_s, what):
    """Time the oracle on a bounded prefix of `jobs`, each a pair (prepare() -> data, solve(data,
    nthreads)) whose preparation (input generation) stays outside the timed region: first on every
    host core at once (one solve per core, a pool of threads; the oracle releases the interpreter
    lock), then one after another on one core; returns the cpu_baseline record."""
    from concurrent.futures import ThreadPoolExecutor
    info = cpu_info()
    ncores = info["nproc"]
    t_par = 0.0
    done_units = done_jobs = 0
    with ThreadPoolExecutor(max_workers=ncores) as ex:
        idx = 0
        while idx < len(jobs) and t_par < budget_s:
            # batches of several solves per core, scheduled dynamically, so that the uneven
            # solve sizes leave few cores idle at the end of a batch
            batch = list(range(idx, min(len(jobs), idx + 4 * ncores)))
            data = list(ex.map(lambda i: jobs[i][0](), batch))
            t0 = time.perf_counter()
            list(ex.map(lambda q: jobs[batch[q]][1](data[q], 1), range(len(batch))))
            t_par += time.perf_counter() - t0
            done_units += sum(unit_per_job[i] for i in batch)
            done_jobs += len(batch)
            idx += len(batch)
    t_ser = 0.0
    s_units = s_jobs = 0
    while s_jobs < len(jobs) and t_ser < serial_budget_s:
        d = jobs[s_jobs][0]()
        t0 = time.perf_counter()
        jobs[s_jobs][1](d, 1)
        t_ser += time.perf_counter() - t0
        s_units += unit_per_job[s_jobs]
        s_jobs += 1
    log("cpu oracle: %d solves on %d cores in %.1f s; %d one after another in %.1f s" %
        (done_jobs, ncores, t_par, s_jobs, t_ser))
    return {"value": round(done_units / t_par, 3), "unit": "RHS/s", "cores": ncores, "kind": "port",
            "sample": "%s: the first %d solves on all %d host cores (%.1f s); the first %d one after another on "
                      "one core (%.1f s)" % (what, done_jobs, ncores, t_par, s_jobs, t_ser),
            "serial_value": round(s_units / t_ser, 3), "serial_cores": 1, **info}


def stage_roofline(avg_ms, counts, passes, groups, config, kernel_names):
    """Every stage against HBM and the FP32 (factor stages) / FP64 (residual) peaks, per launch."""
    traffic = profile_table(config, "traffic")
    mfma = profile_table(config, "mfma")
    out = {}
    for st_name, st_ms in avg_ms.items():
        if st_ms <= 0:
            continue
        reps = passes if st_name in PER_PASS else 1
        nl = groups * reps
        F, B = counts[st_name]
        F, B = F * reps / nl, B * reps / nl          # per launch (the batch's systems split over groups)
        t = st_ms / nl * 1e-3
        pk = PEAK_TFLOPS["f64"] if st_name == "residual" else PEAK_TFLOPS["f32"]
        kn = kernel_names[st_name]
        rec = {"launches": nl, "ms_per_launch": round(st_ms / nl, 4), "kernel": kn,
               "algorithmic_bytes": B, "algorithmic_flops": F,
               "GB_s": round(B / t / 1e9, 2), "hbm_frac": round(B / t / 1e9 / HBM_PEAK_GBS, 4),
               "TFLOP_s": round(F / t / 1e12, 3), "mfma_frac": round(F / t / 1e12 / pk, 4),
               "bound": "mfma" if F / B >= pk * 1e12 / (HBM_PEAK_GBS * 1e9) else "hbm"}
        if kn in traffic:
            rec["traffic"] = int(traffic[kn])
            rec["traffic_over_algorithmic"] = round(traffic[kn] / B, 3)
        if kn in mfma:
            rec["mfma_busy"] = mfma[kn]
        out[st_name] = rec
    return out


def roofline_of(stage_roof, avg_ms):
    """The stage with the largest time per step: achieved algorithmic rate per launch on the side
    its intensity puts it."""
    dom = max(avg_ms, key=avg_ms.get)
    r = stage_roof[dom]
    if r["bound"] == "mfma":
        pk = PEAK_TFLOPS["f64"] if dom == "residual" else PEAK_TFLOPS["f32"]
        roof = {"bound": "mfma", "achieved": r["TFLOP_s"], "peak": pk, "unit": "TFLOP/s", "frac": r["mfma_frac"]}
    else:
        roof = {"bound": "hbm", "achieved": r["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["hbm_frac"]}
    roof.update({"traffic": r.get("traffic"), "kernel": r["kernel"], "stage": dom, "launches_per_step": r["launches"],
                 "algorithmic_flops": r["algorithmic_flops"], "algorithmic_bytes": r["algorithmic_bytes"],
                 "ms_per_launch": r["ms_per_launch"]})
    if "mfma_busy" in r:
        roof["mfma_busy"] = r["mfma_busy"]
    return roof


def kernel_names_for(nts, nrhs_max, one_copy=()):
    """Instantiations the library launches for block sizes 16*NT (FP32 factors) and up to nrhs_max
    RHS columns per chunk (column tiles per sweep workgroup CT with CT * NT <= 16); several block
    sizes are joined with '+'.  `one_copy`: block sizes whose factor launch has more partitions than
    the device runs at once, so it takes the variant with one LDS copy of R's rows (FP32 b = 64)."""
    m = min(nrhs_max, 64)
    ct_max = lambda nt: 4 if 4 * nt * 4 <= 64 else (2 if 2 * nt * 4 <= 64 else 1)   # FP32 sweep tiles
    ct_vw = lambda nt: min(ct_max(nt), nt)           # a spike has at most b = 16 NT columns
    ct_y = lambda nt: min(ct_max(nt), 1 if m <= 16 else (2 if m <= 32 else 4))
    need = (m + 1) // 2
    tw = 4 if need <= 4 else (8 if need <= 8 else 16)
    return {"factor": "+".join("k_factor<float,%d,%d>" % (nt, 1 if (nt in one_copy and nt == 4) else 2) for nt in nts),
            "spikes": "+".join("k_sweep<float,%d,%d,true>" % (nt, ct_vw(nt)) for nt in nts),
            "ysweep": "+".join("k_sweep<float,%d,%d,false>" % (nt, ct_y(nt)) for nt in nts),
            "recover": "+".join("k_recover<float,%d,%d>" % (nt, ct_y(nt)) for nt in nts),
            "residual": "k_residual<double,%d,2>" % tw}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS) + ["c4", "c5"])
    ap.add_argument("--nsys", type=int, default=C4["nsys"], help="c4: number of systems in the batch")
    ap.add_argument("--nreq", type=int, default=C5["nreq"], help="c5: number of requests")
    ap.add_argument("--rate", type=float, default=C5["rate"], help="c5: Poisson arrival rate (requests/s)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison runs")
    ap.add_argument("--probe-copy", action="store_true",
                    help="only run a plain device copy of 256 MiB three times (calibrates the HBM byte counters)")
    ap.add_argument("--probe-stream", action="store_true",
                    help="only measure the streaming bandwidth plain torch kernels reach (read, copy, write)")
    args = ap.parse_args()
    if args.probe_stream:
        return probe_stream()
    if args.probe_copy:
        import torch
        x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        log("copy probe: 3 x (256 MiB read + 256 MiB written)")
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; multi-GPU SPIKE (one partition range per "
              "GPU, tip exchange over xGMI) is a design note, not a benchmark here", file=sys.stderr)
        return 2

    import torch
    import hspike_amd as hs

    if args.config == "c4":
        return bench_batch(args, torch, hs)
    if args.config == "c5":
        return bench_serve(args, hs)
    return bench_single(args, torch, hs)


def probe_stream():
    """Streaming bandwidth of plain torch kernels over 1, 4 and 16 GiB: a read-only reduction
    (sum of FP32, max of int32), a copy (read + write) and a fill (write).  HIP events around 10
    launches after one untimed launch; one JSON line per case.  It is the practical ceiling the
    stage rates are read against (the roofline itself uses the published 8 TB/s)."""
    import torch
    dev = torch.device("cuda:0")

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps

    for gib in (1, 4, 16):
        n = gib * 2 ** 30 // 4
        x = torch.ones(n, dtype=torch.float32, device=dev)
        xi = torch.ones(n, dtype=torch.int32, device=dev)
        y = torch.empty_like(x)
        for name, fn, nbytes in (("read (sum f32)", lambda: x.sum(), 4 * n), ("read (max i32)", lambda: xi.amax(), 4 * n),
                                 ("copy f32 (read + write)", lambda: y.copy_(x), 8 * n), ("write (fill f32)", lambda: y.fill_(1.0), 4 * n)):
            ms = timed(fn)
            print(json.dumps({"probe": name, "gib": gib, "ms": round(ms, 4), "GB_s": round(nbytes / ms / 1e6, 1)}), flush=True)
        del x, xi, y
    return 0


def bench_single(args, torch, hs):
    """C1 / C2 / C3: one system, factor + solve per step."""
    cfg = CONFIGS[args.config]
    n, k, nrhs = cfg["n"], cfg["k"], cfg["nrhs"]
    t_gen = time.time()
    A = hs.gen_banded(n, k, k, fill=cfg["fill"], dom=cfg["dom"], diag_add=cfg["diag_add"], diag_floor=1.0,
                      seed=SEED_A)
    B = hs.gen_dense(n, nrhs, seed=SEED_B)
    t_gen = time.time() - t_gen
    dev = torch.device("cuda:0")
    d_rowptr = torch.from_numpy(A.rowptr).to(dev)
    d_colind = torch.from_numpy(A.colind).to(dev)
    d_aij = torch.from_numpy(A.aij).to(dev)
    d_b = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)       # column-major n x nrhs
    d_x = torch.zeros_like(d_b)
    torch.cuda.synchronize()

    def run(precision, partitions, steps, warmup, graph=False):
        s = hs.Solver(precision, partitions=partitions, tol=cfg["tol"], max_refine=8, graph=graph)
        s.set_matrix_dev(n, d_rowptr, d_colind, d_aij)
        lat, stats = [], []
        for _ in range(warmup):
            s.factor()
            s.solve_dev(d_b, d_x, nrhs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            s.factor()
            s.solve_dev(d_b, d_x, nrhs)
            lat.append((time.perf_counter() - ts) * 1e3)
            stats.append(s.stats())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        work = s.spike_work()                          # spike chains of the last factorization
        tr = time.perf_counter()                       # resident factors (repeat-RHS solves)
        for _ in range(max(2, steps // 2)):
            s.solve_dev(d_b, d_x, nrhs)
        torch.cuda.synchronize()
        t_res = (time.perf_counter() - tr) / max(2, steps // 2)
        x = d_x.cpu().numpy().T.copy(order="F")
        return dict(wall=wall, lat=np.array(lat), stats=stats, last=stats[-1], x=x, resident_ms=t_res * 1e3,
                    work=work)

    log("generated %s: n=%d nnz=%d in %.2f s" % (args.config, n, A.nnz, t_gen))
    # timed steps: factor + the solve replayed from its captured graph; per-stage times from
    # eager steps of the same solve (a replay does not time its stages)
    mixed = run("mixed", 0, args.steps, max(2, args.warmup), graph=True)
    eager = run("mixed", 0, max(2, args.steps // 2), 1)
    log("mixed: %.3f ms/step replayed (%d of %d steps), %.3f ms/step eager, P=%d" %
        (mixed["wall"] / args.steps * 1e3, sum(x["graph_replayed"] for x in mixed["stats"]), args.steps,
         eager["wall"] / len(eager["lat"]) * 1e3, mixed["last"]["P"]))
    P_auto = mixed["last"]["P"]
    rhs_s = args.steps * nrhs / mixed["wall"]

    # the same mixed solve from host memory, as a caller of the reference's host entry points sees
    # it: CSR and B copied in from pageable arrays, factor + solve, X copied back (PCIe-inclusive;
    # never the bench value).  Best of three calls after one warm call.
    hts = []
    hsol = hs.Solver("mixed", tol=cfg["tol"], max_refine=8)
    for r in range(4):
        t0 = time.perf_counter()
        hsol.set_matrix(A)
        hsol.solve(B)
        hts.append(time.perf_counter() - t0)
    hsol.close()
    t_host = min(hts[1:])
    host_bytes = (n + 1) * 4 + A.nnz * 12 + 2 * n * nrhs * 8
    host_path = {"rhs_per_s": round(nrhs / t_host, 2), "ms": round(t_host * 1e3, 3), "bytes_moved": host_bytes,
                 "note": "CSR and B from pageable host memory, factor + solve, X back (Solver.set_matrix + solve)"}
    log("host path: %.3f ms per call (%.1f MB moved)" % (t_host * 1e3, host_bytes / 1e6))

    out_fp64, x64 = None, None
    if not args.no_fp64:
        f64 = run("fp64", 0, max(3, args.steps // 2), 2, graph=True)
        out_fp64 = {"rhs_per_s": round(len(f64["lat"]) * nrhs / f64["wall"], 2),
                    "ms_per_step": round(f64["wall"] / len(f64["lat"]) * 1e3, 3), "partitions": f64["last"]["P"],
                    "max_rel_residual": f64["last"]["max_rel_residual"], "refine_iters": f64["last"]["refine_iters"]}
        if cfg["P_cfg"]:
            f64c = run("fp64", cfg["P_cfg"], 2, 1)
            out_fp64["configured_partitions"] = cfg["P_cfg"]
            out_fp64["rhs_per_s_configured_partitions"] = round(len(f64c["lat"]) * nrhs / f64c["wall"], 2)
            out_fp64["max_rel_residual_configured_partitions"] = f64c["last"]["max_rel_residual"]
            mixc = run("mixed", cfg["P_cfg"], 2, 1)
            out_fp64["mixed_rhs_per_s_configured_partitions"] = round(len(mixc["lat"]) * nrhs / mixc["wall"], 2)
        x64 = f64["x"]
        log("fp64: %s" % out_fp64)

    avg = {st: float(np.mean([x[key] for x in eager["stats"]])) for st, key in STAGES.items()}
    avg_other = {kk: float(np.mean([x[kk] for x in eager["stats"]])) for kk in
                 ("ms_convert", "ms_rtop", "ms_red_setup", "ms_reduced", "ms_norms", "ms_factor_total",
                  "ms_solve_total")}
    passes = max(1, eager["last"]["solve_passes"])
    counts = {st: algorithmic_counts(st, n, A.nnz, k, k, nrhs, P_auto, 4) for st in STAGES}
    counts["spikes"] = spike_counts(eager["work"], 4)
    names = kernel_names_for([mixed["last"]["b"] // 16], nrhs)
    stage_roof = stage_roofline(avg, counts, passes, 1, args.config, names)
    roof = roofline_of(stage_roof, avg)

    cpu = None
    if not args.no_cpu:          # the oracle at the configured partitions, on all cores and on one
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.build()
        Ao = oracle.CSR(A.n, A.rowptr, A.colind, A.aij)
        Pc = cfg["P_cfg"] or 32
        info = cpu_info()
        t0 = time.perf_counter()
        Xo, _ = oracle.spike_solve(Ao, B, Pc, nthreads=info["nproc"], want_reduced=False)
        t_par = time.perf_counter() - t0
        x64 = Xo
        # one core: the same solve on a leading slice of the rows (partitions of the same size,
        # about 10 s of work), scaled by the rows solved
        per_row = t_par * min(info["nproc"], Pc) / n
        ns = int(min(n, max(1000, 10.0 / max(per_row, 1e-12))))
        Ps = max(1, ns // max(1, n // Pc))
        Asl = oracle.extract_csr(Ao, 0, ns, 0, ns)
        t0 = time.perf_counter()
        oracle.spike_solve(Asl, np.asfortranarray(B[:ns]), Ps, nthreads=1, want_reduced=False)
        t_ser = time.perf_counter() - t0
        log("cpu oracle: %.2f s on %d cores; %d rows on one core in %.2f s" % (t_par, info["nproc"], ns, t_ser))
        cpu = {"value": round(nrhs / t_par, 3), "unit": "RHS/s", "cores": info["nproc"], "kind": "port",
               "sample": "full workload: one oracle SPIKE solve (FP64, %d partitions, %d RHS) on all %d host cores in "
                         "%.2f s; one core: the first %d rows (%d partitions of the same size) in %.2f s, scaled by rows"
                         % (Pc, nrhs, info["nproc"], t_par, ns, Ps, t_ser),
               "serial_value": round(nrhs / (t_ser * n / ns), 3), "serial_cores": 1, **info}
    err64 = float(np.linalg.norm(mixed["x"] - x64) / np.linalg.norm(x64)) if x64 is not None else None

    lat = mixed["lat"]
    line = {
        "metric": METRIC, "value": round(rhs_s, 2), "unit": "RHS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(mixed["wall"] / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": args.config.upper() + ": " + cfg["desc"], "n": n, "half_bandwidth": k, "nnz": A.nnz,
                   "nrhs": nrhs, "partitions": P_auto, "block": mixed["last"]["b"], "tol": cfg["tol"],
                   "seeds": [SEED_A, SEED_B], "precision": "mixed: fp32 factor/spikes, fp64 residual/reduced/update"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "accuracy": {"max_rel_residual": mixed["last"]["max_rel_residual"], "refine_iters": mixed["last"]["refine_iters"],
                     "reduced_iters": mixed["last"]["reduced_iters"], "rel_err_vs_fp64": err64},
        "fp64": out_fp64,
        "mixed_vs_fp64": round(rhs_s / out_fp64["rhs_per_s"], 3) if out_fp64 else None,
        "resident_factors_rhs_per_s": round(nrhs / (mixed["resident_ms"] * 1e-3), 2),
        "host_path": host_path,
        "graph": {"replayed_steps": int(sum(x["graph_replayed"] for x in mixed["stats"])),
                  "eager_rhs_per_s": round(len(eager["lat"]) * nrhs / eager["wall"], 2),
                  "eager_ms_per_step": round(eager["wall"] / len(eager["lat"]) * 1e3, 3)},
        "stage_ms": {**{kk: round(v, 4) for kk, v in avg.items()}, **{kk: round(v, 4) for kk, v in avg_other.items()}},
        "roofline": roof,
        "stage_roofline": stage_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


def bench_batch(args, torch, hs):
    """C4: one batched factor + solve of many independent systems per step (spike_batch_solve_dev)."""
    dev = torch.device("cuda:0")
    spec = c4_systems(args.nsys, C4["seed"])
    t_gen = time.time()
    systems, meta = [], []
    for (n, kk, nrhs, sa, sb) in spec:
        A = hs.gen_banded(n, kk, kk, fill=C4["fill"], dom=(C4["dom"], C4["dom"]), diag_floor=1.0, seed=sa)
        B = hs.gen_dense(n, nrhs, seed=sb)
        systems.append(dict(n=n, nrhs=nrhs, rowptr=torch.from_numpy(A.rowptr).to(dev),
                            colind=torch.from_numpy(A.colind).to(dev), aij=torch.from_numpy(A.aij).to(dev),
                            b=torch.from_numpy(np.ascontiguousarray(B.T)).to(dev),
                            x=torch.zeros((nrhs, n), dtype=torch.float64, device=dev)))
        meta.append((n, A.nnz, kk, nrhs, sa, sb))
        if len(meta) % 1024 == 0:
            log("c4: %d of %d systems generated and uploaded" % (len(meta), len(spec)))
    torch.cuda.synchronize()
    rows = sum(m[0] for m in meta)
    nnz = sum(m[1] for m in meta)
    rhs = sum(m[3] for m in meta)
    log("generated c4: %d systems, %d rows, nnz %d, %d RHS in %.1f s" % (len(systems), rows, nnz, rhs, time.time() - t_gen))

    def run(precision, steps, warmup, graph=False):
        bt = hs.Batch(precision, tol=C4["tol"], graph=graph, tune="throughput")
        prep = bt.prepare(systems)                     # descriptors of the batch, built once
        for _ in range(warmup):
            bt.solve_dev(prep)
        torch.cuda.synchronize()
        lat, rels, sts = [], None, []
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            rels, st = bt.solve_dev(prep)
            lat.append((time.perf_counter() - ts) * 1e3)
            sts.append(st)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        work = bt.spike_work()
        bt.close()
        return dict(wall=wall, lat=np.array(lat), rel=rels, stats=sts, work=work)

    # timed steps: factor + the solve replayed from its captured graph; per-stage times from
    # eager steps (a replay does not time its stages)
    mixed = run("mixed", args.steps, max(2, args.warmup), graph=True)
    eager = run("mixed", max(2, args.steps // 2), 1)
    log("mixed: %.2f ms/step replayed (%d of %d steps), %.2f ms/step eager, max rel residual %.3e" %
        (mixed["wall"] / args.steps * 1e3, sum(x["graph_replayed"] for x in mixed["stats"]), args.steps,
         eager["wall"] / len(eager["lat"]) * 1e3, mixed["rel"].max()))
    f64 = None
    if not args.no_fp64:
        r = run("fp64", max(2, args.steps // 2), 1, graph=True)
        f64 = {"rhs_per_s": round(len(r["lat"]) * rhs / r["wall"], 2), "ms_per_step": round(r["wall"] / len(r["lat"]) * 1e3, 3),
               "max_rel_residual": float(r["rel"].max()),
               "note": "the whole batch's FP64 factors exceed HBM: solved in windows of systems"}
        log("fp64: %s" % f64)
    rhs_s = args.steps * rhs / mixed["wall"]
    avg = {st: float(np.mean([x[key] for x in eager["stats"]])) for st, key in STAGES.items()}
    avg_other = {kk: float(np.mean([x[kk] for x in eager["stats"]])) for kk in
                 ("ms_convert", "ms_rtop", "ms_red_setup", "ms_reduced", "ms_norms", "ms_factor_total",
                  "ms_solve_total")}
    last = eager["stats"][-1]
    passes, groups = max(1, last["solve_passes"]), max(1, last["launch_groups"])
    counts = {st: [0.0, 0.0] for st in STAGES}
    nts = set()
    for (n, z, kk, nrhs, _, _) in meta:
        P = hs.autotune_partitions(n, kk, kk, nrhs, precision=hs.SPIKE_PREC_MIXED, tune="throughput")
        nts.add(1 if kk <= 16 else (2 if kk <= 32 else 4))
        for st in STAGES:
            f, b = algorithmic_counts(st, n, z, kk, kk, nrhs, P, 4)
            counts[st][0] += f
            counts[st][1] += b
    counts["spikes"] = spike_counts(eager["work"], 4)
    # the b = 64 group of the batch has thousands of partitions: the one-copy factor variant
    names = kernel_names_for(sorted(nts), C4["nrhs_hi"], one_copy=(4,))
    stage_roof = stage_roofline(avg, counts, passes, groups, "c4", names)
    roof = roofline_of(stage_roof, avg)

    cpu = None
    if not args.no_cpu:          # the oracle on a bounded prefix of the batch, one system per core
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.build()

        def job(i):
            n, _, kk, nrhs, sa, sb = meta[i]

            def prep():
                A = hs.gen_banded(n, kk, kk, fill=C4["fill"], dom=(C4["dom"], C4["dom"]), diag_floor=1.0, seed=sa)
                return oracle.CSR(A.n, A.rowptr, A.colind, A.aij), hs.gen_dense(n, nrhs, seed=sb)

            def go(d, nthreads):
                oracle.spike_solve(d[0], d[1], max(1, min(4, n // (4 * kk + 4))), nthreads=nthreads, want_reduced=False)
            return prep, go
        cpu = oracle_baseline([job(i) for i in range(len(meta))], [m[3] for m in meta], 20.0, 10.0,
                              "oracle SPIKE FP64 (up to 4 partitions, the reference's P) per system, in batch order")
    lat = mixed["lat"]
    line = {
        "metric": METRIC, "value": round(rhs_s, 2), "unit": "RHS/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(mixed["wall"] / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "C4: %d independent banded systems in one batched solve, n log-uniform [1e3, 2e5], "
                               "half-bandwidth in {4,8,16,32,64}, 30%% fill, dominance 1.1, nrhs 1-8" % len(systems),
                   "systems": len(systems), "rows": rows, "nnz": nnz, "nrhs_total": rhs, "tol": C4["tol"],
                   "seed": C4["seed"], "precision": "mixed: fp32 factor/spikes, fp64 residual/reduced/update",
                   "partitions": "autotuned for throughput (about 96 blocks per partition)"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "accuracy": {"max_rel_residual": float(mixed["rel"].max()), "solve_passes": passes},
        "fp64": f64, "mixed_vs_fp64": round(rhs_s / f64["rhs_per_s"], 3) if f64 else None,
        "graph": {"replayed_steps": int(sum(x["graph_replayed"] for x in mixed["stats"])),
                  "eager_rhs_per_s": round(len(eager["lat"]) * rhs / eager["wall"], 2),
                  "eager_ms_per_step": round(eager["wall"] / len(eager["lat"]) * 1e3, 3)},
        "stage_ms": {**{kk: round(v, 4) for kk, v in avg.items()}, **{kk: round(v, 4) for kk, v in avg_other.items()}},
        "roofline": roof, "stage_roofline": stage_roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


def bench_serve(args, hs):
    """C5: Poisson request stream through the request server (spike_server_*); throughput is the
    RHS solved over the span from the first arrival to the last completion."""
    import threading
    reqs, mats = c5_stream(args.nreq, C5["seed"])
    t_gen = time.time()
    A = []
    for j, (cplx, n, k, sd) in enumerate(mats):
        A.append(hs.gen_banded(n, k, k, fill=C5["fill"], dom=(C5["dom"], C5["dom"]), diag_floor=1.0, seed=sd,
                               complex=cplx))
        if (j + 1) % 16 == 0:
            log("c5: %d of %d matrices generated" % (j + 1, len(mats)))
    Bs = [hs.gen_dense(A[mid].n, nrhs, seed=sb, complex=mats[mid][0]) for (_, mid, nrhs, sb) in reqs]
    nnz = sum(a.nnz for a in A)
    rhs = sum(r[2] for r in reqs)
    log("generated c5: %d requests, %d distinct matrices (nnz %d), %d RHS in %.1f s" %
        (len(reqs), len(mats), nnz, rhs, time.time() - t_gen))

    def run(precision):
        srv = hs.Server(precision, tol=C5["tol"], workers=3, max_batch=8,
                        verbose=os.environ.get("HSPIKE_VERBOSE", "") == "1")
        tickets = [None] * len(reqs)
        t0 = time.perf_counter()

        def submitter():
            for i, (ta, mid, nrhs, _) in enumerate(reqs):
                dt = t0 + ta / args.rate - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                tickets[i] = srv.submit(A[mid], Bs[i], key=mid + 1)     # the client names its matrices

        th = threading.Thread(target=submitter)
        th.start()
        infos = []
        for i in range(len(reqs)):
            while tickets[i] is None:
                time.sleep(0.0005)
            X, info = srv.wait(tickets[i])
            infos.append(info)
            if (i + 1) % 40 == 0:
                log("c5 %s: %d of %d requests done" % (precision, i + 1, len(reqs)))
        th.join()
        span = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        bad = [x for x in infos if x["status"] != 0]
        if bad:
            codes = {}
            for x in bad:
                codes[x["status"]] = codes.get(x["status"], 0) + 1
            log("c5 %s: failed requests by status %s" % (precision, codes))
        lat = np.array([x["latency_ms"] for x in infos])
        return dict(span=span, stats=st, failed=len(bad), lat=lat, maxrel=max(x["max_rel_residual"] for x in infos),
                    hits=sum(x["resident_hit"] for x in infos))

    mixed = run("mixed")
    log("mixed: span %.2f s, p50 %.1f ms, p99 %.1f ms, failed %d, oom retries %d, estimate coverage %.3f" %
        (mixed["span"], np.percentile(mixed["lat"], 50), np.percentile(mixed["lat"], 99), mixed["failed"],
         mixed["stats"]["oom_retries"], mixed["stats"]["est_ratio_max"]))
    f64 = None
    if not args.no_fp64:
        r = run("fp64")
        f64 = {"rhs_per_s": round(rhs / r["span"], 2), "p50_ms": round(float(np.percentile(r["lat"], 50)), 3),
               "p99_ms": round(float(np.percentile(r["lat"], 99)), 3), "failed": r["failed"],
               "max_rel_residual": r["maxrel"], "oom_retries": r["stats"]["oom_retries"]}
        log("fp64: %s" % f64)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle.build()
        order = [i for i, (_, mid, _, _) in enumerate(reqs) if A[mid].nnz <= 20_000_000]

        def job(i):
            _, mid, nrhs, _ = reqs[i]
            a, k = A[mid], mats[mid][2]
            return (lambda: oracle.CSR(a.n, a.rowptr, a.colind, a.aij),
                    lambda d, nthreads: oracle.spike_solve(d, Bs[i], max(1, min(4, a.n // (4 * k + 4))),
                                                           nthreads=nthreads, want_reduced=False))
        cpu = oracle_baseline([job(i) for i in order], [reqs[i][2] for i in order], 20.0, 10.0,
                              "requests in arrival order up to 2e7 nonzeros, oracle SPIKE FP64 (up to 4 partitions)")
    rhs_s = rhs / mixed["span"]
    line = {
        "metric": METRIC, "value": round(rhs_s, 2), "unit": "RHS/s", "n_gpus": 1, "steps": len(reqs), "warmup": 0,
        "ms_per_step": round(mixed["span"] * 1e3 / len(reqs), 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "C5: Poisson request stream, %.0f req/s, 80%% real / 20%% complex128, n log-uniform "
                               "[1e4, 5e6], half-bandwidth 2-256, nrhs 1-64, 30%% repeat-RHS" % args.rate,
                   "requests": len(reqs), "distinct_matrices": len(mats), "nnz_total": nnz, "nrhs_total": rhs,
                   "arrival_rate_per_s": args.rate, "fill": C5["fill"], "seed": C5["seed"], "workers": 3, "max_batch": 8,
                   "precision": "mixed: fp32 factor/spikes, fp64 residual/reduced/update"},
        "latency_ms": {"p50": round(float(np.percentile(mixed["lat"], 50)), 3),
                       "p99": round(float(np.percentile(mixed["lat"], 99)), 3)},
        "accuracy": {"max_rel_residual": mixed["maxrel"], "failed": mixed["failed"]},
        "serving": {"resident_hits": mixed["hits"], "peak_device_bytes": mixed["stats"]["peak_bytes"],
                    "budget_bytes": mixed["stats"]["budget_bytes"], "span_s": round(mixed["span"], 3),
                    "oom_retries": mixed["stats"]["oom_retries"], "estimate_coverage_max": mixed["stats"]["est_ratio_max"]},
        "fp64": f64, "mixed_vs_fp64": round(rhs_s / f64["rhs_per_s"], 3) if f64 else None,
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
