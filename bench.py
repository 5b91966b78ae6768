"""Benchmark: UCHOL factor + solve to relative residual 1e-10 on one MI355X.

Workload (default, --config c3): 3D linear elasticity, 40^3 trilinear hex
elements (41^3 nodes x 3 DOF = 206,763 equations), Young's modulus
lognormal(0, 0.5) per element, Poisson ratio 0.3, x = 0 face clamped through
LineCross, 8 right-hand sides of nodal loads N(0,1).  One step = numeric
factorization of the resident matrix values + solve of the 8 right-hand sides
to max relative residual <= 1e-10:
  * headline: FP32 factor (Schur updates on FP32 MFMA) + FP64 iterative refinement;
  * reported beside it: plain FP64 factor + solve on the GPU, and the CPU
    restatement of the reference (oracle/, FP64, threaded with the reference's
    schedule) timed on the whole workload on the host's cores;
  * extra keys: the batched workload (1,024 uneven 2D systems, --config c4)
    and the learned-preconditioner PCG (--config c5), each also runnable alone.
Inputs (matrix values, right-hand sides) are resident in HBM before timing;
the symbolic phase (host, bit-exact AMD etc.) is done once and reported
separately, as the reference's benchmark does (UtduSolverBenchmark.cs:74-77).

Prints ONE JSON line (rank 0).  Only one GPU is used: --gpus above 1 exits.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm_gbs": 8000.0, "fp32_mfma_tflops": 157.3, "fp64_tflops": 78.6}
# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs and scaled by the same counters on a copy of known size (--calibrate-copy)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic.json")
CALIBRATION_BYTES = 256 << 20


def measured_traffic(config, kernel):
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)[config][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


GRAPH_MODES = {0: False, 1: "factorization and solves", 2: "solves"}
COLD_BYTES = 512 << 20  # written before every cold step: larger than the 256 MB Infinity Cache


def cold_ms(step_fn, steps, stream):
    """Mean time of `steps` steps, each after writing COLD_BYTES of scratch (caches hold no
    operand of the step when it starts)."""
    import torch
    scratch = torch.empty(COLD_BYTES // 4, dtype=torch.float32, device="cuda:0")
    times = []
    for i in range(steps):
        scratch.fill_(float(i))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step_fn()
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    del scratch
    return sum(times) / len(times)


def calibrate_copy():
    """One plain device copy of CALIBRATION_BYTES (read and written once each): the reference
    point that scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run."""
    import torch
    src = torch.ones(CALIBRATION_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_system(config):
    import magla_amd as M
    from magla_amd import generators as G
    if config == "c3":
        n, r, c, v, clamped, b = G.config_c3()
        a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
        a.line_cross(np.zeros(n), 0.0, clamped)
        desc = ("3D elasticity 40^3 hex8 (n=206763), E~lognormal(0,0.5), nu=0.3, x=0 clamped, 8 RHS N(0,1); "
                "factor+solve to relres<=1e-10")
    elif config == "c2":
        n, r, c, v, b = G.config_c2()
        a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
        desc = "3D 7-point diffusion 48^3 (n=110592), k~lognormal(0,1), 4 RHS U[-1,1]; factor+solve to relres<=1e-10"
    else:
        n, r, c, v, b = G.config_c1()
        b = b[:, None]
        a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
        desc = "2D 5-point Laplacian 64x64 (n=4096), 1 RHS; factor+solve to relres<=1e-10"
    return a, np.asfortranarray(b), desc


def run_batch(args):
    """1,024 independent 2D 5/9-point systems (side 16..128, shift U[0,0.1], 1..32 RHS each):
    one step = batched FP32 factorization + solve with FP64 refinement of every system to 1e-10.
    Returns the result dict."""
    import torch

    import magla_amd as M
    from magla_amd import generators as G
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    t0 = time.time()
    sysl = G.config_c4(args.batch_count)
    mats = [M.SymmetricCSRMatrix.from_coo(n, r, c, v) for n, r, c, v, b in sysl]
    log(f"[bench] {len(mats)} systems assembled ({time.time() - t0:.1f}s)")
    t0 = time.time()
    b32 = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP32, device_graph=bool(args.graph)))
    symbolic_s = time.time() - t0
    info = b32.info()
    b64 = M.UtduBatch(mats, M.UtduSolverOptions(precision=M.FP64, device_graph=bool(args.graph)))
    nt = int(b32.row_begin[-1])
    k = max(b.shape[1] for *_, b in sysl)
    nrhs_total = sum(b.shape[1] for *_, b in sysl)
    log(f"[bench] batch symbolic {symbolic_s:.2f}s (host threads): n={nt} rhs slots={k} rhs={nrhs_total} "
        f"ops={info['factor_operation_count']:.3e}")
    vals = torch.from_numpy(b32.values()).to(dev)
    B = torch.zeros((k, nt), dtype=torch.float64, device=dev)
    for t, (n, *_, b) in enumerate(sysl):
        r0 = int(b32.row_begin[t])
        B[:b.shape[1], r0:r0 + n] = torch.from_numpy(np.ascontiguousarray(b.T)).to(dev)
    X = torch.zeros_like(B)
    torch.cuda.synchronize()

    def step(batch):
        batch.factorize_dev(vals.data_ptr(), stream=sp)
        return batch.solve_dev(B.data_ptr(), X.data_ptr(), k, tol=args.tol, max_refine=30, stream=sp)

    def timed(batch, steps):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        reps = [step(batch) for _ in range(steps)]
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps, reps

    tiles = {}
    for name, batch in (("fp32", b32), ("fp64", b64)):
        for _ in range(args.warmup):
            rep = step(batch)
        if args.autotune:
            t, tms = batch.autotune_dev(vals.data_ptr(), stream=sp)
            tiles[name] = t.tolist()
            log(f"[bench] autotuned Schur tiles {name}: {t.tolist()} ({tms:.2f} ms factor)")
    ms32, reps32 = timed(b32, args.steps)
    log(f"[bench] batch fp32+IR: {ms32:.2f} ms/step {reps32[-1]}")
    ms64, reps64 = timed(b64, args.steps)
    log(f"[bench] batch fp64: {ms64:.2f} ms/step {reps64[-1]}")
    cold32 = cold_ms(lambda: step(b32), args.steps, stream)
    cold64 = cold_ms(lambda: step(b64), args.steps, stream)
    log(f"[bench] batch cold caches: fp32+IR {cold32:.2f} ms/step, fp64 {cold64:.2f} ms/step")
    M.profile_enable(True)
    step(b32)
    torch.cuda.synchronize()
    prof = M.profile_read()
    M.profile_enable(False)
    name, t = max(prof.items(), key=lambda kv: kv[1]["ms"])
    roof = roofline_of(name, t, "c4")
    cpu = None
    if args.cpu_seconds > 0:
        # oracle numeric factorization of the systems in order until the time budget is spent,
        # scaled to the whole batch by FactorOperationCount (symbolic phase not timed, solves omitted)
        import oracle as O
        num_s, ops_done, k_done = 0.0, 0, 0
        for n_, r_, c_, v_, _b in sysl:
            oc = O.Csr.from_coo(n_, r_, c_, v_)
            osym = O.Symbolic(oc)
            c0 = time.time()
            O.Numeric(osym, oc)
            num_s += time.time() - c0
            ops_done += osym.factor_ops
            k_done += 1
            if num_s >= args.cpu_seconds:
                break
        total_ops = info["factor_operation_count"]
        cpu = {"value": round(num_s * total_ops / ops_done * 1e3, 3), "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"oracle (serial FP64 C restatement of UtduNumericFactorization) factored the first "
                         f"{k_done}/{len(sysl)} systems = {100.0 * ops_done / total_ops:.2f}% of FactorOperationCount "
                         f"in {num_s:.1f}s; value = that time scaled by operation count to the whole batch "
                         f"(solves omitted)"}
    out = {
        "metric": "factor+solve time to 1e-10 residual, FP32-MFMA+IR vs FP64, 1 MI355X",
        "value": round(ms32, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms32, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(mats)} independent 2D 5/9-point Laplacians, side U{{16..128}}, shift "
                               f"U[0,0.1], 1..32 RHS each (padded to {k}); batched factor+solve to relres<=1e-10 "
                               f"per system and column",
                   "n_total": nt, "rhs_total": nrhs_total, "factor_operation_count": info["factor_operation_count"],
                   "hipgraph": bool(args.graph), "schur_tiles": tiles},
        "fp64_ms_per_step": round(ms64, 3), "speedup_fp32ir_vs_fp64": round(ms64 / ms32, 3),
        "cold_ms_per_step": round(cold32, 3), "cold_fp64_ms_per_step": round(cold64, 3),
        "fp32_relres": reps32[-1]["relative_residual"], "fp32_refinement_steps": reps32[-1]["refinement_steps"],
        "symbolic_s": round(symbolic_s, 3), "roofline": roof,
        "kernels": kernel_table(prof), "cpu_baseline": cpu,
    }
    return out


def bench_batch(args):
    print(json.dumps(run_batch(args)), flush=True)


def run_learned(args):
    """Learned-preconditioner PCG on 3D 7-point diffusion m^3 with kappa = exp(GRF): pairs from
    Jacobi-PCG snapshots and the direct solver, GPU training, then iterations and wall time to
    1e-8 against Jacobi-PCG.  Returns the result dict."""
    import torch

    import magla_amd as M
    from magla_amd import generators as G
    from magla_amd import learned
    sp = torch.cuda.current_stream().cuda_stream
    t0 = time.time()
    n, r, c, v, b = G.config_c5(args.grid)
    a = M.SymmetricCSRMatrix.from_coo(n, r, c, v)
    dm = M.DeviceMatrix(a)
    log(f"[bench] c5 system n={n} built ({time.time() - t0:.1f}s)")
    t0 = time.time()
    R, E = learned.make_pairs(a, dm, args.pairs, seed=G.SEED_BASE + 5 + 2000, max_iterations=args.snapshot_its)
    pairs_s = time.time() - t0
    log(f"[bench] {args.pairs} pairs in {pairs_s:.1f}s")
    net = M.LearnedPreconditioner(seed=G.SEED_BASE + 5 + 3000)
    t0 = time.time()
    losses = learned.train(net, dm, R, E, epochs=args.epochs, batch=16, learning_rate=args.lr, kappa=args.kappa,
                           seed=G.SEED_BASE + 5 + 4000)
    train_s = time.time() - t0
    log(f"[bench] trained {args.epochs} epochs in {train_s:.1f}s: loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    del R, E
    dev = torch.device("cuda:0")
    bd = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    xd = torch.zeros_like(bd)
    jac = M.ConjugateGradientGaussPreSolver(relative_tolerance=1e-8)
    lrn = M.LearnedConjugateGradientSolver(net, relative_tolerance=1e-8)

    def run(s):
        torch.cuda.synchronize()
        w0 = time.time()
        res = s.solve_dev(dm, bd.data_ptr(), xd.data_ptr(), stream=sp)
        torch.cuda.synchronize()
        return (time.time() - w0) * 1e3, res

    for s in (jac, lrn):
        for _ in range(args.warmup):
            run(s)
    tj = [run(jac) for _ in range(args.steps)]
    tl = [run(lrn) for _ in range(args.steps)]
    msj, resj = min(x[0] for x in tj), tj[-1][1]
    msl, resl = min(x[0] for x in tl), tl[-1][1]
    log(f"[bench] jacobi-pcg: {resj.iterations} it, {msj:.1f} ms; learned: {resl.iterations} it, {msl:.1f} ms "
        f"(fell back: {lrn.fell_back})")
    # the network runs once per matrix (BF16 inference of the factor G); setting the parameters
    # again drops the cached factor, so this solve pays for the inference as well
    net.params = net.params
    ms_inf, _ = run(lrn)
    log(f"[bench] learned solve including the per-matrix inference: {ms_inf:.1f} ms")

    # roofline of the learned PCG iteration (HBM-bound): full-pattern SpMV, the products with G
    # and G' (FP32 value + int32 column per entry, two FP64 vectors per row), and the 11
    # n-vector streams of a PCG iteration
    import scipy.sparse as sps
    full = sps.coo_matrix((np.ones(len(r)), (r, c)), shape=(n, n)).tocsr()
    full = ((full + full.T) != 0).astype(np.float64)
    nnz_full = full.nnz
    two_hop = sps.tril((full @ full) + full, format="csr")
    two_hop.data[:] = 1.0
    g_entries = int(two_hop.nnz)
    bytes_it = 12.0 * nnz_full + 4.0 * (n + 1) + 16.0 * n + 2 * (8.0 * g_entries + 16.0 * n) + 88.0 * n
    achieved = bytes_it * resl.iterations / (msl * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
            "frac": achieved / PEAK["hbm_gbs"], "traffic": None, "kernel": "learned PCG iteration",
            "bytes_per_iteration": bytes_it, "g_entries": g_entries, "nnz_full": int(nnz_full)}

    cpu = None
    if args.cpu_seconds > 0:
        import oracle as O
        oc = O.Csr.from_coo(n, r, c, v)
        c0 = time.time()
        _, ores = O.pcg(oc, np.asarray(b, dtype=np.float64).reshape(-1), tol=1e-8)
        cpu_s = time.time() - c0
        cpu = {"value": round(cpu_s * 1e3, 3), "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"oracle (serial C restatement of ConjugateGradientGaussPreSolver + JacobiPreconditioner), "
                         f"the whole Jacobi-PCG solve of this system: {ores['iterations']} iterations to 1e-8"}
    out = {
        "metric": "PCG time to 1e-8 relative residual, learned vs Jacobi preconditioner, 1 MI355X",
        "value": round(msl, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(msl, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"3D 7-point diffusion {args.grid}^3 (n={n}), kappa=exp(GRF l=0.1 var=2), "
                               f"PCG to 1e-8; 3-layer MPNN width 32 -> two-hop factor G, z = G'G r, trained on {args.pairs} "
                               f"(r, A^-1 r) pairs (kappa {args.kappa})",
                   "pairs": args.pairs, "epochs": args.epochs},
        "jacobi_iterations": resj.iterations, "jacobi_ms": round(msj, 3),
        "learned_iterations": resl.iterations, "learned_ms": round(msl, 3),
        "iteration_ratio": round(resj.iterations / max(resl.iterations, 1), 3),
        "time_ratio": round(msj / msl, 3), "fell_back": lrn.fell_back,
        "learned_relres": resl.relative_residual, "jacobi_relres": resj.relative_residual,
        "train_losses": [round(float(x), 5) for x in losses], "pairs_s": round(pairs_s, 2),
        "train_s": round(train_s, 2), "learned_ms_incl_inference": round(ms_inf, 3),
        "roofline": roof, "cpu_baseline": cpu,
    }
    return out


def bench_learned(args):
    print(json.dumps(run_learned(args)), flush=True)


def cpu_threads():
    """Host threads for the CPU baseline: the CPU share this process may use (affinity), at
    most 16 (the host share of one GPU on the benchmark machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, min(n, 16))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_bicgstab(args):
    """ILU(0)-preconditioned BiCGSTAB on 2D upwind convection-diffusion (grid^2, natural order):
    the ILU(0) factorization once, then the solve to 1e-8 (timed), against the oracle's serial
    restatement of BiCgStabSolver on the host."""
    import torch

    import magla_amd as M
    from magla_amd import generators as G
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    m = args.grid
    n, r, c, v = G.convection_diffusion_2d(m)
    a = M.CSRMatrix.from_coo(n, n, r, c, v)
    b = G.uniform_rhs(n, 1, G.SEED_BASE + 6 + 1000).reshape(-1)
    dev = torch.device("cuda:0")
    bd = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    xd = torch.zeros_like(bd)
    torch.cuda.synchronize()
    t0 = time.time()
    pre = M.Ilu0Preconditioner(a)
    ilu_ms = (time.time() - t0) * 1e3
    s = M.BiCgStabSolver(relative_tolerance=1e-8)

    def run():
        torch.cuda.synchronize()
        w0 = time.time()
        res = s.solve_dev(a, bd.data_ptr(), xd.data_ptr(), preconditioner=pre, stream=sp)
        torch.cuda.synchronize()
        return (time.time() - w0) * 1e3, res

    for _ in range(args.warmup):
        run()
    times = [run() for _ in range(args.steps)]
    ms, res = min(t[0] for t in times), times[-1][1]
    x = xd.cpu().numpy()
    relres = float(np.linalg.norm(b - a.multiply(x)) / np.linalg.norm(b))
    log(f"[bench] bicgstab+ilu0 {m}^2: {res.iterations} it, {ms:.2f} ms (ILU(0) {ilu_ms:.1f} ms), relres {relres:.2e}")
    nnz = a.nonzero_count
    # per iteration: two SpMVs (12 B per entry + row pointers + x read + y written) and two
    # ILU(0) applications (12 B per factor entry, diagonal positions, three vector streams),
    # plus the 14 n-vector streams of the BiCGSTAB updates and dot products
    spmv_b = 12.0 * nnz + 4.0 * (n + 1) + 16.0 * n
    ilu_b = 12.0 * nnz + 8.0 * (n + 1) + 24.0 * n
    bytes_it = 2 * spmv_b + 2 * ilu_b + 14 * 8.0 * n
    achieved = bytes_it * res.iterations / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
            "frac": achieved / PEAK["hbm_gbs"], "traffic": None, "kernel": "BiCGSTAB+ILU(0) iteration",
            "bytes_per_iteration": bytes_it, "levels_note": f"ILU(0) solves walk {2 * m - 1} dependent levels"}
    cpu = None
    if args.cpu_seconds > 0:
        import oracle as O
        o = O.GCsr.from_coo(n, n, r, c, v)
        c0 = time.time()
        _, ores = O.bicgstab(o, b, tol=1e-8)
        cpu = {"value": round((time.time() - c0) * 1e3, 3), "unit": "ms", "cores": 1, "kind": "port",
               "sample": f"oracle (serial C restatement of BiCgStabSolver + Ilu0Preconditioner), the whole solve "
                         f"including the ILU(0) factorization: {ores['iterations']} iterations to 1e-8"}
    out = {
        "metric": "BiCGSTAB+ILU(0) time to 1e-8 relative residual, 1 MI355X",
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2D upwind convection-diffusion {m}x{m} (n={n}, nnz={nnz}), natural order, "
                               f"Peclet 2, ILU(0)-BiCGSTAB to 1e-8"},
        "iterations": res.iterations, "relative_residual": relres, "ilu0_ms": round(ilu_ms, 3),
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


MFMA_PHASES = {"schur"}  # the Schur update runs on the matrix cores; every other phase streams HBM


def roofline_of(name, t, config=None):
    avg_ms = t["ms"] / max(t["launches"], 1)
    if name in MFMA_PHASES:
        achieved = t["flops"] / (t["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK["fp32_mfma_tflops"], "unit": "TFLOP/s",
                "frac": achieved / PEAK["fp32_mfma_tflops"], "traffic": None}
    else:
        achieved = t["bytes"] / (t["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                "frac": achieved / PEAK["hbm_gbs"], "traffic": None}
    if config is not None:
        roof["traffic"] = measured_traffic(config, name)
    roof.update(kernel=name, launches=t["launches"], avg_launch_ms=avg_ms, kernel_ms_per_step=t["ms"])
    return roof


def kernel_table(prof):
    return {kn: {"launches": kv["launches"], "ms": round(kv["ms"], 3),
                 "TFLOP/s": round(kv["flops"] / max(kv["ms"], 1e-9) / 1e9, 2),
                 "GB/s": round(kv["bytes"] / max(kv["ms"], 1e-9) / 1e6, 1)} for kn, kv in prof.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "bicg"])
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="c4: bounded CPU-baseline sample in seconds (0 = skip)")
    ap.add_argument("--cpu", type=int, default=1, help="c3: time the threaded CPU restatement on the whole workload")
    ap.add_argument("--extras", type=int, default=1, help="c3: add the batched (c4) and learned-PCG (c5) results")
    ap.add_argument("--graph", type=int, default=1, choices=(0, 1, 2),
                    help="hipGraph replay: 1 = factorization and solves, 2 = solves only")
    ap.add_argument("--autotune", type=int, default=1, help="pick Schur tile shapes per front-size bucket first")
    ap.add_argument("--calibrate-copy", type=int, default=0, help="run one plain copy of 256 MiB first (PMC scale)")
    ap.add_argument("--batch-count", type=int, default=1024, help="c4: number of systems")
    ap.add_argument("--grid", type=int, default=64, help="c5 / bicg: grid edge")
    ap.add_argument("--pairs", type=int, default=2000, help="c5: training pairs")
    ap.add_argument("--epochs", type=int, default=20, help="c5: training epochs")
    ap.add_argument("--lr", type=float, default=1e-2, help="c5: Adam learning rate")
    ap.add_argument("--kappa", type=float, default=1.0, help="c5: weight of the Kaporin condition-number term")
    ap.add_argument("--snapshot-its", type=int, default=256, help="c5: largest Jacobi-PCG snapshot iteration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
    if args.config == "c4":
        return bench_batch(args)
    if args.config == "c5":
        return bench_learned(args)
    if args.config == "bicg":
        return bench_bicgstab(args)

    import torch

    import magla_amd as M
    assert torch.cuda.is_available(), "no GPU visible"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    t0 = time.time()
    a, b, desc = build_system(args.config)
    n, k = b.shape
    log(f"[bench] system built: n={n} nnz_upper={a.nonzero_count} rhs={k} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    solver64 = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP64, device_graph=args.graph))
    solver32 = M.SymmetricUtduSolver(M.UtduSolverOptions(precision=M.FP32, device_graph=args.graph))
    sym = solver64.analyze(a)
    symbolic_s = time.time() - t0
    info = sym.info()
    log(f"[bench] symbolic {symbolic_s:.2f}s: supernodes={info['supernode_count']} max_front={info['max_front_size']} "
        f"factor_entries={info['factor_entry_count']} ops={info['factor_operation_count']:.3e}")

    d_vals = torch.from_numpy(a.values).to(dev)
    d_b = torch.from_numpy(np.ascontiguousarray(b.T)).to(dev)  # (k, n): column c at offset c*n
    d_x = torch.zeros_like(d_b)
    nnz = a.nonzero_count
    torch.cuda.synchronize()

    num = {}

    def step(prec):
        solver = solver32 if prec == M.FP32 else solver64
        num[prec] = solver.factorize_dev(sym, d_vals.data_ptr(), nnz, numeric=num.get(prec), stream=sp)
        return num[prec].solve_dev(d_b.data_ptr(), d_x.data_ptr(), k, tol=args.tol, max_refine=30, stream=sp)

    def timed(prec, steps):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0 = time.time()
        e0.record(stream)
        reps = [step(prec) for _ in range(steps)]
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps, (time.time() - w0) * 1e3 / steps, reps

    tiles = {}
    for prec in (M.FP32, M.FP64):
        for _ in range(args.warmup):
            rep = step(prec)
            log(f"[bench] warmup prec={prec} done: {rep}")
        if args.autotune:
            if prec not in num:
                step(prec)
            t, tms = num[prec].autotune_dev(d_vals.data_ptr(), stream=sp)
            tiles["fp32" if prec == M.FP32 else "fp64"] = t.tolist()
            log(f"[bench] autotuned Schur tiles prec={prec}: {t.tolist()} ({tms:.2f} ms factor)")
    ms32, wall32, reps32 = timed(M.FP32, args.steps)
    log(f"[bench] fp32+IR: {ms32:.2f} ms/step (wall {wall32:.2f}) {reps32[-1]}")
    ms64, wall64, reps64 = timed(M.FP64, args.steps)
    log(f"[bench] fp64: {ms64:.2f} ms/step (wall {wall64:.2f}) {reps64[-1]}")
    cold32 = cold_ms(lambda: step(M.FP32), args.steps, stream)
    cold64 = cold_ms(lambda: step(M.FP64), args.steps, stream)
    log(f"[bench] cold caches: fp32+IR {cold32:.2f} ms/step, fp64 {cold64:.2f} ms/step")

    # forward error of the FP32+IR solution against the FP64 factor's solution
    step(M.FP64)
    torch.cuda.synchronize()
    x64 = d_x.clone()
    step(M.FP32)
    torch.cuda.synchronize()
    fwd_err = float((torch.linalg.norm(d_x - x64) / torch.linalg.norm(x64)).item())

    # instrumented pass: per-kernel event timing with algorithmic work
    M.profile_enable(True)
    step(M.FP32)
    torch.cuda.synchronize()
    prof = M.profile_read()
    M.profile_enable(False)
    name, t = max(prof.items(), key=lambda kv: kv[1]["ms"])
    roof = roofline_of(name, t, args.config)
    kernels = kernel_table(prof)
    log(f"[bench] kernels: {json.dumps(kernels)}")

    cpu = None
    fwd_err_ref = None
    if args.cpu:
        # the reference's FP64 factorization and solve restated in C (oracle/), threaded with the
        # reference's schedule, timed whole on the host's cores: no extrapolation
        import oracle as O
        threads = cpu_threads()
        oc = O.Csr.from_csr(n, a.row_pointers, a.column_indices, a.values)
        osym = O.Symbolic(oc)
        c0 = time.time()
        onum = O.Numeric(osym, oc, threads=threads)
        fact_s = time.time() - c0
        c0 = time.time()
        x_ref = onum.solve_many(b, threads=threads)
        solve_s = time.time() - c0
        rel_ref = max(np.linalg.norm(b[:, j] - oc.spmv(x_ref[:, j])) / np.linalg.norm(b[:, j]) for j in range(k))
        x32 = d_x.cpu().numpy().T  # the FP32+IR solution of the forward-error pair above
        fwd_err_ref = float(np.linalg.norm(x32 - x_ref) / np.linalg.norm(x_ref))
        cpu = {"value": round((fact_s + solve_s) * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
               "cpu": cpu_model(),
               "sample": (f"the whole workload: oracle (C restatement of UtduNumericFactorization, -O3 AVX2, "
                          f"the reference's tree schedule and parallel trailing updates over {threads} threads) "
                          f"FP64 factorization {fact_s:.2f}s ({osym.factor_ops / fact_s / 1e9:.1f} GFLOP/s) + "
                          f"{k} solves {solve_s:.2f}s (one thread per right-hand side); relres {rel_ref:.2e}")}
        log(f"[bench] cpu baseline: {cpu}")

    gflops32 = info["factor_operation_count"] / (ms32 * 1e-3) / 1e9
    out = {
        "metric": "factor+solve time to 1e-10 residual, FP32-MFMA+IR vs FP64, 1 MI355X",
        "value": round(ms32, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "n": int(n), "nnz_upper": int(nnz), "rhs": int(k),
                   "supernodes": info["supernode_count"], "max_front": info["max_front_size"],
                   "factor_operation_count": info["factor_operation_count"], "tol": args.tol,
                   "precision": "FP32 factor (MFMA Schur) + FP64 residual/refinement", "schur_tiles": tiles,
                   "hipgraph": GRAPH_MODES[args.graph]},
        "fp64_ms_per_step": round(ms64, 3),
        "speedup_fp32ir_vs_fp64": round(ms64 / ms32, 3),
        "cold_ms_per_step": round(cold32, 3),
        "cold_fp64_ms_per_step": round(cold64, 3),
        "fp32_relres": reps32[-1]["relative_residual"],
        "fp32_refinement_steps": reps32[-1]["refinement_steps"],
        "fp64_relres": reps64[-1]["relative_residual"],
        "forward_error_vs_fp64": fwd_err,
        "forward_error_vs_reference": fwd_err_ref,
        "factor_gflops_fp32": round(gflops32, 1),
        "symbolic_s": round(symbolic_s, 3),
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    if args.extras:
        # the other two benchmark workloads, as extra keys of this line (their own lines: --config c4 / c5)
        sub = argparse.Namespace(**vars(args))
        sub.cpu_seconds = 0
        sub.graph = 1
        c4 = run_batch(sub)
        out["c4_batched"] = {k_: c4[k_] for k_ in ("value", "fp64_ms_per_step", "speedup_fp32ir_vs_fp64",
                                                   "fp32_relres", "fp32_refinement_steps", "cold_ms_per_step")}
        out["c4_batched"].update(workload=c4["config"]["workload"], hipgraph=True, unit="ms")
        sub.graph = 0
        c5 = run_learned(sub)
        out["c5_learned_pcg"] = {k_: c5[k_] for k_ in ("jacobi_iterations", "jacobi_ms", "learned_iterations",
                                                       "learned_ms", "learned_ms_incl_inference", "iteration_ratio",
                                                       "time_ratio", "fell_back", "learned_relres", "pairs_s",
                                                       "train_s")}
        out["c5_learned_pcg"].update(workload=c5["config"]["workload"], unit="ms")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
