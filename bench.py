"""Benchmark: LSQR time-to-tol on the 67^3-grid / 2,000,000-ray system, one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed]
                    [--precond none|col] [--refine-adaptive] [--learned]

A step is one complete LSQR solve from x = 0 to test2 <= 1e-6 (test2 as
lsqr.c:742-744) on the device-resident system; W untimed solves, then K timed
solves between device synchronisations.  value = mean seconds per solve.
The default line is plain FP64 LSQR (the reference's method); beside it
("mixed_colscale") the same b solved in mixed precision on A D (column
scaling) with adaptive FP64 checks, stopping on the same test2 (the FP64
solve's anorm), K timed solves as well.

Also measured inside the run, on the handle's stream with HIP events: the
average duration of one aprod of each mode, priced against the algorithmic
bytes of that launch (each stored value and index read once, each pointer
array once, each dense input once, each dense output read and written once).
The CPU baseline is the reference itself (compiled from its sources, one
MPI rank), or the oracle (the C restatement of the reference, one thread)
when that binary is absent, timed on a bounded sample of the same solve and
scaled to the plain FP64 solve's iteration count.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "LSQR time-to-tol 1e-6 (s) + aprod HBM GB/s, mixed/learned vs plain FP64"
HBM_PEAK_GBS = 8000.0


def load_pkg():
    name = "plsqr_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = os.path.join(ROOT, "plsqr3.2013.09_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


class Hip:
    """The few HIP runtime calls the benchmark needs (device buffers, events)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        L = self.lib
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        L.hipEventSynchronize.argtypes = [C.c_void_p]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.hipEventDestroy.argtypes = [C.c_void_p]

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def free(self, p):
        self.lib.hipFree(C.c_void_p(p))

    def h2d(self, dst, arr):
        self.check(self.lib.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1),
                   "hipMemcpy")

    def d2h(self, arr, src):
        self.check(self.lib.hipMemcpy(arr.ctypes.data_as(C.c_void_p), C.c_void_p(src), arr.nbytes, 2),
                   "hipMemcpy")

    def sync(self):
        self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def event(self):
        e = C.c_void_p()
        self.check(self.lib.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e

    def record(self, e, stream):
        self.check(self.lib.hipEventRecord(e, C.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, e0, e1):
        self.check(self.lib.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        self.check(self.lib.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        return ms.value


def aprod_bytes(m_k, m_d, n, nnz_k, nnz_d, s_v=8, s_x=8, s_y=8):
    """Algorithmic bytes of one aprod call (mode 1, mode 2)."""
    b1 = (nnz_k * (s_v + 4) + 4 * (m_k + 1) + s_x * n + 2 * s_y * m_k
          + nnz_d * (s_v + 4) + 4 * (m_d + 1) + 2 * s_y * m_d)
    b2 = (nnz_k * (s_v + 4) + 4 * (n + 1) + s_y * m_k
          + nnz_d * (s_v + 4) + 4 * (n + 1) + s_y * m_d + 2 * s_x * n)
    return b1, b2


def time_aprod(hip, op, mode, x_dev, y_dev, reps, f32=False, internal=True):
    """Average ms of one aprod launch (FP64, or the mixed solver's FP32
    operator).  internal: on vectors in the handle's order, i.e. the SpMV
    kernels alone; otherwise through the caller-order entry point, which adds
    the vector reordering at the API boundary."""
    if internal:
        call = op.aprod_device_internal_f32 if f32 else op.aprod_device_internal
    else:
        call = op.aprod_device_f32 if f32 else op.aprod_device
    stream = op.lib.plsqr_op_get_stream(op.h)
    for _ in range(3):
        call(mode, x_dev, y_dev)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(reps):
        call(mode, x_dev, y_dev)
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1) / reps
    hip.lib.hipEventDestroy(e0)
    hip.lib.hipEventDestroy(e1)
    return ms


def read_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summaries."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


REF_BIN = os.path.join(ROOT, "oracle", "_ref", "plsqr3_ref_fp64")
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    """Cores this process may run on (the box's share), capped at 16 ranks."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def run_reference(s, names, work, sample_iters, nproc):
    """PLSQR3 itself on nproc MPI ranks (its even column partition,
    plsqr3.c:227-250); returns its "Parallel LSQR time" (the loop alone,
    lsqr.c:476-842) and the wall time of the run."""
    import subprocess
    cmd = [MPIEXEC, "-np", str(nproc), REF_BIN, "-dir", work]
    for key, flag in (("ker_f", "-ker_f"), ("ker_i", "-ker_i"), ("damp_f", "-damp_f"), ("damp_i", "-damp_i"),
                      ("damp_f_bycol", "-damp_f_bycol"), ("damp_i_bycol", "-damp_i_bycol"), ("b_k", "-b_k")):
        cmd += [flag, names[key]]
    cmd += ["-row_k", str(s.m_k), "-row_d", str(s.m_d), "-col", str(s.n), "-itn", str(sample_iters)]
    env = dict(os.environ, OMP_NUM_THREADS="1")
    t0 = time.perf_counter()
    r = subprocess.run(cmd, cwd=work, capture_output=True, text=True, timeout=600, env=env)
    wall = time.perf_counter() - t0
    t = re.findall(r"Parallel LSQR time=(\S+) s", r.stdout)
    if r.returncode != 0 or not t:
        raise RuntimeError(f"reference run on {nproc} ranks failed ({r.returncode}): {r.stderr[-300:]}")
    return float(t[-1]), wall


def cpu_baseline_reference(P, s, iters_to_tol, sample_iters, sample_iters_all=None):
    """The reference itself (oracle/_ref: the PLSQR3 sources compiled here at
    -O3, FP64 configuration): the system written in its file formats, a run
    of sample_iters iterations on one MPI rank and one on every host core
    (up to 16 ranks), each scaled from its own "Parallel LSQR time" (which
    excludes loading) to the GPU solve's iteration count.  The all-core run
    is the baseline; the one-rank run is reported inside it.  None when the
    binary or mpiexec is missing or the one-rank run fails."""
    import shutil
    import subprocess
    import tempfile
    if not (os.path.exists(REF_BIN) and os.path.exists(MPIEXEC)):
        return None
    work = tempfile.mkdtemp(prefix="plsqr_ref_", dir=os.environ.get("TMPDIR", "/tmp"))
    ncore = host_cores()
    sample_all = sample_iters_all or sample_iters
    try:
        t0 = time.perf_counter()
        names = s.write_reference_files(work)
        write_s = time.perf_counter() - t0
        lsqr_1, wall_1 = run_reference(s, names, work, sample_iters, 1)
        all_line = None
        if ncore > 1:
            try:
                lsqr_p, wall_p = run_reference(s, names, work, sample_all, ncore)
                per_p = lsqr_p / (sample_all + 0.5)
                all_line = (lsqr_p, wall_p, per_p)
            except (OSError, RuntimeError, subprocess.SubprocessError) as e:
                print(f"[bench] reference run on {ncore} ranks failed ({e})", file=sys.stderr, flush=True)
    except (OSError, RuntimeError, subprocess.SubprocessError) as e:
        print(f"[bench] reference run failed ({e}); using the port", file=sys.stderr, flush=True)
        return None
    finally:
        shutil.rmtree(work, ignore_errors=True)
    per_1 = lsqr_1 / (sample_iters + 0.5)
    single = {"value": per_1 * (iters_to_tol + 0.5), "unit": "s", "cores": 1,
              "sample": f"mpiexec -np 1, {sample_iters} iterations, Parallel LSQR time {lsqr_1:.1f} s "
                        f"({per_1:.4f} s/iteration; run {wall_1:.0f} s incl. loading)"}
    common = (f"PLSQR3 itself (compiled from its sources at -O3, FP64) on the same system from its files "
              f"(written in {write_s:.0f} s), its own Parallel LSQR time scaled from the sample to "
              f"{iters_to_tol} iterations (the plain FP64 solve's count); host CPU: {cpu_model()}")
    if all_line is None:
        return dict(single, kind="reference", sample=common + "; " + single["sample"], cpu_model=cpu_model(),
                    single_core=single)
    lsqr_p, wall_p, per_p = all_line
    return {"value": per_p * (iters_to_tol + 0.5), "unit": "s", "cores": ncore, "kind": "reference",
            "sample": common + f"; mpiexec -np {ncore} (its even column partition), {sample_all} iterations, "
                               f"Parallel LSQR time {lsqr_p:.1f} s ({per_p:.4f} s/iteration; run {wall_p:.0f} s "
                               f"incl. loading)",
            "cpu_model": cpu_model(), "single_core": single}


def cpu_baseline(P, s, iters_to_tol, sample_iters):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    orc = oracle.Oracle(s)
    t0 = time.perf_counter()
    orc.lsqr(s.b, sample_iters, trace=False)
    t = time.perf_counter() - t0
    # the sample is the start-up (one A^T u, about half an iteration) plus
    # sample_iters iterations; scale to the GPU solve's iteration count
    per_it = t / (sample_iters + 0.5)
    return {"value": per_it * (iters_to_tol + 0.5), "unit": "s", "cores": 1, "kind": "port",
            "sample": f"oracle LSQR, {sample_iters} iterations of the same system from x=0 "
                      f"({t:.1f} s measured, {per_it:.3f} s/iteration), scaled to "
                      f"{iters_to_tol} iterations"}


def bench_batch(P, hip, a):
    """Batched uneven solves: every (problem, right-hand side) pair of the
    workload solved to test2 <= tol in one batch; a step is one such batch."""
    specs = P.batch_specs(a.workload)
    t0 = time.perf_counter()
    systems, B, _ = P.generate_batch(specs=specs)
    gen_s = time.perf_counter() - t0
    K = len(specs[0]["rhs"])
    nnz = sum(s.nnz_k for s in systems)
    print(f"[bench] generated {a.workload}: {len(systems)} problems, {K} rhs each, "
          f"n={sum(s.n for s in systems)} rays={sum(s.m_k for s in systems)} nnz_k={nnz} "
          f"({gen_s:.1f} s)", file=sys.stderr, flush=True)
    prec = P.PREC_FP64 if a.precision == "fp64" else P.PREC_MIXED
    t0 = time.perf_counter()
    bt = P.Batch(systems, K, flags=P.BATCH_TUNE_MIXED if prec == P.PREC_MIXED else 0)
    setup_s = time.perf_counter() - t0
    tune = bt.get_tuning(prec)
    print(f"[bench] batch on device ({setup_s:.1f} s incl. autotune, {bt.device_bytes / 2**30:.2f} GiB); "
          f"shape {tune}", file=sys.stderr, flush=True)
    b = bt.pack_b(B)
    b_dev = hip.malloc(8 * b.size)
    x_dev = hip.malloc(8 * sum(s.n for s in systems) * K)
    hip.h2d(b_dev, b)
    pc, ref64 = P.PRECOND_NONE, None
    if a.precond == "col":
        if prec != P.PREC_MIXED:
            sys.exit("--precond col needs --precision mixed")
        # every pair's scaled solve stops as strictly as its plain FP64 solve:
        # its test2 takes the anorm that solve reports
        ref64 = bt.solve_device(b_dev, x_dev, a.max_iters, tol=a.tol)
        bt.set_anorm_refs([[i["anorm"] for i in row] for row in ref64[0]])
        pc = P.PRECOND_COLSCALE
        print(f"[bench] plain FP64 batch: {ref64[1]:.1f} ms", file=sys.stderr, flush=True)
    kw = dict(tol=a.tol, precision=prec, refine_every=a.refine_every, precond=pc, precond_power=a.precond_power,
              refine_adaptive=a.refine_adaptive)
    for _ in range(a.warmup):
        bt.solve_device(b_dev, x_dev, a.max_iters, **kw)
    hip.sync()
    runs = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        runs.append(bt.solve_device(b_dev, x_dev, a.max_iters, **kw))
    hip.sync()
    wall = time.perf_counter() - t0
    infos, tot = runs[-1]
    per_problem = np.array([max(i["elapsed_ms"] for i in row) for row in infos])
    iters = np.array([[i["itn"] for i in row] for row in infos])
    reached = all(i["arnorm"] <= a.tol for row in infos for i in row)
    pairs = len(systems) * K
    print(f"[bench] {a.steps} batches: {1000 * wall / a.steps:.1f} ms each (device {tot:.1f} ms), "
          f"iterations {iters.min()}..{iters.max()}, reached={reached}", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu:
        # oracle on the median-size problem's first right-hand side, scaled by
        # nonzeros and by each pair's iteration count
        order = np.argsort([s.nnz_k for s in systems])
        ref = systems[int(order[len(order) // 2])]
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        orc = oracle.Oracle(ref)
        t1 = time.perf_counter()
        orc.lsqr(B[int(order[len(order) // 2])][0], a.cpu_iters, trace=False)
        tc = time.perf_counter() - t1
        per_it_nnz = tc / (a.cpu_iters + 0.5) / (ref.nnz_k + ref.nnz_d)
        total = sum(per_it_nnz * (s.nnz_k + s.nnz_d) * (int(iters[p].sum()) + 0.5 * K)
                    for p, s in enumerate(systems))
        cpu = {"value": round(pairs / total, 4), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"oracle LSQR, {a.cpu_iters} iterations of the median problem "
                         f"(nnz {ref.nnz_k + ref.nnz_d}, {tc:.1f} s), scaled by nonzeros and by the "
                         f"iterations each pair took on the GPU ({total:.0f} s estimated)"}
    out = {
        "metric": "batched LSQR solves/s to test2 <= %g" % a.tol,
        "value": round(pairs / (wall / a.steps), 3) if reached else None, "unit": "solves/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1000 * wall / a.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == P.PREC_FP64 else "f32/f64",
        "data": "synthetic (seeded straight-ray generator, inputs resident in HBM)",
        "config": {"workload": a.workload, "problems": len(systems), "rhs_per_problem": K,
                   "grids": [sp["grid"][0] for sp in specs], "nnz_k": nnz, "precision": a.precision,
                   "refine_every": a.refine_every if prec == P.PREC_MIXED else None, "tol": a.tol,
                   "precond": a.precond, "precond_power": a.precond_power if a.precond == "col" else None,
                   "tol_reached": bool(reached), "parallelism": "single GPU"},
        "per_problem_ms": {"p50": round(float(np.percentile(per_problem, 50)), 2),
                           "p99": round(float(np.percentile(per_problem, 99)), 2),
                           "max": round(float(per_problem.max()), 2)},
        "iterations": {"min": int(iters.min()), "median": int(np.median(iters)), "max": int(iters.max())},
        "device_ms": round(tot, 2),
        "plain_fp64_batch_ms": None if ref64 is None else round(ref64[1], 2),
        "tuning": tune,
        "cpu_baseline": cpu,
        "setup_s": {"generate": round(gen_s, 2), "batch_create": round(setup_s, 2)},
    }
    print(json.dumps(out), flush=True)
    bt.close()
    hip.free(b_dev)
    hip.free(x_dev)


def bench_learned(P, hip, a):
    """Learned warm start on the 67^3 / 2M-ray system: train the network on
    (A^T b, x*) pairs of fresh blob fields and noise, then time LSQR to tol on
    held-out right-hand sides from zero (plain FP64, the reference method)
    and from the network's x0."""
    import importlib
    W = importlib.import_module(P.__name__ + ".warmstart")
    cfg = P.CONFIGS[a.workload]
    t0 = time.perf_counter()
    s = P.generate_config(a.workload)
    gen_s = time.perf_counter() - t0
    grid = (cfg["nx"], cfg["ny"], cfg["nz"])
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    log(f"[bench] generated {a.workload} ({gen_s:.1f} s); {a.train_pairs} training pairs, {a.eval_pairs} held out")
    res, net = W.run(s, grid, a.train_pairs, a.eval_pairs, cfg["sigma"], 500000, a.target_tol, a.epochs, a.tol,
                     a.refine_every, log=log, colscale=a.precond == "col", refine_adaptive=a.refine_adaptive,
                     precond_power=a.precond_power, input_kind=a.net_input)
    if a.save_params:
        np.save(a.save_params, net.get_params())
    net.close()
    S = res["summary"]
    methods = [k for k in S if isinstance(S[k], dict) and k != "fp64_zero"]
    best = min(methods, key=lambda k: S[k]["ms_mean"])
    # the headline: mixed precision from the network's x0 (on A D with --precond col)
    head = "mixed_warm_cs" if a.precond == "col" else "mixed_warm"
    out = {
        "metric": METRIC, "value": round(S[head]["ms_mean"] / 1000.0, 6), "unit": "s",
        "n_gpus": 1, "steps": a.eval_pairs, "warmup": 0, "ms_per_step": round(S[head]["ms_mean"], 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64 + bf16 network",
        "data": "synthetic (seeded straight-ray generator; fixed ray geometry, fresh blob fields and noise "
                "per pair)",
        "config": {"workload": a.workload + "+learned_warm_start", "train_pairs": a.train_pairs,
                   "eval_pairs": a.eval_pairs, "target_tol": a.target_tol, "epochs": a.epochs, "tol": a.tol,
                   "method": head, "precond": a.precond, "refine_adaptive": a.refine_adaptive,
                   "precond_power": a.precond_power if a.precond == "col" else None,
                   "net_input": a.net_input, "refine_every": a.refine_every, "parallelism": "single GPU"},
        "summary": S, "best_method": best, "times": res["times"], "loss": res["loss"], "targets": res["targets"],
    }
    print(json.dumps(out), flush=True)


def converged_solution(op, s, tol=1e-12, max_iters=20000):
    """Plain FP64 LSQR run far past the benchmark's tolerance (untimed): the
    reference solution the per-solve errors are measured against.  Returns
    x*, ||r*|| (true FP64 residual) and the solve's info."""
    xs, info = op.lsqr(s.b, max_iters, tol=tol)
    rn, an, fro = op.residual_norms(s.b, xs)
    return xs, rn, info, an / (info["anorm"] * rn) if rn > 0 else 0.0


def solution_errors(op, b, x, xstar, rstar, anorm):
    """x and residual error of a solve against the converged FP64 solution,
    and its true FP64 test2 = ||A^T r|| / (anorm ||r||) (lsqr.c:742-744)."""
    rn, an, _ = op.residual_norms(b, x)
    return {"x_rel_err_vs_converged": float(np.linalg.norm(x - xstar) / np.linalg.norm(xstar)),
            "r_rel_excess_vs_converged": float((rn - rstar) / rstar) if rstar > 0 else None,
            "true_test2": float(an / (anorm * rn)) if rn > 0 and anorm > 0 else 0.0}


def bench_learned_line(P, hip, op, s, a, log):
    """The learned warm start on held-out right-hand sides of the 67^3 system
    with the committed parameters: per right-hand side, plain FP64 LSQR from
    zero (the reference's method; its anorm makes the other stops as strict),
    mixed precision on A D from zero, and the learned path timed end to end
    from b resident in HBM: A^T b on the device, the Jacobi-scaled network
    input, the network's x0, r0 = [b;0] - A x0 with the fallback test, and
    mixed precision on A D from x0."""
    import importlib
    W = importlib.import_module(P.__name__ + ".warmstart")
    cfg = P.CONFIGS[a.workload]
    grid = (cfg["nx"], cfg["ny"], cfg["nz"])
    wpath = os.path.join(ROOT, "plsqr3.2013.09_amd", "weights", "warmstart_grid67_rays2m.npy")
    if not os.path.exists(wpath) or a.workload != "grid67_rays2m":
        return None
    t0 = time.perf_counter()
    seeds = [a.learned_seed0 + i for i in range(a.learned_pairs)]
    with P.Batch([s], 4, grids=[grid], flags=P.BATCH_NO_TUNE) as bt:
        B, _ = W.make_pairs(bt, s, grid, seeds, cfg["sigma"])
    colss = W.column_ssq(s)
    inv = np.where(colss > 0, 1.0 / np.where(colss > 0, colss, 1.0), 1.0)
    net = P.WarmStartNet(grid, max_batch=1)
    net.set_params(np.load(wpath, allow_pickle=False))
    m, n = s.m_k + s.m_d, s.n
    b_dev, y_dev, atb_dev = hip.malloc(8 * s.m_k), hip.malloc(8 * m), hip.malloc(8 * n)
    x0_dev, x_dev = hip.malloc(8 * n), hip.malloc(8 * n)
    setup_s = time.perf_counter() - t0
    log(f"[bench] learned: {len(seeds)} held-out right-hand sides ready ({setup_s:.1f} s)")
    kw = dict(precision=P.PREC_MIXED, refine_every=a.refine_every, precond=P.PRECOND_COLSCALE,
              refine_adaptive=True, precond_power=a.precond_power)
    zeros_n = np.zeros(n)
    rows = []
    try:
        for j, b in enumerate(B):
            y = np.zeros(m)
            y[:s.m_k] = b
            hip.h2d(b_dev, np.ascontiguousarray(b))
            hip.h2d(y_dev, y)
            hip.sync()
            # plain FP64 from zero
            t1 = time.perf_counter()
            i64 = op.lsqr_device(b_dev, x_dev, a.max_iters, tol=a.tol)
            hip.sync()
            t64 = time.perf_counter() - t1
            x64 = np.empty(n)
            hip.d2h(x64, x_dev)
            # mixed precision on A D from zero
            t1 = time.perf_counter()
            iz = op.lsqr_x0_device(b_dev, 0, x_dev, a.max_iters, tol=a.tol, anorm_ref=i64["anorm"], **kw)
            hip.sync()
            tz = time.perf_counter() - t1
            xz = np.empty(n)
            hip.d2h(xz, x_dev)
            # learned: everything from b on the device
            t1 = time.perf_counter()
            hip.h2d(atb_dev, zeros_n)
            op.aprod_device(2, atb_dev, y_dev)
            hip.sync()  # the operator's stream does not order with the copy below
            atb = np.empty(n)
            hip.d2h(atb, atb_dev)
            x0 = net.predict((atb * inv)[None, :])[0]
            hip.h2d(x0_dev, x0)
            iw = op.lsqr_x0_device(b_dev, x0_dev, x_dev, a.max_iters, tol=a.tol, anorm_ref=i64["anorm"], **kw)
            hip.sync()
            tw = time.perf_counter() - t1
            xw = np.empty(n)
            hip.d2h(xw, x_dev)
            r = {"fp64_zero": {"s": t64, "itn": i64["itn"], "test2": i64["arnorm"]},
                 "mixed_colscale_zero": {"s": tz, "itn": iz["itn"], "test2": iz["arnorm"],
                                         "x_rel_diff_vs_fp64": float(np.linalg.norm(xz - x64) / np.linalg.norm(x64))},
                 "learned": {"s": tw, "itn": iw["itn"], "test2": iw["arnorm"], "used_x0": iw["used_x0"],
                             "x0_rel_diff_vs_fp64": float(np.linalg.norm(x0 - x64) / np.linalg.norm(x64)),
                             "x_rel_diff_vs_fp64": float(np.linalg.norm(xw - x64) / np.linalg.norm(x64))}}
            if j < a.learned_converged:
                # the converged FP64 solution of this b (untimed)
                xs, inf_s = op.lsqr(b, 20000, tol=1e-12)
                rs, _, _ = op.residual_norms(b, xs)
                for key, xx in (("fp64_zero", x64), ("mixed_colscale_zero", xz), ("learned", xw)):
                    r[key].update(solution_errors(op, b, xx, xs, rs, i64["anorm"]))
            rows.append(r)
            if (j + 1) % 8 == 0 or j + 1 == len(B):
                log(f"[bench] learned: {j + 1} of {len(B)} right-hand sides")
    finally:
        net.close()
        for p in (b_dev, y_dev, atb_dev, x0_dev, x_dev):
            hip.free(p)

    def stat(key, field):
        v = np.array([r[key][field] for r in rows], float)
        return {"mean": float(v.mean()), "median": float(np.median(v)), "max": float(v.max())}

    out = {"method": "A^T b / diag(A^T A) -> 3-D CNN (committed parameters) -> x0, fallback to x0 = 0 if "
                     "||[b;0] - A x0|| > ||[b;0]||, then mixed precision on A D (column scaling) with adaptive FP64 "
                     "checks; timed from b resident in HBM, including A^T b, the network and r0",
           "right_hand_sides": len(rows), "seeds": [seeds[0], seeds[-1]], "tol": a.tol}
    for key in ("fp64_zero", "mixed_colscale_zero", "learned"):
        out[key] = {"s": stat(key, "s"), "iterations": stat(key, "itn"), "test2_max": stat(key, "test2")["max"]}
        if key != "fp64_zero":
            out[key]["x_rel_diff_vs_fp64"] = stat(key, "x_rel_diff_vs_fp64")
            out[key]["speedup_vs_fp64"] = round(out["fp64_zero"]["s"]["mean"] / out[key]["s"]["mean"], 3)
        conv = [r[key] for r in rows if "x_rel_err_vs_converged" in r[key]]
        if conv:
            out[key]["vs_converged"] = {f: [c[f] for c in conv] for f in ("x_rel_err_vs_converged",
                                                                           "r_rel_excess_vs_converged",
                                                                           "true_test2")}
    out["learned"]["used_x0"] = int(sum(r["learned"]["used_x0"] for r in rows))
    out["learned"]["x0_rel_diff_vs_fp64"] = stat("learned", "x0_rel_diff_vs_fp64")
    out["learned"]["tol_reached"] = bool(all(r["learned"]["test2"] <= a.tol for r in rows))
    out["setup_s"] = round(setup_s, 2)
    return out


def bench_batch_line(P, hip, a, log):
    """The batched workload (24 problems of 20^3..48^3 cells, 4 right-hand
    sides each) beside the single-system line: one FP64 batch solve and one
    mixed-precision batch solve on A_p D_p, every pair to test2 <= tol."""
    name = "batch24_grid20to48_rhs4"
    specs = P.batch_specs(name)
    t0 = time.perf_counter()
    systems, B, _ = P.generate_batch(specs=specs)
    gen_s = time.perf_counter() - t0
    K = len(specs[0]["rhs"])
    t0 = time.perf_counter()
    bt = P.Batch(systems, K, flags=P.BATCH_TUNE_MIXED)
    setup_s = time.perf_counter() - t0
    log(f"[bench] batch: generated ({gen_s:.1f} s), on device with autotuning ({setup_s:.1f} s)")
    try:
        b = bt.pack_b(B)
        b_dev = hip.malloc(8 * b.size)
        x_dev = hip.malloc(8 * sum(x.n for x in systems) * K)
        hip.h2d(b_dev, b)
        res = {}
        try:
            for label, kw in (("fp64", dict()),
                              ("mixed_colscale", dict(precision=P.PREC_MIXED, refine_every=a.refine_every,
                                                      precond=P.PRECOND_COLSCALE, precond_power=a.precond_power,
                                                      refine_adaptive=True))):
                if label == "mixed_colscale":
                    bt.set_anorm_refs([[i["anorm"] for i in row] for row in res["fp64"]["infos"]])
                bt.solve_device(b_dev, x_dev, a.max_iters, tol=a.tol, **kw)  # warm-up
                hip.sync()
                t1 = time.perf_counter()
                infos, tot = bt.solve_device(b_dev, x_dev, a.max_iters, tol=a.tol, **kw)
                hip.sync()
                wall = time.perf_counter() - t1
                per = np.array([max(i["elapsed_ms"] for i in row) for row in infos])
                its = np.array([[i["itn"] for i in row] for row in infos])
                res[label] = {"infos": infos, "s_per_batch": wall, "device_ms": tot,
                              "solves_per_s": len(systems) * K / wall,
                              "iterations": {"min": int(its.min()), "max": int(its.max())},
                              "per_problem_ms": {"p50": float(np.percentile(per, 50)),
                                                 "p99": float(np.percentile(per, 99))},
                              "tol_reached": bool(all(i["arnorm"] <= a.tol for row in infos for i in row))}
        finally:
            hip.free(b_dev)
            hip.free(x_dev)
    finally:
        bt.close()
    out = {"workload": name, "problems": len(systems), "rhs_per_problem": K,
           "nnz_k": int(sum(x.nnz_k for x in systems)), "generate_s": round(gen_s, 1),
           "create_and_autotune_s": round(setup_s, 1)}
    for label, r in res.items():
        r.pop("infos")
        out[label] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
    if "fp64" in out and "mixed_colscale" in out:
        out["mixed_colscale"]["speedup_vs_fp64"] = round(res["fp64"]["s_per_batch"] /
                                                         res["mixed_colscale"]["s_per_batch"], 3)
    return out


def row_block_system(P, s, k, nblk):
    """Row block k of nblk of the system as one GPU of a row-sharded solve
    would hold it: a contiguous block of kernel rows (balanced by nonzeros)
    and the matching share of the damping rows, all n columns (x replicated)."""
    import scipy.sparse as sp
    Ak = sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)).tocsr()
    cnt = np.cumsum(np.diff(Ak.indptr))
    lo = 0 if k == 0 else int(np.searchsorted(cnt, s.nnz_k * k / nblk)) + 1
    hi = int(np.searchsorted(cnt, s.nnz_k * (k + 1) / nblk)) + 1 if k + 1 < nblk else s.m_k
    D = sp.csr_matrix((s.d_val, s.d_col_idx, s.d_row_ptr), shape=(s.m_d, s.n))
    dlo, dhi = s.m_d * k // nblk, s.m_d * (k + 1) // nblk
    Kb = Ak[lo:hi].tocsc()
    Kb.sort_indices()
    Db = D[dlo:dhi]
    Dc = Db.tocsc()
    Dc.sort_indices()
    return P.System.from_arrays(hi - lo, dhi - dlo, s.n, Kb.indptr, Kb.indices, Kb.data, Db.indptr, Db.indices,
                                Db.data, Dc.indptr, Dc.indices, Dc.data, np.asarray(s.b)[lo:hi], grid=s.grid)


def bench_row_block(P, hip, s, a, log):
    """One GPU's share of the row-sharded design at 8 GPUs: the operator of
    one nonzero-balanced block of kernel rows (2M / 8 rays) and 1/8 of the
    damping rows, all columns, timed per aprod mode with HIP events."""
    t0 = time.perf_counter()
    sb = row_block_system(P, s, 0, 8)
    with P.Operator(sb) as ob:
        setup_s = time.perf_counter() - t0
        m = sb.m_k + sb.m_d
        x_dev, y_dev = hip.malloc(8 * sb.n), hip.malloc(8 * m)
        try:
            hip.h2d(x_dev, np.random.default_rng(0).standard_normal(sb.n))
            hip.h2d(y_dev, np.random.default_rng(1).standard_normal(m) * 1e-3)
            ms1 = time_aprod(hip, ob, 1, x_dev, y_dev, a.aprod_reps)
            ms2 = time_aprod(hip, ob, 2, x_dev, y_dev, a.aprod_reps)
        finally:
            hip.free(x_dev)
            hip.free(y_dev)
    b1, b2 = aprod_bytes(sb.m_k, sb.m_d, sb.n, sb.nnz_k, sb.nnz_d)
    log(f"[bench] row block 1/8: {sb.m_k} rays, mode 1 {ms1:.3f} ms, mode 2 {ms2:.3f} ms")
    return {"block": "rows 1 of 8 (nonzero-balanced kernel rows, 1/8 of the damping rows, all columns)",
            "m_k": sb.m_k, "m_d": sb.m_d, "n": sb.n, "nnz_k": int(sb.nnz_k), "nnz_d": int(sb.nnz_d),
            "mode1_ms": round(ms1, 4), "mode2_ms": round(ms2, 4), "mode1_bytes": int(b1), "mode2_bytes": int(b2),
            "mode1_gbs": round(b1 / ms1 / 1e6, 1), "mode2_gbs": round(b2 / ms2 / 1e6, 1),
            "collective_not_timed": "one all-reduce of n doubles (A^T u partials) per iteration",
            "setup_s": round(setup_s, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="grid67_rays2m")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--refine-every", type=int, default=25,
                    help="mixed precision: inner FP32 iterations per FP64 residual refinement")
    ap.add_argument("--precond", default="none", choices=["none", "col"],
                    help="mixed precision: LSQR on A D with D the inverse column norms (col); the FP64 "
                         "test2 then uses the anorm of an untimed plain FP64 solve of the same b")
    ap.add_argument("--precond-power", type=float, default=1.5,
                    help="column scaling D = diag(||A e_j||^-p); 1.5 measured fastest of 1, 1.25, 1.5 and 2 "
                         "across the 24^3, 40^3 and 67^3 top-face systems")
    ap.add_argument("--refine-adaptive", action="store_true",
                    help="mixed precision: FP64 checks scheduled from the inner test2 estimate (first after "
                         "--refine-every iterations, then when the estimate predicts tol, at most 4x apart)")
    ap.add_argument("--max-iters", type=int, default=5000)
    ap.add_argument("--aprod-reps", type=int, default=20)
    ap.add_argument("--cpu-iters", type=int, default=40,
                    help="iterations of the oracle (C port) timed for the CPU baseline")
    ap.add_argument("--cpu-kind", default="auto", choices=["auto", "reference", "port"],
                    help="CPU baseline: the reference compiled from its sources (oracle/_ref), the port, "
                         "or the reference when it is there")
    ap.add_argument("--ref-iters", type=int, default=40,
                    help="iterations of the reference run timed for the CPU baseline")
    ap.add_argument("--ref-iters-all", type=int, default=100,
                    help="iterations of the reference run on every host core")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="FP64 line only: skip the mixed-precision column-scaled solve timed beside it")
    ap.add_argument("--aprod-only", action="store_true",
                    help="only time aprod launches and a device copy of known size "
                         "(the workload for rocprofv3 counter passes)")
    ap.add_argument("--exact-order", action="store_true",
                    help="operator summing every row/column in the reference's order")
    ap.add_argument("--learned", action="store_true",
                    help="learned warm start: train the network, then time LSQR from its x0 against "
                         "plain FP64 LSQR from zero on held-out right-hand sides")
    ap.add_argument("--train-pairs", type=int, default=512)
    ap.add_argument("--eval-pairs", type=int, default=16)
    ap.add_argument("--target-tol", type=float, default=1e-4,
                    help="test2 the FP64 training targets x* are solved to")
    ap.add_argument("--epochs", type=float, default=20)
    ap.add_argument("--save-params", default=None, help="write the trained parameters (.npy)")
    ap.add_argument("--learned-pairs", type=int, default=32,
                    help="default line: held-out right-hand sides of the learned warm-start sub-object")
    ap.add_argument("--learned-seed0", type=int, default=600000,
                    help="first seed of those right-hand sides (training used 500000-500511)")
    ap.add_argument("--learned-converged", type=int, default=2,
                    help="of them, how many also get a converged FP64 solution for the error figures")
    ap.add_argument("--converged-tol", type=float, default=1e-12,
                    help="test2 of the untimed FP64 solution the per-solve errors are measured against")
    ap.add_argument("--no-converged", action="store_true",
                    help="skip the converged FP64 solution and the error figures measured against it")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-row-block", action="store_true")
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="skip the optional sub-objects whose estimated time would pass this many seconds")
    ap.add_argument("--net-input", default="jacobi", choices=["atb", "jacobi"],
                    help="network input: A^T b divided by the diagonal of A^T A (x0 error 0.28 at 67^3), "
                         "or A^T b itself (0.43)")
    a = ap.parse_args()
    t_start = time.perf_counter()
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(the row-sharded multi-GPU design is described but not run)", file=sys.stderr)
        sys.exit(2)
    P = load_pkg()
    hip = Hip()
    if a.workload in P.BATCH_CONFIGS:
        bench_batch(P, hip, a)
        return
    if a.learned:
        bench_learned(P, hip, a)
        return
    t0 = time.perf_counter()
    s = P.generate_config(a.workload)
    gen_s = time.perf_counter() - t0
    print(f"[bench] generated {a.workload}: m_k={s.m_k} m_d={s.m_d} n={s.n} nnz_k={s.nnz_k} "
          f"nnz_d={s.nnz_d} ({gen_s:.1f} s)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    op = P.Operator(s, P.OP_EXACT_ORDER if a.exact_order else P.OP_DEFAULT)
    setup_s = time.perf_counter() - t0
    print(f"[bench] operator on device ({setup_s:.1f} s, {op.device_bytes / 2**30:.2f} GiB)",
          file=sys.stderr, flush=True)
    prec = P.PREC_FP64 if a.precision == "fp64" else P.PREC_MIXED
    b_dev = hip.malloc(8 * s.m_k)
    x_dev = hip.malloc(8 * s.n)
    y_dev = hip.malloc(8 * (s.m_k + s.m_d))
    hip.h2d(b_dev, np.ascontiguousarray(s.b))
    hip.h2d(x_dev, np.random.default_rng(0).standard_normal(s.n))
    hip.h2d(y_dev, np.random.default_rng(1).standard_normal(s.m_k + s.m_d) * 1e-3)
    x32_dev = hip.malloc(4 * s.n)
    y32_dev = hip.malloc(4 * (s.m_k + s.m_d))
    hip.h2d(x32_dev, np.random.default_rng(0).standard_normal(s.n).astype(np.float32))
    hip.h2d(y32_dev, (np.random.default_rng(1).standard_normal(s.m_k + s.m_d) * 1e-3).astype(np.float32))

    if a.aprod_only:
        # a plain device-to-device copy of 1 GiB and a streaming sum over
        # 1 GiB calibrate the byte counters (each reads 1 GiB; the copy
        # writes 1 GiB)
        cp_src, cp_dst = hip.malloc(1 << 30), hip.malloc(1 << 30)
        hip.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        hip.check(hip.lib.hipMemcpy(C.c_void_p(cp_dst), C.c_void_p(cp_src), 1 << 30, 3), "hipMemcpy")
        hip.sync()
        try:
            import torch
            t = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
            float(t.sum())
            del t
            torch.cuda.synchronize()
        except Exception as e:  # calibration only
            print(f"[bench] streaming-sum calibration skipped: {e}", file=sys.stderr)
        ms1 = time_aprod(hip, op, 1, x_dev, y_dev, a.aprod_reps)
        ms2 = time_aprod(hip, op, 2, x_dev, y_dev, a.aprod_reps)
        cm1 = time_aprod(hip, op, 1, x_dev, y_dev, a.aprod_reps, internal=False)
        cm2 = time_aprod(hip, op, 2, x_dev, y_dev, a.aprod_reps, internal=False)
        b1, b2 = aprod_bytes(s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d)
        line = {"aprod_only": True, "mode1_ms": ms1, "mode2_ms": ms2, "mode1_bytes": b1, "mode2_bytes": b2,
                "mode1_ms_caller_order": cm1, "mode2_ms_caller_order": cm2}
        if not a.exact_order:
            f1 = time_aprod(hip, op, 1, x32_dev, y32_dev, a.aprod_reps, f32=True)
            f2 = time_aprod(hip, op, 2, x32_dev, y32_dev, a.aprod_reps, f32=True)
            fb1, fb2 = aprod_bytes(s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d, 4, 4, 4)
            line.update({"mode1_ms_f32": f1, "mode2_ms_f32": f2, "mode1_bytes_f32": fb1, "mode2_bytes_f32": fb2})
        print(json.dumps(line), flush=True)
        hip.free(cp_src)
        hip.free(cp_dst)
        for p in (b_dev, x_dev, y_dev, x32_dev, y32_dev):
            hip.free(p)
        op.close()
        return
    pc, anorm_ref, ref64 = P.PRECOND_NONE, 0.0, None
    if a.precond == "col":
        if prec != P.PREC_MIXED:
            sys.exit("--precond col needs --precision mixed")
        # the stop of the scaled solve is made as strict as the plain FP64
        # solve's: its test2 takes the anorm that solve reports
        ref64 = op.lsqr_device(b_dev, x_dev, a.max_iters, tol=a.tol)
        x_fp64 = np.empty(s.n)
        hip.d2h(x_fp64, x_dev)
        pc, anorm_ref = P.PRECOND_COLSCALE, ref64["anorm"]
        print(f"[bench] plain FP64 solve: {ref64['itn']} iterations, {ref64['elapsed_ms']:.1f} ms, "
              f"anorm {anorm_ref:.6e}", file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        op.lsqr_device(b_dev, x_dev, a.max_iters, tol=a.tol, precision=prec, refine_every=a.refine_every,
                       precond=pc, anorm_ref=anorm_ref, refine_adaptive=a.refine_adaptive,
                       precond_power=a.precond_power)
    hip.sync()
    infos = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        infos.append(op.lsqr_device(b_dev, x_dev, a.max_iters, tol=a.tol, precision=prec,
                                    refine_every=a.refine_every, precond=pc, anorm_ref=anorm_ref,
                                    refine_adaptive=a.refine_adaptive, precond_power=a.precond_power))
    hip.sync()
    wall = time.perf_counter() - t0
    if ref64 is not None:
        x_last = np.empty(s.n)
        hip.d2h(x_last, x_dev)
        ref64["xdiff"] = float(np.linalg.norm(x_last - x_fp64) / np.linalg.norm(x_fp64))
    iters = infos[-1]["itn"]
    reached = infos[-1]["arnorm"] <= a.tol
    ms_step = 1000.0 * wall / a.steps
    print(f"[bench] {a.steps} solves: {ms_step:.2f} ms each, {iters} iterations, "
          f"test2={infos[-1]['arnorm']:.3e} (device loop {infos[-1]['elapsed_ms']:.2f} ms)",
          file=sys.stderr, flush=True)
    mixed_line = None
    if prec == P.PREC_FP64 and not a.exact_order and not a.no_mixed:
        # beside the plain FP64 line: the mixed-precision solve of the same b
        # on A D (column scaling) with adaptive FP64 checks, stopping on the
        # same test2 (the anorm of the FP64 solve just timed)
        x64 = np.empty(s.n)
        hip.d2h(x64, x_dev)
        kw = dict(tol=a.tol, precision=P.PREC_MIXED, refine_every=a.refine_every, precond=P.PRECOND_COLSCALE,
                  anorm_ref=infos[-1]["anorm"], refine_adaptive=True, precond_power=a.precond_power)
        op.lsqr_device(b_dev, x_dev, a.max_iters, **kw)
        hip.sync()
        minfos = []
        t1 = time.perf_counter()
        for _ in range(a.steps):
            minfos.append(op.lsqr_device(b_dev, x_dev, a.max_iters, **kw))
        hip.sync()
        mwall = (time.perf_counter() - t1) / a.steps
        xm = np.empty(s.n)
        hip.d2h(xm, x_dev)
        mi = minfos[-1]
        x_mixed = xm
        mixed_line = {"method": f"mixed precision on A D, D = diag(||A e_j||^-{a.precond_power:g}) "
                                f"(column scaling), adaptive FP64 checks",
                      "s_per_solve": round(mwall, 6), "iterations_to_tol": mi["itn"],
                      "tol_reached": bool(mi["arnorm"] <= a.tol), "fp64_checks": mi["refinements"],
                      "ms_per_iteration": round(mi["elapsed_ms"] / max(mi["itn"], 1), 4),
                      "speedup_vs_fp64": round((wall / a.steps) / mwall, 3),
                      "x_rel_diff_vs_fp64": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))}
        print(f"[bench] mixed + column scaling: {1000 * mwall:.2f} ms per solve, {mi['itn']} iterations, "
              f"{mixed_line['speedup_vs_fp64']}x the FP64 solve", file=sys.stderr, flush=True)
        # the same scaled solve with the FP64 operator: what the scaling gains
        # without FP32 arithmetic (3 timed solves)
        kw64 = dict(kw, precision=P.PREC_FP64)
        op.lsqr_device(b_dev, x_dev, a.max_iters, **kw64)
        hip.sync()
        t1 = time.perf_counter()
        for _ in range(3):
            fi = op.lsqr_device(b_dev, x_dev, a.max_iters, **kw64)
        hip.sync()
        fwall = (time.perf_counter() - t1) / 3
        x_fp64cs = np.empty(s.n)
        hip.d2h(x_fp64cs, x_dev)
        mixed_line["fp64_same_scaling"] = {
            "s_per_solve": round(fwall, 6), "iterations_to_tol": fi["itn"], "fp64_checks": fi["refinements"],
            "tol_reached": bool(fi["arnorm"] <= a.tol),
            "ms_per_iteration": round(fi["elapsed_ms"] / max(fi["itn"], 1), 4),
            "speedup_vs_fp64": round((wall / a.steps) / fwall, 3),
            "x_rel_diff_vs_fp64": float(np.linalg.norm(x_fp64cs - x64) / np.linalg.norm(x64)),
            "note": "column scaling alone; the mixed line's gain over this is FP32 arithmetic"}
        print(f"[bench] FP64 + column scaling: {1000 * fwall:.2f} ms per solve, {fi['itn']} iterations",
              file=sys.stderr, flush=True)

    # errors of the timed solves against a converged FP64 solution (untimed),
    # and the sub-objects of the other workloads
    extra = {}
    if prec == P.PREC_FP64 and not a.exact_order and not a.aprod_only and not a.no_converged:
        t1 = time.perf_counter()
        xstar, rstar, istar, t2star = converged_solution(op, s, a.converged_tol)
        log(f"[bench] converged FP64 solution: {istar['itn']} iterations, test2 {istar['arnorm']:.2e} "
            f"(true {t2star:.2e}), {time.perf_counter() - t1:.1f} s")
        extra["converged_fp64"] = {"iterations": istar["itn"], "test2": istar["arnorm"], "true_test2": t2star,
                                   "rnorm": rstar}
        x64_last = np.empty(s.n)
        if mixed_line is not None:
            x64_last = x64
        else:
            op.lsqr_device(b_dev, x_dev, a.max_iters, tol=a.tol)
            hip.d2h(x64_last, x_dev)
        extra["fp64_errors"] = solution_errors(op, s.b, x64_last, xstar, rstar, infos[-1]["anorm"])
        if mixed_line is not None:
            mixed_line.update(solution_errors(op, s.b, x_mixed, xstar, rstar, infos[-1]["anorm"]))
        budget_left = lambda: a.budget_s - (time.perf_counter() - t_start)  # noqa: E731
        for name, enabled, need_s, fn in (
                ("learned", not a.no_learned, 3.5 * a.learned_pairs + 30,
                 lambda: bench_learned_line(P, hip, op, s, a, log)),
                ("row_block", not a.no_row_block, 30, lambda: bench_row_block(P, hip, s, a, log)),
                ("batch", not a.no_batch, 150, lambda: bench_batch_line(P, hip, a, log))):
            if not enabled:
                continue
            if budget_left() < need_s:
                extra[name] = {"skipped": f"about {need_s:.0f} s needed, {budget_left():.0f} s of the budget left"}
                continue
            t1 = time.perf_counter()
            try:
                extra[name] = fn()
            except Exception as e:  # a sub-object never takes the main line down with it
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] {name}: {time.perf_counter() - t1:.1f} s")

    # aprod kernels, timed with events on the handle's stream: on vectors in
    # the handle's order (the SpMV kernels alone, as in the solve loop), and
    # through the caller-order entry point (plus the reordering at the API
    # boundary)
    if prec == P.PREC_MIXED:
        # the mixed solve's inner iterations run the FP32 operator
        ms1 = time_aprod(hip, op, 1, x32_dev, y32_dev, a.aprod_reps, f32=True)
        ms2 = time_aprod(hip, op, 2, x32_dev, y32_dev, a.aprod_reps, f32=True)
        cm1 = time_aprod(hip, op, 1, x32_dev, y32_dev, a.aprod_reps, f32=True, internal=False)
        cm2 = time_aprod(hip, op, 2, x32_dev, y32_dev, a.aprod_reps, f32=True, internal=False)
        b1, b2 = aprod_bytes(s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d, 4, 4, 4)
    else:
        ms1 = time_aprod(hip, op, 1, x_dev, y_dev, a.aprod_reps)
        ms2 = time_aprod(hip, op, 2, x_dev, y_dev, a.aprod_reps)
        cm1 = time_aprod(hip, op, 1, x_dev, y_dev, a.aprod_reps, internal=False)
        cm2 = time_aprod(hip, op, 2, x_dev, y_dev, a.aprod_reps, internal=False)
        b1, b2 = aprod_bytes(s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d)
    gbs1, gbs2 = b1 / ms1 / 1e6, b2 / ms2 / 1e6
    gbs = (b1 + b2) / (ms1 + ms2) / 1e6
    print(f"[bench] aprod mode1 {ms1:.3f} ms {gbs1:.0f} GB/s; mode2 {ms2:.3f} ms {gbs2:.0f} GB/s",
          file=sys.stderr, flush=True)
    dom_mode, dom_ms, dom_b = (2, ms2, b2) if ms2 >= ms1 else (1, ms1, b1)
    mixed = prec == P.PREC_MIXED
    traffic = read_traffic(a.workload)
    tr_val = None
    tr_key = "bytes_per_launch_f32" if mixed else "bytes_per_launch"
    if traffic and not a.exact_order and str(dom_mode) in traffic.get(tr_key, {}):
        tr_val = traffic[tr_key][str(dom_mode)]
    sfx = "32" if mixed else ""
    if a.exact_order:
        kname = "k_rows_exact: reference row order" if dom_mode == 1 else "k_cols_exact: reference column order"
    elif dom_mode == 1:
        kname = (f"k_rows_tiled_c{sfx} + k_rows_tiled_combine{sfx}: y += A x as ray segments per tile of "
                 "voxels then per-ray sums")
    else:
        kname = f"k_cols_tiled{sfx} + k_cols_combine{sfx}: x += A^T y on voxel bricks"
    roofline = {"bound": "hbm", "achieved": round(dom_b / dom_ms / 1e6, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(dom_b / dom_ms / 1e6 / HBM_PEAK_GBS, 4),
                "traffic": tr_val,
                "kernel": f"aprod mode {dom_mode} ({kname})",
                "algorithmic_bytes": dom_b, "avg_ms": round(dom_ms, 4)}
    cpu = None
    # the reference runs plain LSQR: its time-to-tol is at the plain FP64 solve's iteration count
    iters_ref = ref64["itn"] if ref64 is not None else iters
    if not a.no_cpu:
        if a.cpu_kind in ("auto", "reference") and not a.exact_order:
            cpu = cpu_baseline_reference(P, s, iters_ref, a.ref_iters, a.ref_iters_all)
        if cpu is None and a.cpu_kind != "reference":
            cpu = cpu_baseline(P, s, iters_ref, a.cpu_iters)
    out = {
        "metric": METRIC, "value": round(wall / a.steps, 6) if reached else None, "unit": "s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32/f64",
        "data": "synthetic (seeded straight-ray generator, inputs resident in HBM)",
        "config": {"workload": a.workload, "grid": [67, 67, 67] if "67" in a.workload else None,
                   "rays": s.m_k, "nnz_k": s.nnz_k, "m_d": s.m_d, "nnz_d": s.nnz_d,
                   "precision": a.precision, "precond": a.precond, "refine_adaptive": a.refine_adaptive,
                   "precond_power": a.precond_power if a.precond == "col" else None,
                   "tol": a.tol, "iterations_to_tol": iters, "fp64_checks": infos[-1]["refinements"],
                   "tol_reached": bool(reached), "parallelism": "single GPU",
                   "operator": "exact-order" if a.exact_order else "default"},
        "ms_per_iteration": round(infos[-1]["elapsed_ms"] / max(iters, 1), 4),
        "plain_fp64": None if ref64 is None else {
            "iterations": ref64["itn"], "device_ms": round(ref64["elapsed_ms"], 2), "anorm": ref64["anorm"],
            "speedup_device": round(ref64["elapsed_ms"] / max(infos[-1]["elapsed_ms"], 1e-9), 3),
            "x_rel_diff": ref64["xdiff"]},
        "mixed_colscale": mixed_line,
        **extra,
        "aprod_gbs": {"mode1": round(gbs1, 1), "mode2": round(gbs2, 1), "both": round(gbs, 1)},
        "aprod_caller_order_ms": {"mode1": round(cm1, 4), "mode2": round(cm2, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "setup_s": {"generate": round(gen_s, 2), "operator_create": round(setup_s, 2)},
    }
    print(json.dumps(out), flush=True)
    op.close()
    for p in (b_dev, x_dev, y_dev, x32_dev, y32_dev):
        hip.free(p)


if __name__ == "__main__":
    main()
