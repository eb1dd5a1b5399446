#!/usr/bin/env python3
"""Benchmark of the Weighted MinHash hashing path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batched|plain]

A step is one hashing pass over one batch of synthetic CSR input already
resident in HBM (device entry points of the C ABI, launched on torch's current
stream).  Prints ONE JSON line: the metric nnz*samples/s, the roofline of the
dominant kernel (HBM-bound parameter gathers) and the CPU oracle baseline
timed on this host.

Workloads (128 samples, parameters from the device Philox generator):
  batched  (default) 64 independent CSR matrices, dim 2^20, rows per matrix
           log-uniform in [1k, 200k], nnz/row Zipf(1.2) clipped to [1, 20000],
           lognormal(0, 1.5) weights.  `value` is the batched persistent
           kernel (one launch for all 64 matrices); the plain kernel (one
           launch per matrix) is timed on the same data beside it.
  plain    100,000 x 2^20 CSR, nnz/row ~ Poisson(100), lognormal(0, 1)
           weights, plain kernel (one launch, one wavefront per row).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "minhashcuda_amd")
HBM_PEAK_GBS = 8000.0
METRIC = "nnz*samples/sec (128 samples, bit-exact) batched vs plain; p50/p99 serve latency"
SAMPLES = 128


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--workload", default="batched", choices=["batched", "plain", "serve", "large", "c1", "copy"],
                   help="copy: a plain 4 GiB device-to-device copy (calibrates counters)")
    p.add_argument("--qps", default="1000,3000,6000,9000,12000",
                   help="serve: comma-separated arrival rates (requests/s)")
    p.add_argument("--serve-seconds", type=float, default=3.0,
                   help="serve: duration of each arrival-rate run")
    p.add_argument("--slots", type=int, default=4,
                   help="serve: launches in flight at once (one stream and worker each)")
    p.add_argument("--pinned", action="store_true",
                   help="serve: requests live in pinned host memory (direct DMA)")
    p.add_argument("--spl", type=int, default=0, help="samples per lane (0: default)")
    p.add_argument("--seg", type=int, default=0, help="nonzeros per work item (0: default)")
    p.add_argument("--wpc", type=int, default=0, help="persistent waves per CU (0: default)")
    p.add_argument("--stripe-mib", type=int, default=-1,
                   help="persistent column-stripe size in MiB (0: off, -1: library default)")
    p.add_argument("--no-plain", action="store_true", help="batched: skip the plain comparison")
    p.add_argument("--kernel", default="plain", choices=["plain", "persistent"],
                   help="plain workload: the kernel to time (default: the plain kernel)")
    p.add_argument("--no-graph", action="store_true",
                   help="launch eagerly instead of replaying a captured hipGraph")
    p.add_argument("--no-autotune", action="store_true",
                   help="batched: keep the default tuning instead of autotuning in warm-up")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="target duration of the single-thread CPU oracle sample")
    p.add_argument("--count", type=int, default=64, help="batched: number of matrices")
    return p.parse_args()


def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, m, samples, target_s, gpu_hashes, what):
    """Times this repository's C oracle of the hashing path on a bounded row
    prefix of `m`: single thread, and OpenMP on up to 16 of the cores this
    process may use; checks the sample's hashes against the GPU's."""
    import numpy as np
    import oracle
    rs, ln_cs, betas = params
    n0 = min(m.shape[0], 200)
    while n0 < m.shape[0] and m[:n0].nnz * samples < 2e7:
        n0 = min(m.shape[0], n0 * 2)
    sub = m[:n0]
    t = time.perf_counter()
    oracle.calc_csr(rs, ln_cs, betas, sub)
    rate = max(sub.nnz * samples / max(time.perf_counter() - t, 1e-6), 1.0)
    want = target_s * rate / samples
    n1 = int(np.searchsorted(m.indptr, want)) if want < m.nnz else m.shape[0]
    n1 = max(1, min(m.shape[0], n1))
    sub = m[:n1]
    t = time.perf_counter()
    h1 = oracle.calc_csr(rs, ln_cs, betas, sub)
    t1 = time.perf_counter() - t
    threads = min(cpu_threads(), 16)
    want2 = sub.nnz * threads
    n2 = int(np.searchsorted(m.indptr, want2)) if want2 < m.nnz else m.shape[0]
    n2 = max(n1, min(m.shape[0], n2))
    sub2 = m[:n2]
    t = time.perf_counter()
    h2, used = oracle.calc_omp(rs, ln_cs, betas, sub2.data, sub2.indices, sub2.indptr, threads)
    t2 = time.perf_counter() - t
    ok = bool(np.array_equal(h2, gpu_hashes[:n2]) and np.array_equal(h1, gpu_hashes[:n1]))
    # information only: how often FP64 arithmetic would pick another argmin
    n3 = max(1, min(n1, 2000))
    h64 = oracle.calc_f64(rs, ln_cs, betas, m.data, m.indices, m.indptr[:n3 + 1])
    flips = float(np.mean(np.any(h64 != gpu_hashes[:n3], axis=2)))
    return {
        "value": sub2.nnz * samples / t2, "unit": "nnz*samples/s", "cores": int(used),
        "kind": "port", "cpu_model": cpu_model(),
        "sample": "%s: first %d rows (%d nnz), OpenMP C oracle, %.1f s" % (what, n2, sub2.nnz, t2),
        "single_thread_value": sub.nnz * samples / t1,
        "single_thread_sample": "%s: first %d rows (%d nnz), %.1f s" % (what, n1, sub.nnz, t1),
        "sample_bit_exact_vs_gpu": ok,
        "fp64_argmin_flip_rate": flips,
        "fp64_flip_sample": "%s: first %d rows, (row, sample) hashes where an FP64 evaluation "
                            "of the same formula selects another nonzero or t" % (what, n3),
    }


class Engine:
    def __init__(self, dim, seed, args):
        import libMHCUDA
        from minhashcuda_amd import capi as capi_mod
        self.L = libMHCUDA
        self.capi_mod = capi_mod
        self.capi = capi_mod.load(os.path.join(PKG, "libMHCUDA.so"))
        self.gen = libMHCUDA.minhash_cuda_init(dim, SAMPLES, seed=seed, devices=1)
        self.g = ctypes.c_void_p(self.gen)
        capi_mod.check(self.capi.mhcuda_set_tuning(self.g, args.spl, args.seg, args.wpc), "tuning")
        if getattr(args, "stripe_mib", -1) >= 0:
            capi_mod.check(self.capi.mhcuda_set_stripes(self.g, args.stripe_mib), "stripes")
        capi_mod.check(self.capi.mhcuda_set_timing(self.g, 1), "timing")
        capi_mod.check(self.capi.mhcuda_set_graph(self.g, 0 if args.no_graph else 1), "graph")
        self.graph = not args.no_graph

    def tuning(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.capi.mhcuda_get_tuning(self.g, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        mib, k, n = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_uint32()
        self.capi.mhcuda_get_stripes(self.g, ctypes.byref(mib), ctypes.byref(k), ctypes.byref(n))
        return {"samples_per_lane": a.value, "segment": b.value, "waves_per_cu": c.value,
                "stripe_mib": mib.value, "stripes": k.value, "striped_row_min_nnz": n.value,
                "hip_graph": self.graph}

    def kernel_ms(self):
        ms = ctypes.c_float()
        self.capi_mod.check(self.capi.mhcuda_kernel_time(self.g, ctypes.byref(ms)), "kernel_time")
        return ms.value

    def batch(self, descs, kernel, stream):
        self.capi_mod.check(self.capi.mhcuda_calc_batch_device(
            self.g, len(descs), descs, kernel, ctypes.c_void_p(stream.cuda_stream)), "calc_batch")

    def close(self):
        self.L.minhash_cuda_fini(self.gen)


def measured_traffic(workload, kernel, alg_bytes, stripes=None):
    """HBM bytes per launch of `kernel` on this workload as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded in
    profiles/hbm_traffic.json (scaled by a copy-kernel calibration); None when
    no measurement of this exact configuration (input size, kernel and, for
    the persistent kernel, number of column stripes) is recorded."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)["workloads"][workload]
    except (OSError, KeyError, ValueError):
        return None
    for rec in recs:
        if rec.get("kernel") != kernel or rec.get("algorithmic_bytes") != alg_bytes:
            continue
        if stripes is not None and rec.get("stripes") != stripes:
            continue
        return rec["traffic_bytes"]
    return None


def explicit_tuning(args):
    """True when the command line fixes a tuning knob (then nothing is autotuned)."""
    return bool(args.spl or args.seg or args.wpc) or args.stripe_mib >= 0


def autotune(engine, descs, stream):
    """mhcuda_autotune on the batch; returns (kernel, result dict)."""
    cm = engine.capi_mod
    t0 = time.perf_counter()
    tr = cm.TuneResult()
    cm.check(engine.capi.mhcuda_autotune(engine.g, len(descs), descs,
                                         ctypes.c_void_p(stream.cuda_stream),
                                         ctypes.byref(tr)), "autotune")
    tuned = dict(tr.as_dict(), seconds=time.perf_counter() - t0)
    log("autotuned: %s" % tuned)
    return tr.kernel, tuned


def timed_steps(torch, run, engine, steps, warmup):
    """Runs `run` warmup + steps times; returns (wall seconds of the timed
    steps, per-step kernel milliseconds)."""
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
        kms.append(engine.kernel_ms())
    torch.cuda.synchronize()
    return time.perf_counter() - t0, kms


def serve_requests(n, dim, seed):
    """Serving requests: 1-4096 rows (log-uniform), nnz/row lognormal with
    median 200 and sigma 1 clipped to [1, 50000], lognormal(0, 1) weights."""
    import numpy as np
    from minhashcuda_amd import workloads
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        rows = int(np.exp(rng.uniform(0, np.log(4097))))
        rows = max(1, min(4096, rows))
        nnz = np.clip(np.rint(rng.lognormal(np.log(200), 1.0, rows)), 1, 50000).astype(np.int64)
        reqs.append(workloads.batched_csr(nnz, dim=dim, sigma=1.0, seed=seed * 100000 + i))
    return reqs


def serve_bench(args, torch):
    """Continuous-batching serving loop under Poisson arrivals at each rate of
    --qps: p50/p99 latency (submit -> hashes in host memory) and throughput."""
    import threading
    import numpy as np
    import libMHCUDA
    from minhashcuda_amd import serving
    dim, samples = 1 << 22, 256
    log("serve: generating request pool")
    pool = serve_requests(256, dim, seed=5)
    pinned_pool = []
    if args.pinned:
        for m in pool:
            arrs = []
            for a in (m.data.astype(np.float32), m.indices.astype(np.uint32),
                      m.indptr.astype(np.uint32)):
                t = torch.empty(a.shape, dtype=torch.float32 if a.dtype == np.float32
                                else torch.int32, pin_memory=True)
                v = t.numpy().view(a.dtype)
                v[...] = a
                arrs.append((t, v))
            ot = torch.empty((m.shape[0], samples, 2), dtype=torch.int32, pin_memory=True)
            arrs.append((ot, ot.numpy().view(np.uint32)))
            pinned_pool.append(arrs)
    gen = libMHCUDA.minhash_cuda_init(dim, samples, seed=1, devices=1)
    if explicit_tuning(args):
        from minhashcuda_amd import capi as capi_mod
        capi = capi_mod.load(os.path.join(PKG, "libMHCUDA.so"))
        g = ctypes.c_void_p(gen)
        capi_mod.check(capi.mhcuda_set_tuning(g, args.spl, args.seg, args.wpc), "tuning")
        if args.stripe_mib >= 0:
            capi_mod.check(capi.mhcuda_set_stripes(g, args.stripe_mib), "stripes")
    server = serving.Server(gen, memory_budget=4 << 30, slots=args.slots, max_batch_requests=256)
    # warm-up: every buffer reaches its working size
    for t in [server.submit(m) for m in pool[:64]]:
        server.wait(t)
    server.stats(reset=True)
    results = []
    rng = np.random.default_rng(11)
    for qps in [float(q) for q in args.qps.split(",")]:
        n = max(20, int(qps * args.serve_seconds))
        gaps = rng.exponential(1.0 / qps, n)
        tickets = []
        done = threading.Event()

        def waiter():
            i = 0
            while i < n:
                while i >= len(tickets):
                    time.sleep(0.0002)
                server.wait(tickets[i])
                i += 1
            done.set()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        due = t0
        for i in range(n):
            due += gaps[i]
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                if due - now > 0.0005:
                    time.sleep(due - now - 0.0003)
            if args.pinned:
                a = pinned_pool[i % len(pool)]
                tickets.append(server.submit_arrays(a[0][1], a[1][1], a[2][1], a[3][1]))
            else:
                tickets.append(server.submit(pool[i % len(pool)]))
        submit_span = time.perf_counter() - t0
        th.join()
        st = server.stats(reset=True)
        nnz = sum(pool[i % len(pool)].nnz for i in range(n))
        rows = sum(pool[i % len(pool)].shape[0] for i in range(n))
        results.append({"offered_qps": qps, "requests": n, "achieved_qps": n / st["window_s"],
                        "submit_qps": n / submit_span, "p50_ms": st["p50_ms"],
                        "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"], "max_ms": st["max_ms"],
                        "nnz_samples_per_sec": st["nnz_samples_per_sec"],
                        "mean_requests_per_batch": st["mean_requests_per_batch"],
                        "launches": st["batches"], "mean_rows_per_request": rows / n,
                        "mean_nnz_per_request": nnz / n})
        log("qps %.0f: p50 %.2f ms p99 %.2f ms, %.3g nnz*S/s, %.1f req/launch" % (
            qps, st["p50_ms"], st["p99_ms"], st["nnz_samples_per_sec"],
            st["mean_requests_per_batch"]))
    server.stop()
    libMHCUDA.minhash_cuda_fini(gen)
    best = max(results, key=lambda r: r["nnz_samples_per_sec"])
    print(json.dumps({
        "metric": METRIC, "workload": "serve", "unit": "nnz*samples/s",
        "value": best["nnz_samples_per_sec"], "higher_is_better": True, "n_gpus": 1,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "serving: Poisson arrivals, requests of 1-4096 rows "
                               "(log-uniform), dim 2^22, nnz/row lognormal(median 200, sigma 1) "
                               "clipped to 50000, lognormal(0,1) weights, 256 samples, "
                               "parameters resident (12.9 GB), %d slots, 4 GiB budget, "
                               "request arrays in %s host memory"
                               % (args.slots, "pinned" if args.pinned else "pageable"),
                   "pool_requests": len(pool)},
        "sweep": results}), flush=True)


def c1_bench(args, torch):
    """The reference's CPU-runnable case (test.py:93-97, README.md:85-90):
    6400 x 130 CSR from numpy seed 1 (values in [0,100), 80% zeroed), 128
    samples, parameters assigned through assign_random_vars.  Times the whole
    6400 rows on the CPU oracle (single thread and OpenMP) and on the GPU
    (device-resident input), and checks that the hashes agree."""
    import numpy as np
    import scipy.sparse as sp
    import oracle
    from minhashcuda_amd import capi as cm
    data = oracle.c1_data(1)
    m = sp.csr_matrix(data, dtype=np.float32)
    rs, ln_cs, betas = oracle.datasketch_style_params(130, SAMPLES, seed=1)
    steps = args.steps or 200
    engine = Engine(130, 1, args)
    engine.L.minhash_cuda_assign_vars(engine.gen, rs, ln_cs, betas)
    stream = torch.cuda.current_stream()
    w = torch.from_numpy(m.data.astype(np.float32)).cuda()
    c = torch.from_numpy(m.indices.astype(np.int32)).cuda()
    r = torch.from_numpy(m.indptr.astype(np.int64).astype(np.int32)).cuda()
    out = torch.empty((m.shape[0], SAMPLES, 2), dtype=torch.int32, device="cuda")
    desc = (cm.Matrix * 1)(cm.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), m.shape[0],
                                     out.data_ptr()))
    res = {}
    for name, kernel in (("plain", cm.KERNEL_PLAIN), ("persistent", cm.KERNEL_PERSISTENT)):
        wall, kms = timed_steps(torch, lambda: engine.batch(desc, kernel, stream), engine, steps, 5)
        res[name] = {"value": m.nnz * SAMPLES * steps / wall, "ms_per_step": wall * 1e3 / steps,
                     "avg_kernel_ms": sum(kms) / len(kms)}
    gpu_h = out.cpu().numpy().view(np.uint32)
    t = time.perf_counter()
    h1 = oracle.calc_csr(rs, ln_cs, betas, m)
    t1 = time.perf_counter() - t
    threads = min(cpu_threads(), 16)
    t = time.perf_counter()
    h2, used = oracle.calc_omp(rs, ln_cs, betas, m.data, m.indices, m.indptr, threads)
    t2 = time.perf_counter() - t
    best = max(res, key=lambda k: res[k]["value"])
    print(json.dumps({
        "metric": METRIC, "workload": "c1", "unit": "nnz*samples/s", "value": res[best]["value"],
        "steps": steps, "higher_is_better": True, "n_gpus": 1, "dtype": "f32",
        "data": "synthetic (test.py recipe, seed 1)",
        "config": {"workload": "c1: 6400 x 130 CSR (nnz %d), 128 samples, assigned parameters; "
                               "GPU value = best of plain / persistent" % m.nnz},
        "gpu": res,
        "cpu_baseline": {"value": m.nnz * SAMPLES / t2, "unit": "nnz*samples/s", "cores": used,
                         "kind": "port", "cpu_model": cpu_model(),
                         "sample": "all 6400 rows, OpenMP C oracle",
                         "single_thread_value": m.nnz * SAMPLES / t1,
                         "single_thread_sample": "all 6400 rows, %.3f s" % t1},
        "bit_exact": bool(np.array_equal(gpu_h, h1) and np.array_equal(h1, h2)),
    }), flush=True)
    engine.close()


def large_bench(args, torch):
    """1,000,000 x 12,000,000 CSR, 16,800 nonzeros per row (density 0.0014),
    uniform columns, weights uniform in (0, 100], 128 samples, parameters
    resident (18.4 GB).  Two figures:
      * kernel: the hashing kernel (autotuned on the chunk unless a tuning
        knob is given) on one device-resident 65,536-row chunk (1.1e9
        nonzeros) -- the roofline figure, reported as `value`;
      * end to end: all 1,000,000 rows through mhcuda_calc_batch from a pinned
        host copy of that chunk (16 matrices of <= 65,536 rows, the last one
        partial), with double-buffered uploads overlapping the hashing --
        PCIe-inclusive, reported beside it."""
    import numpy as np
    import scipy.sparse as sp
    from minhashcuda_amd import workloads
    dim, per_row, chunk_rows, total_rows = 12_000_000, 16_800, 65_536, 1_000_000
    steps = args.steps or 3
    engine = Engine(dim, 1, args)
    cm = engine.capi_mod
    stream = torch.cuda.current_stream()
    log("large: generating one 65,536-row chunk on the GPU")
    w, c, r = workloads.batched_csr_torch(np.full(chunk_rows, per_row), dim, seed=7,
                                          weights="uniform100")
    nnz = int(w.numel())
    out = torch.empty((chunk_rows, SAMPLES, 2), dtype=torch.int32, device="cuda")
    desc = (cm.Matrix * 1)(cm.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), chunk_rows,
                                     out.data_ptr()))
    torch.cuda.synchronize()
    kernel, tuned = cm.KERNEL_PERSISTENT, None
    if not args.no_autotune and not explicit_tuning(args):
        kernel, tuned = autotune(engine, desc, stream)
    kname = "wmh_persistent_kernel" if kernel == cm.KERNEL_PERSISTENT else "wmh_plain_kernel"
    wall, kms = timed_steps(torch, lambda: engine.batch(desc, kernel, stream), engine, steps, 1)
    avg_ms = sum(kms) / len(kms)
    alg = workloads.algorithmic_bytes(nnz, chunk_rows, SAMPLES)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    # pinned host copy of the chunk, then the whole 1M rows end to end
    log("large: pinned host copy")
    hw = torch.empty(nnz, dtype=torch.float32, pin_memory=True)
    hc = torch.empty(nnz, dtype=torch.int32, pin_memory=True)
    hr = torch.empty(chunk_rows + 1, dtype=torch.int32, pin_memory=True)
    hw.copy_(w)
    hc.copy_(c)
    hr.copy_(r)
    hout = torch.empty((total_rows, SAMPLES, 2), dtype=torch.int32, pin_memory=True)
    lengths = [chunk_rows] * (total_rows // chunk_rows)
    if total_rows % chunk_rows:
        lengths.append(total_rows % chunk_rows)
    mats = (cm.Matrix * len(lengths))()
    row0 = 0
    for i, n in enumerate(lengths):
        mats[i] = cm.Matrix(hw.data_ptr(), hc.data_ptr(), hr.data_ptr(), n,
                            hout[row0:].data_ptr())
        row0 += n
    torch.cuda.synchronize()
    ends = []
    for _ in range(2):
        t0 = time.perf_counter()
        cm.check(engine.capi.mhcuda_calc_batch(engine.g, len(lengths), mats,
                                               cm.KERNEL_PERSISTENT), "calc_batch")
        ends.append(time.perf_counter() - t0)
    rp = r.cpu().numpy().view(np.uint32).astype(np.int64)
    e2e_nnz = sum(int(rp[n]) for n in lengths)
    e2e_units = e2e_nnz * SAMPLES
    # checks: the end-to-end hashes of every chunk equal the resident chunk's
    dev_h = out.cpu().numpy().view(np.uint32)
    host_h = hout.numpy().view(np.uint32)
    row0, same = 0, True
    for n in lengths:
        same = same and np.array_equal(host_h[row0:row0 + n], dev_h[:n])
        row0 += n
    result = {
        "metric": METRIC, "workload": "large", "unit": "nnz*samples/s",
        "value": nnz * SAMPLES * steps / wall, "steps": steps, "warmup": 1,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "n_gpus": 1,
        "scaling": "weak", "dtype": "f32", "data": "synthetic", "vs_baseline": None,
        "config": dict({"workload": "large: 1,000,000 x 12,000,000 CSR, 16,800 nnz/row, "
                                    "uniform(0,100] weights, 128 samples, parameters resident "
                                    "(18.4 GB); value = the autotuned hashing kernel on one "
                                    "resident 65,536-row chunk (1.1e9 nnz)",
                        "chunk_rows": chunk_rows, "chunk_nnz": nnz}, **engine.tuning()),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("large", kname, alg,
                                                 engine.tuning()["stripes"]
                                                 if kname == "wmh_persistent_kernel" else None),
                     "algorithmic_bytes_per_launch": alg, "avg_kernel_ms": avg_ms},
        "end_to_end": {"rows": total_rows, "nnz": e2e_nnz,
                       "seconds": min(ends), "value": e2e_units / min(ends),
                       "unit": "nnz*samples/s (PCIe-inclusive, pinned host input)",
                       "identical_to_resident_chunk": bool(same)},
    }
    if tuned:
        result["autotune"] = tuned
    if not args.no_cpu_baseline:
        log("large: cpu baseline (retrieving 18.4 GB of parameters)")
        params = engine.L.minhash_cuda_retrieve_vars(engine.gen)
        cut = 600
        m0 = sp.csr_matrix((hw[:rp[cut]].numpy(), hc[:rp[cut]].numpy().view(np.uint32),
                            rp[:cut + 1]), shape=(cut, dim))
        result["cpu_baseline"] = cpu_baseline(params, m0, SAMPLES, args.cpu_seconds,
                                              dev_h[:cut], "first rows of the chunk")
    engine.close()
    print(json.dumps(result), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine hashes on a single MI355X "
              "(multi-GPU row sharding is not built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch  # before the engine: one HIP runtime per process
    sys.path.insert(0, ROOT)
    sys.path.insert(0, PKG)
    from minhashcuda_amd import workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    if args.workload == "copy":
        n = 1 << 30  # 4 GiB of float32
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"workload": "copy", "bytes_read_per_copy": 4 * n,
                          "bytes_written_per_copy": 4 * n, "copies": 3}), flush=True)
        return
    if args.workload == "serve":
        serve_bench(args, torch)
        return
    if args.workload == "large":
        large_bench(args, torch)
        return
    if args.workload == "c1":
        c1_bench(args, torch)
        return
    dim = 1 << 20
    engine = Engine(dim, 1, args)
    cm = engine.capi_mod
    result = {"metric": METRIC, "unit": "nnz*samples/s", "n_gpus": 1,
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
              "dtype": "f32", "data": "synthetic"}

    if args.workload == "plain":
        steps = args.steps or 20
        warmup = args.warmup if args.warmup is not None else 3
        log("generating plain workload")
        m = workloads.plain_csr(100000, dim, 100.0, 1.0, seed=1)
        mats = [m]
        tensors = [(torch.from_numpy(m.data.astype(np.float32)).cuda(),
                    torch.from_numpy(m.indices.astype(np.int32)).cuda(),
                    torch.from_numpy(m.indptr.astype(np.int64).astype(np.int32)).cuda())]
        if args.kernel == "persistent":
            kernel = cm.KERNEL_PERSISTENT
            kname = "wmh_persistent_kernel"
        else:
            kernel = cm.KERNEL_PLAIN
            kname = "wmh_plain_kernel"
        desc = ("plain: 100000 x 2^20 CSR, nnz/row~Poisson(100), lognormal(0,1) weights, "
                "128 samples, one launch (%s kernel)" % args.kernel)
    else:
        steps = args.steps or 3
        warmup = args.warmup if args.warmup is not None else 1
        log("generating batched workload on the GPU")
        layout = workloads.batched_layout(count=args.count, seed=1)
        tensors = []
        for i, n in enumerate(layout):
            tensors.append(workloads.batched_csr_torch(n, dim, 1.5, seed=100 + i))
        mats = None
        kernel = cm.KERNEL_PERSISTENT
        kname = "wmh_persistent_kernel"
        desc = ("batched: %d CSR matrices, dim 2^20, rows/matrix log-uniform [1k,200k], "
                "nnz/row Zipf(1.2) clipped [1,20000], lognormal(0,1.5) weights, 128 samples, "
                "one persistent launch" % args.count)
    lengths = [int(t[2].numel()) - 1 for t in tensors]
    nnzs = [int(t[0].numel()) for t in tensors]
    rows, nnz = sum(lengths), sum(nnzs)
    outs = [torch.empty((n, SAMPLES, 2), dtype=torch.int32, device="cuda") for n in lengths]
    descs = (cm.Matrix * len(tensors))()
    for i, ((w, c, r), o) in enumerate(zip(tensors, outs)):
        descs[i] = cm.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), lengths[i], o.data_ptr())
    torch.cuda.synchronize()
    tuned = None
    if args.workload == "batched" and not args.no_autotune and not explicit_tuning(args):
        kernel, tuned = autotune(engine, descs, stream)
        kname = "wmh_persistent_kernel" if kernel == cm.KERNEL_PERSISTENT else "wmh_plain_kernel"
    log("rows %d nnz %d; running %d warmup + %d timed steps" % (rows, nnz, warmup, steps))

    wall, kms = timed_steps(torch, lambda: engine.batch(descs, kernel, stream), engine, steps,
                            warmup)
    units = nnz * SAMPLES
    avg_ms = sum(kms) / len(kms)
    alg = workloads.algorithmic_bytes(nnz, rows, SAMPLES)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    result.update({
        "value": units * steps / wall, "steps": steps, "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps,
        "config": dict({"workload": desc, "rows": rows, "nnz": nnz, "samples": SAMPLES,
                        "matrices": len(tensors), "parallelism": "single GPU"},
                       **engine.tuning()),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, kname, alg,
                                                 engine.tuning()["stripes"]
                                                 if kname == "wmh_persistent_kernel" else None),
                     "algorithmic_bytes_per_launch": alg, "avg_kernel_ms": avg_ms,
                     "timed_span": ("hipGraph replay: descriptor upload, plan, init, %s, "
                                    "fix-up (HIP events around the graph launch)" % kname
                                    if engine.graph and kernel == cm.KERNEL_PERSISTENT else
                                    "HIP events around the %s launch(es)" % kname)},
    })
    if tuned is not None:
        result["autotune"] = tuned
    if args.workload == "batched" and not args.no_plain:
        ref_out = [o.clone() for o in outs]
        for o in outs:
            o.fill_(0)
        pwall, pkms = timed_steps(torch, lambda: engine.batch(descs, cm.KERNEL_PLAIN, stream),
                                  engine, steps, warmup)
        same = all(torch.equal(a, b) for a, b in zip(ref_out, outs))
        result["plain"] = {"value": units * steps / pwall, "ms_per_step": pwall * 1e3 / steps,
                           "avg_kernel_ms": sum(pkms) / len(pkms),
                           "launches_per_step": len(tensors),
                           "identical_to_batched": bool(same)}
        result["batched_over_plain"] = result["value"] / result["plain"]["value"]
        del ref_out
    if not args.no_cpu_baseline:
        log("cpu baseline")
        params = engine.L.minhash_cuda_retrieve_vars(engine.gen)
        if mats is None:
            # the CPU sample: the first rows of the first matrix, copied back
            w, c, r = tensors[0]
            rp = r.cpu().numpy().view(np.uint32).astype(np.int64)
            cut = int(np.searchsorted(rp, 60_000_000))
            cut = max(1, min(lengths[0], cut))
            import scipy.sparse as sp
            m0 = sp.csr_matrix((w[:rp[cut]].cpu().numpy(),
                                c[:rp[cut]].cpu().numpy().view(np.uint32),
                                rp[:cut + 1]), shape=(cut, dim))
            gh = outs[0][:cut].cpu().numpy().view(np.uint32)
            result["cpu_baseline"] = cpu_baseline(params, m0, SAMPLES, args.cpu_seconds, gh,
                                                  "matrix 0 of the batch")
        else:
            gh = outs[0].cpu().numpy().view(np.uint32)
            result["cpu_baseline"] = cpu_baseline(params, mats[0], SAMPLES, args.cpu_seconds,
                                                  gh, "plain workload")
    engine.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
