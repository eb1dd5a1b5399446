#!/usr/bin/env python3
"""Benchmark of the Weighted MinHash hashing path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batched|plain|large|serve|c1]

A step is one hashing pass over one batch of synthetic CSR input already
resident in HBM (device entry points of the C ABI, launched on torch's current
stream).  Prints ONE JSON line with the metric nnz*samples/s, the roofline of
the dominant kernel against the resource that binds it (VALU issue or HBM, from
the rocprofv3 counters committed under profiles/), and the CPU baseline timed
on this host.

Workloads (parameters from the device Philox generator):
  batched  (default) 64 independent CSR matrices, dim 2^20, rows per matrix
           log-uniform in [1k, 200k], nnz/row Zipf(1.2) clipped to [1, 20000],
           lognormal(0, 1.5) weights, 128 samples.  `value` is the batched
           persistent kernel (one launch for all 64 matrices, autotuned,
           hipGraph replay); the plain kernel (one launch per matrix) is timed
           on the same data beside it.  The same line also carries, under
           `large`, the 1,000,000 x 12,000,000 share and, under `serve`, the
           serving loop's p50/p99 latency and throughput (see below).
  plain    100,000 x 2^20 CSR, nnz/row ~ Poisson(100), lognormal(0, 1)
           weights, 128 samples, plain kernel (one launch, one wavefront per row).
  large    1,000,000 x 12,000,000 CSR, 16,800 nnz/row, uniform(0, 100]
           weights, 128 samples, 18.4 GB of parameters resident: one resident
           65,536-row chunk, plus all 1M rows end to end from pinned host memory.
  serve    continuous-batching serving loop: Poisson arrivals at a QPS sweep,
           requests of 1-4096 rows, dim 2^22, 256 samples, 12.9 GB resident;
           plus the hashing kernel on the largest launch the loop makes (the
           256 pool requests coalesced, device-resident) and its roofline.
  serve_kernel  that kernel measurement alone (for profiling).
  c1       the reference's CPU-runnable case (6400 x 130, test.py recipe).
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
# VALU issue ceiling in lane operations per second: the published FP32 vector
# peak (157.3 TFLOP/s) counts an FMA as two operations and two operations per
# lane per cycle; one VALU instruction per lane per cycle is a quarter of it
# (256 CUs x 64 lanes x 2.4 GHz).
VALU_PEAK_LANE_OPS = 157.3e12 / 4
METRIC = "nnz*samples/sec (128 samples, bit-exact) batched vs plain; p50/p99 serve latency"
SAMPLES = 128


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--workload", default="batched",
                   choices=["batched", "plain", "serve", "serve_kernel", "large", "c1", "copy"],
                   help="copy: a plain 4 GiB device-to-device copy (calibrates counters)")
    p.add_argument("--qps", default=None,
                   help="serve: comma-separated arrival rates (requests/s)")
    p.add_argument("--serve-seconds", type=float, default=None,
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
    p.add_argument("--seeds", type=int, default=-1,
                   help="heavy seeds of striped rows (0: off, 1: on, -1: library default)")
    p.add_argument("--no-skip", action="store_true",
                   help="hash without the skip bound (mhcuda_set_skip(0))")
    p.add_argument("--no-plain", action="store_true", help="batched: skip the plain comparison")
    p.add_argument("--no-seed2", action="store_true",
                   help="batched: skip the second layout seed")
    p.add_argument("--no-large", action="store_true",
                   help="batched: leave the 1M x 12M share out of the line")
    p.add_argument("--no-serve", action="store_true",
                   help="batched: leave the serving sweep out of the line")
    p.add_argument("--kernel", default="plain", choices=["plain", "persistent"],
                   help="plain workload: the kernel to time (default: the plain kernel)")
    p.add_argument("--no-graph", action="store_true",
                   help="launch eagerly instead of replaying a captured hipGraph")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the default tuning instead of autotuning in warm-up")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=8.0,
                   help="target duration of each timed CPU baseline sample")
    p.add_argument("--count", type=int, default=64, help="batched: number of matrices")
    return p.parse_args(argv)


# ---------------------------------------------------------------------------
# host facts and the CPU baseline

def cpu_threads():
    """Threads of the CPU baseline: the host's CPU share for this process
    (OMP_NUM_THREADS when the environment sets it, as the GPU hosts do for one
    GPU's share of their cores), else every CPU this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(affinity())


def affinity():
    try:
        return sorted(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return list(range(os.cpu_count() or 1))


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


def host_facts():
    return {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity_cpus": len(affinity()),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def rows_for(m, units_target, samples):
    """Smallest row prefix of CSR `m` holding about units_target nnz*samples."""
    import numpy as np
    want = units_target / samples
    n = int(np.searchsorted(m.indptr, want)) if want < m.nnz else m.shape[0]
    return max(1, min(m.shape[0], n))


def cpu_baseline(params, m, samples, target_s, gpu_hashes, what):
    """Times the CPU port of the hashing formula (oracle/wmh_cpu_fast.c: the
    same binary32 operations in the same order, tables transposed so the loop
    over samples vectorises, rows over OpenMP threads) on bounded row prefixes
    of `m` -- one thread and the host's CPU share -- plus the scalar C oracle
    (the reference's [samples][dim] layout) on one thread, and checks every
    sample's hashes against the GPU's."""
    import numpy as np
    import oracle
    params_t = oracle.transpose_params(*params)
    threads = cpu_threads()
    out = {}

    def timed(fn, rows):
        sub = m[:rows]
        t = time.perf_counter()
        h = fn(sub)
        dt = time.perf_counter() - t
        return sub, h, dt

    # calibrate on a small prefix, then time about target_s of work
    probe = rows_for(m, 2e8, samples)
    sub, _, dt = timed(lambda s: oracle.calc_fast(params_t, s.data, s.indices, s.indptr, 1)[0],
                       probe)
    rate1 = max(sub.nnz * samples / max(dt, 1e-6), 1.0)
    n1 = rows_for(m, rate1 * target_s, samples)
    sub1, h1, t1 = timed(lambda s: oracle.calc_fast(params_t, s.data, s.indices, s.indptr,
                                                    1)[0], n1)
    n2 = rows_for(m, rate1 * threads * target_s, samples)
    used = [0]

    def omp(s):
        h, used[0] = oracle.calc_fast(params_t, s.data, s.indices, s.indptr, threads)
        return h

    sub2, h2, t2 = timed(omp, n2)
    n3 = rows_for(m, rate1 * min(target_s, 3.0) / 20.0, samples)  # the scalar oracle is ~20x slower
    sub3, h3, t3 = timed(lambda s: oracle.calc_csr(*params, s), n3)
    ok = bool(np.array_equal(h1, gpu_hashes[:n1]) and np.array_equal(h2, gpu_hashes[:n2]) and
              np.array_equal(h3, gpu_hashes[:n3]))
    # information only: how often FP64 arithmetic would pick another argmin
    n4 = max(1, min(n3, 2000))
    h64 = oracle.calc_f64(*params, m.data, m.indices, m.indptr[:n4 + 1])
    flips = float(np.mean(np.any(h64 != gpu_hashes[:n4], axis=2)))
    out.update({
        "value": sub2.nnz * samples / t2, "unit": "nnz*samples/s", "cores": int(used[0]),
        "kind": "port",
        "sample": "%s: first %d rows (%d nnz), OpenMP + AVX2 port of the formula on %d threads, "
                  "%.1f s" % (what, n2, sub2.nnz, used[0], t2),
        "single_thread_value": sub1.nnz * samples / t1,
        "single_thread_sample": "%s: first %d rows (%d nnz), the same port on 1 thread, %.1f s"
                                % (what, n1, sub1.nnz, t1),
        "scalar_oracle_value": sub3.nnz * samples / t3,
        "scalar_oracle_sample": "%s: first %d rows (%d nnz), the scalar C oracle (reference "
                                "[samples][dim] layout), 1 thread, %.1f s"
                                % (what, n3, sub3.nnz, t3),
        "sample_bit_exact_vs_gpu": ok,
        "fp64_argmin_flip_rate": flips,
        "fp64_flip_sample": "%s: first %d rows, (row, sample) hashes where an FP64 evaluation "
                            "of the same formula selects another nonzero or t" % (what, n4),
        "threads_note": "threads = the host's CPU share for one GPU (OMP_NUM_THREADS when "
                        "set), not nproc",
    }, **host_facts())
    return out


# ---------------------------------------------------------------------------
# the engine through the C ABI

class Engine:
    def __init__(self, dim, seed, args, samples=SAMPLES):
        import libMHCUDA
        from minhashcuda_amd import capi as capi_mod
        self.L = libMHCUDA
        self.capi_mod = capi_mod
        self.capi = capi_mod.load(os.path.join(PKG, "libMHCUDA.so"))
        self.gen = libMHCUDA.minhash_cuda_init(dim, samples, seed=seed, devices=1)
        self.g = ctypes.c_void_p(self.gen)
        capi_mod.check(self.capi.mhcuda_set_tuning(self.g, args.spl, args.seg, args.wpc), "tuning")
        if getattr(args, "stripe_mib", -1) >= 0:
            capi_mod.check(self.capi.mhcuda_set_stripes(self.g, args.stripe_mib), "stripes")
        if getattr(args, "no_skip", False):
            capi_mod.check(self.capi.mhcuda_set_skip(self.g, 0), "skip")
        if getattr(args, "seeds", -1) >= 0:
            capi_mod.check(self.capi.mhcuda_set_seeds(self.g, args.seeds), "seeds")
        capi_mod.check(self.capi.mhcuda_set_timing(self.g, 1), "timing")
        capi_mod.check(self.capi.mhcuda_set_graph(self.g, 0 if args.no_graph else 1), "graph")
        self.graph = not args.no_graph

    def tuning(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.capi.mhcuda_get_tuning(self.g, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        mib, k, n = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_uint32()
        self.capi.mhcuda_get_stripes(self.g, ctypes.byref(mib), ctypes.byref(k), ctypes.byref(n))
        en, act = ctypes.c_int(), ctypes.c_int()
        self.capi.mhcuda_get_skip(self.g, ctypes.byref(en), ctypes.byref(act))
        sd = ctypes.c_int()
        self.capi.mhcuda_get_seeds(self.g, ctypes.byref(sd))
        return {"samples_per_lane": a.value, "segment": b.value, "waves_per_cu": c.value,
                "stripe_mib": mib.value, "stripes": k.value, "striped_row_min_nnz": n.value,
                "skip_bound": bool(act.value), "heavy_seeds": bool(sd.value),
                "hip_graph": self.graph}

    def kernel_ms(self):
        ms = ctypes.c_float()
        self.capi_mod.check(self.capi.mhcuda_kernel_time(self.g, ctypes.byref(ms)), "kernel_time")
        return ms.value

    def eager_kernel_ms(self, run, n=2):
        """The persistent kernel's own duration (HIP events around its launch
        on the launch stream), from n launches without graph replay: a
        replayed graph is timed as a whole (plan, init, heavy seeds, the
        kernel and the fix-up).  None when graphs are off (the timed steps
        then measured the kernel itself)."""
        if not self.graph:
            return None
        self.capi_mod.check(self.capi.mhcuda_set_graph(self.g, 0), "graph")
        try:
            kms = []
            for _ in range(n):
                run()
                kms.append(self.kernel_ms())
        finally:
            self.capi_mod.check(self.capi.mhcuda_set_graph(self.g, 1), "graph")
        return sum(kms) / len(kms)

    def batch(self, descs, kernel, stream):
        self.capi_mod.check(self.capi.mhcuda_calc_batch_device(
            self.g, len(descs), descs, kernel, ctypes.c_void_p(stream.cuda_stream)), "calc_batch")

    def close(self):
        self.L.minhash_cuda_fini(self.gen)


# ---------------------------------------------------------------------------
# rooflines from the committed rocprofv3 counter records

def counter_record(workload, kernel, alg_bytes, stripes=None, seeds=False):
    """The rocprofv3 record (profiles/hbm_traffic.json) of `kernel` on this
    workload's input (identified by its algorithmic bytes) at this stripe
    count and heavy-seed setting, or None."""
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
        if bool(rec.get("seeds", False)) != bool(seeds):
            continue
        return rec
    return None


def measured_traffic(workload, kernel, alg_bytes, stripes=None, seeds=False):
    """HBM bytes per launch of `kernel` on this workload as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded in
    profiles/hbm_traffic.json (scaled by a copy-kernel calibration); None when
    no measurement of this exact configuration (input size, kernel and, for
    the persistent kernel, number of column stripes and heavy-seed setting)
    is recorded."""
    rec = counter_record(workload, kernel, alg_bytes, stripes, seeds)
    return None if rec is None else rec.get("traffic_bytes")


def roofline(workload, kname, alg_bytes, avg_ms, stripes, seeds=False):
    """The dominant kernel against the resource that binds it.  Two ceilings
    are computed from the counter record of the same kernel, input, stripe
    count and heavy-seed setting: VALU issue (SQ_INSTS_VALU x 64 lane
    operations per launch over the launch time, against VALU_PEAK_LANE_OPS)
    and HBM (FETCH_SIZE + WRITE_SIZE bytes per launch over the launch time,
    against 8 TB/s).  `bound` is the
    one closer to its peak; the algorithmic bytes over the launch time are
    reported beside it as hbm_algorithmic_x (above 1 where the caches serve
    re-read parameters)."""
    t = avg_ms * 1e-3
    persistent = kname == "wmh_persistent_kernel"
    rec = counter_record(workload, kname, alg_bytes, stripes if persistent else None,
                         seeds and persistent)
    traffic = None if rec is None else rec.get("traffic_bytes")
    insts = None if rec is None else rec.get("valu_insts")
    hbm_frac = traffic / t / (HBM_PEAK_GBS * 1e9) if traffic else None
    valu_frac = insts * 64 / t / VALU_PEAK_LANE_OPS if insts else None
    out = {"kernel": kname, "avg_kernel_ms": avg_ms, "traffic": traffic,
           "algorithmic_bytes_per_launch": alg_bytes,
           "hbm_algorithmic_x": alg_bytes / t / (HBM_PEAK_GBS * 1e9),
           "hbm_traffic_frac": hbm_frac, "valu_frac": valu_frac,
           "valu_lane_ops_per_launch": insts * 64 if insts else None,
           "counters": None if rec is None else rec.get("profile")}
    if valu_frac is not None and (hbm_frac is None or valu_frac >= hbm_frac):
        out.update({"bound": "valu", "achieved": insts * 64 / t / 1e12,
                    "peak": VALU_PEAK_LANE_OPS / 1e12, "unit": "T lane-ops/s",
                    "frac": valu_frac})
    elif hbm_frac is not None:
        out.update({"bound": "hbm", "achieved": traffic / t / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": hbm_frac})
    else:
        # no counter record of this configuration: the HBM ceiling in
        # algorithmic bytes, which cache reuse can exceed
        out.update({"bound": "hbm", "achieved": alg_bytes / t / 1e9, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s (algorithmic bytes; no counter record of this tuning)",
                    "frac": alg_bytes / t / (HBM_PEAK_GBS * 1e9)})
    return out


def explicit_tuning(args):
    """True when the command line fixes a tuning knob (then nothing is autotuned)."""
    return bool(args.spl or args.seg or args.wpc) or args.stripe_mib >= 0 or args.seeds >= 0


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


def second_seed(args, torch, engine, kernel, stream, steps):
    """The uneven batch again with another draw of its layout (row counts and
    nnz/row from seed 2, matrix contents from seeds 200 + i), hashed with the
    engine's current tuning: a check that the headline figure is not a
    property of one draw.  Returns its value and timings."""
    from minhashcuda_amd import workloads
    cm = engine.capi_mod
    dim = 1 << 20
    layout = workloads.batched_layout(count=args.count, seed=2)
    tensors = [workloads.batched_csr_torch(n, dim, 1.5, seed=200 + i)
               for i, n in enumerate(layout)]
    lengths = [int(t[2].numel()) - 1 for t in tensors]
    nnz = sum(int(t[0].numel()) for t in tensors)
    outs = [torch.empty((n, SAMPLES, 2), dtype=torch.int32, device="cuda") for n in lengths]
    descs = (cm.Matrix * len(tensors))()
    for i, ((w, c, r), o) in enumerate(zip(tensors, outs)):
        descs[i] = cm.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), lengths[i], o.data_ptr())
    torch.cuda.synchronize()
    wall, kms = timed_steps(torch, lambda: engine.batch(descs, kernel, stream), engine, steps, 1)
    res = {"layout_seed": 2, "rows": sum(lengths), "nnz": nnz, "steps": steps,
           "value": nnz * SAMPLES * steps / wall, "ms_per_step": wall * 1e3 / steps,
           "avg_kernel_ms": sum(kms) / len(kms), "tuning": engine.tuning()}
    del tensors, outs, descs
    torch.cuda.empty_cache()
    return res


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


# ---------------------------------------------------------------------------
# serving

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


SERVE_DIM, SERVE_SAMPLES, SERVE_POOL = 1 << 22, 256, 256


def serve_engine(args):
    """The serving generator (dim 2^22, 256 samples, seed 1) with the command
    line's tuning, if any."""
    import libMHCUDA
    gen = libMHCUDA.minhash_cuda_init(SERVE_DIM, SERVE_SAMPLES, seed=1, devices=1)
    if explicit_tuning(args):
        from minhashcuda_amd import capi as capi_mod
        capi = capi_mod.load(os.path.join(PKG, "libMHCUDA.so"))
        g = ctypes.c_void_p(gen)
        capi_mod.check(capi.mhcuda_set_tuning(g, args.spl, args.seg, args.wpc), "tuning")
        if args.stripe_mib >= 0:
            capi_mod.check(capi.mhcuda_set_stripes(g, args.stripe_mib), "stripes")
    return gen


def serve_kernel(args, torch, gen, pool, steps=5):
    """The hashing pass of the largest launch the serving loop makes: all
    requests of the pool coalesced (the server's max_batch_requests),
    device-resident, one persistent launch with the generator's tuning.
    Returns its rate, kernel time and roofline."""
    import numpy as np
    from minhashcuda_amd import capi as cm, workloads
    capi = cm.load(os.path.join(PKG, "libMHCUDA.so"))
    g = ctypes.c_void_p(gen)
    cm.check(capi.mhcuda_set_timing(g, 1), "timing")
    tens = []
    for m in pool:
        tens.append((torch.from_numpy(m.data.astype(np.float32)).cuda(),
                     torch.from_numpy(m.indices.astype(np.int32)).cuda(),
                     torch.from_numpy(m.indptr.astype(np.int64).astype(np.int32)).cuda(),
                     m.shape[0]))
    outs = [torch.empty((t[3], SERVE_SAMPLES, 2), dtype=torch.int32, device="cuda")
            for t in tens]
    descs = (cm.Matrix * len(tens))()
    for i, ((w, c, r, n), o) in enumerate(zip(tens, outs)):
        descs[i] = cm.Matrix(w.data_ptr(), c.data_ptr(), r.data_ptr(), n, o.data_ptr())
    stream = torch.cuda.current_stream()

    def run():
        cm.check(capi.mhcuda_calc_batch_device(g, len(tens), descs, cm.KERNEL_PERSISTENT,
                                               ctypes.c_void_p(stream.cuda_stream)), "calc")

    class _Timer:
        def kernel_ms(self):
            ms = ctypes.c_float()
            cm.check(capi.mhcuda_kernel_time(g, ctypes.byref(ms)), "kernel_time")
            return ms.value

    torch.cuda.synchronize()
    wall, kms = timed_steps(torch, run, _Timer(), steps, 2)
    cm.check(capi.mhcuda_set_timing(g, 0), "timing")
    nnz = sum(m.nnz for m in pool)
    rows = sum(m.shape[0] for m in pool)
    avg_ms = sum(kms) / len(kms)
    mib, k = ctypes.c_int(), ctypes.c_uint32()
    capi.mhcuda_get_stripes(g, ctypes.byref(mib), ctypes.byref(k), None)
    alg = workloads.algorithmic_bytes(nnz, rows, SERVE_SAMPLES)
    res = {"requests": len(pool), "rows": rows, "nnz": nnz, "steps": steps,
           "value": nnz * SERVE_SAMPLES * steps / wall, "ms_per_step": wall * 1e3 / steps,
           "avg_kernel_ms": avg_ms,
           "roofline": roofline("serve", "wmh_persistent_kernel", alg, avg_ms, k.value)}
    del tens, outs, descs
    torch.cuda.empty_cache()
    return res


def serve_kernel_bench(args, torch):
    import libMHCUDA
    log("serve: generating request pool")
    pool = serve_requests(SERVE_POOL, SERVE_DIM, seed=5)
    gen = serve_engine(args)
    res = serve_kernel(args, torch, gen, pool, steps=args.steps or 5)
    libMHCUDA.minhash_cuda_fini(gen)
    print(json.dumps(dict({"metric": METRIC, "workload": "serve_kernel",
                           "unit": "nnz*samples/s", "n_gpus": 1, "dtype": "f32",
                           "data": "synthetic"}, **res)), flush=True)


def serve_sweep(args, torch, qps_list, seconds, with_kernel=False):
    """Continuous-batching serving loop under Poisson arrivals at each rate:
    p50/p99 latency (submit -> hashes in host memory) and throughput; with
    with_kernel, also serve_kernel on the same pool and generator."""
    import threading
    import numpy as np
    import libMHCUDA
    from minhashcuda_amd import serving
    dim, samples = SERVE_DIM, SERVE_SAMPLES
    log("serve: generating request pool")
    pool = serve_requests(SERVE_POOL, dim, seed=5)
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
    gen = serve_engine(args)
    kernel = None
    if with_kernel:
        log("serve: hashing kernel on the %d pool requests coalesced" % len(pool))
        kernel = serve_kernel(args, torch, gen, pool)
    server = serving.Server(gen, memory_budget=4 << 30, slots=args.slots,
                            max_batch_requests=SERVE_POOL)
    # warm-up: every buffer reaches its working size
    for t in [server.submit(m) for m in pool[:64]]:
        server.wait(t)
    server.stats(reset=True)
    results = []
    rng = np.random.default_rng(11)
    for qps in qps_list:
        n = max(20, int(qps * seconds))
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
                        "mean_nnz_per_request": nnz / n,
                        "phases_ms": {k: st[k] for k in ("queue_ms", "stage_in_ms",
                                                         "device_ms", "stage_out_ms")}})
        log("qps %.0f: p50 %.2f ms p99 %.2f ms, %.3g nnz*S/s, %.1f req/launch; queue %.2f, "
            "per launch: stage in %.2f, device %.2f, stage out %.2f ms" % (
                qps, st["p50_ms"], st["p99_ms"], st["nnz_samples_per_sec"],
                st["mean_requests_per_batch"], st["queue_ms"], st["stage_in_ms"],
                st["device_ms"], st["stage_out_ms"]))
    server.stop()
    libMHCUDA.minhash_cuda_fini(gen)
    config = ("serving: Poisson arrivals, requests of 1-4096 rows (log-uniform), dim 2^22, "
              "nnz/row lognormal(median 200, sigma 1) clipped to 50000, lognormal(0,1) weights, "
              "256 samples, parameters resident (12.9 GB), %d slots, 4 GiB budget, request "
              "arrays in %s host memory" % (args.slots, "pinned" if args.pinned else "pageable"))
    return results, config, len(pool), kernel


def serve_summary(results):
    """The highest offered rate the loop keeps up with (achieved >= 0.98 x
    offered) and its latencies, plus the saturation throughput of the sweep."""
    stable = [r for r in results if r["achieved_qps"] >= 0.98 * r["offered_qps"]]
    top = max(stable, key=lambda r: r["offered_qps"]) if stable else None
    sat = max(results, key=lambda r: r["nnz_samples_per_sec"])
    return {
        "stable_offered_qps": top["offered_qps"] if top else None,
        "p50_ms": top["p50_ms"] if top else None,
        "p99_ms": top["p99_ms"] if top else None,
        "nnz_samples_per_sec_at_stable": top["nnz_samples_per_sec"] if top else None,
        "saturation_nnz_samples_per_sec": sat["nnz_samples_per_sec"],
        "saturation_offered_qps": sat["offered_qps"],
        "saturation_p99_ms": sat["p99_ms"],
    }


SERVE_QPS = "2000,4000,6000,8000,10000,12000,14000,16000"


def serve_bench(args, torch):
    qps = [float(q) for q in (args.qps or SERVE_QPS).split(",")]
    results, config, pool, kernel = serve_sweep(args, torch, qps, args.serve_seconds or 3.0,
                                                with_kernel=True)
    best = max(results, key=lambda r: r["nnz_samples_per_sec"])
    print(json.dumps({
        "metric": METRIC, "workload": "serve", "unit": "nnz*samples/s",
        "value": best["nnz_samples_per_sec"], "higher_is_better": True, "n_gpus": 1,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": config, "pool_requests": pool},
        "summary": serve_summary(results), "kernel": kernel, "sweep": results}), flush=True)


# ---------------------------------------------------------------------------
# the reference's CPU-runnable case

def c1_bench(args, torch):
    """The reference's CPU-runnable case (test.py:93-97, README.md:85-90):
    6400 x 130 CSR from numpy seed 1 (values in [0,100), 80% zeroed), 128
    samples, parameters assigned through assign_random_vars.  Times the whole
    6400 rows on the CPU (the scalar oracle and the OpenMP port) and on the GPU
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
    threads = cpu_threads()
    pt = oracle.transpose_params(rs, ln_cs, betas)
    t = time.perf_counter()
    h2, used = oracle.calc_fast(pt, m.data, m.indices, m.indptr, threads)
    t2 = time.perf_counter() - t
    best = max(res, key=lambda k: res[k]["value"])
    print(json.dumps({
        "metric": METRIC, "workload": "c1", "unit": "nnz*samples/s", "value": res[best]["value"],
        "steps": steps, "higher_is_better": True, "n_gpus": 1, "dtype": "f32",
        "data": "synthetic (test.py recipe, seed 1)",
        "config": {"workload": "c1: 6400 x 130 CSR (nnz %d), 128 samples, assigned parameters; "
                               "GPU value = best of plain / persistent" % m.nnz},
        "gpu": res,
        "cpu_baseline": dict({"value": m.nnz * SAMPLES / t2, "unit": "nnz*samples/s",
                              "cores": used, "kind": "port",
                              "sample": "all 6400 rows, OpenMP + AVX2 port",
                              "scalar_oracle_value": m.nnz * SAMPLES / t1,
                              "scalar_oracle_sample": "all 6400 rows, %.3f s" % t1},
                             **host_facts()),
        "bit_exact": bool(np.array_equal(gpu_h, h1) and np.array_equal(h1, h2)),
    }), flush=True)
    engine.close()


# ---------------------------------------------------------------------------
# the 1M x 12M share

def large_run(args, torch, steps, warmup, with_cpu):
    """1,000,000 x 12,000,000 CSR, 16,800 nonzeros per row (density 0.0014),
    uniform columns, weights uniform in (0, 100], 128 samples, parameters
    resident (18.4 GB).  Two figures:
      * kernel: the hashing kernel (autotuned on the chunk unless a tuning
        knob is given) on one device-resident 65,536-row chunk (1.1e9
        nonzeros) -- the roofline figure;
      * end to end: all 1,000,000 rows through mhcuda_calc_batch from a pinned
        host copy of that chunk (16 matrices of <= 65,536 rows, the last one
        partial), with double-buffered uploads overlapping the hashing --
        PCIe-inclusive, reported beside it.
    Returns the result dict."""
    import numpy as np
    import scipy.sparse as sp
    from minhashcuda_amd import workloads
    dim, per_row, chunk_rows, total_rows = 12_000_000, 16_800, 65_536, 1_000_000
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
    wall, kms = timed_steps(torch, lambda: engine.batch(desc, kernel, stream), engine, steps,
                            warmup)
    span_ms = sum(kms) / len(kms)
    avg_ms = span_ms
    if kernel == cm.KERNEL_PERSISTENT:
        avg_ms = engine.eager_kernel_ms(lambda: engine.batch(desc, kernel, stream)) or span_ms
    alg = workloads.algorithmic_bytes(nnz, chunk_rows, SAMPLES)
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
    tuning = engine.tuning()
    result = {
        "value": nnz * SAMPLES * steps / wall, "unit": "nnz*samples/s", "steps": steps,
        "warmup": warmup, "ms_per_step": wall * 1e3 / steps,
        "config": dict({"workload": "large: 1,000,000 x 12,000,000 CSR, 16,800 nnz/row, "
                                    "uniform(0,100] weights, 128 samples, parameters resident "
                                    "(18.4 GB); value = the autotuned hashing kernel on one "
                                    "resident 65,536-row chunk (1.1e9 nnz)",
                        "chunk_rows": chunk_rows, "chunk_nnz": nnz}, **tuning),
        "roofline": dict(roofline("large", kname, alg, avg_ms, tuning["stripes"],
                                  tuning["heavy_seeds"]), graph_span_ms=span_ms),
        "end_to_end": {"rows": total_rows, "nnz": e2e_nnz,
                       "seconds": min(ends), "value": e2e_units / min(ends),
                       "unit": "nnz*samples/s (PCIe-inclusive, pinned host input)",
                       "identical_to_resident_chunk": bool(same)},
    }
    if tuned:
        result["autotune"] = tuned
    if with_cpu:
        log("large: cpu baseline (retrieving 18.4 GB of parameters)")
        params = engine.L.minhash_cuda_retrieve_vars(engine.gen)
        cut = 600
        m0 = sp.csr_matrix((hw[:rp[cut]].numpy(), hc[:rp[cut]].numpy().view(np.uint32),
                            rp[:cut + 1]), shape=(cut, dim))
        result["cpu_baseline"] = cpu_baseline(params, m0, SAMPLES, args.cpu_seconds,
                                              dev_h[:cut], "first rows of the chunk")
    engine.close()
    del w, c, r, out, hw, hc, hr, hout
    torch.cuda.empty_cache()
    return result


def large_bench(args, torch):
    res = large_run(args, torch, args.steps or 3, args.warmup if args.warmup is not None else 1,
                    not args.no_cpu_baseline)
    print(json.dumps(dict({"metric": METRIC, "workload": "large", "n_gpus": 1,
                           "higher_is_better": True, "scaling": "weak", "dtype": "f32",
                           "data": "synthetic", "vs_baseline": None}, **res)), flush=True)


# ---------------------------------------------------------------------------
# the uneven batch (default) and the plain configuration

def main(argv=None):
    args = parse(argv)
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
    if args.workload == "serve_kernel":
        serve_kernel_bench(args, torch)
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
    span_ms = sum(kms) / len(kms)
    avg_ms = span_ms
    graphed = engine.graph and kernel == cm.KERNEL_PERSISTENT
    if graphed:
        avg_ms = engine.eager_kernel_ms(lambda: engine.batch(descs, kernel, stream))
    alg = workloads.algorithmic_bytes(nnz, rows, SAMPLES)
    tuning = engine.tuning()
    roof = roofline(args.workload, kname, alg, avg_ms, tuning["stripes"], tuning["heavy_seeds"])
    if graphed:
        roof["timed_span"] = ("HIP events around the %s launch alone (2 launches without "
                              "graph replay after the timed steps)" % kname)
        roof["graph_span_ms"] = span_ms
        roof["graph_span"] = ("hipGraph replay of a timed step: descriptor upload, plan, init, "
                              "heavy seeds when on, %s, fix-up" % kname)
    else:
        roof["timed_span"] = "HIP events around the %s launch(es)" % kname
    result.update({
        "value": units * steps / wall, "steps": steps, "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps,
        "config": dict({"workload": desc, "rows": rows, "nnz": nnz, "samples": SAMPLES,
                        "matrices": len(tensors), "parallelism": "single GPU"}, **tuning),
        "roofline": roof,
    })
    if tuned is not None:
        result["autotune"] = tuned
    if args.workload == "plain":
        # the reference's own entry point on the same matrix: numpy arrays in
        # host memory through libMHCUDA.minhash_cuda_calc (uploads, hashing,
        # downloads), PCIe-inclusive
        import numpy as np
        engine.L.minhash_cuda_calc(engine.gen, m)
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            host_h = engine.L.minhash_cuda_calc(engine.gen, m)
            times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        result["host_path"] = {
            "value": m.nnz * SAMPLES / t, "ms_per_call": t * 1e3, "calls": len(times),
            "unit": "nnz*samples/s (PCIe-inclusive: numpy inputs and output, "
                    "libMHCUDA.minhash_cuda_calc, median call)",
            "identical_to_device_resident": bool(np.array_equal(
                host_h, outs[0].cpu().numpy().view(np.uint32)))}
    if args.workload == "batched" and not args.no_plain:
        # the plain kernel is timed on fewer steps (it is about 5x slower and
        # its launches vary little)
        psteps = min(steps, 5)
        ref_out = [o.clone() for o in outs]
        for o in outs:
            o.fill_(0)
        pwall, pkms = timed_steps(torch, lambda: engine.batch(descs, cm.KERNEL_PLAIN, stream),
                                  engine, psteps, 1)
        same = all(torch.equal(a, b) for a, b in zip(ref_out, outs))
        result["plain"] = {"value": units * psteps / pwall, "steps": psteps,
                           "ms_per_step": pwall * 1e3 / psteps,
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
            cut = int(np.searchsorted(rp, 120_000_000))
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
    del tensors, outs, descs
    torch.cuda.empty_cache()
    if args.workload == "batched" and not args.no_seed2:
        log("second layout seed")
        result["seed2"] = second_seed(args, torch, engine, kernel, stream, min(steps, 3))
    engine.close()
    torch.cuda.empty_cache()
    if args.workload == "batched" and not args.no_large:
        log("large share")
        result["large"] = large_run(args, torch, 3, 1, False)
    if args.workload == "batched" and not args.no_serve:
        log("serving sweep")
        qps = [float(q) for q in (args.qps or SERVE_QPS).split(",")]
        sweep, config, pool, kernel = serve_sweep(args, torch, qps, args.serve_seconds or 2.0,
                                                  with_kernel=True)
        result["serve"] = dict(serve_summary(sweep), config=config, pool_requests=pool,
                               kernel=kernel, sweep=sweep)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
