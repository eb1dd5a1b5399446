#!/usr/bin/env python3
"""Benchmark of the Weighted MinHash hashing path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain]

A step is one hashing pass over one batch of synthetic CSR input that is
already resident in HBM (device entry point of the C ABI, launched on torch's
current stream).  Prints ONE JSON line with the metric nnz*samples/s, the
roofline of the dominant kernel (HBM-bound parameter gathers) and the CPU
oracle baseline timed on this host.

Workloads:
  plain    100,000 x 2^20 CSR, nnz/row ~ Poisson(100), lognormal(0,1) weights,
           128 samples, plain kernel (one launch, one wavefront per row).
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="plain", choices=["plain"])
    p.add_argument("--spl", type=int, default=0, help="samples per lane (0: default)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="target duration of the single-thread CPU oracle sample")
    return p.parse_args()


def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def cpu_baseline(params, m, samples, target_s, gpu_hashes):
    """Times the C oracle (this repository's CPU restatement of the hashing
    path) on a bounded row sample: single thread, and OpenMP on every core
    this process may use.  Also checks the sample's hashes against the GPU's."""
    import numpy as np
    import oracle
    rs, ln_cs, betas = params
    # calibrate on a small prefix, then size the single-thread sample
    n0 = min(m.shape[0], 500)
    sub = m[:n0]
    t = time.perf_counter()
    oracle.calc_csr(rs, ln_cs, betas, sub)
    rate = max(sub.nnz * samples / max(time.perf_counter() - t, 1e-6), 1.0)
    n1 = int(min(m.shape[0], max(n0, n0 * target_s * rate / max(sub.nnz * samples, 1))))
    sub = m[:n1]
    t = time.perf_counter()
    h1 = oracle.calc_csr(rs, ln_cs, betas, sub)
    t1 = time.perf_counter() - t
    threads = min(cpu_threads(), 16)
    n2 = int(min(m.shape[0], n1 * threads))
    sub2 = m[:n2]
    t = time.perf_counter()
    h2, used = oracle.calc_omp(rs, ln_cs, betas, sub2.data, sub2.indices, sub2.indptr, threads)
    t2 = time.perf_counter() - t
    ok = bool(np.array_equal(h2, gpu_hashes[:n2]) and np.array_equal(h1, gpu_hashes[:n1]))
    return {
        "value": sub2.nnz * samples / t2, "unit": "nnz*samples/s", "cores": int(used),
        "kind": "port",
        "sample": "first %d of %d rows (%d nnz), OpenMP C oracle, %.1f s" % (n2, m.shape[0], sub2.nnz, t2),
        "single_thread_value": sub.nnz * samples / t1,
        "single_thread_sample": "first %d rows (%d nnz), %.1f s" % (n1, sub.nnz, t1),
        "sample_bit_exact_vs_gpu": ok,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine hashes on a single MI355X "
              "(multi-GPU row sharding is described in DESIGN.md but not built); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch  # before the engine: one HIP runtime per process
    sys.path.insert(0, ROOT)
    sys.path.insert(0, PKG)
    import libMHCUDA
    from minhashcuda_amd import capi as capi_mod
    from minhashcuda_amd import workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    capi = capi_mod.load(os.path.join(PKG, "libMHCUDA.so"))
    samples = 128
    dim = 1 << 20
    log("generating workload", args.workload)
    m = workloads.plain_csr(100000, dim, 100.0, 1.0, seed=1)
    rows, nnz = m.shape[0], m.nnz
    gen = libMHCUDA.minhash_cuda_init(dim, samples, seed=1, devices=1)
    if args.spl:
        capi_mod.check(capi.mhcuda_set_tuning(ctypes.c_void_p(gen), args.spl, 0, 0), "set_tuning")
    spl = ctypes.c_int()
    capi.mhcuda_get_tuning(ctypes.c_void_p(gen), ctypes.byref(spl), None, None)
    w = torch.from_numpy(m.data.astype(np.float32)).cuda()
    c = torch.from_numpy(m.indices.astype(np.int32)).cuda()
    r = torch.from_numpy(m.indptr.astype(np.int64).astype(np.int32)).cuda()
    out = torch.empty((rows, samples, 2), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    sptr = ctypes.c_void_p(stream.cuda_stream)

    def step():
        capi_mod.check(capi.mhcuda_calc_device(ctypes.c_void_p(gen), w.data_ptr(), c.data_ptr(),
                                               r.data_ptr(), rows, out.data_ptr(), sptr),
                       "mhcuda_calc_device")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    units = nnz * samples
    value = units * args.steps / wall
    alg = workloads.algorithmic_bytes(nnz, rows, samples)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    gpu_hashes = out.cpu().numpy().view(np.uint32)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "nnz*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain: %d x %d CSR, nnz/row~Poisson(100), lognormal(0,1) weights, "
                        "%d samples, one launch (plain kernel)" % (rows, dim, samples),
            "rows": rows, "dim": dim, "nnz": nnz, "samples": samples,
            "samples_per_lane": spl.value, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "kernel": "wmh_plain_kernel", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": alg,
            "avg_kernel_ms": avg_ms,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline")
        params = libMHCUDA.minhash_cuda_retrieve_vars(gen)
        result["cpu_baseline"] = cpu_baseline(params, m, samples, args.cpu_seconds, gpu_hashes)
    libMHCUDA.minhash_cuda_fini(gen)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
