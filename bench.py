"""Throughput of the convolution hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

One step is one convolution call over one batch of synthetic input already
resident in HBM.  The default workload is the uneven batch: 4096 rank-1 f32
signal/kernel pairs (signal lengths log-uniform in [256, 262144], kernel
lengths log-uniform in [3, 1023], Uniform[-1, 1), ConvMode::Same,
PaddingMode::Circular) in one persistent launch (a HIP-graph replay of the
same step is timed beside it), the autotuner choosing BF16x3 / BF16x6 matrix-core direct, FP32 direct or
FP32 FFT per pair.  Other workloads (--workload): fft2d, fft2d_large, fft1d,
fft3d, fft3d_c64, direct2d_i32, direct2d_f32 (with --algo bf16 / bf16x3), proc2d (the FFT
processor's forward + backward round trip of a 256x5120 f32 array, --proc-shape for others), copy.  The conv_fft
workloads take --baked (the kernel spectrum made once before timing).

Prints ONE JSON line with the throughput, the error against the FP64 oracle
(normwise relative and max abs; bit-exactness for integers), the roofline of
the dominant kernel (its algorithmic bytes or flops per launch over its
average duration, timed with the HIP events its launches carry in a second
pass over the same steps) and a CPU baseline: the oracle's float32
restatement of the crate's algorithm (scipy pocketfft) on this host's cores,
on a bounded sample, with its single-thread figure (the serial conv_fft
stand-in) beside it.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ndarray-conv_amd"))

METRIC = "conv outputs/s + % HBM/MFMA roofline; BF16x3 vs FP32 at ≤1e-5 rel err"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3


def workloads():
    import ndconv as nc
    B = nc.BorderType
    return {
        "fft2d": dict(
            name="2-D f32 conv_fft x200x5000 k11x31 dilation 2 Same Custom[Reflect,Circular]",
            seed=2, op="fft", dtype=np.float32, x=(200, 5000), k=(11, 31), dil=2, mode=nc.ConvMode.Same,
            pad=nc.PaddingMode.Custom([B.Reflect, B.Circular]), dist="uniform",
            cpu=("same", ("custom", ["reflect", "circular"]))),
        "fft2d_large": dict(
            name="2-D f32 conv_fft x4096x32768 k11x31 dilation 2 Same Custom[Reflect,Circular]",
            seed=2, op="fft", dtype=np.float32, x=(4096, 32768), k=(11, 31), dil=2, mode=nc.ConvMode.Same,
            pad=nc.PaddingMode.Custom([B.Reflect, B.Circular]), dist="uniform",
            cpu=("same", ("custom", ["reflect", "circular"])), cpu_budget=20.0),
        "fft1d": dict(
            name="1-D f32 conv_fft x5000 k31 Same Zeros", seed=1, op="fft", dtype=np.float32, x=(5000,), k=(31,),
            dil=1, mode=nc.ConvMode.Same, pad=nc.PaddingMode.Zeros, dist="uniform", cpu=("same", "zeros")),
        "fft3d": dict(
            name="3-D f32 conv_fft x10x100x200 k5x11x31 Full Zeros", seed=3, op="fft", dtype=np.float32,
            x=(10, 100, 200), k=(5, 11, 31), dil=1, mode=nc.ConvMode.Full, pad=nc.PaddingMode.Zeros,
            dist="uniform", cpu=("full", "zeros")),
        "fft3d_c64": dict(
            name="3-D c64 conv_fft x10x100x200 k5x11x31 Full Zeros", seed=3, op="fft", dtype=np.complex64,
            x=(10, 100, 200), k=(5, 11, 31), dil=1, mode=nc.ConvMode.Full, pad=nc.PaddingMode.Zeros,
            dist="normal", cpu=("full", "zeros")),
        "direct2d_i32": dict(
            name="2-D i32 conv x200x5000 k11x31 Same Custom[Reflect,Circular]", seed=2, op="conv",
            dtype=np.int32, x=(200, 5000), k=(11, 31), dil=1, mode=nc.ConvMode.Same,
            pad=nc.PaddingMode.Custom([B.Reflect, B.Circular]), dist="int8", cpu=None,
            oracle=("same", ("custom", ["reflect", "circular"]))),
        "direct2d_f32": dict(
            name="2-D f32 conv x4096x4096 k31x31 dilation 2 stride 2 Valid Replicate", seed=4, op="conv",
            dtype=np.float32, x=(4096, 4096), k=(31, 31), dil=2, mode=nc.ConvMode.Custom([0, 0], [2, 2]),
            pad=nc.PaddingMode.Replicate, dist="normal", cpu=None,
            oracle=(("custom", [0, 0], [2, 2]), "replicate")),
    }


def uneven_lengths(seed=5, count=4096):
    """Signal lengths log-uniform in [256, 262144] and kernel lengths
    log-uniform in [3, 1023] (floor of exp of a uniform draw, clipped)."""
    rng = np.random.default_rng(seed)
    n = np.clip(np.floor(np.exp(rng.uniform(np.log(256), np.log(262145), count))), 256, 262144).astype(np.int64)
    k = np.clip(np.floor(np.exp(rng.uniform(np.log(3), np.log(1024), count))), 3, 1023).astype(np.int64)
    return n, k, rng


def uneven_batch(seed=5, count=4096):
    """The uneven batch: lengths as above, values Uniform[-1, 1), packed."""
    n, k, rng = uneven_lengths(seed, count)
    x = rng.uniform(-1.0, 1.0, int(n.sum())).astype(np.float32)
    kv = rng.uniform(-1.0, 1.0, int(k.sum())).astype(np.float32)
    return n, k, x, kv


def make_inputs(w):
    rng = np.random.default_rng(w["seed"])
    dt = np.dtype(w["dtype"])
    if w["dist"] == "uniform":
        return rng.random(w["x"]).astype(dt), rng.random(w["k"]).astype(dt)
    if w["dist"] == "int8":
        return rng.integers(-8, 9, w["x"]).astype(dt), rng.integers(-8, 9, w["k"]).astype(dt)
    if dt.kind == "c":
        x = rng.standard_normal(w["x"]) + 1j * rng.standard_normal(w["x"])
        k = rng.standard_normal(w["k"]) + 1j * rng.standard_normal(w["k"])
        return x.astype(dt), k.astype(dt)
    return rng.standard_normal(w["x"]).astype(dt), rng.standard_normal(w["k"]).astype(dt)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def note(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` in `workload`, measured with rocprofv3
    --pmc FETCH_SIZE and WRITE_SIZE and scaled by what the same counters read
    for a plain device copy of known size (the `copy` workload); None when
    that measurement has not been recorded."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["kernels"][workload][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def run_copy(args):
    """A plain device-to-device copy of 512 MiB per step: the reference
    kernel for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 128 << 20
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print(json.dumps({"workload": "copy", "bytes_read_per_step": 4 * n, "bytes_written_per_step": 4 * n,
                      "ms_per_step": ms, "GB/s": 8 * n / (ms * 1e-3) / 1e9}))


def cpu_baseline_one(w, x, k, threads, ref, budget_s):
    """The oracle's f32 restatement of conv_fft (scipy pocketfft in single
    precision, a fresh plan per call as conv_fft makes) on `threads` workers:
    1 is the serial conv_fft stand-in, more the conv_fft_par stand-in."""
    from oracle import ndconv_oracle as O
    mode, pm = w["cpu"]
    f = lambda: O.conv_fft(x, k, mode, pm, w["dil"], True, precision="f32", workers=threads)
    y = f()
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end and len(times) < 2000:
        t0 = time.perf_counter()
        f()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    model, nproc, _ = host_cpu()
    out = {"value": int(y.size) / med, "unit": "outputs/s", "cores": threads, "kind": "port",
           "sample": f"{len(times)} full calls of the same workload (median {med * 1e3:.3f} ms); the oracle's "
                     f"float32 restatement of conv_fft with scipy pocketfft, workers={threads}, on {model}, "
                     f"nproc {nproc}"}
    if ref is not None:
        out["max_rel_err"], out["max_abs_err"] = O.relerr(y, ref)
    return out


def cpu_baseline(w, x, k, ref):
    """All-core (the box's one-GPU share) and single-thread CPU baselines."""
    if w["cpu"] is None:
        return None
    _, _, threads = host_cpu()
    budget = w.get("cpu_budget", 10.0)
    out = cpu_baseline_one(w, x, k, threads, ref, budget)
    out["single_thread"] = cpu_baseline_one(w, x, k, 1, ref, budget)
    return out


def oracle_result(w, x, k):
    """The FP64 oracle's result (floats) or the exact wrapping result
    (integers) of a single-problem workload."""
    from oracle import ndconv_oracle as O
    mode, pm = w["cpu"] if w["cpu"] is not None else w["oracle"]
    if np.dtype(w["dtype"]).kind in "iu":
        return O.conv(x, k, mode, pm, w["dil"], True)
    wide = np.complex128 if np.dtype(w["dtype"]).kind == "c" else np.float64
    return O.conv_fft(x.astype(wide), k.astype(wide), mode, pm, w["dil"], True)


def host_cpu():
    """CPU model, nproc, and the threads a CPU baseline may use here: this
    process's affinity, at most 16 (a one-GPU share of the host)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, os.cpu_count() or 1, max(1, min(16, avail))


def pair_offsets(*lens):
    return [np.concatenate([[0], np.cumsum(a)]).astype(np.int64) for a in lens]


def batch_oracle_errors(n, k, x, kv, ys, y_lens, threads):
    """Per-pair normwise relative and max abs error of each packed output in
    `ys` against the FP64 oracle's conv_fft of that pair, over every pair."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import ndconv_oracle as O
    xo, ko, yo = pair_offsets(n, k, y_lens)

    def one(i):
        ref = O.conv_fft(x[xo[i]:xo[i + 1]].astype(np.float64), kv[ko[i]:ko[i + 1]].astype(np.float64), "same",
                         "circular")
        return [O.relerr(y[yo[i]:yo[i + 1]], ref) for y in ys]

    with ThreadPoolExecutor(threads) as ex:
        per = list(ex.map(one, range(len(n))))
    out = []
    for j in range(len(ys)):
        rel = np.array([p[j][0] for p in per])
        ab = np.array([p[j][1] for p in per])
        out.append({"max_rel_err": float(rel.max()), "max_abs_err": float(ab.max()),
                    "median_rel_err": float(np.median(rel)), "worst_pair": int(rel.argmax())})
    return out


def batch_cpu_baseline(n, k, x, kv, threads, budget_s=10.0):
    """The oracle's f32 restatement of conv_fft per pair (scipy pocketfft in
    single precision, a fresh transform plan per call as conv_fft makes), on a
    bounded random sample of the batch: `threads` = 1 is the serial conv_fft
    stand-in, more threads (pairs in flight) the conv_fft_par stand-in.  Also
    the restatement's own error against the FP64 oracle on the sampled pairs."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import ndconv_oracle as O
    xo, ko = pair_offsets(n, k)
    order = np.random.default_rng(0).permutation(len(n))
    got = {}

    def one(i):
        got[int(i)] = O.conv_fft(x[xo[i]:xo[i + 1]], kv[ko[i]:ko[i + 1]], "same", "circular", precision="f32")
        return got[int(i)].size

    done, outs = 0, 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < budget_s and done < len(order):
            chunk = order[done:done + 4 * threads]
            outs += sum(ex.map(one, chunk))
            done += len(chunk)
    el = time.perf_counter() - t0
    rel, ab = 0.0, 0.0
    for i, y in got.items():
        ref = O.conv_fft(x[xo[i]:xo[i + 1]].astype(np.float64), kv[ko[i]:ko[i + 1]].astype(np.float64), "same",
                         "circular")
        r, a = O.relerr(y, ref)
        rel, ab = max(rel, r), max(ab, a)
    model, nproc, _ = host_cpu()
    return {"value": outs / el, "unit": "outputs/s", "cores": threads, "kind": "port",
            "sample": f"{done} of {len(n)} pairs drawn at random ({outs} outputs, {el:.1f} s); the oracle's float32 "
                      f"restatement of conv_fft per pair (scipy pocketfft), {threads} pair(s) in flight on "
                      f"{model}, nproc {nproc}",
            "max_rel_err": rel, "max_abs_err": ab}


def run_proc(args):
    """FftProcessor (processor/mod.rs:79-121, real.rs:105-155,233-281): one step is
    forward then backward of an f32 array on the device (--proc-shape, default
    256x5120, the processor shape of the 2-D configuration)."""
    import torch
    import ndconv as nc
    from ndconv import _lib as C
    from oracle import ndconv_oracle as O

    shape = tuple(int(v) for v in args.proc_shape.split("x"))
    if len(shape) != 2 or min(shape) < 1:
        sys.exit("--proc-shape takes ROWSxCOLS")
    spec_shape = (shape[1] // 2 + 1, shape[0])
    x = np.random.default_rng(7).standard_normal(shape).astype(np.float32)
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(x).to(dev)
    st = torch.empty(spec_shape, dtype=torch.complex64, device=dev)
    bt = torch.empty(shape, dtype=torch.float32, device=dev)
    lib = C.lib()
    proc = nc.get_fft_processor(np.float32)
    h = proc.handle
    stream = torch.cuda.current_stream(dev)
    nc._check(lib.ndc_set_stream(h, ctypes.c_void_p(stream.cuda_stream)), h)
    sh = (ctypes.c_int64 * 2)(*shape)
    ssh = (ctypes.c_int64 * 2)(*spec_shape)
    px, ps, pb = ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(bt.data_ptr())

    def step():
        nc._check(lib.ndc_fft_forward(h, C.F32, 2, sh, px, ps, 0), h)
        nc._check(lib.ndc_fft_backward(h, C.F32, 2, ssh, ps, pb, 0), h)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    outputs = int(np.prod(shape))
    nc._check(lib.ndc_profile_enable(h, 1), h)
    for _ in range(args.steps):
        step()
    buf = ctypes.create_string_buffer(1 << 16)
    nc._check(lib.ndc_profile_report(h, buf, len(buf)), h)
    nc._check(lib.ndc_profile_enable(h, 0), h)
    prof = json.loads(buf.value.decode())
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bpl = dom["bytes"] / dom["launches"]
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "kernel": dom_name, "avg_ms": avg_ms, "bytes_per_launch": bpl}
    kernels = {n: {"avg_ms": v["total_ms"] / v["launches"], "launches_per_step": v["launches"] / args.steps,
                   "GB/s": (v["bytes"] / v["launches"]) / (v["total_ms"] / v["launches"] * 1e-3) / 1e9}
               for n, v in prof.items()}
    note("checking against the oracle")
    ref = O.processor_forward(x.astype(np.float64))
    fwd_err = O.relerr(st.cpu().numpy(), ref)
    rt_err = O.relerr(bt.cpu().numpy(), x)
    cpu = None
    if not args.no_cpu:
        note("CPU baseline")
        reps, t = 0, time.perf_counter()
        while time.perf_counter() - t < 10.0 and reps < 200:
            O.processor_backward(O.processor_forward(x), shape[1])
            reps += 1
        cpu_s = (time.perf_counter() - t) / reps
        cpu = {"value": outputs / cpu_s, "unit": "outputs/s", "cores": 1, "kind": "port",
               "sample": f"{reps} round trips of the whole array: the oracle's numpy restatement (FP64 pocketfft), one thread"}
    line = {"metric": METRIC, "value": outputs / (ms * 1e-3), "unit": "outputs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"FftProcessor forward + backward, {shape[0]}x{shape[1]} f32 "
                                   f"(spectrum {spec_shape[0]}x{spec_shape[1]} c64)",
                       "outputs_per_step": outputs, "seed": 7},
            "forward_max_rel_err": fwd_err[0], "roundtrip_max_rel_err": rt_err[0], "roofline": roof,
            "cpu_baseline": cpu, "kernels": kernels, "wall_s": wall}
    print(json.dumps(line))


def run_batch(args):
    """The uneven batch: 4096 rank-1 f32 pairs, one persistent launch per
    step (launched directly; a HIP-graph replay of the same step is timed
    beside it and reported as graph_value)."""
    import torch
    import ndconv as nc
    from ndconv import _lib as C

    n, k, x, kv = uneven_batch(5, 4096)
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(x).to(dev)
    kt = torch.from_numpy(kv).to(dev)
    proc = nc.get_fft_processor()
    h = proc.handle
    lib = C.lib()
    t0 = time.perf_counter()
    bp = nc.BatchPlan(n, k, nc.ConvMode.Same, nc.PaddingMode.Circular, algo=args.batch_algo, processor=proc)
    plan_s = time.perf_counter() - t0
    yt = torch.empty(bp.y_total, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def timed(plan, graph, steps, warmup):
        for _ in range(warmup):
            plan.run(xt, kt, yt, graph=graph)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            plan.run(xt, kt, yt, graph=graph)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps, time.perf_counter() - t

    note(f"batch plan ready ({plan_s:.1f} s); timing {args.steps} direct and graph-replayed steps")
    ms_step, wall = timed(bp, False, args.steps, args.warmup)
    ms_graph, _ = timed(bp, True, args.steps, args.warmup)
    outputs = int(bp.y_total)
    y_head = yt.cpu().numpy()

    # dominant kernel: HIP events carried by each launch (its dispatch's start/end)
    nc._check(lib.ndc_profile_enable(h, 1), h)
    for _ in range(args.steps):
        bp.run(xt, kt, yt)
    buf = ctypes.create_string_buffer(1 << 16)
    nc._check(lib.ndc_profile_report(h, buf, len(buf)), h)
    nc._check(lib.ndc_profile_enable(h, 0), h)
    prof = json.loads(buf.value.decode())
    dom_name, dom = max(prof.items(), key=lambda kv_: kv_[1]["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bpl = dom["bytes"] / dom["launches"]
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("batch1d", dom_name), "kernel": dom_name,
            "avg_ms": avg_ms, "bytes_per_launch": bpl}
    kernels = {nm: {"avg_ms": v["total_ms"] / v["launches"], "launches_per_step": v["launches"] / args.steps,
                    "GB/s": (v["bytes"] / v["launches"]) / (v["total_ms"] / v["launches"] * 1e-3) / 1e9}
               for nm, v in prof.items()}
    report = bp.report()

    # the FP32-only autotuned form (direct and FFT items only) on the same batch
    ys, names = [y_head], [args.batch_algo]
    fp32 = None
    if args.batch_algo != "auto":
        note("timing the FP32-only form")
        bf = nc.BatchPlan(n, k, nc.ConvMode.Same, nc.PaddingMode.Circular, algo="auto", processor=proc)
        fp32, _ = timed(bf, False, args.steps, args.warmup)
        ys.append(yt.cpu().numpy())
        names.append("auto")
        bf.close()

    # the unbatched path: one single-problem conv_fft launch sequence per pair
    plain = None
    if args.plain_steps > 0:
        note(f"timing {args.plain_steps} steps of the per-pair path")
        bq = nc.BatchPlan(n, k, nc.ConvMode.Same, nc.PaddingMode.Circular, algo="plain", processor=proc)
        plain, _ = timed(bq, False, args.plain_steps, 1)
        bq.close()

    _, _, threads = host_cpu()
    note(f"checking every pair against the FP64 oracle ({threads} threads)")
    errs = dict(zip(names, batch_oracle_errors(n, k, x, kv, ys, bp.y_lens, threads)))
    cpu = None
    if not args.no_cpu:
        note("CPU baseline, all threads then one")
        cpu = batch_cpu_baseline(n, k, x, kv, threads)
        cpu["single_thread"] = batch_cpu_baseline(n, k, x, kv, 1)
    value = outputs / (ms_step * 1e-3)
    line = {"metric": METRIC, "value": value, "unit": "outputs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "uneven batch: 4096 rank-1 f32 pairs, n log-uniform [256, 262144], "
                                   "k log-uniform [3, 1023], Uniform[-1,1), Same, Circular",
                       "outputs_per_step": outputs, "seed": 5, "algo": args.batch_algo, "graph_replay": False},
            "max_rel_err": errs[args.batch_algo]["max_rel_err"], "max_abs_err": errs[args.batch_algo]["max_abs_err"],
            "error_vs_fp64": errs,
            "roofline": roof, "cpu_baseline": cpu, "kernels": kernels, "wall_s": wall,
            "graph_value": outputs / (ms_graph * 1e-3), "ms_per_step_graph": ms_graph, "plan_s": plan_s,
            "schedule": {kk: report[kk] for kk in ("items", "lines", "kernel_spectra", "grid", "blocks_per_cu",
                                                   "costs")}}
    if fp32 is not None:
        line["fp32_auto_ms_per_step"] = fp32
        line["fp32_auto_value"] = outputs / (fp32 * 1e-3)
        line["speedup_vs_fp32_auto"] = fp32 / ms_step
    if plain is not None:
        line["plain_ms_per_step"] = plain
        line["plain_value"] = outputs / (plain * 1e-3)
        line["speedup_vs_plain"] = plain / ms_step
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="batch1d")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="also time hipGraph replay of the step")
    ap.add_argument("--baked", action="store_true",
                    help="conv_fft workloads: the kernel's spectrum made once before timing (ndc_conv_fft_bake), "
                         "each step ndc_conv_fft_with_baked")
    ap.add_argument("--algo", default="auto", help="direct workloads: auto (exact FP32/integer), bf16, bf16x3")
    ap.add_argument("--batch-algo", default="auto_bf16x3",
                    help="uneven batch: auto_bf16x3 (autotuned matrix-core/FP32 mix), auto (FP32 only), fft, direct, "
                         "bf16x3, bf16x6")
    ap.add_argument("--plain-steps", type=int, default=2, help="uneven batch: steps of the unbatched path")
    ap.add_argument("--proc-shape", default="256x5120", help="proc2d: the f32 array's ROWSxCOLS")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only (the library has no multi-GPU path); run with --gpus 1"}))
        sys.exit(2)
    if args.workload == "batch1d":
        run_batch(args)
        return
    if args.workload == "copy":
        run_copy(args)
        return
    if args.workload == "proc2d":
        run_proc(args)
        return

    import torch
    import ndconv as nc
    from ndconv import _lib as C

    w = workloads()[args.workload]
    x, k = make_inputs(w)
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(x).to(dev)
    kt = torch.from_numpy(k).to(dev)
    lib = C.lib()
    proc = nc.get_fft_processor()
    h = proc.handle
    d = C.ConvDesc()
    kwd = nc.with_dilation(kt, w["dil"])
    nc._fill_desc(d, xt, kwd, w["mode"], w["pad"], nc._NP_DTYPE[np.dtype(w["dtype"])], host=False)
    d.algo = nc._ALGOS[args.algo]
    op = C.OP_CONV_FFT if w["op"] == "fft" else C.OP_CONV
    info = C.PlanInfo()
    nc._check(lib.ndc_plan(ctypes.byref(d), op, ctypes.byref(info)))
    out_shape = tuple(int(info.out_shape[i]) for i in range(len(w["x"])))
    yt = torch.empty(out_shape, dtype=xt.dtype, device=dev)
    fn = lib.ndc_conv_fft if op == C.OP_CONV_FFT else lib.ndc_conv
    stream = torch.cuda.current_stream(dev)
    nc._check(lib.ndc_set_stream(h, ctypes.c_void_p(stream.cuda_stream)), h)
    px, pk, py = ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(kt.data_ptr()), ctypes.c_void_p(yt.data_ptr())
    dref = ctypes.byref(d)
    baked = None
    if args.baked:
        if op != C.OP_CONV_FFT:
            sys.exit("--baked applies to conv_fft workloads")
        baked = ctypes.c_void_p()
        nc._check(lib.ndc_conv_fft_bake(h, dref, pk, ctypes.byref(baked)), h)

    def step():
        st = lib.ndc_conv_fft_with_baked(baked, px, py) if baked is not None else fn(h, dref, px, pk, py)
        if st:
            nc._check(st, h)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    outputs = int(np.prod(out_shape))
    value = outputs / (ms * 1e-3)

    # per-kernel durations: HIP events carried by every launch (dispatch start/end)
    nc._check(lib.ndc_profile_enable(h, 1), h)
    for _ in range(args.steps):
        step()
    buf = ctypes.create_string_buffer(1 << 16)
    nc._check(lib.ndc_profile_report(h, buf, len(buf)), h)
    nc._check(lib.ndc_profile_enable(h, 0), h)
    prof = json.loads(buf.value.decode())
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    flops_per_launch = dom["flops"] / dom["launches"]
    if w["op"] == "fft":
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS}
    elif args.algo in ("bf16", "bf16x3"):
        # FP32-equivalent algorithmic flops (2 per output and tap) against the
        # dense BF16 matrix-core peak; BF16x3 issues three products per tap
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / BF16_PEAK_TFLOPS, "products_per_tap": 3 if args.algo == "bf16x3" else 1}
    else:
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS}
    roof["traffic"] = measured_traffic(args.workload, dom_name)
    roof["kernel"] = dom_name
    roof["avg_ms"] = avg_ms
    roof["bytes_per_launch"] = bytes_per_launch
    kernels = {n: {"avg_ms": v["total_ms"] / v["launches"], "launches_per_step": v["launches"] / args.steps,
                   "GB/s": (v["bytes"] / v["launches"]) / (v["total_ms"] / v["launches"] * 1e-3) / 1e9}
               for n, v in prof.items()}

    graph_value = None
    if args.graph:
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(dev)
        s.wait_stream(stream)
        with torch.cuda.stream(s):
            nc._check(lib.ndc_set_stream(h, ctypes.c_void_p(s.cuda_stream)), h)
            step()
            torch.cuda.synchronize()
            with torch.cuda.graph(g, stream=s):
                step()
        torch.cuda.synchronize()
        for _ in range(3):
            g.replay()
        torch.cuda.synchronize()
        g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        g0.record()
        for _ in range(args.steps):
            g.replay()
        g1.record()
        torch.cuda.synchronize()
        graph_value = outputs / (g0.elapsed_time(g1) / args.steps * 1e-3)
        nc._check(lib.ndc_set_stream(h, ctypes.c_void_p(stream.cuda_stream)), h)

    y_host = yt.cpu().numpy()
    if baked is not None:
        nc._check(lib.ndc_baked_destroy(baked), h)
    note("checking against the oracle")
    ref = oracle_result(w, x, k)
    if np.dtype(w["dtype"]).kind in "iu":
        acc = {"bit_exact": bool(np.array_equal(y_host, ref))}
    else:
        from oracle import ndconv_oracle as O
        r, a = O.relerr(y_host, ref)
        acc = {"max_rel_err": r, "max_abs_err": a}
    if not args.no_cpu:
        note("CPU baseline")
    cpu = None if args.no_cpu else cpu_baseline(w, x, k, ref if np.dtype(w["dtype"]).kind not in "iu" else None)
    line = {"metric": METRIC, "value": value, "unit": "outputs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": {np.dtype(np.float32): "f32", np.dtype(np.complex64): "c64",
                                           np.dtype(np.int32): "i32"}[np.dtype(w["dtype"])],
            "data": "synthetic", "config": {"workload": w["name"] + (", kernel spectrum baked" if baked is not None else ""),
                                            "outputs_per_step": outputs, "seed": w["seed"], "algo": args.algo},
            **acc, "roofline": roof, "cpu_baseline": cpu, "kernels": kernels, "wall_s": wall}
    if graph_value is not None:
        line["graph_replay_value"] = graph_value
    print(json.dumps(line))


if __name__ == "__main__":
    main()
