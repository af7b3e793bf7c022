#!/usr/bin/env python3
"""Time-to-1e-10 of Jacobi-PCG on one MI355X: FP64 vs mixed-precision refined.

Workload: 3D Laplace stiffness of Q2 hexahedra on a 64^3 unit-cube mesh,
Dirichlet rows eliminated (2,048,383 DOFs, 127,263,527 nonzeros, int32
columns), b ~ U(-1,1) from SplitMix64 seed 1, x0 = 0, stop at relative
residual 1e-10. Inputs are generated on the host and resident in HBM before
timing starts.

One step = one complete solve to 1e-10 with the mixed-precision refined PCG
(the headline value, seconds per solve). The FP64 fused PCG is timed the same
way and reported beside it, with the standalone SpMV bandwidths, the roofline
of the dominant kernel (the FP32 SpMV q = A p inside the mixed solve, timed
with HIP events on the solver stream over the timed steps) and the threaded
CPU restatement of the reference CG on a bounded sample.

Prints one JSON line. --workload assembly | reorder | pmg measures the widened
paths instead (device assembly of the benchmark stiffness, its RCM reordering,
poisson2D's p-multigrid), each with its own roofline and CPU baseline.
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


def spmv_bytes(rows, nnz, value_bytes, index_bytes=4):
    """Algorithmic bytes of one SpMV y = A x over int32 CSR: entries + columns, row extents, x read once,
    y written once (index_bytes 4). With index_bytes 2: the same over the narrow 16-bit column plane."""
    return (value_bytes + index_bytes) * nnz + 4 * (rows + 1) + 2 * value_bytes * rows


def index_bytes(matrix, precision):
    """Bytes per column index the selected SpMV launch reads."""
    kernel, width, _ = matrix.launch(precision)
    if kernel == "tiled":
        return matrix.column_bytes()
    if kernel == "blocks":
        return matrix.column_blocks(width)[1]
    return 4  # int32 columns, or the 32-bit key per sorted or paired block


def format_bytes(matrix, precision):
    """Bytes of the format the selected SpMV launch reads: per entry (CSR) or per column block
    (w values + one column each; a sorted block's column is its 32-bit key), plus row extents, x once and
    y once (tile descriptors, 8-12 B per tile of thousands of blocks, are not counted for any kernel)."""
    value = 8 if precision == 64 else 4
    kernel, width, _ = matrix.launch(precision)
    if kernel == "paired":  # wide and narrow blocks, each with its 32-bit key
        wide, narrow = matrix.paired_blocks(width)
        return ((width // 10) * value + 4) * wide + ((width % 10) * value + 4) * narrow + 4 * (matrix.rows + 1) \
            + 2 * value * matrix.rows
    if kernel in ("blocks", "sorted"):
        blocks, column = matrix.column_blocks(width)
        column = 4 if kernel == "sorted" else column
        return (width * value + column) * blocks + 4 * (matrix.rows + 1) + 2 * value * matrix.rows
    return spmv_bytes(matrix.rows, matrix.nnz, value, index_bytes(matrix, precision))


def pin_launches(matrix, spec):
    """Sets the launches of --launch: KERNEL,WIDTH,UNROLL for both precisions, or FP64/FP32."""
    specs = spec.split("/")
    for precision, one in zip((64, 32), specs if len(specs) == 2 else specs * 2):
        kernel, width, unroll = one.split(",")
        matrix.set_launch(precision, kernel, int(width), int(unroll))


def time_spmv(torch, linalg, matrix, precision, repetitions=50):
    import ctypes

    lib = linalg.library()
    dtype = torch.float64 if precision == 64 else torch.float32
    x = torch.ones(matrix.columns, dtype=dtype, device="cuda:0")
    y = torch.zeros(matrix.rows, dtype=dtype, device="cuda:0")
    fn = lib.cie_hip_csr_product_f64 if precision == 64 else lib.cie_hip_csr_product_f32
    stream = ctypes.c_void_p()
    lib.cie_hip_context_stream(matrix.context.handle, ctypes.byref(stream))
    torch.cuda.synchronize()
    ext_stream = torch.cuda.ExternalStream(stream.value)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(3):
        fn(matrix.context.handle, matrix.handle, 0.0, ctypes.c_void_p(x.data_ptr()), matrix.columns, 1.0,
           ctypes.c_void_p(y.data_ptr()), matrix.rows)
    start.record(ext_stream)
    for _ in range(repetitions):
        fn(matrix.context.handle, matrix.handle, 0.0, ctypes.c_void_p(x.data_ptr()), matrix.columns, 1.0,
           ctypes.c_void_p(y.data_ptr()), matrix.rows)
    stop.record(ext_stream)
    stop.synchronize()
    return start.elapsed_time(stop) * 1e-3 / repetitions


def host_cpu():
    """The host the CPU baseline runs on: usable cores (nproc), and lscpu's model, sockets, cores and threads."""
    info = {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}
    fields = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
              "Thread(s) per core": "threads_per_core"}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
    except (OSError, subprocess.SubprocessError):
        text = ""
    for line in text.splitlines():
        key, _, value = line.partition(":")
        if key.strip() in fields:
            value = value.strip()
            info[fields[key.strip()]] = int(value) if value.isdigit() else value
    return info


def cpu_threads(args):
    return args.cpu_threads or max(1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))


def cpu_baseline_sampled(workload, iterations_to_tolerance, sample_iterations, threads):
    """A bounded sample of the oracle's threaded FP64 PCG, scaled to the GPU's FP64 iteration count (for the
    workloads whose full CPU solve would take minutes)."""
    import oracle

    _, stats, _, seconds = oracle.pcg(workload.matrix, workload.rhs, max_iterations=sample_iterations,
                                      relative_tolerance=1e-30, threads=threads)
    per_iteration = seconds / max(1, stats.iterationCount)
    return {
        "value": per_iteration * (iterations_to_tolerance + 1),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "measured": "sample",
        "sample": (f"oracle threaded FP64 Jacobi-PCG (restated ConjugateGradients_impl.hpp:49-201) on the same "
                   f"system, first {stats.iterationCount} iterations in {seconds:.2f} s on {threads} threads, "
                   f"scaled to the {iterations_to_tolerance + 1} loop trips FP64 PCG needs"),
        "host": host_cpu(),
    }


def cpu_baseline_full(workload, tolerance, threads, one_thread_iterations):
    """The oracle's threaded FP64 Jacobi-PCG (restated ConjugateGradients_impl.hpp:49-201, thread pool capped
    by OMP_NUM_THREADS as ThreadPoolBase.cpp:52-72 caps it) timed on a full solve of the same system to the
    same tolerance, once on all `threads`; a 1-thread sample of `one_thread_iterations` iterations; and the
    CPU configuration (2D 5-point, 256^2, tol 1e-8) solved completely, median of 3, on all threads and on 1.
    Returns the baseline record, the oracle's solution at `tolerance` and its continuation to an absolute
    residual of 1e-14 ||b|| (the tight reference the GPU results' errors are measured against)."""
    import numpy as np

    import oracle
    from cie_amd import systems

    csr, b = workload.matrix, workload.rhs
    x, stats, _, seconds = oracle.pcg(csr, b, relative_tolerance=tolerance, max_iterations=100000, threads=threads)
    print(f"[bench] CPU oracle: {stats.iterationCount} iterations in {seconds:.2f} s on {threads} threads",
          file=sys.stderr, flush=True)
    # Untimed: continue from the oracle's solution to an absolute residual of 1e-14 ||b||.
    tight, tight_stats, _, tight_seconds = oracle.pcg(csr, b, x0=x, absolute_tolerance=1e-14 * np.linalg.norm(b),
                                                      relative_tolerance=0.0, max_iterations=100000, threads=threads)
    record = {
        "value": seconds,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "measured": "full solve",
        "sample": (f"oracle threaded FP64 Jacobi-PCG (restated ConjugateGradients_impl.hpp:49-201) solving the "
                   f"same system to relative residual {tolerance:g} once on {threads} threads: "
                   f"{stats.iterationCount} iterations (reference counting), {seconds:.2f} s including the Jacobi "
                   f"set-up"),
        "iterations": int(stats.iterationCount),
        "relative_residual": stats.relativeResidual,
        "tight_continuation": {"iterations": int(tight_stats.iterationCount), "seconds": tight_seconds},
        "host": host_cpu(),
    }
    if one_thread_iterations > 0:
        _, s1, _, t1 = oracle.pcg(csr, b, max_iterations=one_thread_iterations, relative_tolerance=1e-30, threads=1)
        per_iteration = t1 / max(1, s1.iterationCount)
        record["one_thread"] = {
            "iterations_timed": int(s1.iterationCount), "seconds": t1, "seconds_per_iteration": per_iteration,
            "projected_solve_seconds": per_iteration * (int(stats.iterationCount) + 1),
            "note": "a sample: the 1-thread solve is projected from it, not timed"}
    c1 = systems.five_point_2d(256)
    for label, count in (("c1_all_threads", threads), ("c1_one_thread", 1)):
        runs = [oracle.pcg(c1.matrix, c1.rhs, relative_tolerance=c1.relative_tolerance, threads=count)
                for _ in range(3)]
        record[label] = {"seconds_median_of_3": statistics.median(r[3] for r in runs), "threads": count,
                         "iterations": int(runs[0][1].iterationCount),
                         "relative_residual": runs[0][1].relativeResidual,
                         "workload": "2D 5-point Laplacian, 256^2 interior, b ~ U(-1,1) seed 0, tol 1e-8"}
    return record, x, tight


def main():
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=5)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--elements", type=int, default=None,
                        help="hexahedra per dimension (default 64 for pcg, 32 for mixed-q4: the configurations)")
    parser.add_argument("--tolerance", type=float, default=1e-10)
    parser.add_argument("--cpu-iterations", type=int, default=10,
                        help="CPU baseline: iterations of the 1-thread sample (the all-thread solve is complete; "
                             "for the other workloads the sample itself); 0 disables the CPU leg")
    parser.add_argument("--cpu-threads", type=int, default=0)
    parser.add_argument("--no-profile", action="store_true", help="skip the in-solve SpMV event timing")
    parser.add_argument("--launch", default=None,
                        help="pcg / mixed-q4 / adjoint / pmg: pin the SpMV launches instead of autotuning and "
                             "selecting them (profiling runs): "
                             "KERNEL,WIDTH,UNROLL for both precisions, or FP64/FP32 (e.g. sorted,3,4/sorted,3,8)")
    parser.add_argument("--batch-tiles", choices=("auto", "tiled", "blocks", "sorted", "paired"), default="auto",
                        help="batch workload: the tile format (default: the autotuner's)")
    parser.add_argument("--calibration-copy", action="store_true",
                        help="end with a 1 GiB device-to-device copy (scale reference for rocprofv3 byte counters)")
    parser.add_argument("--workload", choices=("pcg", "mixed-q4", "batch", "adjoint", "assembly", "reorder", "pmg"),
                        default="pcg",
                        help="pcg: the headline solve (3D Q2 Laplace); mixed-q4: the same solves on 3D Q4 with "
                             "lognormal conductivity; batch / adjoint: the added forms; assembly / reorder / pmg: "
                             "the widened paths; one JSON line each")
    args = parser.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: the solve path runs on "
              "one device (a system would shard by contiguous row blocks with a halo exchange of the search "
              "direction per product; batches by whole systems)", file=sys.stderr)
        sys.exit(2)

    if args.workload not in ("pcg", "mixed-q4"):
        {"batch": bench_batch, "adjoint": bench_adjoint, "assembly": bench_assembly, "reorder": bench_reorder,
         "pmg": bench_pmg}[args.workload](args)
        if args.calibration_copy:
            _calibration_copy()
        return

    import torch

    from cie_amd import linalg, systems
    from cie_amd._lib import Statistics

    q4 = args.workload == "mixed-q4"
    if args.elements is None:
        args.elements = 32 if q4 else 64
    t0 = time.perf_counter()
    if q4:
        workload = systems.q4_lognormal_3d(args.elements, tolerance=args.tolerance)
        description = (f"3D Q4 Laplace, {args.elements}^3 hexahedra, k_e = exp(g_e), g_e ~ N(0,1) seed 2, "
                       f"Dirichlet eliminated, Jacobi-PCG to relative residual {args.tolerance:g}")
    else:
        workload = systems.q2_laplace_3d(args.elements, tolerance=args.tolerance)
        description = (f"3D Q2 Laplace, {args.elements}^3 hexahedra, Dirichlet eliminated, Jacobi-PCG to "
                       f"relative residual {args.tolerance:g}")
    generate_seconds = time.perf_counter() - t0
    csr = workload.matrix
    rows, nnz = csr.rows, csr.nnz
    print(f"[bench] generated {workload.name}: {rows} rows, {nnz} nonzeros in {generate_seconds:.1f}s",
          file=sys.stderr, flush=True)

    context = linalg.Context.default()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    matrix = linalg.CSRMatrix(csr, context)
    context.synchronize()
    upload_seconds = time.perf_counter() - t0
    # Set-up beside the solve: the upload, then the autotuner, which builds every derived format it times
    # (narrow columns, column blocks, sorted tiles) and frees the ones no launch keeps.
    t0 = time.perf_counter()
    if args.launch:  # a pinned launch (profiling runs), built but not timed against others
        pin_launches(matrix, args.launch)
        tune64 = time_spmv(torch, linalg, matrix, 64, repetitions=3)
    else:
        tune64 = matrix.autotune(64, 10, kernels=("staged", "vector", "tiled", "blocks", "sorted", "paired"))
    context.synchronize()
    autotune64_seconds = time.perf_counter() - t0
    t0 = time.perf_counter()
    if args.launch:
        tune32 = time_spmv(torch, linalg, matrix, 32, repetitions=3)
    else:
        tune32 = matrix.autotune(32, 10, kernels=("staged", "vector", "tiled", "blocks", "sorted", "paired"))
    context.synchronize()
    autotune32_seconds = time.perf_counter() - t0
    b = torch.from_numpy(workload.rhs).to("cuda:0")
    x = torch.zeros_like(b)
    settings = Statistics(100000, 0.0, args.tolerance)
    fp64 = linalg.FusedPCG(matrix, settings)
    mixed = linalg.MixedPCG(matrix, settings)
    print(f"[bench] uploaded in {upload_seconds:.1f}s; launch fp64 {matrix.launch(64)} ({tune64 * 1e3:.3f} ms), "
          f"fp32 {matrix.launch(32)} ({tune32 * 1e3:.3f} ms)", file=sys.stderr, flush=True)

    def solve(solver):
        x.zero_()
        solver.product(0, b, 1, x)
        return solver.info

    # Launch selection by solve time: the autotuner times the standalone SpMV; the solves run the product
    # kernel with the fused p.q, whose best tile size can differ. Each precision keeps whichever of its
    # autotuned launch, the sorted 1 x 3 tiles of 1024, 2048 and 4096 blocks and the paired 1 x 5 / 1 x 3
    # tiles of 512 and 1024 blocks per width solves fastest (FP64 first: the mixed solve's start and
    # refinement steps use the FP64 launch).
    selection = {}
    t0 = time.perf_counter()
    if not args.launch:
        for precision, solver in ((64, fp64), (32, mixed)):
            candidates = [tuple(matrix.launch(precision))]
            candidates += [c for c in (("sorted", 3, 4), ("sorted", 3, 8), ("sorted", 3, 16), ("paired", 53, 2),
                                       ("paired", 53, 4)) if c not in candidates]
            times = {}
            for candidate in candidates:
                matrix.set_launch(precision, *candidate)
                solve(solver)  # builds the candidate's tiles
                torch.cuda.synchronize()
                start = time.perf_counter()
                solve(solver)
                torch.cuda.synchronize()
                times[candidate] = time.perf_counter() - start
            best = min(times, key=times.get)
            matrix.set_launch(precision, *best)
            selection[f"fp{precision}"] = {",".join(map(str, c)): t for c, t in times.items()}
            print(f"[bench] fp{precision} solve launch {best} ({times[best]:.4f} s of "
                  f"{', '.join(f'{t:.4f}' for t in times.values())})", file=sys.stderr, flush=True)
    selection_seconds = time.perf_counter() - t0

    for _ in range(args.warmup):
        solve(mixed)
        solve(fp64)
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # Timed region: K mixed solves (the value), in-solve SpMV events on.
    if not args.no_profile:
        linalg.library().cie_hip_context_set_profiling(context.handle, 1)
    torch.cuda.synchronize()
    start = time.perf_counter()
    mixed_infos = [solve(mixed) for _ in range(args.steps)]
    torch.cuda.synchronize()
    mixed_wall = (time.perf_counter() - start) / args.steps
    linalg.library().cie_hip_context_set_profiling(context.handle, 0)
    print(f"[bench] mixed: {mixed_wall:.4f} s/solve", file=sys.stderr, flush=True)
    mixed_x = x.cpu().numpy()

    torch.cuda.synchronize()
    start = time.perf_counter()
    fp64_infos = [solve(fp64) for _ in range(args.steps)]
    torch.cuda.synchronize()
    fp64_wall = (time.perf_counter() - start) / args.steps
    fp64_x = x.cpu().numpy()
    print(f"[bench] fp64: {fp64_wall:.4f} s/solve", file=sys.stderr, flush=True)

    import numpy as np

    # The CPU baseline: the oracle's FP64 PCG on the same system, a full solve on all host threads for the
    # headline configuration (a bounded sample for the larger Q4 one). Its solution, and its continuation to
    # 1e-14 ||b||, are what the GPU results' errors are measured against.
    cpu, oracle_x, oracle_tight = None, None, None
    if args.cpu_iterations > 0:
        if q4:
            cpu = cpu_baseline_sampled(workload, int(fp64_infos[-1].stats.iterationCount), max(20, args.cpu_iterations),
                                       cpu_threads(args))
        else:
            cpu, oracle_x, oracle_tight = cpu_baseline_full(workload, args.tolerance, cpu_threads(args),
                                                            args.cpu_iterations)

    def error(v, reference):
        if reference is None:
            return None
        return float(np.linalg.norm(v - reference) / np.linalg.norm(reference))

    spmv64 = time_spmv(torch, linalg, matrix, 64)
    spmv32 = time_spmv(torch, linalg, matrix, 32)
    # Algorithmic bytes (int32 CSR: the roofline's measure) and the bytes of the format each launch
    # actually reads (narrow 16-bit columns, or column blocks with their padding).
    csr64 = spmv_bytes(rows, nnz, 8)
    csr32 = spmv_bytes(rows, nnz, 4)
    bytes64 = format_bytes(matrix, 64)
    bytes32 = format_bytes(matrix, 32)

    m = mixed_infos[-1]
    f = fp64_infos[-1]
    product_seconds = statistics.mean(i.productSeconds for i in mixed_infos if i.productSamples) \
        if any(i.productSamples for i in mixed_infos) else spmv32
    achieved = csr32 / product_seconds / 1e9
    traffic = measured_traffic(matrix.launch(32), f"spmv_{'q4' if q4 else 'q2'}_{args.elements}_traffic.json")
    rocprof_mean = _profiled_mean_seconds(f"pcg_{'q4' if q4 else 'q2'}_{args.elements}_kernel_stats.md",
                                          f"tileProductKernel<{_walker(matrix.launch(32), 'float')} >")

    result = {
        "metric": "time-to-1e-10 residual (s) + SpMV HBM GB/s, FP64 vs mixed-refined PCG, 1 GPU",
        "value": mixed_wall,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed_wall * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement",
        "data": f"synthetic (structured {'Q4' if q4 else 'Q2'} stiffness, SplitMix64 rhs)",
        "config": {
            "workload": description,
            "rows": rows, "nnz": nnz, "rhs_seed": 1, "x0": "zero",
            "parallelism": "single GPU",
        },
        "mixed": {
            "seconds": mixed_wall, "device_seconds": m.seconds,
            "device_seconds_median": statistics.median(i.seconds for i in mixed_infos),
            "iterations": int(m.stats.iterationCount),
            "refinements": int(m.refinements), "relative_residual_fp64": m.stats.relativeResidual,
            "true_relative_residual": m.trueRelativeResidual,
            "error_vs_oracle_tight": error(mixed_x, oracle_tight), "error_vs_oracle": error(mixed_x, oracle_x),
        },
        "fp64": {
            "seconds": fp64_wall, "device_seconds": f.seconds,
            "device_seconds_median": statistics.median(i.seconds for i in fp64_infos),
            "iterations": int(f.stats.iterationCount),
            "relative_residual": f.stats.relativeResidual, "true_relative_residual": f.trueRelativeResidual,
            "error_vs_oracle_tight": error(fp64_x, oracle_tight), "error_vs_oracle": error(fp64_x, oracle_x),
        },
        "errors_note": ("error_vs_oracle: relative 2-norm distance to the oracle's FP64 solution at the same "
                        "tolerance; error_vs_oracle_tight: to that solution continued to an absolute residual "
                        "of 1e-14 ||b||; null when the CPU leg is off"),
        "setup": {"upload_seconds": upload_seconds, "autotune_fp64_seconds": autotune64_seconds,
                  "autotune_fp32_seconds": autotune32_seconds, "solve_selection_seconds": selection_seconds,
                  "seconds": upload_seconds + autotune64_seconds + autotune32_seconds,
                  "note": "once per matrix: host-to-device copy of the CSR, then the autotuner (builds and times "
                          "every candidate format and launch); not part of value. solve_selection_seconds: the "
                          "bench's own choice among a few launches by timing whole solves (two per candidate)"},
        "solve_launch_selection": selection or "pinned",
        "speedup_mixed_vs_fp64": fp64_wall / mixed_wall,
        "spmv": {
            "fp64_seconds": spmv64, "fp64_GBps": csr64 / spmv64 / 1e9, "fp64_bytes": csr64,
            "fp64_format_bytes": bytes64, "fp64_format_GBps": bytes64 / spmv64 / 1e9,
            "fp32_seconds": spmv32, "fp32_GBps": csr32 / spmv32 / 1e9, "fp32_bytes": csr32,
            "fp32_format_bytes": bytes32, "fp32_format_GBps": bytes32 / spmv32 / 1e9,
            "launch_fp64": list(matrix.launch(64)), "launch_fp32": list(matrix.launch(32)),
            "index_bytes_fp64": index_bytes(matrix, 64), "index_bytes_fp32": index_bytes(matrix, 32),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "measured_GBps": None if traffic is None or rocprof_mean is None else traffic / rocprof_mean / 1e9,
            "rocprof_mean_launch_seconds": rocprof_mean,
            "measured_note": "measured_GBps: calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per launch over the "
                             "rocprofv3 mean launch time, both from the committed profiles of this workload and "
                             "launch (null when the autotuner picked a launch they do not cover)",
            "kernel": f"tileProductKernel {list(matrix.launch(32))} (q = A p, FP32 plane, fused p.q) inside the mixed solve",
            "bytes_per_launch": csr32, "bytes_rule": "int32 CSR: (4 + 4) nnz + 4 (N + 1) + 2 * 4 N",
            "mean_launch_seconds": product_seconds,
            "format_bytes_per_launch": bytes32, "format_GBps": bytes32 / product_seconds / 1e9,
        },
        "setup_launch": "pinned " + args.launch if args.launch else "autotuned, then selected by solve time",
    }
    if cpu is not None:
        result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)
    if args.calibration_copy:
        _calibration_copy()


def _calibration_copy():
    """The 1 GiB device-to-device copy that ends a run profiled for HBM traffic: FETCH_SIZE and WRITE_SIZE of
    this copy scale every kernel's counters to bytes (cie_amd/profiling.py)."""
    import torch

    source = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    torch.empty_like(source).copy_(source)
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# The widened paths: device assembly, RCM reordering, p-multigrid.
# ---------------------------------------------------------------------------

def _library_stream_events(torch, linalg, context):
    """A timer of work on the library context's stream (HIP events recorded on that stream)."""
    from ctypes import byref, c_void_p

    pointer = c_void_p()
    linalg.library().cie_hip_context_stream(context.handle, byref(pointer))
    stream = torch.cuda.ExternalStream(pointer.value)

    def timed(call, repetitions):
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(repetitions):
            call()
        stop.record(stream)
        stop.synchronize()
        return start.elapsed_time(stop) * 1e-3 / repetitions

    return timed


def _kernel_seconds(linalg, context, call, repetitions):
    """Mean main-kernel duration of `call` from the context's profiling events (HIP events on its stream)."""
    from ctypes import byref, c_double

    library = linalg.library()
    library.cie_hip_context_set_profiling(context.handle, 1)
    seconds = []
    try:
        for _ in range(repetitions):
            call()
            value = c_double()
            library.cie_hip_context_kernel_seconds(context.handle, byref(value))
            seconds.append(value.value)
    finally:
        library.cie_hip_context_set_profiling(context.handle, 0)
    return statistics.mean(seconds)


def _line(args, metric, value, unit, dtype, config, roofline, cpu, extra=None):
    result = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (structured meshes, SplitMix64 inputs)", "config": config,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    result.update(extra or {})
    print(json.dumps(result), flush=True)


def _roofline(kernel, bytes_per_launch, seconds, note, traffic=None):
    achieved = bytes_per_launch / seconds / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic, "kernel": kernel,
            "bytes_per_launch": bytes_per_launch, "mean_launch_seconds": seconds, "bytes": note}


def _walker(launch, value="double"):
    """Kernel-name form of the SpMV walker a launch (kernel, width, unroll) selects."""
    kernel, width, unroll = launch
    if kernel == "tiled":
        return f"BlockTiles<{value}, {value}, {unroll}, {width}>"
    if kernel == "blocks":
        return f"ColumnBlockTiles<{value}, {value}, {unroll}, 256, {width}>"
    if kernel == "sorted":
        return f"SortedBlockTiles<{value}, {value}, {unroll}, 256, {width}>"
    if kernel == "paired":
        return f"PairedBlockTiles<{value}, {value}, {unroll}, 256, {width // 10}, {width % 10}>"
    return f"{'StagedRows' if kernel == 'staged' else 'VectorRows'}<{value}, {value}, {width}, {unroll}, 256>"


def _profiled_traffic(profile, kernel):
    """HBM bytes per dispatch of `kernel` (its short name) in the committed calibrated rocprofv3 summary
    `profile` of this workload, else None."""
    path = ROOT / "profiles" / profile
    if not path.exists():
        return None
    entry = json.loads(path.read_text()).get("kernels", {}).get(kernel)
    return None if entry is None else entry["traffic_bytes"]


def _profiled_mean_seconds(stats, kernel):
    """Mean duration in seconds of the working launches of `kernel` (its short name) in the committed rocprofv3
    --kernel-trace --stats summary `stats` of this workload (a markdown table, means in microseconds), else None."""
    path = ROOT / "profiles" / stats
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.split("|")]
        if len(cells) >= 8 and cells[1] == f"`{kernel}`" and cells[5]:
            return float(cells[5]) * 1e-6  # the mean of the long launches (trips past convergence left out)
    return None


_BATCH_WALKER = {"tiled": 0, "blocks": 3, "sorted": -3, "paired": 53}  # the walker code of the batch's tile format
_BATCH_TILES = {"tiled": "entries", "blocks": "1 x 3 column blocks", "sorted": "1 x 3 column blocks, column-sorted",
                "paired": "paired 1 x 5 / 1 x 3 column blocks, column-sorted"}


def bench_batch(args):
    """256 uneven 2D Q2 systems (n_i in [16, 256], seed 3) solved together, mixed-refined to 1e-10."""
    import numpy as np
    import torch

    import oracle
    from cie_amd import linalg, systems
    from cie_amd._lib import Statistics

    w = systems.q2_quad_batch(256)
    offsets, rows, nnz = w.extra["offsets"], w.matrix.rows, w.matrix.nnz
    context = linalg.Context.default()
    matrix = linalg.CSRMatrix(w.matrix, context)
    batch = linalg.Batch(matrix, offsets)
    batch.autotune(32, 5)
    if args.batch_tiles != "auto":  # a fixed tile format (profiling passes that must see the same kernels)
        batch.set_tile_kernel(args.batch_tiles)
    b = torch.from_numpy(w.rhs).cuda()
    x = torch.zeros_like(b)
    settings = Statistics(100000, 0.0, w.relative_tolerance)

    def solve(mixed=True, graph=True):
        x.zero_()
        return batch.solve(b, x, settings, mixed=mixed, graph=graph)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        stats, status, refinements, summary = solve()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - start) / args.steps
    assert np.all(status == 1)
    iterations = np.array([st.iterationCount for st in stats], dtype=np.int64)
    fp64_stats, _, _, fp64_summary = solve(mixed=False)
    fp64_iterations = np.array([st.iterationCount for st in fp64_stats], dtype=np.int64)
    # Dominant kernel: the batched FP32 product; the first 8 launches (every system still runs) timed with events.
    linalg.library().cie_hip_context_set_profiling(context.handle, 1)
    _, _, _, profiled = solve(graph=False)
    linalg.library().cie_hip_context_set_profiling(context.handle, 0)
    assert iterations.min() >= 8 and profiled.productSamples > 0
    algorithmic = spmv_bytes(rows, nnz, 4)
    # CPU baseline: the reference solves each system with its own ConjugateGradients; the oracle's FP64 PCG on
    # systems in order until the sample time is spent, scaled by nonzeros x FP64 iterations.
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    extents = w.matrix.row_extents
    sampled_work, sampled_seconds, sampled = 0, 0.0, 0
    for i in range(256):
        lo, hi = int(offsets[i]), int(offsets[i + 1])
        sub = systems.CSR(hi - lo, hi - lo, (extents[lo:hi + 1] - extents[lo]).astype(np.int32),
                          (w.matrix.column_indices[extents[lo]:extents[hi]] - lo).astype(np.int32),
                          w.matrix.entries[extents[lo]:extents[hi]])
        _, st, _, secs = oracle.pcg(sub, w.rhs[lo:hi], relative_tolerance=w.relative_tolerance, max_iterations=100000,
                                    threads=threads)
        sampled_work += sub.nnz * (int(st.iterationCount) + 1)
        sampled_seconds += secs
        sampled += 1
        if sampled_seconds > 15.0:
            break
    system_nnz = np.diff(extents[offsets])
    total_work = int(np.sum(system_nnz * (fp64_iterations + 1)))
    _line(args, "batched solve time (s), 256 uneven 2D Q2 systems to 1e-10, mixed-refined", wall, "s",
          "f32 inner / f64 refinement",
          {"workload": "256 2D Q2 Laplace systems, n_i ~ U{16..256} (seed 3), k_e lognormal(0, 0.5), b ~ U(-1,1), "
                       "relative residual 1e-10 each", "rows": rows, "nnz": nnz, "systems": 256},
          _roofline(f"batchProductKernel<float, {_BATCH_WALKER[batch.tile_kernel()]}> (all systems active; tiles of "
                    f"{_BATCH_TILES[batch.tile_kernel()]})", algorithmic,
                    profiled.productSeconds, "B_SpMV(FP32 plane, int32 CSR) over the whole block-diagonal matrix",
                    _profiled_traffic("batch_q2_256_traffic.json",
                                      f"batchProductKernel<float, {_BATCH_WALKER[batch.tile_kernel()]}>")),
          {"value": sampled_seconds * total_work / sampled_work, "unit": "s", "cores": threads, "kind": "port",
           "sample": f"oracle FP64 Jacobi-PCG (ConjugateGradients_impl.hpp:49-201) on the first {sampled} systems "
                     f"in {sampled_seconds:.1f} s, scaled by nonzeros x FP64 iterations to all 256"},
          {"device_seconds": summary.seconds, "max_iterations": int(iterations.max()),
           "tile_format": batch.tile_kernel(),
           "tile_rows": {str(r): int(c) for r, c in zip((32, 64, 128, 256),
                                                        np.bincount(batch.tile_rows(), minlength=257)[[32, 64, 128, 256]])},
           "refinements": int(refinements.sum()), "fp64": {"device_seconds": fp64_summary.seconds,
                                                           "max_iterations": int(fp64_iterations.max())}})


def bench_adjoint(args):
    """dJ/dg for J = 0.5 ||u(g) - u_t||^2 on 3D Q2 48^3 (k = exp(g)): assembly, forward and adjoint PCG, gather."""
    import numpy as np
    import torch

    import oracle
    from cie_amd import adjoint, linalg, systems

    n = 48
    w = systems.q2_adjoint_3d(n)
    g, gt = w.extra["log_conductivity"], w.extra["target_log_conductivity"]
    problem = adjoint.LagrangeProblem(3, 2, n, w.rhs, np.zeros(w.matrix.rows), tolerance=w.relative_tolerance)
    problem.assemble(gt)
    if args.launch:  # pinned (profiling runs)
        pin_launches(problem.matrix, args.launch)
    else:
        problem.matrix.autotune(64, 5, kernels=("tiled", "blocks", "sorted", "paired"))
        problem.matrix.autotune(32, 5, kernels=("tiled", "blocks", "sorted", "paired"))  # the mixed variant's FP32 plane
    ut, _ = problem.solve(problem.rhs)
    problem.target = ut
    context = problem.context

    def gradient():
        g_t = torch.tensor(g, dtype=torch.float64, device="cuda:0", requires_grad=True)
        j = adjoint.Objective.apply(g_t, problem)
        j.backward()
        return g_t.grad

    for _ in range(args.warmup):
        gradient()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        gradient()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - start) / args.steps
    fp64_gradient = gradient()
    # The same gradient with the forward and adjoint solves mixed-refined (stopped on the FP64 true residual).
    problem.mixed = True
    for _ in range(args.warmup):
        gradient()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        mixed_gradient = gradient()
    torch.cuda.synchronize()
    mixed_wall = (time.perf_counter() - start) / args.steps
    problem.mixed = False
    mixed_difference = float(torch.linalg.norm(mixed_gradient - fp64_gradient) / torch.linalg.norm(fp64_gradient))
    linalg.library().cie_hip_context_set_profiling(context.handle, 1)
    problem.assemble(g)
    u, info = problem.solve(problem.rhs)
    linalg.library().cie_hip_context_set_profiling(context.handle, 0)
    lam, info_adjoint = problem.solve(u - problem.target)
    rows, nnz = w.matrix.rows, w.matrix.nnz
    algorithmic = spmv_bytes(rows, nnz, 8)
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sample = 20
    _, _, _, secs = oracle.pcg(w.matrix, w.rhs, relative_tolerance=0.0, max_iterations=sample, threads=threads)
    total_iterations = int(info.stats.iterationCount) + int(info_adjoint.stats.iterationCount) + 2
    t0 = time.perf_counter()
    oracle.element_gradient(problem.dofs_host, problem.element_matrix_host, lam.cpu().numpy(), u.cpu().numpy())
    gather = time.perf_counter() - t0
    _line(args, "adjoint gradient time (s), dJ/dg over every element, forward + adjoint FP64 PCG to 1e-11", wall,
          "s", "f64",
          {"workload": f"3D Q2 Laplace {n}^3 hexahedra, k_e = exp(g_e), g ~ N(0, 0.5) seed 4, target from seed 5, "
                       f"b ~ U(-1,1) seed 1", "rows": rows, "nnz": nnz, "elements": int(problem.count)},
          _roofline(f"tileProductKernel {list(problem.matrix.launch(64))} (q = A p) inside the forward solve",
                    algorithmic, info.productSeconds, "B_SpMV(FP64, int32 CSR)",
                    next((t for t in (_profiled_traffic("adjoint_q2_48_traffic.json",
                                                        f"{form}<{_walker(problem.matrix.launch(64))} >")
                                      for form in ("tileProductKernel", "productKernel", "csrProductKernel"))
                          if t is not None), None)),
          {"value": secs * total_iterations / sample + gather, "unit": "s", "cores": threads, "kind": "port",
           "sample": f"oracle FP64 Jacobi-PCG, {sample} iterations in {secs:.2f} s, scaled to the "
                     f"{total_iterations} forward + adjoint iterations, plus the oracle's element gradient "
                     f"({gather:.2f} s)"},
          {"forward_iterations": int(info.stats.iterationCount),
           "adjoint_iterations": int(info_adjoint.stats.iterationCount),
           "mixed_solves": {"seconds": mixed_wall, "speedup": wall / mixed_wall,
                            "gradient_relative_difference": mixed_difference,
                            "note": "the same gradient with the forward and adjoint solves mixed-refined to the "
                                    "same FP64 true-residual tolerance (LagrangeProblem(mixed=True))"}})


def bench_assembly(args):
    """K = sum_e k_e P_e^T K_e0 P_e on the benchmark pattern (3D Q2 64^3, k_e = exp(N(0,1))), device assembly."""
    import numpy as np
    import torch

    import oracle
    from cie_amd import linalg, systems

    n = args.elements or 64
    count = systems.lagrange_size(3, 2, n)[2]
    k = np.exp(systems.normal(2, 1.0, count))
    pattern = systems.lagrange(3, 2, n, threads=0)
    dofs = systems.element_dofs(3, 2, n)
    element = systems.element_matrix(3, 2, n)
    context = linalg.Context.default()
    matrix = linalg.CSRMatrix(pattern, context)
    assembly = linalg.Assembly(matrix, dofs)
    element_dev, k_dev = torch.from_numpy(element).cuda(), torch.from_numpy(k).cuda()
    timed = _library_stream_events(torch, linalg, context)
    for _ in range(args.warmup):
        assembly.assemble(element_dev, k_dev)
    seconds = timed(lambda: assembly.assemble(element_dev, k_dev), args.steps)
    kernel = _kernel_seconds(linalg, context, lambda: assembly.assemble(element_dev, k_dev), args.steps)
    rows, nnz = pattern.rows, pattern.nnz
    kept = np.count_nonzero(dofs >= 0, axis=1).astype(np.int64)
    contributions = int(np.sum(kept * kept))  # one per (element, kept local row, kept local column)
    algorithmic = 12 * nnz + 4 * (nnz + 1) + 6 * contributions + 8 * count
    t0 = time.perf_counter()
    oracle.assemble(pattern, dofs, element, k)
    cpu = time.perf_counter() - t0
    _line(args, "device stiffness assembly time (s), K = sum_e k_e P_e^T K_e0 P_e, FP64 + FP32 planes", seconds,
          "s", "f64", {"workload": f"3D Q2 Laplace {n}^3 hexahedra ({count} elements), k_e = exp(N(0,1)) seed 2, "
                       f"benchmark pattern", "rows": rows, "nnz": nnz, "contributions": contributions},
          _roofline("assembleSlotsKernel<uint16_t> (one launch per assembly)", algorithmic, kernel,
                    "12 nnz (FP64 + FP32 values written) + 4 (nnz + 1) (slot offsets) + 6 per contribution "
                    "(element, local pair) + 8 E (k_e); K_e0 stays in cache",
                    _profiled_traffic("assembly_q2_64_traffic.json", "assembleSlotsKernel<unsigned short>")),
          {"value": cpu, "unit": "s", "cores": 1, "kind": "port",
           "sample": "oracle_assemble (the Serial addContribution scatter, Assembler_impl.hpp:156-194) over all "
                     "elements, one thread"})


def bench_reorder(args):
    """RCM of the benchmark matrix numbered at random: host map, device CSR permutation, SpMV before/after."""
    import numpy as np
    import torch

    import oracle
    from cie_amd import linalg, systems

    n = args.elements or 64
    natural = systems.lagrange(3, 2, n, threads=0)
    rows, nnz = natural.rows, natural.nnz
    shuffle = np.random.default_rng(11).permutation(rows).astype(np.int32)
    context = linalg.Context.default()
    base = linalg.CSRMatrix(natural, context)
    shuffled = base.reorder(shuffle)
    host_shuffled = shuffled.copy_out()
    t0 = time.perf_counter()
    mapping = linalg.makeReordering(host_shuffled, "reverse-cuthill-mckee")
    map_seconds = time.perf_counter() - t0
    timed = _library_stream_events(torch, linalg, context)
    results = []

    def permute():
        results.append(shuffled.reorder(mapping))

    for _ in range(args.warmup):
        permute()
    seconds = timed(permute, args.steps)
    kernel = _kernel_seconds(linalg, context, permute, args.steps)
    reordered = results.pop()
    for extra in results:
        extra.close()
    spmv = {name: m.autotune(64, repetitions=10) for name, m in
            (("natural", base), ("random", shuffled), ("rcm", reordered))}
    algorithmic = 2 * (4 + 8 + 4) * nnz + 2 * 4 * (rows + 1) + 3 * 4 * rows
    t0 = time.perf_counter()
    oracle.make_reordering(host_shuffled, "reverse-cuthill-mckee")
    cpu_map = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.reorder_csr(host_shuffled, mapping)
    cpu_permute = time.perf_counter() - t0
    _line(args, "device CSR permutation time (s) by the reverse Cuthill-McKee map", seconds, "s", "int32 + f64",
          {"workload": f"3D Q2 Laplace {n}^3 hexahedra, rows numbered at random (seed 11), reverse Cuthill-McKee",
           "rows": rows, "nnz": nnz},
          _roofline("permuteRowsKernel (one launch per permutation; the value above is the whole call, with its "
                    "allocations, map upload, device preparation of the maps and row extents, and extents download)", algorithmic, kernel,
                    "(4 + 8 + 4) nnz read + written (columns, FP64, FP32) + 2 x 4 (N + 1) extents + 3 x 4 N maps",
                    _profiled_traffic("reorder_q2_64_traffic.json", "permuteRowsKernel")),
          {"value": cpu_permute, "unit": "s", "cores": 1, "kind": "port",
           "sample": f"oracle_reorder_csr (reorder.cpp:215-365) of the whole matrix, one thread; the oracle's RCM "
                     f"map took {cpu_map:.3f} s"},
          {"rcm_map_host_seconds": map_seconds,
           "spmv_fp64_seconds": spmv, "spmv_rcm_vs_random": spmv["random"] / spmv["rcm"]})


def bench_pmg(args):
    """solveMultigrid on 2D Q4 integrated Legendre, 512^2 elements, to 1e-6 of ||b||."""
    import numpy as np
    import torch

    import oracle
    from cie_amd import linalg, systems
    from cie_amd._lib import Statistics

    order, n = 4, 512
    csr = systems.lagrange(2, order, n, basis="legendre", threads=0)
    mask_values = systems.ansatz_mask(2, order, n)
    b = systems.uniform(1, csr.rows)
    rows, nnz = csr.rows, csr.nnz
    context = linalg.Context.default()
    matrix = linalg.CSRMatrix(csr, context)
    if args.launch:  # pinned (profiling runs)
        pin_launches(matrix, args.launch)
    else:  # the finest grid's sweeps and updates use this launch
        matrix.autotune(64, 5, kernels=("tiled", "blocks", "sorted", "paired"))
    mask = linalg.Mask(mask_values, context)
    t0 = time.perf_counter()
    pmg = linalg.PMultigrid(matrix, mask, order)
    setup = time.perf_counter() - t0
    b_dev = torch.from_numpy(b).cuda()
    x = torch.zeros_like(b_dev)
    # The finest grid's launch by solve time, as for the headline solve: the autotuner times the plain SpMV, the
    # cycle runs the sweep and update kernels, whose best shape can differ.
    def cycle_time():
        x.zero_()
        torch.cuda.synchronize()
        start = time.perf_counter()
        pmg.solve(b_dev, x)
        torch.cuda.synchronize()
        return time.perf_counter() - start

    def select(target, candidates):
        candidates = [tuple(target.launch(64))] + [c for c in candidates if c != tuple(target.launch(64))]
        times = {}
        for candidate in candidates:
            target.set_launch(64, *candidate)
            cycle_time()  # builds the candidate's format
            times[candidate] = min(cycle_time(), cycle_time())
        target.set_launch(64, *min(times, key=times.get))
        return {",".join(map(str, c)): t for c, t in times.items()}

    t0 = time.perf_counter()
    # Then each coarser grid's launch the same way (pmg.grid(g).set_launch): the masked grids 1 and 2 and the
    # renumbered coarse matrix; the finest grid of order 4 is the matrix itself.
    if args.launch:
        selection = "pinned " + args.launch
    else:
        selection = {"finest": select(matrix, (("sorted", 3, 4), ("sorted", 3, 8), ("sorted", 3, 16),
                                               ("paired", 53, 2), ("paired", 53, 4)))}
        grid_candidates = (("tiled", 256, 8), ("sorted", 3, 4), ("paired", 53, 2), ("paired", 53, 4),
                           ("paired", 32, 2), ("paired", 32, 4))
        for index, name in ((1, "grid 1"), (2, "grid 2"), (-1, "coarse")):
            grid = pmg.grid(index)
            if grid is not None and grid.nnz != matrix.nnz:
                selection[name] = select(grid, grid_candidates)
    selection_seconds = time.perf_counter() - t0
    for _ in range(args.warmup):
        x.zero_()
        pmg.solve(b_dev, x)
    infos = []
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        x.zero_()
        infos.append(pmg.solve(b_dev, x))
    torch.cuda.synchronize()
    wall = (time.perf_counter() - start) / args.steps
    info = infos[-1]
    # Dominant kernel: a Jacobi sweep on the finest grid (the whole matrix), from 7-sweep minus 1-sweep runs.
    space = linalg.HIPSpace()
    jacobi = {k: linalg.JacobiOperator(space, rows, matrix, linalg.makeDiagonalOperator(matrix, space), k, 2.0 / 3.0)
              for k in (1, 7)}
    y = torch.zeros_like(b_dev)
    timed = _library_stream_events(torch, linalg, context)
    one = timed(lambda: jacobi[1].product(0, b_dev, 1, y), 5)
    seven = timed(lambda: jacobi[7].product(0, b_dev, 1, y), 5)
    sweep = (seven - one) / 6
    algorithmic = spmv_bytes(rows, nnz, 8) + 2 * 8 * rows
    solver = linalg.FusedPCG(matrix, Statistics(100000, 0.0, 1e-6))
    xp = torch.zeros_like(b_dev)
    solver.product(0, b_dev, 1, xp)
    sample_cycles = 3
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    oracle.pmg(csr, mask_values, order, b, max_cycles=sample_cycles, threads=threads)
    cpu_sample = time.perf_counter() - t0
    _line(args, "p-multigrid time to ||r|| <= 1e-6 ||b|| (s), poisson2D's solveMultigrid", wall, "s", "f64",
          {"workload": f"2D Q{order} integrated Legendre Laplace, {n}^2 elements, Dirichlet eliminated, "
                       f"b ~ U(-1,1) seed 1, x0 = 0", "rows": rows, "nnz": nnz},
          _roofline(f"jacobiSweepKernel {list(matrix.launch(64))} on the finest grid", algorithmic, sweep,
                    "B_SpMV(FP64, int32 CSR) + 2 x 8 N (right-hand side and inverse diagonal)",
                    _profiled_traffic("pmg_q4_legendre_512_traffic.json",
                                      f"jacobiSweepKernel<{_walker(matrix.launch(64))} >")),
          {"value": cpu_sample * info.cycles / sample_cycles, "unit": "s", "cores": threads, "kind": "port",
           "sample": f"oracle_pmg (solver.cpp:111-232 restated) for {sample_cycles} cycles in {cpu_sample:.2f} s, "
                     f"scaled to the {info.cycles} cycles of the solve"},
          {"solve_launch_selection": selection, "solve_selection_seconds": selection_seconds,
           "cycles": int(info.cycles), "coarse_cg_steps": int(info.coarseIterations),
           "device_seconds": info.seconds, "setup_seconds": setup, "tracked_relative_residual":
               info.residual / info.initialResidual,
           "jacobi_pcg_to_1e-6": {"iterations": int(solver.info.stats.iterationCount),
                                  "seconds": solver.info.seconds}})


def measured_traffic(launch, profile):
    """HBM bytes per launch of the in-solve FP32 product kernel for this launch shape, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (`profile`, calibrated on a
    1 GiB copy), else None."""
    path = ROOT / "profiles" / profile
    if not path.exists():
        return None
    kernel, width, unroll = launch
    if kernel == "tiled":
        walker = f"BlockTiles<float, float, {unroll}, {width}> >"
    elif kernel == "blocks":
        walker = f"ColumnBlockTiles<float, float, {unroll}, 256, {width}> >"
    elif kernel == "sorted":
        walker = f"SortedBlockTiles<float, float, {unroll}, 256, {width}> >"
    elif kernel == "paired":
        walker = f"PairedBlockTiles<float, float, {unroll}, 256, {width // 10}, {width % 10}> >"
    else:
        walker = f"{'StagedRows' if kernel == 'staged' else 'VectorRows'}<float, float, {width}, {unroll}, 256> >"
    kernels = json.loads(path.read_text()).get("kernels", {})
    # the in-solve product kernel; else the standalone SpMV of the same shape (the autotuner times every shape)
    entry = (kernels.get(f"tileProductKernel<{walker}") or kernels.get(f"productKernel<{walker}")
             or kernels.get(f"csrProductKernel<{walker}"))
    return None if entry is None else entry["traffic_bytes"]


if __name__ == "__main__":
    main()
