"""Benchmark of the matrix-free Pauli MVP on one MI355X.

Workload (default): 28 qubits, 4 x 7 open Heisenberg XXZ lattice (Jz = 1,
Jxy ~ U[0.5, 1.5]) with a random Z field h ~ U[-1, 1] and 1,885 random
weight-1..4 terms (N(0, 0.1)), 2,048 raw terms, seed 20260801; the state is
2^28 complex128 Gaussian amplitudes normalised to 1, generated on the GPU.
One step is one y = H v through the library's C ABI with v and y resident in
HBM.  ``value`` is the grouped-gather algorithmic bytes of one MVP divided by
its time (GB/s); term*amp/s is reported beside it.

Printed as one JSON line with ``roofline`` (HBM bound, 8 TB/s peak) and
``cpu_baseline`` (the C oracle's TermDirect row gather -- the reference's
default lazy algorithm -- on a bounded row sample, all host threads).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "28q MVP HBM GB/s & term·amp/s; Davidson s-to-1e-8 learned vs Jacobi"
HBM_PEAK_GBS = 8000.0
SEED = 20260801


def parse_args() -> argparse.Namespace:
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--nqubits", type=int, default=28, help="28 = the 4x7 lattice workload; else random")
    p.add_argument("--random-terms", type=int, default=1885)
    p.add_argument("--dtype", choices=("complex128", "complex64"), default="complex128")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "mvp_tile_28q_traffic.json"))
    return p.parse_args()


def build_operator(args):
    import tencirpauli_amd as tcp
    from tencirpauli_amd import lattice

    if args.nqubits == 28:
        codes, coef = lattice.heisenberg_4x7_with_random_terms(SEED, args.random_terms)
        name = f"heisenberg_4x7_open_field_plus_{args.random_terms}_random_28q"
    else:
        codes, coef = lattice.random_pauli_sum(args.nqubits, 1024, SEED)
        name = f"random_pauli_{args.nqubits}q_1024"
    return tcp.PauliOperator.from_code_arrays(codes, coef), name, int(codes.shape[0])


def cpu_baseline(op, v_host: np.ndarray, y_dev_host_rows, args, bytes_alg: int, dim: int):
    """Time the oracle's TermDirect gather on a contiguous row sample."""
    import oracle

    threads = max(1, min(16, os.cpu_count() or 1))
    codes, coef = op.code_arrays()
    masks = oracle.matrix_terms(op.nqubits, codes, coef)
    probe = np.arange(4096, dtype=np.uint64)
    t0 = time.perf_counter()
    oracle.termdirect_rows(op.nqubits, codes, coef, v_host, probe, nthreads=threads, masks=masks)
    rate = 4096 / max(time.perf_counter() - t0, 1e-6)
    nrows = int(min(dim, max(4096, rate * args.cpu_seconds)))
    rows = np.arange(nrows, dtype=np.uint64)
    t0 = time.perf_counter()
    y_rows = oracle.termdirect_rows(op.nqubits, codes, coef, v_host, rows, nthreads=threads, masks=masks)
    elapsed = time.perf_counter() - t0
    full_seconds = elapsed * dim / nrows
    gpu_rows = y_dev_host_rows(nrows)
    err = float(np.linalg.norm(gpu_rows - y_rows) / max(np.linalg.norm(y_rows), 1e-300))
    return {
        "value": round(bytes_alg / full_seconds / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"TermDirect rows 0..{nrows - 1} of {dim} ({elapsed:.1f} s, extrapolated to one full MVP: "
                  f"{full_seconds:.1f} s)",
        "term_amp_per_s": round(op.term_count * nrows / elapsed, 1),
    }, {"rows": nrows, "rel_err": err}


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the sharded design is described in DESIGN.md)"}))
        return 2
    import torch

    import tencirpauli_amd as tcp

    torch.cuda.set_device(0)
    op, name, raw_terms = build_operator(args)
    t0 = time.perf_counter()
    plan = op.native_mvp_plan()
    plan_seconds = time.perf_counter() - t0
    n = op.nqubits
    dim = 1 << n
    cdtype = torch.complex128 if args.dtype == "complex128" else torch.complex64
    gen = torch.Generator(device="cuda").manual_seed(SEED + 1)
    v = torch.randn(dim, dtype=torch.complex128, device="cuda", generator=gen)
    v /= torch.linalg.vector_norm(v)
    v = v.to(cdtype)
    y = torch.empty_like(v)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        plan.apply_into(v, y)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        plan.apply_into(v, y)
    stop.record()
    torch.cuda.synchronize()
    ms_per_step = start.elapsed_time(stop) / args.steps

    tag = "complex64" if cdtype == torch.complex64 else "complex128"
    bytes_alg = plan.bytes_alg(tag)
    achieved = bytes_alg / (ms_per_step * 1e-3) / 1e9
    traffic = None
    if os.path.exists(args.traffic_json):
        with open(args.traffic_json) as fh:
            recorded = json.load(fh)
        traffic = recorded.get(tag, {}).get("hbm_bytes_per_launch")
    record = {
        "metric": METRIC,
        "value": round(achieved, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if tag == "complex128" else "f32-storage/f64-accumulate",
        "data": "synthetic (seeded lattice operator, Gaussian state generated on the GPU)",
        "config": {
            "workload": name,
            "nqubits": n,
            "raw_terms": raw_terms,
            "terms": plan.term_count,
            "x_groups": plan.group_count,
            "tile_classes": plan.class_count,
            "amplitudes": tag,
            "bytes_alg": bytes_alg,
        },
        "term_amp_per_s": round(plan.term_count * dim / (ms_per_step * 1e-3), 1),
        "plan_build_s": round(plan_seconds, 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        v_host = v.to(torch.complex128).cpu().numpy()
        y_full = y.to(torch.complex128)

        def gpu_rows(count: int) -> np.ndarray:
            return y_full[:count].cpu().numpy()

        record["cpu_baseline"], record["parity"] = cpu_baseline(op, v_host, gpu_rows, args, bytes_alg, dim)
    print(json.dumps(record), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
