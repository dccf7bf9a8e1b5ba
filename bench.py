"""Benchmark: amplitudes/s of the tensor-network contraction path on one MI355X.

Workload (default): 5x5-grid random circuit, depth 20, complex128, saved plan with
no sliced bonds, 256 seeded bitstrings (qxtools.jl_amd/workloads/rqc_5x5_d20.*).
A step is one `qx_amplitudes` call over all 256 bitstrings, the plan's tensors
already resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--precision c128|c64_4m|c64_3m] [--cpu-seconds S [--cpu-threads T] | --no-cpu]
"""

from __future__ import annotations

import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "qxtools.jl_amd"))

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"c128": 78.6, "c64_4m": 157.3, "c64_3m": 157.3, "c64_bf16x3": 2500.0}
ESIZE = {"c128": 16, "c64_4m": 8, "c64_3m": 8, "c64_bf16x3": 8}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_summary(stats: dict, prec: str) -> dict:
    """Per kernel family: launches, time, algorithmic bytes and flops (per-step
    counts from the plan; bytes = e*B*(MK+KN+MN), flops = 8*B*M*N*K)."""
    fam: dict = {}
    batch = next(b for b in stats["batches"] if b.get("active"))
    for st in batch["steps"]:
        calls = st.get("calls", 0)
        if not calls:
            continue
        f = fam.setdefault(st["variant"], {"launches": 0, "ms": 0.0, "bytes": 0.0, "flops": 0.0})
        f["launches"] += calls
        f["ms"] += st["ms"]
        f["bytes"] += st["bytes"] * calls
        f["flops"] += st["flops"] * calls
    return fam


def plan_roofline(stats: dict, prec: str, ms_per_step: float, steps: int) -> float:
    """sum_s max(flops_s/P, bytes_s/BW) over the launches of one step / measured."""
    batch = next(b for b in stats["batches"] if b.get("active"))
    t = 0.0
    for st in batch["steps"]:
        calls = st.get("calls", 0) / max(steps, 1)
        t += calls * max(st["flops"] / (PEAK_TFLOPS[prec] * 1e12), st["bytes"] / (HBM_PEAK_GBS * 1e9))
    return t / (ms_per_step * 1e-3)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(dsl: str, data: dict, bits: list, seconds: float, threads: int = 0, gpu_amps=None) -> dict:
    """The oracle (numpy TTGT, OpenBLAS zgemm) contracting whole amplitudes one
    bitstring at a time, as the reference's per-bitstring loop does, on `threads`
    BLAS threads (0: the host's cores, at most 16).  The amplitudes it computes are
    kept and compared with the GPU's for the same bitstrings (`gpu_amps`)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import qxoracle
    from threadpoolctl import threadpool_limits

    cores = threads if threads > 0 else min(os.cpu_count() or 1, 16)
    done = 0
    ref = []
    t0 = time.perf_counter()
    with threadpool_limits(limits=cores, user_api="blas"):
        while done == 0 or (time.perf_counter() - t0 < seconds and done < len(bits)):
            # the saved tensor's elements in column-major order, as the GPU returns them
            ref.append(np.asarray(qxoracle.amplitudes(dsl, data, [bits[done]]))[0].reshape(-1, order="F"))
            done += 1
            log(f"cpu baseline: {done} amplitude(s) in {time.perf_counter() - t0:.1f} s")
    el = time.perf_counter() - t0
    out = {"value": done / el, "unit": "amplitudes/s", "cores": cores, "kind": "port",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
           "sample": f"{done} of the workload's bitstrings, each contracted over the whole plan "
                     f"by the oracle ({el:.1f} s)"}
    if gpu_amps is not None:
        g = np.asarray(gpu_amps).reshape(len(bits), -1)[:done]
        r = np.stack(ref)
        out["gpu_vs_oracle_max_rel_err"] = float(np.max(np.abs(g - r) / np.abs(r)))
        out["compared_amplitudes"] = int(r.size)
    return out


def golden_check(workload: str, bits: list, amps) -> dict | None:
    """Per-amplitude relative error of the GPU result against the committed FP64
    statevector amplitudes of the workload, where such a fixture exists."""
    path = os.path.join(ROOT, "tests", "golden", f"{workload}_amplitudes.npz")
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        if list(z["bits"]) != list(bits):
            return None
        ref = z["sv_amps"]
    rel = np.abs(np.asarray(amps).reshape(-1) - ref) / np.abs(ref)
    return {"reference": "FP64 statevector (2^25 entries)", "amplitudes": int(ref.size),
            "max_rel_err": float(rel.max()), "p99_rel_err": float(np.quantile(rel, 0.99)),
            "median_rel_err": float(np.median(rel))}


def measured_traffic(workload: str, prec: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes scaled by a copy of known size), or None."""
    import glob
    import re

    here = os.path.dirname(os.path.abspath(__file__))
    for path in glob.glob(os.path.join(here, "profiles", "pmc_traffic_*.md")):
        m = re.search(r"```json\s*(\{.*?\})\s*```", open(path).read(), re.S)
        if not m:
            continue
        d = json.loads(m.group(1))
        if d.get("workload") == workload and d.get("precision") == prec:
            return d["traffic_bytes_per_launch"].get(kernel)
    return None


def c64_comparison(dsl: str, data: dict, bits: list, ref: np.ndarray, a, precs=("c64_4m", "c64_3m")) -> dict:
    """Other precisions on the same plan: throughput and per-amplitude relative error
    against `ref` (the complex128 GPU result).  A precision whose distinguishing product
    ran on no step of the plan (3M: the matrix-core GEMMs; split bf16: plain steps with
    M, N, K >= 1024) is the 4M path there and is left out of the line."""
    res = {}
    for prec in precs:
        amps, v, ms, var = precision_line(dsl, data, bits, prec, a.steps, a.warmup, False, graph=a.graph,
                                          autotune=a.autotune)
        if prec == "c64_3m" and not any(x.startswith(("k_plain_gemm", "k_contract_gemm_mfma")) for x in var):
            log("c64_3m: no step ran a three-product GEMM on this plan (same as 4M); left out")
            continue
        if prec == "c64_bf16x3" and not any("bf16x3" in x or x.startswith("rocblas") for x in var):
            log("c64_bf16x3: no step met the split product's size (same as 4M); left out")
            continue
        res[prec] = {"value": v, "ms_per_step": ms, "kernels": var, **error_stats(amps, ref.reshape(-1))}
        log(f"{prec}: {res[prec]['value']:.0f} amplitudes/s, max rel err {res[prec]['rel_err_max']:.2e}")
    return res


def c128_reference(dsl, data, bits, w):
    from qxhip.simulation import Executor

    ex = Executor(dsl, data, precision="c128", device=0, batch=min(256, len(bits)))
    r = np.asarray(ex.amplitudes(bits))
    ex.close()
    return r


def precision_line(dsl, data, bits, prec, steps, warmup, batch_last, graph=1, autotune=1, **opts) -> tuple:
    """(amplitudes, amplitudes/s, ms per call, kernel variants run) of one precision on
    a plan: an autotuning call, warm-up calls, then `steps` timed calls."""
    import torch

    from qxhip.simulation import Executor

    ex = Executor(dsl, data, precision=prec, device=0, batch=min(256, len(bits)), graph=graph, autotune=autotune,
                  batch_last=batch_last, **opts)
    try:
        if autotune:
            ex.amplitudes(bits)
        for _ in range(warmup):
            ex.amplitudes(bits)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            amps = ex.amplitudes(bits)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        batch = next(b for b in ex.stats()["batches"] if b.get("active"))
        variants = sorted({st["variant"] for st in batch["steps"]})
        per_call = len(bits) * ex.plan.result_elems
    finally:
        ex.close()
    return np.asarray(amps).reshape(-1), per_call * steps / el, 1e3 * el / steps, variants


def error_stats(amps, ref) -> dict:
    rel = np.abs(amps - ref) / np.abs(ref)
    return {"rel_err_max": float(rel.max()), "rel_err_p99": float(np.quantile(rel, 0.99)),
            "rel_err_median": float(np.median(rel)),
            "normwise_err": float(np.linalg.norm(amps - ref) / np.linalg.norm(ref))}


def batching_bench(a) -> dict:
    """The batching workload: 7x7 grid, depth 12 (the stated depth 24 has no plan that
    fits one GPU), 2^10 slices, 5 open output legs x 32 bitstring prefixes = 1024
    amplitudes per call.  Plain complex128, complex64 4M and the split bf16 product,
    each with per-amplitude error against complex128; and complex128 with one launch
    per contraction (no chain passes, no level batching, no graph) for comparison."""
    from qxhip import workloads

    w = workloads.WORKLOADS["rqc_7x7_d12_s10_open5"]
    dsl, data, bits = w.load()
    steps, warm = max(1, min(a.steps, 5)), 1
    out = {"workload": w.name, "qubits": w.qubits, "depth": w.depth, "sliced_bonds": w.sliced,
           "amplitudes_per_call": None, "note": w.note}
    ref, v, ms, var = precision_line(dsl, data, bits, "c128", steps, warm, w.batch_last)
    out["amplitudes_per_call"] = int(ref.size)
    out["c128"] = {"value": v, "ms_per_step": ms, "kernels": var}
    log(f"batching workload c128: {v:.0f} amplitudes/s")
    for prec in ("c64_4m", "c64_bf16x3"):
        amps, v, ms, var = precision_line(dsl, data, bits, prec, steps, warm, w.batch_last)
        out[prec] = {"value": v, "ms_per_step": ms, "vs_c128": v / out["c128"]["value"], "kernels": var,
                     **error_stats(amps, ref)}
        log(f"batching workload {prec}: {v:.0f} amplitudes/s, max rel err {out[prec]['rel_err_max']:.2e}")
    amps, v, ms, var = precision_line(dsl, data, bits, "c128", steps, 0, w.batch_last, graph=0, autotune=0,
                                      chains=0)
    out["c128_one_launch_per_contraction"] = {"value": v, "ms_per_step": ms,
                                              "batched_speedup": out["c128"]["value"] / v,
                                              "max_rel_diff": float(np.max(np.abs(amps - ref) / np.abs(ref)))}
    log(f"batching workload c128 one launch per contraction: {v:.0f} amplitudes/s")
    return out


def grad_bench(a) -> dict:
    """Config-5 style: loss = sum |amp|^2 over the bitstrings of a parameterised
    circuit; forward amplitudes/s and gradient / forward time."""
    import torch

    from qxhip import gradients, workloads

    w = workloads.PARAM_WORKLOAD
    prec = a.precision or w.precision
    pc = gradients.ParamCircuit(w.circuit(), planner="greedy", nslice=w.sliced, precision=prec,
                                batch=min(256, w.num_bitstrings), **w.plan_kwargs)
    bits = w.bitstrings()
    for _ in range(a.warmup):   # both paths: the forward alone and loss + gradient
        pc.amplitudes(bits)
        pc.loss_and_grad(bits)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        pc.amplitudes(bits)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / a.steps
    t0 = time.perf_counter()
    for _ in range(a.steps):
        loss, grad, _ = pc.loss_and_grad(bits)
    torch.cuda.synchronize()
    both = (time.perf_counter() - t0) / a.steps
    amps = pc.amplitudes(bits)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        kg = pc.executor.vjp(bits, 2.0 * amps)
    torch.cuda.synchronize()
    pull = (time.perf_counter() - t0) / a.steps
    t0 = time.perf_counter()
    for _ in range(a.steps):
        pc.param_grad(kg)
    chain = (time.perf_counter() - t0) / a.steps
    return {"metric": "gradient time / forward time", "value": (both - fwd) / fwd, "unit": "ratio",
            "higher_is_better": False, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "forward_ms": 1e3 * fwd, "forward_plus_gradient_ms": 1e3 * both,
            "pullback_ms": 1e3 * pull, "host_chain_rule_ms": 1e3 * chain,
            "forward_amplitudes_per_s": len(bits) / fwd, "params": int(pc.num_params),
            "dtype": "c128" if prec == "c128" else "c64", "data": "synthetic",
            "config": {"workload": w.name, "qubits": w.qubits, "depth": w.depth, "bitstrings": len(bits),
                       "sliced_bonds": w.sliced, "note": w.note}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rqc_5x5_d20")
    ap.add_argument("--precision", default=None)
    ap.add_argument("--batch", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="BLAS threads of the CPU baseline (default: the host's cores, at most 16)")
    ap.add_argument("--graph", type=int, default=1, help="replay each call from a captured hipGraph")
    ap.add_argument("--no-c64", action="store_true", help="skip the complex64 comparison")
    ap.add_argument("--grad", action="store_true", help="gradient / forward time of the parameterised workload")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batching workload and the gradient ratio in the headline line")
    ap.add_argument("--batch-last", type=int, default=None,
                    help="lay the bitstring batch leg out slowest in intermediates (default: per workload)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="time every candidate kernel per step in one untimed call first, keep the fastest")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the steps, copy this many MiB device to device (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the contraction path on a single "
              "MI355X (multi-GPU slice sharding is a design, not a measured path)", file=sys.stderr)
        return 2

    import numpy as np  # noqa: F811
    import torch

    from qxhip import workloads
    from qxhip.simulation import Executor, context

    if a.grad:
        print(json.dumps(grad_bench(a)), flush=True)
        return 0
    w = workloads.WORKLOADS[a.workload]
    prec = a.precision or w.precision
    dsl, data, bits = w.load()
    log(f"workload {w.name}: {len(bits)} bitstrings, precision {prec}")
    batch_last = w.batch_last if a.batch_last is None else a.batch_last
    ex = Executor(dsl, data, precision=prec, device=0, batch=a.batch or min(256, len(bits)), graph=a.graph,
                  autotune=a.autotune, batch_last=batch_last)
    if a.autotune:
        t_tune = time.perf_counter()
        ex.amplitudes(bits)
        log(f"autotuning call done in {time.perf_counter() - t_tune:.2f} s")
    for i in range(a.warmup):
        ex.amplitudes(bits)
        log(f"warmup {i + 1}/{a.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        amps = ex.amplitudes(bits)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    log(f"timed {a.steps} steps in {el:.3f} s (graph replay {'on' if a.graph else 'off'})")
    ms_step = 1e3 * el / a.steps
    per_call = len(bits) * ex.plan.result_elems      # open output legs: 2^m amplitudes per bitstring
    value = per_call * a.steps / el
    # kernel durations: the same steps launched eagerly with a HIP event pair around
    # every launch on the library's stream (graph replay records no events)
    ex.plan.set_option("profile", 1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(a.steps):
        ex.amplitudes(bits)
    torch.cuda.synchronize()
    el_ev = time.perf_counter() - t1
    log(f"event-timed pass: {a.steps} steps in {el_ev:.3f} s")
    stats = ex.stats()

    fam = kernel_summary(stats, prec)
    dom_name, dom = max(fam.items(), key=lambda kv: kv[1]["ms"])
    t_hbm = dom["bytes"] / (HBM_PEAK_GBS * 1e9)
    t_mfma = dom["flops"] / (PEAK_TFLOPS[prec] * 1e12)
    if t_hbm >= t_mfma:
        ach = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
    else:
        ach = dom["flops"] / (dom["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                "frac": ach / PEAK_TFLOPS[prec]}
    roof.update({"traffic": measured_traffic(w.name, prec, dom_name), "kernel": dom_name, "launches": dom["launches"],
                 "avg_launch_us": 1e3 * dom["ms"] / dom["launches"],
                 "algorithmic_bytes_per_launch": dom["bytes"] / dom["launches"],
                 "algorithmic_flops_per_launch": dom["flops"] / dom["launches"]})
    kernels = {k: {"launches": v["launches"], "ms_per_step": v["ms"] / a.steps} for k, v in fam.items()}

    out = {
        "metric": "amplitudes/s", "value": value, "unit": "amplitudes/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128" if prec == "c128" else "c64",
        "data": "synthetic: seeded random grid circuit and seeded bitstrings",
        "config": {"workload": w.name, "qubits": w.qubits, "depth": w.depth, "bitstrings": len(bits),
                   "amplitudes_per_step": per_call, "note": w.note or None,
                   "sliced_bonds": w.sliced, "precision": prec, "autotune": bool(a.autotune),
                   "batch_leg_last": bool(batch_last),
                   "batch": next(b for b in stats["batches"] if b.get("active"))["batch"]},
        "roofline": roof,
        "plan_roofline_frac": plan_roofline(stats, prec, ms_step, a.steps),
        "ms_per_step_eager_with_events": 1e3 * el_ev / a.steps,
        "kernels": kernels,
    }
    # the device's own ceilings beside the published peaks: a plain copy of known size
    # (its HBM counters also calibrate those of the contraction kernels: bytes read =
    # bytes written = size) and register-only matrix-core loops
    mib = a.copy_mib or 1024
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    # read + write bytes per copy / time: the practical HBM ceiling of a stream
    copy_gbps = 2 * (mib << 20) * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    ctx = context(0)
    out["measured_peaks"] = {"copy_gbps": copy_gbps, "copy_mib": mib,
                             "fp64_mfma_tflops": ctx.measure_mfma("c128"),
                             "fp32_mfma_tflops": ctx.measure_mfma("c64_4m")}
    if a.copy_mib:
        out["copy_calibration_bytes"] = a.copy_mib << 20
        out["copy_gbps"] = copy_gbps
    if not a.no_c64 and prec == "c128":
        out["c64"] = c64_comparison(dsl, data, bits, np.asarray(amps), a, precs=("c64_4m", "c64_3m", "c64_bf16x3"))
    elif not a.no_c64:
        ref = c128_reference(dsl, data, bits, w)
        out["error_vs_c128"] = c64_comparison(dsl, data, bits, ref, a, precs=(prec,))[prec]
        out["error_vs_c128"].pop("value"), out["error_vs_c128"].pop("ms_per_step")
    if not a.no_extras:
        out["batching"] = batching_bench(a)
        g = grad_bench(a)
        out["gradient"] = {k: g[k] for k in ("value", "forward_ms", "forward_plus_gradient_ms", "pullback_ms",
                                              "params", "config")}
        out["gradient"]["metric"] = g["metric"]
        log(f"gradient / forward: {g['value']:.2f}")
    gold = golden_check(w.name, bits, amps)
    if gold is not None:
        out["parity_vs_statevector"] = gold
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(dsl, data, bits, a.cpu_seconds, a.cpu_threads, gpu_amps=amps)
    print(json.dumps(out), flush=True)
    # the plain complex128 path must match the FP64 references it is checked against
    if prec == "c128":
        errs = [gold["max_rel_err"]] if gold else []
        if "cpu_baseline" in out:
            errs.append(out["cpu_baseline"]["gpu_vs_oracle_max_rel_err"])
        if errs and max(errs) > 1e-10:
            log(f"parity failure: GPU amplitudes differ from the FP64 references by {max(errs):.3e}")
            return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
