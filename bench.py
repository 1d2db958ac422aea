#!/usr/bin/env python
"""Throughput of the ensemble membrane solid step on one MI355X.

Workload (default): the 102,400-node tube (ntheta 256 x nz 400, R 0.5 cm, L 6 cm, 204,288
triangles, 715,776 3x3 blocks) with 128 GMRF Young's-modulus samples, Rayleigh beta = 1e-4,
dt = 1.2e-8 s, FP64, stored block-CSR with sample-fastest values.  A "step" is one fused kernel
launch advancing all 128 samples by one time step.

Output: one JSON line with sample-steps/s (``value``), the roofline of the step kernel (algorithmic
HBM bytes per launch / average launch time from HIP events on the launch stream), the matrix-free
variant measured the same way, and the CPU baseline (the C/OpenMP restatement of the reference CPU
step timed on a bounded sample on this host) plus a full-size parity check of the GPU state against
it after the baseline's steps.  ``roofline.traffic`` is the HBM traffic per launch measured with
rocprofv3 counters (``profiles/pmc_traffic.json``, scaled by the same counters on a plain copy).

Other workloads (``--workload``):
  batch    the uneven 64-tube ensemble advanced by one persistent launch per step (HIP-graph blocks
           of 99 steps, autotuned grid), against per-mesh launches of the plain stored step and of
           the matrix-free step (``--batch-fit``: sample counts scaled so all stored values fit);
  adjoint  forward + adjoint gradient of the 49,152-node tube (32 samples, 2,000 steps, 20
           observations, checkpoints every 100 steps) and the adjoint / forward time ratio;
  combined the stenosed 512,000-node tapered tube with 256 samples and beta = 1e-4: stored and
           matrix-free steps timed, then 2,000 forward steps and the adjoint gradient with two-level
           checkpoints (every 200 steps, then every 20 inside a segment) and the observations in
           host memory, sized to fit 288 GB;
  copy     the hand-written streaming kernels (16-byte non-temporal and ordinary reads, a copy)
           over ``--copy-bytes``: the streaming rate the box reaches, and the calibration kernels
           for the traffic counters;
  updated  the updated-geometry step (stiffness and follower load on X + u, re-formed every step
           as the shipped CVFES cfgs do, and every 10 steps) of the 102,400-node tube, each kernel
           form timed, with HBM and FP64 rooflines.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def reads_upper_blocks(sysm):
    """Whether the stored step of ``sysm`` reads the symmetric (upper-block) storage: the system
    holds only that storage, or its step kernel is one of the symmetric variants."""
    from cvfes_amd import _lib
    kern = int(sysm.lib.cvfes_step_kernel(sysm._sys, _lib.METHOD_STORED))
    return sysm.values is None or kern in (_lib.VARIANT_SYM, _lib.VARIANT_SYM_NT, _lib.VARIANT_SYM_NTL)


def algorithmic_bytes(sysm, method):
    """HBM bytes one step kernel must move at minimum (each datum once)."""
    N, S, nnzb, E = sysm.n_nodes, sysm.n_smp, sysm.nnzb, sysm.n_elms
    state = 3 * (3 * N * S * 8)                 # read u, read up, write u+
    lm = sysm.lm.numel() * 8
    small = 3 * N * 8 + 4 * (N + 1) + 4 * nnzb + N   # f1, indptr, indices, fixed
    if getattr(sysm, "row_order", None) is not None:
        small += 4 * N                                 # the visiting order
    if method == "stored" and reads_upper_blocks(sysm):                  # symmetric storage
        op = sysm.nnzu * 9 * S * 8 + 4 * nnzb                          # upper values, vblk
    elif method == "stored":
        op = nnzb * 9 * S * 8
    else:
        op = E * S * 8 + 9 * E * 9 * 8 + 4 * 9 * E + 4 * (nnzb + 1)   # coef, khat, celem, cptr
    return state + lm + small + op


def kernel_name(sysm, method):
    """Name of the step kernel ``cvfes_step`` launches for ``sysm`` and ``method``."""
    from cvfes_amd import _lib
    kern = int(sysm.lib.cvfes_step_kernel(sysm._sys, _lib.METHODS[method]))
    names = {v: k for k, v in _lib.VARIANTS.items()}
    return f"{method} step, variant {names.get(kern, kern)}"


def traffic_key(method, storage, variant, order):
    """Key of a step kernel's counter summary in ``profiles/pmc_traffic.json``."""
    key = method if method == "matfree" else f"stored-{storage}"
    if variant != "auto":
        key += f"-{variant}"
    return key if order == "natural" else f"{key}-{order}"


def measured_traffic(workload, kernel):
    """HBM bytes per launch from the committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
        return table[workload][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def time_steps(sysm, load, dt, beta, method, warmup, steps):
    import torch
    u0 = np.zeros(3 * sysm.n_nodes)
    sysm.initial_state(u0, u0, load, dt, beta=beta, method=method)
    sysm.step(load, dt, beta=beta, method=method, offset=0, nsteps=warmup)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[2 * k].record(stream)
        sysm.step(load, dt, beta=beta, method=method, offset=warmup + k, nsteps=1)
        ev[2 * k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([ev[2 * k].elapsed_time(ev[2 * k + 1]) for k in range(steps)]))
    u = sysm.u
    if not bool(torch.isfinite(u).all()):
        raise RuntimeError(f"{method}: non-finite displacement after {warmup + steps} steps")
    return wall, kern_ms


def cpu_baseline(case, sysm, nodes, el, bdy, coefK, dt, beta, steps, threads, warmup=2,
                 one_thread_steps=3):
    """C/OpenMP restatement of the reference CPU step on a bounded sample; returns (result, u_cpu).

    ``steps`` steps of the recurrence on ``threads`` threads, the first ``warmup`` untimed, then
    ``one_thread_steps`` more steps of the same recurrence on one thread."""
    from oracle import cpu_oracle as co
    from oracle import solid_oracle as so
    co.build()
    S = case.n_smp
    t_asm = time.perf_counter()
    K = co.assemble_K(nodes, el, so.d_matrix(case.nu), coefK, sysm.indptr_h, sysm.indices_h, threads)
    t_asm = time.perf_counter() - t_asm
    LM = sysm.lumped_mass_dofs()
    f1 = sysm.f1.cpu().numpy().ravel()
    p = case.pressures(0, steps + one_thread_steps, dt)
    ndof = 3 * sysm.n_nodes
    # u0 = 0 and du0 = 0: K u0 = 0, so u_-1 = dt^2/2 * f_0 / LM (physicsSolver.py:217-225).
    u = np.zeros((ndof, S))
    up = 0.5 * dt * dt * np.divide(p[0] * f1[:, None] * np.ones((1, S)), LM,
                                   out=np.zeros((ndof, S)), where=LM != 0)
    LHS = LM
    bdy_dofs = (3 * bdy[:, None] + np.arange(3)).ravel()
    times = []
    for n in range(steps):
        t0 = time.perf_counter()
        unew = co.step(sysm.indptr_h, sysm.indices_h, K, LM, LHS, f1, p[n], u, up, dt, beta=beta,
                       bdy_dofs=bdy_dofs, threads=threads)
        times.append(time.perf_counter() - t0)
        up, u = u, unew
    timed = times[warmup:] if len(times) > warmup else times
    rate = S / float(np.mean(timed))
    res = {"value": rate, "unit": "sample-steps/s", "cores": int(threads),
           "kind": "port",
           "sample": f"{case.name}: same mesh, {S} samples and inputs; {len(timed)} timed steps after "
                     f"{len(times) - len(timed)} warm-up ({float(np.sum(times)):.1f} s of CPU work); "
                     f"K assembled in the reference layout [nnzb][S][3][3] in {t_asm:.1f} s (not timed)",
           "step_s": [float(t) for t in timed]}
    if threads != 1 and one_thread_steps > 0:
        # the next steps of the recurrence on one thread (results discarded)
        a, b, t1 = u, up, []
        for n in range(steps, steps + one_thread_steps):
            t0 = time.perf_counter()
            c = co.step(sysm.indptr_h, sysm.indices_h, K, LM, LHS, f1, p[n], a, b, dt, beta=beta,
                        bdy_dofs=bdy_dofs, threads=1)
            t1.append(time.perf_counter() - t0)
            a, b = c, a
        res["one_thread"] = {"value": S / float(np.mean(t1)), "unit": "sample-steps/s", "cores": 1,
                             "steps": len(t1), "step_s": [float(t) for t in t1]}
    return res, u


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="plain",
                    choices=["plain", "batch", "adjoint", "combined", "copy", "updated"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--batch-fit", action="store_true",
                    help="batch: scale the sample counts so all stored values fit at once")
    ap.add_argument("--case", default="tube-256x400-s128")
    ap.add_argument("--method", default="stored", choices=["stored", "matfree"])
    ap.add_argument("--storage", default="sym", choices=["sym", "full"],
                    help="stored values: upper blocks only (sym) or every block (full)")
    ap.add_argument("--variant", default="auto", help="step kernel variant (cvfes_amd._lib.VARIANTS)")
    ap.add_argument("--waves-per-simd", default="auto", choices=["auto", "0", "2", "3", "4", "5", "6", "7"],
                    help="occupancy cap of the timed step (auto: the fastest of none, 6, 5, 4, 3, 2)")
    ap.add_argument("--row-order", default="auto", choices=["auto", "natural", "strips"],
                    help="row visiting order of the timed step (auto: the fastest of natural and "
                         "strips 32/64 on 20 steps each; the result does not depend on it)")
    ap.add_argument("--strip-width", type=int, default=32, help="strip width of --row-order strips")
    ap.add_argument("--cpu-steps", type=int, default=22,
                    help="CPU baseline steps (the first two are warm-up)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all OpenMP threads")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-matfree", action="store_true")
    ap.add_argument("--adjoint-block", type=int, default=64,
                    help="combined workload: reverse sweep by sample blocks of this size, one checkpoint "
                         "level (0 = the whole ensemble, two levels)")
    ap.add_argument("--adjoint-every", type=int, default=0, help="combined workload: checkpoint interval")
    ap.add_argument("--adjoint-host", action="store_true", help="combined workload: checkpoints in host memory")
    ap.add_argument("--adjoint-sequential", action="store_true",
                    help="combined workload: gradient steps on the main stream (default: beside the "
                         "adjoint steps)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint/forward ratio of the differentiation workload")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the ensemble step is sample-parallel with no per-step "
              "exchange, so an N-GPU run is N independent sample shards; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "adjoint":
        return run_adjoint(args)
    if args.workload == "copy":
        return run_copy(args)
    if args.workload == "combined":
        return run_combined(args)
    if args.workload == "updated":
        return run_updated(args)
    args.steps = 200 if args.steps is None else args.steps
    args.warmup = 20 if args.warmup is None else args.warmup

    import torch
    from cvfes_amd.configs import get_case
    from cvfes_amd.system import MembraneSystem
    from oracle import solid_oracle as so

    case = get_case(args.case)
    dt, beta = case.dt, case.beta
    t_setup = time.perf_counter()
    nodes, el, bdy = case.geometry()
    from cvfes_amd.gmrf import sample_element_E
    gm = sample_element_E(el, case.n_smp, seed=case.seed, E0=case.E0, cov=case.cov, device="cuda")
    N, S = nodes.shape[0], case.n_smp
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, case.thickness), elmE=gm.E,
                          density=case.density, nu=case.nu, store_values=True, storage=args.storage)
    sysm.set_variant(args.variant)
    total = args.warmup + args.steps
    load = sysm.loads(case.pressures(0, max(total, args.cpu_steps + 1), dt))
    if args.waves_per_simd == "auto":
        occ = sysm.autotune_occupancy(load, dt, beta, method=args.method, steps=20)
    else:
        sysm.set_occupancy(int(args.waves_per_simd))
        occ = {"waves_per_simd": int(args.waves_per_simd)}
    if args.row_order == "auto":
        order = sysm.autotune_row_order(load, dt, beta, method=args.method)
    else:
        sysm.set_row_order(args.row_order, width=args.strip_width)
        order = {"order": args.row_order, "width": args.strip_width if args.row_order == "strips" else 0}
    order_name = order["order"] + (str(order["width"]) if order["order"] == "strips" else "")
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] setup {t_setup:.1f} s: N={N} E={el.shape[0]} nnzb={sysm.nnzb} S={S} dt={dt}",
          file=sys.stderr, flush=True)

    wall, kern_ms = time_steps(sysm, load, dt, beta, args.method, args.warmup, args.steps)
    abytes = algorithmic_bytes(sysm, args.method)
    achieved = abytes / (kern_ms * 1e-3) / 1e9
    value = S * args.steps / wall
    print(f"[bench] {args.method}: {kern_ms:.4f} ms/kernel, {achieved:.0f} GB/s, "
          f"{value:.0f} sample-steps/s", file=sys.stderr, flush=True)

    out = {
        "metric": "ensemble sample-steps/s at 102k nodes x 128 samples; achieved HBM GB/s",
        "value": value, "unit": "sample-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded tube mesh and per-element GMRF Young's modulus samples",
        "config": {"workload": case.name, "n_nodes": N, "n_elements": int(el.shape[0]),
                   "nnzb": sysm.nnzb, "n_samples": S, "dt": dt, "beta": beta,
                   "operator": args.method, "variant": args.variant, "row_order": order_name,
                   "waves_per_simd": occ["waves_per_simd"],
                   "storage": args.storage if args.method == "stored" else None,
                   "step_kernel": kernel_name(sysm, args.method),
                   "layout": ("upper blocks uvalues [nnzu][9][S]" if args.storage == "sym"
                              else "values [nnzb][9][S]") + ", state [N][3][S]"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(case.name, traffic_key(args.method, args.storage,
                                                                        args.variant, order_name)),
                     "algorithmic_bytes_per_launch": abytes, "kernel_ms": kern_ms,
                     "kernel": kernel_name(sysm, args.method)},
    }
    # achievable streaming rate of this box with the step kernels' load forms (calibration kernels)
    ceil = stream_ceiling()
    out["stream_ceiling"] = ceil
    out["roofline"]["frac_of_stream_read_nt"] = achieved / ceil["read_nt_GBs"]
    print(f"[bench] streaming ceiling: {ceil['read_nt_GBs']:.0f} GB/s non-temporal reads, "
          f"{ceil['read_GBs']:.0f} GB/s reads, {ceil['copy_nt_GBs']:.0f} GB/s copy",
          file=sys.stderr, flush=True)
    if not args.no_matfree:
        # the other operator, the other storage and order of the stored one, the element-by-element
        # form, each measured the same way: (operator, storage, variant, row order)
        if args.method == "stored":
            other_storage = "full" if args.storage == "sym" else "sym"
            alts = [("matfree", None, "auto", "natural"), ("stored", other_storage, "auto", "natural")]
            if order_name != "natural":
                alts.append(("stored", args.storage, args.variant, "natural"))
            alts.append(("matfree", None, "ebe", "natural"))
        else:
            alts = [("stored", args.storage, "auto", "natural")]
        out["alternatives"] = []
        main_occ = sysm.waves_per_simd
        for other, storage, variant, aorder in alts:
            if storage is not None and storage != sysm.storage:
                sysm.set_storage(storage)
            sysm.set_variant(variant)
            sysm.set_row_order(aorder)
            aocc = sysm.autotune_occupancy(load, dt, beta, method=other, steps=20)
            msteps = max(20, args.steps // 2)
            mwall, mk = time_steps(sysm, load, dt, beta, other, args.warmup, msteps)
            mbytes = algorithmic_bytes(sysm, other)
            key = traffic_key(other, storage, variant, aorder)
            out["alternatives"].append({
                "operator": other, "storage": storage, "variant": variant, "row_order": aorder,
                "kernel": kernel_name(sysm, other),
                "waves_per_simd": aocc["waves_per_simd"], "value": S * msteps / mwall,
                "unit": "sample-steps/s", "kernel_ms": mk, "achieved_GBs": mbytes / (mk * 1e-3) / 1e9,
                "frac": mbytes / (mk * 1e-3) / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": mbytes,
                "traffic": measured_traffic(case.name, key)})
            print(f"[bench] {key}: {mk:.4f} ms/kernel, {S * msteps / mwall:.0f} sample-steps/s",
                  file=sys.stderr, flush=True)
        if sysm.storage != args.storage:
            sysm.set_storage(args.storage)
        sysm.set_variant(args.variant)
        sysm.set_occupancy(main_occ)
        sysm.set_row_order(order["order"], width=order["width"] or 32)
    if not args.no_cpu:
        from oracle import cpu_oracle as co
        threads = args.cpu_threads or co.max_threads()
        elmGE, elmGThick = so.gauss_properties(np.full((N, S), case.E0), np.full((N, S), case.thickness), el)
        coefK = so.coef_k_element_E(elmGThick, gm.E)
        del elmGE, elmGThick
        cres, u_cpu = cpu_baseline(case, sysm, nodes, el, bdy, coefK, dt, beta, args.cpu_steps, threads)
        out["cpu_baseline"] = cres
        # Full-size parity: the GPU state after the same number of steps from the same start.
        u0 = np.zeros(3 * N)
        sysm.initial_state(u0, u0, load, dt, beta=beta, method=args.method)
        sysm.step(load, dt, beta=beta, method=args.method, offset=0, nsteps=args.cpu_steps)
        ug = sysm.u.cpu().numpy().reshape(-1, S)
        out["parity_vs_cpu"] = {"steps": args.cpu_steps,
                                "rel_max_err": float(np.max(np.abs(ug - u_cpu)) / np.max(np.abs(u_cpu)))}
        one = cres.get("one_thread", {}).get("value")
        print(f"[bench] cpu baseline {cres['value']:.1f} sample-steps/s on {threads} threads"
              + (f" ({one:.1f} on one)" if one else "") + "; "
              f"parity {out['parity_vs_cpu']['rel_max_err']:.2e}", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        # the differentiation workload's cost ratio (adjoint gradient / forward solve)
        r = adjoint_measure("tube-192x256-s32", "stored")
        out["adjoint_over_forward"] = {k: r[k] for k in (
            "workload", "n_samples", "steps", "observations", "checkpoint_every", "ratio",
            "forward_s", "adjoint_s", "ratio_separate_kernels", "fused_steps", "fused_equals_separate")}
        print(f"[bench] adjoint / forward {r['ratio']:.2f} ({r['forward_s']:.3f} s forward)",
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)



def stream_ceiling(nbytes=4 << 30, reps=5):
    """Streaming rates of the hand-written calibration kernels (``cvfes_stream_read`` / ``_copy``,
    the step kernels' 16-byte load forms) over ``nbytes``, timed with HIP events on the launch
    stream: GB/s for non-temporal and ordinary reads and for a copy (read + write counted)."""
    import torch
    from cvfes_amd._lib import check, lib
    h = lib()
    n = (nbytes // 16) * 2
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    part = torch.empty(int(h.cvfes_stream_blocks(n)), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    sh = st.cuda_stream
    out = {"bytes": 8 * n}
    for name, fn, nt, moved in (("read_nt", h.cvfes_stream_read, 1, 8 * n),
                                ("read", h.cvfes_stream_read, 0, 8 * n),
                                ("copy_nt", h.cvfes_stream_copy, 1, 16 * n)):
        target = part if fn is h.cvfes_stream_read else dst
        for _ in range(2):
            check(fn(sh, src.data_ptr(), n, nt, target.data_ptr()))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            check(fn(sh, src.data_ptr(), n, nt, target.data_ptr()))
        e1.record(st)
        torch.cuda.synchronize()
        out[f"{name}_GBs"] = moved / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    if abs(float(part.sum()) - n) > 1e-6 * n or not bool(torch.equal(dst, src)):
        raise RuntimeError("streaming calibration kernels returned wrong results")
    del src, dst, part
    torch.cuda.empty_cache()
    return out


def run_copy(args):
    """Calibration: the hand-written streaming kernels over ``--copy-bytes`` (see
    :func:`stream_ceiling`), and torch's ``dst = src * 1`` for comparison."""
    import torch
    res = {"workload": "copy"}
    res.update(stream_ceiling(args.copy_bytes))
    n = args.copy_bytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = (torch.cuda.Event(enable_timing=True) for _ in range(2))
    e0.record(st)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    e1.record(st)
    torch.cuda.synchronize()
    res["torch_mul_copy_GBs"] = 16 * n / (e0.elapsed_time(e1) / 5e3) / 1e9
    print(json.dumps(res), flush=True)


def run_batch(args):
    """Uneven 64-tube ensemble: batched persistent launches vs per-mesh launches.

    Default: the configured sample counts (8..256 per mesh).  The batched form is matrix-free
    (the stored values of all meshes would not fit HBM together), so the per-mesh plain (stored)
    step is timed mesh by mesh, each mesh with its own assembled values, and summed: that is the
    cost of per-mesh plain launches.  ``--batch-fit`` instead scales the sample counts so that the
    stored values of every mesh fit at once (150 GB) and times all forms side by side."""
    import torch
    from cvfes_amd.batch import BatchedEnsemble
    from cvfes_amd.configs import batch_cases
    from cvfes_amd.gmrf import sample_element_E
    from cvfes_amd.system import MembraneSystem

    steps = 297 if args.steps is None else args.steps
    block = 99
    if steps % block:
        raise SystemExit("--steps must be a multiple of 99 (graph blocks of 99 steps)")
    ref_steps = max(33, steps // 3)
    fit = args.batch_fit
    t0 = time.perf_counter()
    cases = batch_cases(byte_budget=150e9 if fit else float("inf"))
    systems, loads, dts = [], [], []
    for i, c in enumerate(cases):
        nodes, el, bdy = c.geometry()
        E = sample_element_E(el, c.n_smp, seed=c.seed, device="cuda").E
        s = MembraneSystem(nodes, el, bdy, c.n_smp, vthickness=np.full(nodes.shape[0], c.thickness),
                           elmE=E, density=c.density, nu=c.nu, store_values=fit)
        dt = 0.8 * s.critical_dt(iters=30, method="matfree")[0]
        systems.append(s)
        dts.append(dt)
        loads.append(s.loads(c.pressures(0, 3 * steps + 6 * ref_steps + 600, dt)))
        if i % 16 == 0:
            print(f"[bench] batch mesh {i}: N={c.n_nodes} S={c.n_smp} ({time.perf_counter() - t0:.0f} s)",
                  file=sys.stderr, flush=True)
    ens = BatchedEnsemble(systems, loads, dts=dts, methods="matfree")
    ens.initial_state()
    tuned = ens.autotune()
    print(f"[bench] batch setup {time.perf_counter() - t0:.0f} s, tasks {ens.n_tasks}, tuned {tuned}",
          file=sys.stderr, flush=True)

    def timed(fn, n):
        fn(n)
        torch.cuda.synchronize()
        a = time.perf_counter()
        fn(n)
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / n

    sum_s = sum(c.n_smp for c in cases)
    t_batch = timed(lambda n: ens.run(n, graph=True, block=block), steps)
    print(f"[bench] batched {1e3 * t_batch:.3f} ms/step", file=sys.stderr, flush=True)
    t_mf = timed(lambda n: ens.run_per_mesh(n, method="matfree"), ref_steps)
    print(f"[bench] per-mesh matrix-free {1e3 * t_mf:.3f} ms/step", file=sys.stderr, flush=True)
    if fit:
        t_plain = timed(lambda n: ens.run_per_mesh(n, method="stored"), ref_steps)
    else:
        # mesh by mesh: assemble the mesh's values, time its plain step, release the values
        t_plain = 0.0
        for s, ld, dt in zip(systems, loads, dts):
            s.assemble_values()
            z = np.zeros(3 * s.n_nodes)
            s.initial_state(z, z, ld, dt, method="stored")
            s.step(ld, dt, method="stored", offset=0, nsteps=3)
            torch.cuda.synchronize()
            a = time.perf_counter()
            s.step(ld, dt, method="stored", offset=3, nsteps=20)
            torch.cuda.synchronize()
            t_plain += (time.perf_counter() - a) / 20
            s.release_values()
            torch.cuda.empty_cache()
    print(f"[bench] per-mesh plain {1e3 * t_plain:.3f} ms/step", file=sys.stderr, flush=True)
    # parity: the batched launch against per-mesh launches of the same operator (bit for bit)
    ens.initial_state()
    ens.run(2 * block, graph=True, block=block)
    batched = [s.u.cpu().numpy() for s in systems]
    ens.initial_state()
    ens.run_per_mesh(2 * block, method="matfree")
    worst = max(float(np.max(np.abs(a - s.u.cpu().numpy())) / np.max(np.abs(s.u.cpu().numpy())))
                for a, s in zip(batched, systems))
    out = {
        "metric": "ensemble sample-steps/s, 64 uneven tubes in one persistent launch per step",
        "value": sum_s / t_batch, "unit": "sample-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": steps, "ms_per_step": 1e3 * t_batch, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic: seeded tubes and GMRF samples",
        "config": {"workload": "batch-64-uneven-tubes" + ("-fit" if fit else ""), "meshes": len(cases),
                   "sum_samples": sum_s, "node_samples": int(sum(c.n_smp * c.n_nodes for c in cases)),
                   "tasks": ens.n_tasks, "grid_waves": tuned["n_waves"], "graph_block": block,
                   "operator": "matfree"},
        "per_mesh_plain_ms": 1e3 * t_plain, "per_mesh_matfree_ms": 1e3 * t_mf,
        "per_mesh_plain_timing": "all meshes resident" if fit else "mesh by mesh, own values each",
        "speedup_vs_per_mesh_plain": t_plain / t_batch,
        "speedup_vs_per_mesh_matfree": t_mf / t_batch,
        "parity_vs_per_mesh_matfree": {"steps": 2 * block, "rel_max_err": worst},
    }
    print(json.dumps(out), flush=True)


# FP64 operations of the updated-geometry step per (row node, sample, incident element), FMA = 2,
# square root and division = 1 each: frame (edge vectors 6, two normalisations 2 x (5 + 1 + 1 + 3),
# projection 5 + 6, cross product 9), local coordinates 30, area 12, edge terms 6, element scale 4,
# three blocks applied in the local frame 3 x 46, accumulation 21, follower load 9.
UPDATED_FLOPS_PER_ROW_ELEMENT = 272
UPDATED_FLOPS_PER_ROW = 20          # the central-difference update of the row (3 components)


def updated_algorithmic_bytes(sysm, interval):
    """HBM bytes one updated-geometry step must move at minimum: coefficients, the three state
    arrays, the geometry displacement when it is a saved copy (interval > 1), the reference
    coordinates, connectivity, node-element lists, lumped mass and boundary flags (the element
    blocks and the load are formed from the geometry, not read)."""
    N, S, E = sysm.n_nodes, sysm.n_smp, sysm.n_elms
    b = E * S * 8 + 3 * (3 * N * S * 8)                     # coef; read u, up, write u+
    if interval > 1:
        b += 3 * N * S * 8                                 # ug
    b += 24 * N + 12 * E + 4 * (N + 1) + 12 * E            # nodes, elems, nptr, nent
    b += sysm.lm.numel() * 8 + N                           # lumped mass, fixed
    return b


def run_updated(args):
    """Updated-geometry step (stiffness and follower load on X + u) of the 102,400-node tube with
    128 samples: every kernel form at re-formation interval 1 (every step, what the shipped CVFES
    cfgs select) and the fastest at interval 10; HBM and FP64 rooflines."""
    import torch
    from cvfes_amd.configs import get_case
    from cvfes_amd.gmrf import sample_element_E
    from cvfes_amd.system import MembraneSystem

    case = get_case(args.case)
    nodes, el, bdy = case.geometry()
    N, S, E = nodes.shape[0], case.n_smp, el.shape[0]
    gm = sample_element_E(el, S, seed=case.seed, E0=case.E0, cov=case.cov, device="cuda")
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, case.thickness), elmE=gm.E,
                          density=case.density, nu=case.nu, store_values=False)
    dt, beta = case.dt, case.beta
    steps = 100 if args.steps is None else args.steps
    warm = 5 if args.warmup is None else args.warmup
    load = sysm.loads(case.pressures(0, 2 * (warm + steps) + 10, dt))
    flops = (UPDATED_FLOPS_PER_ROW_ELEMENT * 3 * E + UPDATED_FLOPS_PER_ROW * N) * S
    st = torch.cuda.current_stream()
    res = {}
    for interval in (1, 10):
        forms = ("generic", "rowload1", "rowload") if interval == 1 else (best_form,)
        for form in forms:
            sysm.set_variant(form)
            z = np.zeros(3 * N)
            sysm.initial_state(z, z, load, dt, beta=beta, method="matfree")
            sysm.step_updated(load, dt, interval, beta=beta, offset=0, nsteps=warm + 1)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(st)
            sysm.step_updated(load, dt, interval, beta=beta, offset=warm + 1, nsteps=steps)
            e1.record(st)
            torch.cuda.synchronize()
            if not bool(torch.isfinite(sysm.u).all()):
                raise RuntimeError(f"updated step ({form}): non-finite displacement")
            ms = e0.elapsed_time(e1) / steps
            abytes = updated_algorithmic_bytes(sysm, interval)
            res[f"interval{interval}-{form}"] = {
                "kernel_ms": ms, "sample_steps_per_s": S / (ms * 1e-3),
                "hbm": {"achieved_GBs": abytes / (ms * 1e-3) / 1e9, "frac": abytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_launch": abytes},
                "fp64": {"achieved_TFLOPs": flops / (ms * 1e-3) / 1e12,
                         "frac": flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "flops_per_launch": flops}}
            print(f"[bench] updated interval {interval} {form}: {ms:.4f} ms/step, "
                  f"{flops / (ms * 1e-3) / 1e12:.1f} TFLOP/s", file=sys.stderr, flush=True)
        if interval == 1:
            best_form = min(("generic", "rowload1", "rowload"), key=lambda f: res[f"interval1-{f}"]["kernel_ms"])
    sysm.set_variant("auto")
    best = res[f"interval1-{best_form}"]
    out = {"metric": "updated-geometry ensemble sample-steps/s at 102k nodes x 128 samples",
           "value": best["sample_steps_per_s"], "unit": "sample-steps/s", "n_gpus": 1, "steps": steps,
           "warmup": warm, "ms_per_step": best["kernel_ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic: seeded tube mesh and GMRF samples",
           "config": {"workload": case.name + "-updated", "interval": 1, "kernel_form": best_form,
                      "beta": beta, "dt": dt},
           "roofline": {"bound": "fp64", "achieved": best["fp64"]["achieved_TFLOPs"],
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": best["fp64"]["frac"],
                        "hbm_frac": best["hbm"]["frac"], "flops_per_launch": flops},
           "forms": res}
    print(json.dumps(out), flush=True)


def adjoint_measure(case_name="tube-192x256-s32", method="stored"):
    """Forward solve and adjoint gradient of the displacement misfit (20 observation times,
    checkpoints every 100 steps), timed: the pipelined (fused) reverse sweep and the one with
    separate kernels, whose gradients must agree bit for bit."""
    import torch
    from cvfes_amd.adjoint import AdjointProblem
    from cvfes_amd.configs import get_case
    from cvfes_amd.gmrf import sample_element_E
    from cvfes_amd.system import MembraneSystem

    case = get_case(case_name)
    nodes, el, bdy = case.geometry()
    N, S = nodes.shape[0], case.n_smp
    gm = sample_element_E(el, S, seed=case.seed, E0=case.E0, cov=case.cov, device="cuda")
    theta = np.log(gm.E)
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, case.thickness), elmE=gm.E,
                          density=case.density, nu=case.nu)
    nsteps, dt = case.nsteps, case.dt
    load = sysm.loads(case.pressures(0, nsteps))
    obs = list(range(nsteps // 20, nsteps + 1, nsteps // 20))
    theta_obs = np.log(sample_element_E(el, S, seed=case.seed + 100, E0=case.E0, cov=case.cov,
                                        tau=gm.tau, device="cuda").E)
    sysm.set_log_modulus(torch.as_tensor(theta_obs).cuda())
    u0 = np.zeros(3 * N)
    sysm.initial_state(u0, u0, load, dt, method=method)
    states, done = [], 0
    for k in obs:
        sysm.step(load, dt, method=method, offset=done, nsteps=k - done)
        done = k
        states.append(sysm.u.clone())
    prob = AdjointProblem(sysm, load, nsteps, dt, method=method, obs_steps=obs,
                          u_obs=torch.stack(states), checkpoint_every=100)
    prob.set_theta(torch.as_tensor(theta).cuda())
    prob.forward()
    prob.backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prob.forward()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g = prob.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd, adj = t1 - t0, t2 - t1
    fused_steps = prob.stats.get("fused_steps", 0)
    fusion_timing = prob.stats.get("fusion_timing_ms")
    # the same reverse sweep with recomputation and adjoint steps in separate kernels
    prob.fuse = False
    prob.forward()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    g_sep = prob.backward()
    torch.cuda.synchronize()
    adj_sep = time.perf_counter() - t3
    return {"workload": case.name, "n_nodes": N, "n_samples": S, "steps": nsteps,
            "observations": len(obs), "checkpoint_every": 100, "operator": method,
            "storage": sysm.storage if method == "stored" else None,
            "forward_s": fwd, "adjoint_s": adj, "ratio": adj / fwd,
            "forward_sample_steps_per_s": S * nsteps / fwd,
            "gradient_max_abs": float(g.abs().max()), "fused_steps": fused_steps,
            "fusion_timing_ms": fusion_timing, "adjoint_separate_kernels_s": adj_sep,
            "ratio_separate_kernels": adj_sep / fwd, "fused_equals_separate": bool(torch.equal(g, g_sep))}


def run_adjoint(args):
    """Forward and adjoint gradient of the displacement misfit on the 49,152-node tube."""
    case = args.case if args.case != "tube-256x400-s128" else "tube-192x256-s32"
    r = adjoint_measure(case, args.method)
    out = {
        "metric": "adjoint gradient time / forward solve time (49k-node tube, 32 samples, 2,000 steps)",
        "value": r["ratio"], "unit": "ratio", "n_gpus": 1, "steps": r["steps"], "warmup": r["steps"],
        "ms_per_step": 1e3 * (r["forward_s"] + r["adjoint_s"]) / r["steps"], "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded tube and GMRF samples",
        "config": {k: r[k] for k in ("workload", "n_nodes", "n_samples", "observations",
                                     "checkpoint_every", "operator", "storage")},
    }
    out.update({k: v for k, v in r.items() if k not in out["config"]})
    print(json.dumps(out), flush=True)


def run_combined(args):
    """Stenosed 512,000-node tube, 256 samples: stored vs matrix-free steps, forward + adjoint."""
    import torch
    from cvfes_amd.adjoint import AdjointProblem
    from cvfes_amd.configs import get_case
    from cvfes_amd.gmrf import sample_element_E
    from cvfes_amd.system import MembraneSystem

    log = lambda msg: print(f"[bench] combined {msg}", file=sys.stderr, flush=True)
    t_start = time.perf_counter()
    case = get_case("stenosed-512x1000-s256")
    nodes, el, bdy = case.geometry()
    N, S, E = nodes.shape[0], case.n_smp, el.shape[0]
    gm = sample_element_E(el, S, seed=case.seed, E0=case.E0, cov=case.cov, device="cuda")
    theta = np.log(gm.E)
    log(f"GMRF samples ({time.perf_counter() - t_start:.0f} s)")
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, case.thickness), elmE=gm.E,
                          density=case.density, nu=case.nu, store_values=True)
    beta = case.beta
    dt = case.dt or 0.8 * sysm.critical_dt(beta=beta, iters=40, method="matfree")[0]
    nsteps = case.nsteps if args.steps is None else args.steps
    load = sysm.loads(case.pressures(0, nsteps + 100, dt))
    log(f"assembled N={N} E={E} nnzb={sysm.nnzb} S={S} dt={dt:.4e} ({time.perf_counter() - t_start:.0f} s)")
    per_step = {}
    finals = {}
    for method in ("stored", "matfree"):
        occ = sysm.autotune_occupancy(load, dt, beta, method=method, steps=5)
        wall, kern_ms = time_steps(sysm, load, dt, beta, method, 5, 50)
        per_step[method] = {"kernel_ms": kern_ms, "sample_steps_per_s": S / (wall / 50),
                            "achieved_GBs": algorithmic_bytes(sysm, method) / (kern_ms * 1e-3) / 1e9,
                            "algorithmic_bytes_per_launch": algorithmic_bytes(sysm, method),
                            "waves_per_simd": occ["waves_per_simd"]}
        finals[method] = sysm.u.cpu().numpy()
        log(f"{method}: {kern_ms:.3f} ms/step at waves_per_simd {occ['waves_per_simd']}")
    agree = float(np.max(np.abs(finals["stored"] - finals["matfree"])) / np.max(np.abs(finals["stored"])))
    del finals
    sysm.release_values()
    torch.cuda.empty_cache()
    # observations: the same model under a second GMRF modulus field, kept in pinned host memory
    theta_obs = np.log(sample_element_E(el, S, seed=case.seed + 100, E0=case.E0, cov=case.cov,
                                        tau=gm.tau, device="cuda").E)
    sysm.set_log_modulus(torch.as_tensor(theta_obs).cuda())
    obs = list(range(nsteps // 20, nsteps + 1, nsteps // 20))
    u_obs = torch.empty((len(obs), N, 3, S), dtype=torch.float64, pin_memory=True)
    u0 = np.zeros(3 * N)
    sysm.initial_state(u0, u0, load, dt, beta=beta, method="matfree")
    done = 0
    for i, k in enumerate(obs):
        sysm.step(load, dt, beta=beta, method="matfree", offset=done, nsteps=k - done)
        done = k
        u_obs[i].copy_(sysm.u)
    log(f"observations ({time.perf_counter() - t_start:.0f} s)")
    if args.adjoint_block:
        # one checkpoint level (pairs every 125 steps: 101 GB), the reverse sweep by blocks of 64
        # samples so that a block's recomputed segment (124 states of 0.79 GB) fits beside them
        C, C2, B = args.adjoint_every or 125, None, args.adjoint_block
    else:
        # two levels (pairs every 200 steps, and every 20 inside a segment): the whole ensemble's
        # states are 3.1 GB each, so one level (2,000/C pairs + C states) does not fit 288 GB
        C, C2, B = args.adjoint_every or 200, 20, None
    prob = AdjointProblem(sysm, load, nsteps, dt, beta=beta, method="matfree", obs_steps=obs,
                          u_obs=u_obs, checkpoint_every=C, inner_every=C2, sample_block=B,
                          host_checkpoints=args.adjoint_host, concurrent=not args.adjoint_sequential)
    prob.set_theta(torch.as_tensor(theta).cuda())
    torch.cuda.reset_peak_memory_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = prob.forward()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    log(f"forward {t1 - t0:.1f} s")
    g = prob.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"adjoint {t2 - t1:.1f} s ({prob.stats.get('fused_steps', 0)} fused steps)")
    fwd, adj = t1 - t0, t2 - t1
    peak = torch.cuda.max_memory_allocated() / 1e9
    out = {
        "metric": "forward + adjoint gradient on the stenosed 512k-node tube, 256 samples, 2,000 steps",
        "value": S * nsteps / (fwd + adj), "unit": "sample-steps/s (forward + adjoint)", "n_gpus": 1,
        "steps": nsteps, "warmup": 5, "ms_per_step": 1e3 * (fwd + adj) / nsteps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded stenosed tube and GMRF samples",
        "config": {"workload": case.name, "n_nodes": N, "n_elements": E, "nnzb": sysm.nnzb,
                   "n_samples": S, "dt": dt, "beta": beta, "observations": len(obs),
                   "checkpoint_every": C, "inner_every": C2, "sample_block": B,
                   "checkpoints_in": "host memory" if args.adjoint_host else "device memory",
                   "concurrent_gradient": not args.adjoint_sequential,
                   "adjoint_operator": "matfree"},
        "step_operators": per_step, "stored_vs_matfree_rel_max_err": agree,
        "forward_s": fwd, "adjoint_s": adj, "adjoint_over_forward": adj / fwd,
        "fused_steps": prob.stats.get("fused_steps", 0),
        "fusion_timing_ms": prob.stats.get("fusion_timing_ms"),
        "peak_device_GB": peak, "loss_mean": float(loss.mean()),
        "gradient_max_abs": float(g.abs().max()),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
