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
it after the baseline's steps.
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


def algorithmic_bytes(sysm, method):
    """HBM bytes one step kernel must move at minimum (each datum once)."""
    N, S, nnzb, E = sysm.n_nodes, sysm.n_smp, sysm.nnzb, sysm.n_elms
    state = 3 * (3 * N * S * 8)                 # read u, read up, write u+
    lm = sysm.lm.numel() * 8
    small = 3 * N * 8 + 4 * (N + 1) + 4 * nnzb + N   # f1, indptr, indices, fixed
    if method == "stored":
        op = nnzb * 9 * S * 8
    else:
        op = E * S * 8 + 9 * E * 9 * 8 + 4 * 9 * E + 4 * (nnzb + 1)   # coef, khat, celem, cptr
    return state + lm + small + op


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


def cpu_baseline(case, sysm, nodes, el, bdy, coefK, dt, beta, steps, threads):
    """C/OpenMP restatement of the reference CPU step on a bounded sample; returns (result, u_cpu)."""
    from oracle import cpu_oracle as co
    from oracle import solid_oracle as so
    co.build()
    S = case.n_smp
    t_asm = time.perf_counter()
    K = co.assemble_K(nodes, el, so.d_matrix(case.nu), coefK, sysm.indptr_h, sysm.indices_h, threads)
    t_asm = time.perf_counter() - t_asm
    LM = sysm.lumped_mass_dofs()
    f1 = sysm.f1.cpu().numpy().ravel()
    p = case.pressures(0, steps + 1, dt)
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
    timed = times[1:] if len(times) > 1 else times
    rate = S / float(np.mean(timed))
    res = {"value": rate, "unit": "sample-steps/s", "cores": int(threads),
           "kind": "port",
           "sample": f"{case.name}: same mesh, {S} samples and inputs; {len(timed)} timed steps after "
                     f"1 warm-up ({float(np.sum(times)):.1f} s of CPU work); K assembled in the "
                     f"reference layout [nnzb][S][3][3] in {t_asm:.1f} s (not timed)"}
    return res, u


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--case", default="tube-256x400-s128")
    ap.add_argument("--method", default="stored", choices=["stored", "matfree"])
    ap.add_argument("--cpu-steps", type=int, default=4, help="CPU baseline steps (first is warm-up)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all OpenMP threads")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-matfree", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the ensemble step is sample-parallel with no per-step "
              "exchange, so an N-GPU run is N independent sample shards; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from cvfes_amd.configs import get_case
    from cvfes_amd.system import MembraneSystem
    from oracle import solid_oracle as so

    case = get_case(args.case)
    dt, beta = case.dt, case.beta
    t_setup = time.perf_counter()
    nodes, el, bdy = case.geometry()
    gm = case.sample_E(el)
    N, S = nodes.shape[0], case.n_smp
    sysm = MembraneSystem(nodes, el, bdy, S, vthickness=np.full(N, case.thickness), elmE=gm.E,
                          density=case.density, nu=case.nu, store_values=True)
    total = args.warmup + args.steps
    load = sysm.loads(case.pressures(0, max(total, args.cpu_steps + 1), dt))
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
                   "operator": args.method, "layout": "values [nnzb][9][S], state [N][3][S]"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": abytes, "kernel_ms": kern_ms,
                     "kernel": "k_step (fused block-CSR product + central-difference update)"},
    }
    if not args.no_matfree:
        other = "matfree" if args.method == "stored" else "stored"
        msteps = max(20, args.steps // 2)
        mwall, mk = time_steps(sysm, load, dt, beta, other, args.warmup, msteps)
        mbytes = algorithmic_bytes(sysm, other)
        out["alternative"] = {"operator": other, "value": S * msteps / mwall,
                              "unit": "sample-steps/s", "kernel_ms": mk,
                              "achieved_GBs": mbytes / (mk * 1e-3) / 1e9,
                              "frac": mbytes / (mk * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "algorithmic_bytes_per_launch": mbytes}
        print(f"[bench] {other}: {mk:.4f} ms/kernel, {S * msteps / mwall:.0f} sample-steps/s",
              file=sys.stderr, flush=True)
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
        print(f"[bench] cpu baseline {cres['value']:.1f} sample-steps/s on {threads} threads; "
              f"parity {out['parity_vs_cpu']['rel_max_err']:.2e}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
