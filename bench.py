"""Benchmark: FGMRES solve time per frequency (both polarisations) to a 1e-8 residual, and the
operator-apply HBM bandwidth, on one MI355X.

Workload (default): the 128x128x32 model with 20 seeded random boxes over the 3-layer background
(dx = dy = 200 m, dz 50 -> 400 m), its 8 frequencies 1e-3 .. 10 Hz, FP64 (complex128) throughout,
FGMRES(30) with the inner GMRES(10) right preconditioner, tolerance 1e-8 (the reference's stopping
test on the Arnoldi residual; the true residual is reported beside it).

A step is the solve of one frequency: chi <- chi0 for both polarisations and the side-by-side
FGMRES solve of A chi = chi0 (the reference's SolveCie span, pol X + pol Y). Green tensors and
chi0 of every frequency are built on the GPU before the timed region (the reference accounts
the Green tensor separately, as AtoA_Green). Step s solves frequency s mod 8.

The JSON line also carries the live roofline of the dominant kernel (the Green-tensor
contraction, timed with HIP events on its own stream over the timed region) and a CPU baseline:
the project's FP64 C restatement (the oracle) timed on this host on a bounded sample.
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
METRIC = "FGMRES solve time per freq (2 pol) to 1e-8 residual; op-apply HBM GB/s"

# HBM bytes per launch of the contraction kernel, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (one counter per run) and scaled by the same counters on a 1 GiB device copy
# (FETCH x 2.0000, WRITE x 1.0000); summaries in profiles/<kernel>_<workload>_hbm_traffic.md.
# Keyed by (workload, nz, right-hand sides per launch, Green layout); no entry reports null.
MEASURED_TRAFFIC = {("C2", 32, 2, "folded"): 1.5225e9,   # contract_fold32_kernel<double>
                    ("C5", 64, 2, "folded"): 2.0671e10,  # contract_fold64_kernel<double>
                    ("C2", 32, 2, "full"): 4.7647e9,     # contract_stream_kernel<double, 2>
                    ("C5", 64, 2, "full"): 7.2476e10}    # contract_stream64_kernel<double, 2>


def contract_kernel_name(nz, layout):
    """The FP64 contraction kernel a two-right-hand-side apply launches at this nz and Green layout."""
    if layout == "folded":
        return {32: "contract_fold32_kernel<double>", 64: "contract_fold64_kernel<double>"}.get(
            nz, "contract_kernel<double>")
    return {32: "contract_stream_kernel<double, 2>", 64: "contract_stream64_kernel<double, 2>"}.get(
        nz, "contract_kernel<double>")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c5"],
                    help="c2 (default) is the metric's workload; c5 (256x256x64, 69 GB Green tensor per "
                         "frequency) runs its frequencies one at a time")
    ap.add_argument("--precision", default="fp64", choices=["fp64"])
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision pass (complex64 operator inside FP64 refinement)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-models", action="store_true",
                    help="skip the two-model batching measurement (two seeded models on one Green tensor)")
    ap.add_argument("--cpu-applies", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (the known-size kernel that "
                         "scales FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration()
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(the frequencies of a sweep shard across GPUs with no exchange)"}))
        sys.exit(2)

    import torch
    from extreme_amd import cartesian_model as cm
    from extreme_amd import forward as F
    from extreme_amd import xtreme as X

    if args.config == "c5":
        large_config(cm.config_c5(), args)
        return
    model = cm.config_c2() if args.config == "c2" else cm.config_c1()
    freqs = model.frequencies
    settings = X.ForwardSettings(residual=1e-8, inner_buffer_length=10, outer_buffer_length=30,
                                 max_repeats_number=100, number_of_hankels=10)
    nsolve = min(len(freqs), max(args.steps, 1))
    t0 = time.perf_counter()
    # "mixed" handles keep the complex128 and the complex64 Green tensor: the FP64 solves below use
    # only the former (bit-identical to an FP64 handle), the mixed-precision pass uses both.
    prec = "fp64" if args.no_mixed else "mixed"
    problems = [F.FrequencySolve(model, freqs[i], settings, precision=prec) for i in range(nsolve)]
    setup_s = time.perf_counter() - t0
    green_s = [p.green_seconds for p in problems]
    print(f"# setup {setup_s:.1f}s (Green tensors {sum(green_s):.1f}s for {nsolve} frequencies)", file=sys.stderr,
          flush=True)

    for w in range(args.warmup):
        infos = problems[w % nsolve].solve()
        print(f"# warmup {w}: f={problems[w % nsolve].frequency:.4g} its {[i.number_of_iterations for i in infos]}",
              file=sys.stderr, flush=True)

    for p in problems:
        p.op.set_profiling(True)
    results = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        p = problems[s % nsolve]
        infos = p.solve()
        results.append((p, infos))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps

    # ---- per-stage device time and algorithmic bytes ------------------------------------------
    # Green tensor read once per apply (the blocks the device stores: (Nx+1)(Ny+1) when folded, else
    # 4 Nx Ny), spectrum columns read and written per right-hand side, x, y and R per apply.
    m = model
    nz, L = m.nz, 4 * m.nx * m.ny
    gblk = 4 * (nz * (nz + 1) // 2) + 2 * nz * nz
    n = m.n
    N = m.nx * m.ny * nz
    M = L * 3 * nz
    SG = SV = SR = 16
    stage_ms = {k: 0.0 for k in X.ConvolutionOperator.STAGES}
    rounds = 0
    rhs_applies = 0
    for p in problems:
        st, cnt = p.op.stage_times()
        for k in st:
            stage_ms[k] += st[k]
        rounds += cnt
        p.op.set_profiling(False)
    for p, infos in results:
        rhs_applies += sum(i.matvecs for i in infos)
    layout, nblk = problems[0].op.green_layout()
    green_bytes = SG * nblk * gblk
    contract_bytes = rounds * green_bytes + rhs_applies * 2 * M * SV
    apply_alg_bytes = rounds * green_bytes + rhs_applies * 2 * n * SV + rounds * N * SR
    apply_ms = sum(stage_ms.values())
    contract_avg_ms = stage_ms["contract"] / max(rounds, 1)
    achieved_gbs = contract_bytes / (stage_ms["contract"] * 1e-3) / 1e9
    apply_gbs = apply_alg_bytes / (apply_ms * 1e-3) / 1e9

    iters = [[i.number_of_iterations for i in infos] for _, infos in results]
    matvecs = [[i.matvecs for i in infos] for _, infos in results]
    be = [[i.backward_error for i in infos] for _, infos in results]
    bea = [[i.arnoldi_backward_error for i in infos] for _, infos in results]
    conv = all(i.convergence_achieved for _, infos in results for i in infos)

    # ---- mixed precision: the same solves, complex64 operator inside FP64 refinement ----------
    mixed = None
    if not args.no_mixed:
        mixed = mixed_pass(problems, args.steps, elapsed / args.steps)

    # ---- batching two models that share the Green tensor (4 right-hand sides per apply) --------
    batching = None
    if not args.no_models and args.config == "c2":
        batching = model_batching(problems[min(4, nsolve - 1)], settings)

    # ---- the host-pointer boundary: one apply of both polarisations with PCIe copies ---------
    pcie = host_boundary_apply(problems[0])

    # ---- CPU baseline: the oracle's FP64 operator apply on this host ----------------------
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(problems[0], args.cpu_applies, sum(sum(mv) for mv in matvecs) / args.steps)

    line = {
        "metric": METRIC,
        "value": elapsed / args.steps,
        "unit": "s/freq",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded random-box model, SplitMix64)",
        "config": {"workload": f"{m.name}: {m.nx}x{m.ny}x{m.nz} cells, 3-layer background, 20 boxes, "
                               f"{nsolve} of {len(freqs)} frequencies, both polarisations, FGMRES(30)/GMRES(10), "
                               "tol 1e-8", "nx": m.nx, "ny": m.ny, "nz": m.nz,
                   "frequencies_hz": [round(p.frequency, 6) for p in problems], "precision": args.precision},
        "solves_per_s": 2 * args.steps / elapsed,
        "apply_gbs": apply_gbs,
        "apply_frac_of_hbm_peak": apply_gbs / HBM_PEAK_GBS,
        "stage_ms_per_apply": {k: v / max(rounds, 1) for k, v in stage_ms.items()},
        "fft_tiles": problems[0].op.fft_tiles(),
        "applies": rounds,
        "rhs_applies": rhs_applies,
        "iterations": iters,
        "matvecs": matvecs,
        "arnoldi_backward_error": bea,
        "true_backward_error": be,
        "converged": conv,
        "green_seconds_per_freq": float(np.mean(green_s)),
        "roofline": {"kernel": contract_kernel_name(m.nz, layout), "bound": "hbm", "achieved": achieved_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((m.name, nz, 2, layout)),
                     "avg_launch_ms": contract_avg_ms, "green_layout": layout, "green_blocks": nblk,
                     "algorithmic_bytes_per_launch": contract_bytes / max(rounds, 1)},
        "cpu_baseline": cpu,
        "host_boundary": pcie,
        "mixed_precision": mixed,
        "model_batching": batching,
    }
    print(json.dumps(line), flush=True)


def mixed_pass(problems, steps, fp64_s_per_freq):
    """Times the same frequency solves with FP64 iterative refinement around complex64 FGMRES, and
    states each solution's difference from the FP64 solve's."""
    import torch
    ref = {}
    for p in problems:                       # FP64 solutions to compare against
        p.solve()
        ref[p.frequency] = p.chi.clone()
    results = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        p = problems[s % len(problems)]
        results.append((p, p.solve_mixed()))
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    diffs = []
    for p in problems:
        p.solve_mixed()
        for pol in (0, 1):
            a, b = p.chi[pol], ref[p.frequency][pol]
            diffs.append(float(torch.linalg.norm(a - b) / torch.linalg.norm(b)))
    infos = [i for _, inf in results for i in inf]
    return {"s_per_freq": el, "speedup_vs_fp64": fp64_s_per_freq / el,
            "converged": all(i.converged for i in infos),
            "true_residual_max": max(i.residual for i in infos),
            "refinements": [[i.refinements for i in inf] for _, inf in results],
            "fp32_applies": [[i.matvecs32 for i in inf] for _, inf in results],
            "fp64_applies": [[i.matvecs64 for i in inf] for _, inf in results],
            "solution_rel_diff_vs_fp64_max": max(diffs)}


def model_batching(problem0, settings):
    """Two seeded C2 models (same grid and background, different boxes) at one frequency: solved
    one after the other (2 right-hand sides per apply each) and as one group (4 per apply)."""
    import torch
    from extreme_amd import cartesian_model as cm
    from extreme_amd import forward as F
    f = problem0.frequency
    m1 = cm.config_c2(model_index=1)
    p1 = F.FrequencySolve(m1, f, settings)
    grp = F.ModelGroupSolve([problem0.model, m1], f, settings)
    p1.solve()
    grp.solve()                                  # warm both paths
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    problem0.solve()
    p1.solve()
    torch.cuda.synchronize()
    seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    infos = grp.solve()
    torch.cuda.synchronize()
    grouped = time.perf_counter() - t0
    out = {"frequency_hz": f, "models": 2, "right_hand_sides": 4, "sequential_s": seq, "grouped_s": grouped,
           "speedup": seq / grouped, "iterations": [i.number_of_iterations for i in infos],
           "converged": all(i.convergence_achieved for i in infos)}
    del grp, p1
    torch.cuda.empty_cache()
    return out


def host_boundary_apply(problem, reps=3):
    """Apply of both polarisations through the host-pointer entry point (H2D, apply, D2H) beside
    the same apply on device-resident vectors."""
    import torch
    op = problem.op
    xd = problem.chi0
    xh = xd.cpu().numpy()
    yd = torch.empty_like(xd)
    yh = np.empty_like(xh)
    op.apply(xd, yd)
    op.apply(xh, yh)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        op.apply(xd, yd)
    torch.cuda.synchronize()
    dev = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        op.apply(xh, yh)
    host = (time.perf_counter() - t0) / reps
    return {"apply_device_ms": dev * 1e3, "apply_host_pointers_ms": host * 1e3,
            "bytes_over_pcie_per_apply": int(2 * xh.nbytes)}


def large_config(model, args):
    """A configuration whose per-frequency tensors do not fit several at once: each step builds one
    frequency (complex128 and complex64 Green tensors on the GPU), times its FP64 solve of both
    polarisations and its mixed-precision solve, and frees it. Green assembly is outside the timed
    spans; the first frequency is solved once untimed when warmup > 0."""
    import torch
    from extreme_amd import forward as F
    from extreme_amd import xtreme as X
    settings = X.ForwardSettings(residual=1e-8, inner_buffer_length=10, outer_buffer_length=30,
                                 max_repeats_number=100, number_of_hankels=10)
    freqs = model.frequencies
    rows = []
    for s in range(args.steps):
        f = freqs[s % len(freqs)]
        t0 = time.perf_counter()
        p = F.FrequencySolve(model, f, settings, precision="fp64" if args.no_mixed else "mixed")
        torch.cuda.synchronize()
        build = time.perf_counter() - t0
        if s == 0 and args.warmup > 0:
            p.solve()
        p.op.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = p.solve()
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        st, rounds = p.op.stage_times()
        p.op.set_profiling(False)
        layout, nblk = p.op.green_layout()
        ref = p.chi.clone()
        row = {"frequency_hz": f, "green_build_s": build, "fp64_s": t64, "green_layout": layout, "green_blocks": nblk,
               "contract_ms_total": st["contract"],
               "fp64_iterations": [i.number_of_iterations for i in infos],
               "fp64_matvecs": [i.matvecs for i in infos],
               "fp64_true_residual": [i.backward_error for i in infos],
               "apply_rounds": rounds, "stage_ms_per_apply": {k: v / max(rounds, 1) for k, v in st.items()},
               "fft_tiles": p.op.fft_tiles()}
        if not args.no_mixed:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            minfo = p.solve_mixed()
            torch.cuda.synchronize()
            row.update(mixed_s=time.perf_counter() - t0, mixed_true_residual=[i.residual for i in minfo],
                       mixed_fp32_applies=[i.matvecs32 for i in minfo],
                       mixed_rel_diff_vs_fp64=[float(torch.linalg.norm(p.chi[k] - ref[k]) / torch.linalg.norm(ref[k]))
                                               for k in (0, 1)])
        print(f"# {model.name} f={f:.4g} Hz: build {build:.1f}s fp64 {t64:.2f}s "
              f"its {row['fp64_iterations']}" + (f" mixed {row['mixed_s']:.2f}s" if 'mixed_s' in row else ""),
              file=sys.stderr, flush=True)
        rows.append(row)
        del p, ref
        torch.cuda.empty_cache()
    m = model
    L, nz = 4 * m.nx * m.ny, m.nz
    gblk = 4 * (nz * (nz + 1) // 2) + 2 * nz * nz
    M = L * 3 * nz
    fp64 = sum(r["fp64_s"] for r in rows) / len(rows)
    # Green blocks once per apply, spectrum columns in and out per right-hand side, over all applies
    total_rounds = sum(r["apply_rounds"] for r in rows)
    contract_bytes = sum(r["apply_rounds"] * 16 * r["green_blocks"] * gblk + sum(r["fp64_matvecs"]) * 2 * M * 16
                         for r in rows)
    contract_total_ms = sum(r["contract_ms_total"] for r in rows)
    contract_ms = contract_total_ms / max(total_rounds, 1)
    achieved = contract_bytes / (contract_total_ms * 1e-3) / 1e9
    layout = rows[0]["green_layout"]
    line = {
        "metric": METRIC, "value": fp64, "unit": "s/freq", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000 * fp64, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic (seeded random-box model, SplitMix64)",
        "config": {"workload": f"{m.name}: {m.nx}x{m.ny}x{m.nz} cells, 3-layer background + slab, 60 boxes, "
                               f"{args.steps} of {len(m.frequencies)} frequencies one at a time, both polarisations, "
                               "FGMRES(30)/GMRES(10), tol 1e-8", "nx": m.nx, "ny": m.ny, "nz": m.nz},
        "roofline": {"kernel": contract_kernel_name(nz, layout), "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((m.name, nz, 2, layout)),
                     "avg_launch_ms": contract_ms, "green_layout": layout,
                     "algorithmic_bytes_per_launch": contract_bytes / max(total_rounds, 1)},
        "cpu_baseline": None,
        "mixed_s_per_freq": (sum(r["mixed_s"] for r in rows) / len(rows)) if not args.no_mixed else None,
        "frequencies": rows,
    }
    print(json.dumps(line), flush=True)


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain elementwise copy of nbytes (read nbytes, write nbytes) per launch."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
                      "ms_per_copy": dt * 1e3, "gbs_read_plus_write": 2 * nbytes / dt / 1e9}), flush=True)


def cpu_baseline(problem, applies, mean_rhs_applies_per_freq):
    """Times the oracle's FP64 operator apply (the FFT convolution + packed-block contraction the
    reference runs per matvec) on the host cores, on a bounded sample of `applies` applies, and
    scales to seconds per frequency with the GPU run's operator-apply count."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    m = problem.model
    G = problem.op.green_tensor()
    oop = O.Operator(m, G)
    oop.prepare("A", m.zeta(problem.omega).reshape(-1), m.zeta0(problem.omega))
    x = problem.chi0[0].cpu().numpy()
    oop.apply(x)  # first touch
    t0 = time.perf_counter()
    for _ in range(applies):
        oop.apply(x)
    per_apply = (time.perf_counter() - t0) / applies
    return {"value": per_apply * mean_rhs_applies_per_freq, "unit": "s/freq", "cores": threads,
            "kind": "port",
            "sample": f"{applies} FP64 operator applies of the oracle (C restatement, OpenMP, self-written FFT) "
                      f"at {m.nx}x{m.ny}x{m.nz}, {per_apply:.3f} s each, times the GPU run's "
                      f"{mean_rhs_applies_per_freq:.0f} operator applies per frequency (both pols); "
                      "Krylov vector work not included"}


if __name__ == "__main__":
    main()
