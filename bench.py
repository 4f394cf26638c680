"""Benchmark: FGMRES solve time per frequency (both polarisations) to a 1e-8 residual, and the
operator-apply HBM bandwidth, on one MI355X.

Workload (default, --config c5): the largest single-GPU configuration, 256x256x64 cells
(dx = dy = 100 m, dz 25 -> 800 m geometric) over the 3-layer background with a 1000 ohm-m slab to the
domain bottom, 60 seeded random boxes of 0.3-3000 ohm-m, and its 12 frequencies 1e-4 .. 100 Hz.
FP64 (complex128) throughout, FGMRES(30) with the inner GMRES(10) right preconditioner, tolerance
1e-8 on the reference's stopping test (the Arnoldi residual; the FP64 true residual is reported
beside it). --config c2 runs the 128x128x32 model with its 8 frequencies the same way.

A step is the solve of one frequency: chi <- chi0 for both polarisations and the side-by-side
FGMRES solve of A chi = chi0 (the reference's SolveCie span, pol X + pol Y). Step s solves frequency
s mod F. The Green tensor and chi0 of a step's frequency are built on the GPU before its timed span
(the reference accounts the Green tensor separately, as AtoA_Green); at C5 one frequency's tensors
(17.5 GB complex128 + 8.7 GB complex64, folded) are built per step and freed after it. Each step's
solve is bracketed by device synchronisations; the value is the mean over frequencies of each
frequency's mean solve time, so that a step count that is not a multiple of F does not weight the
cheap or the expensive frequencies twice.

Beside the FP64 value the line carries, per frequency: the mixed-precision solve (complex64 Green
tensor and FFTs inside FP64 iterative refinement, stopping on the FP64 true residual) and the
relative difference of the observed E, H, Z, rho_a and phase at the surface between the two; the
live roofline of the dominant kernel (the Green-tensor contraction, timed with HIP events on its own
stream over the timed FP64 solves); a CPU baseline (the project's FP64 C restatement, the oracle,
timed on this host on a bounded sample and scaled by the measured apply count); and the two-model
grouping of the 128x128x32 model (four right-hand sides per apply) against solving the models one
after the other.
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


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5", choices=["c1", "c2", "c4", "c5"],
                    help="c5 (default): the 256x256x64 model, the largest that fits one GPU; c2: 128x128x32; "
                         "c4: 64x64x16 (model 0 of the preconditioner-training set)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision solves (complex64 operator inside FP64 refinement)")
    ap.add_argument("--no-observe", action="store_true",
                    help="skip the surface fields and responses (E, H, Z, rho_a, phase) of both solves")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-models", action="store_true",
                    help="skip the two-model grouping measurement on the 128x128x32 model")
    ap.add_argument("--cpu-applies", type=int, default=0, help="oracle applies timed (0: 2 at C5, 3 otherwise)")
    ap.add_argument("--c3-models", type=int, default=64,
                    help="models of the batched uneven-grid workload (256 in full; 0 skips it)")
    ap.add_argument("--profile-applies", type=int, default=0,
                    help="only run this many two-right-hand-side FP64 applies, then as many complex64 applies on "
                         "complex64 vectors, of the workload's first frequency on the main thread (for kernel traces)")
    ap.add_argument("--profile-krylov", action="store_true",
                    help="only run one single-right-hand-side FGMRES solve of the workload's first frequency on the "
                         "main thread (for traces and counters of the Krylov kernels)")
    ap.add_argument("--profile-mixed", action="store_true",
                    help="only run one single-right-hand-side mixed-precision solve of the workload's first frequency "
                         "on the main thread (for traces of the complex64 kernels)")
    ap.add_argument("--profile-batch", type=int, default=0,
                    help="only run the uneven-model scheduler on this many of its models on one host thread "
                         "(groups of four, FP64; for kernel traces of the small-nz kernels)")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the HBM stream probe (read-only, the contraction's read/write mix, 1:1 copy)")
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
    if args.steps < 1:
        print(json.dumps({"error": "--steps must be at least 1"}))
        sys.exit(2)

    import torch
    from extreme_amd import cartesian_model as cm

    if args.profile_batch:
        profile_batch(args.profile_batch)
        return
    model = {"c5": cm.config_c5, "c2": cm.config_c2, "c4": cm.config_c4, "c1": cm.config_c1}[args.config]()
    if args.profile_applies or args.profile_krylov or args.profile_mixed:
        profile_run(model, args)
        return
    line = sweep(model, args)
    if not args.no_probe:
        probe = hbm_probe(model.nz)
        line["hbm_probe"] = probe
        rl = line["roofline"]
        mix = probe["contraction_mix"]["gbs"]
        rl["practical_ceiling"] = {"gbs": mix, "frac": rl["achieved"] / mix,
                                   "what": "the contraction's read/write mix streamed by the same copy path "
                                           "without arithmetic (hbm_probe.contraction_mix)"}
    if not args.no_models and args.config != "c1":
        line["model_batching"] = model_batching(line["_settings"])
    if args.c3_models > 0:
        line["batched_uneven_models"] = batched_uneven(args.c3_models, line["_settings"])
    line.pop("_settings")
    torch.cuda.synchronize()
    print(json.dumps(line), flush=True)


def sweep(model, args):
    """Times `steps` frequency solves (step s: frequency s mod F), each on a freshly built operator."""
    import torch
    from extreme_amd import forward as F
    from extreme_amd import xtreme as X
    settings = X.ForwardSettings(residual=1e-8, inner_buffer_length=10, outer_buffer_length=30,
                                 max_repeats_number=100, number_of_hankels=10)
    freqs = model.frequencies
    nf = len(freqs)
    prec = "fp64" if args.no_mixed else "mixed"
    m = model
    nz, L = m.nz, 4 * m.nx * m.ny
    gblk = 4 * (nz * (nz + 1) // 2) + 2 * nz * nz
    n, N, M = m.n, m.nx * m.ny * nz, L * 3 * nz
    SG = SV = SR = 16

    rows = {}          # frequency index -> summary (first occurrence)
    solve_s = {}       # frequency index -> list of timed FP64 solve spans
    stage_ms = {k: 0.0 for k in X.ConvolutionOperator.STAGES}
    rounds = rhs_applies = rounds2 = 0
    contract_bytes = apply_bytes = 0.0
    contract2_ms = 0.0                 # contraction launches with two right-hand sides
    layout = None
    cpu = None
    green_s = []
    mixed_rows = {}
    for s in range(args.steps):
        fi = s % nf
        f = freqs[fi]
        t0 = time.perf_counter()
        p = F.FrequencySolve(model, f, settings, precision=prec)
        torch.cuda.synchronize()
        build = time.perf_counter() - t0
        green_s.append(p.green_seconds)
        if s == 0:
            for w in range(args.warmup):
                infos = p.solve()
                log(f"warmup {w}: f={f:.4g} Hz its {[i.number_of_iterations for i in infos]}")
            if args.warmup > 0 and not args.no_mixed:
                p.solve_mixed()                       # measures the complex64 FFT tiles once
        p.op.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = p.solve()
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        st, cnt = p.op.stage_times()
        st2, cnt2 = p.op.stage_times(2)
        p.op.set_profiling(False)
        contract2_ms += st2["contract"]
        rounds2 += cnt2
        layout, nblk = p.op.green_layout()
        for k in st:
            stage_ms[k] += st[k]
        rounds += cnt
        mv = sum(i.matvecs for i in infos)
        rhs_applies += mv
        contract_bytes += cnt * SG * nblk * gblk + mv * 2 * M * SV
        apply_bytes += cnt * SG * nblk * gblk + mv * 2 * n * SV + cnt * N * SR
        solve_s.setdefault(fi, []).append(t64)
        msg = f"step {s}: f={f:.4g} Hz build {build:.1f}s fp64 {t64:.3f}s its {[i.number_of_iterations for i in infos]}"
        if fi not in rows:
            rows[fi] = {"frequency_hz": f, "green_build_s": p.green_seconds,
                        "fp64_iterations": [i.number_of_iterations for i in infos],
                        "fp64_matvecs": [i.matvecs for i in infos],
                        "fp64_arnoldi_residual": [i.arnoldi_backward_error for i in infos],
                        "fp64_true_residual": [i.backward_error for i in infos],
                        "converged": all(i.convergence_achieved for i in infos)}
            ref = None
            if not args.no_observe:
                t0 = time.perf_counter()
                p.build_level(0.0)
                ref = p.level_responses()
                rows[fi]["atoo_s"] = time.perf_counter() - t0
            if not args.no_mixed:
                chi64 = p.chi.clone()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                minfo = p.solve_mixed()
                torch.cuda.synchronize()
                tm = time.perf_counter() - t0
                mixed_rows[fi] = tm
                rows[fi].update(mixed_s=tm, mixed_true_residual=[i.residual for i in minfo],
                                mixed_converged=all(i.converged for i in minfo),
                                mixed_refinements=[i.refinements for i in minfo],
                                mixed_fp32_applies=[i.matvecs32 for i in minfo],
                                mixed_fp64_applies=[i.matvecs64 for i in minfo],
                                mixed_chi_rel_diff=[float(torch.linalg.norm(p.chi[k] - chi64[k]) /
                                                          torch.linalg.norm(chi64[k])) for k in (0, 1)])
                msg += f" mixed {tm:.3f}s"
                if ref is not None:
                    rows[fi]["mixed_result_rel_diff"] = F.result_errors(p.level_responses(), ref)
                del chi64
            rows[fi]["fft_tiles"] = p.op.fft_tiles()
        log(msg)
        if cpu is None and s == args.steps - 1 and not args.no_cpu_baseline:
            cpu = cpu_baseline(p, args.cpu_applies or (2 if m.name == "C5" else 3),
                               rhs_applies / args.steps)
        del p
        torch.cuda.empty_cache()

    per_freq = [float(np.mean(solve_s[k])) for k in sorted(solve_s)]
    value = float(np.mean(per_freq))
    total = sum(sum(v) for v in solve_s.values())
    for k in rows:
        rows[k]["fp64_s"] = solve_s[k]
    contract_total_ms = stage_ms["contract"]
    achieved_all = contract_bytes / (contract_total_ms * 1e-3) / 1e9
    # the roofline line: launches with two right-hand sides (both polarisations), the launch the
    # algorithmic bytes below describe; rounds after one polarisation has converged carry one
    bytes2 = SG * nblk * gblk + 2 * 2 * M * SV
    launch2_ms = contract2_ms / max(rounds2, 1)
    achieved = bytes2 / (launch2_ms * 1e-3) / 1e9
    apply_ms = sum(stage_ms.values())
    mixed = None
    if mixed_rows:
        ks = sorted(mixed_rows)
        m64 = float(np.mean([np.mean(solve_s[k]) for k in ks]))
        m32 = float(np.mean([mixed_rows[k] for k in ks]))
        worst = {}
        for k in ks:
            for q, e in rows[k].get("mixed_result_rel_diff", {}).items():
                worst[q] = max(worst.get(q, 0.0), e)
        mixed = {"s_per_freq": m32, "fp64_s_per_freq_same_frequencies": m64, "speedup_vs_fp64": m64 / m32,
                 "frequencies": len(ks), "converged": all(rows[k]["mixed_converged"] for k in ks),
                 "true_residual_max": max(max(rows[k]["mixed_true_residual"]) for k in ks),
                 "result_rel_diff_vs_fp64_max": worst or None}
    return {
        "metric": METRIC,
        "value": value,
        "unit": "s/freq",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * total / args.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded random-box model, SplitMix64)",
        "config": {"workload": f"{m.name}: {m.nx}x{m.ny}x{m.nz} cells, 3-layer background, "
                               f"{len(solve_s)} of {nf} frequencies {min(freqs):.4g}-{max(freqs):.4g} Hz, both "
                               "polarisations, FGMRES(30)/GMRES(10), tol 1e-8, FP64",
                   "nx": m.nx, "ny": m.ny, "nz": m.nz, "frequencies_hz": [round(freqs[k], 6) for k in sorted(solve_s)],
                   "precision": "fp64"},
        "value_definition": "mean over the frequencies solved of each frequency's mean FP64 solve time (both "
                            "polarisations); ms_per_step is the plain mean over the steps",
        "solves_per_s": 2.0 / value,
        "converged": all(r["converged"] for r in rows.values()),
        "true_residual_max": max(max(r["fp64_true_residual"]) for r in rows.values()),
        "applies": rounds,
        "rhs_applies": rhs_applies,
        "apply_gbs": apply_bytes / (apply_ms * 1e-3) / 1e9,
        "apply_frac_of_hbm_peak": apply_bytes / (apply_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "apply_gpu_share_of_solve": apply_ms * 1e-3 / total,
        "stage_ms_per_apply": {k: v / max(rounds, 1) for k, v in stage_ms.items()},
        "green_seconds_per_freq": float(np.mean(green_s)),
        "roofline": {"kernel": contract_kernel_name(nz, layout), "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((m.name, nz, 2, layout)),
                     "avg_launch_ms": launch2_ms, "launches": rounds2, "green_layout": layout,
                     "algorithmic_bytes_per_launch": bytes2,
                     "launch": "two right-hand sides (both polarisations), HIP events on the operator's stream",
                     "all_launches": {"launches": rounds, "avg_launch_ms": contract_total_ms / max(rounds, 1),
                                      "gbs": achieved_all,
                                      "note": "includes the launches with one right-hand side after one "
                                              "polarisation converged"}},
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "per_frequency": [rows[k] for k in sorted(rows)],
        "_settings": settings,
    }


def model_batching(settings):
    """Two seeded 128x128x32 models (same grid and background, different boxes) at one frequency:
    solved one after the other (2 right-hand sides per apply each) and as one group (4 per apply)."""
    import torch
    from extreme_amd import cartesian_model as cm
    from extreme_amd import forward as F
    m0, m1 = cm.config_c2(0), cm.config_c2(1)
    f = m0.frequencies[4]
    p0 = F.FrequencySolve(m0, f, settings)
    p1 = F.FrequencySolve(m1, f, settings)
    grp = F.ModelGroupSolve([m0, m1], f, settings)
    p0.solve()
    p1.solve()
    grp.solve()                                  # warm all paths
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    i0 = p0.solve()
    i1 = p1.solve()
    torch.cuda.synchronize()
    seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    infos = grp.solve()
    torch.cuda.synchronize()
    grouped = time.perf_counter() - t0
    same = all(a.number_of_iterations == b.number_of_iterations for a, b in zip(i0 + i1, infos))
    out = {"workload": "C2 models 0 and 1 at 0.193 Hz", "frequency_hz": f, "models": 2, "right_hand_sides": 4,
           "sequential_s": seq, "grouped_s": grouped, "speedup": seq / grouped,
           "iterations": [i.number_of_iterations for i in infos], "iterations_equal_sequential": same,
           "converged": all(i.convergence_achieved for i in infos)}
    del grp, p0, p1
    torch.cuda.empty_cache()
    return out


def batched_uneven(nmodels, settings):
    """The batched uneven-grid workload: seeded models of {32,48,64,96}^2 x {8,12,16,24} cells, each at
    two of {0.01, 0.1, 1} Hz, both polarisations. Solved one model and frequency at a time (the
    reference's order, on one handle at a time) and by the batch scheduler (groups of up to four models
    sharing grid and frequency, four groups in flight), in FP64 and in mixed precision. Green-tensor
    assembly is included in all three."""
    import torch
    from extreme_amd import batch as B
    from extreme_amd import cartesian_model as cm
    models = [cm.config_c3(k) for k in range(nmodels)]
    B.run(models[:2], settings, group_size=1, workers=1)          # warm: FFT tile measurement of two shapes
    torch.cuda.synchronize()
    seq = B.run(models, settings, group_size=1, workers=1)
    torch.cuda.synchronize()
    grp = B.run(models, settings, group_size=4, workers=4)
    torch.cuda.synchronize()
    B.run(models[:2], settings, group_size=1, workers=1, precision="mixed")     # warm the complex64 paths
    torch.cuda.synchronize()
    mix = B.run(models, settings, group_size=4, workers=4, precision="mixed")
    torch.cuda.synchronize()
    same = all(seq.iterations[k] == grp.iterations[k] for k in seq.iterations)
    log(f"batched uneven models: {seq.solves} solves, one at a time {seq.seconds:.2f} s, grouped {grp.seconds:.2f} s")
    return {"workload": f"C3: {nmodels} of 256 seeded models, (Nx, Ny) in {{32,48,64,96}}^2, Nz in {{8,12,16,24}}, "
                        "2 of {0.01, 0.1, 1} Hz each, both polarisations, FP64, tol 1e-8",
            "right_hand_sides": seq.solves, "one_at_a_time_s": seq.seconds, "grouped_s": grp.seconds,
            "solves_per_s_one_at_a_time": seq.solves / seq.seconds, "solves_per_s_grouped": grp.solves / grp.seconds,
            "speedup": seq.seconds / grp.seconds, "groups": grp.groups, "group_size": 4, "workers": 4,
            "iterations_equal": same, "converged": seq.converged and grp.converged,
            "true_residual_max": max(seq.max_true_residual, grp.max_true_residual),
            "mixed_precision": {"grouped_s": mix.seconds, "solves_per_s": mix.solves / mix.seconds,
                                "speedup_vs_fp64_grouped": grp.seconds / mix.seconds, "converged": mix.converged,
                                "true_residual_max": mix.max_true_residual,
                                "what": "the same groups solved by FP64 refinement around complex64 corrections "
                                        "(complex64 tensor, FFTs, contraction and Krylov vectors)"}}


def profile_run(model, args):
    """Kernel-trace workloads on one host thread: operator applies of both polarisations' chi0 (FP64,
    then complex64 on complex64 vectors), one FGMRES solve of one polarisation, or one mixed-precision
    solve of one polarisation."""
    import torch
    from extreme_amd import forward as F
    from extreme_amd import xtreme as X
    f = model.frequencies[0]
    p = F.FrequencySolve(model, f, precision="mixed")
    torch.cuda.synchronize()
    if args.profile_mixed:
        x = p.chi0[0].clone()
        t0 = time.perf_counter()
        info = p.solver.solve_refined_batch([p.chi0[0]], [x])[0]
        log(f"profile: one mixed-precision solve at {f:g} Hz, {info.refinements} refinements, {info.matvecs32} "
            f"complex64 and {info.matvecs64} FP64 applies, residual {info.residual:.2e}, {time.perf_counter() - t0:.3f} s")
        return
    if args.profile_krylov:
        x = p.chi0[0].clone()
        t0 = time.perf_counter()
        info = X.AnomalyCurrentFgmresSolver(p.op).solve(p.chi0[0], x)
        log(f"profile: one solve at {f:g} Hz, {info.number_of_iterations} iterations, {info.matvecs} applies, "
            f"{time.perf_counter() - t0:.3f} s")
        return
    y = torch.empty_like(p.chi0)
    t0 = time.perf_counter()
    for _ in range(args.profile_applies):
        p.op.apply(p.chi0, y)
    torch.cuda.synchronize()
    t64 = (time.perf_counter() - t0) / args.profile_applies
    # complex64 applies on complex64 vectors, through the mixed handle's FP32 path
    x32 = p.chi0.to(torch.complex64)
    y32 = torch.empty_like(x32)
    t0 = time.perf_counter()
    for _ in range(args.profile_applies):
        p.op.apply_c64(x32, y32)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.profile_applies
    log(f"profile: {args.profile_applies} applies at {f:g} Hz, {t64 * 1e3:.3f} ms FP64, {t32 * 1e3:.3f} ms complex64")


def profile_batch(nmodels):
    """The uneven-model scheduler on the first nmodels models on one host thread (groups of up to
    four models sharing grid and frequency, FP64), after a warm-up on two of them."""
    import torch
    from extreme_amd import batch as B
    from extreme_amd import cartesian_model as cm
    from extreme_amd import xtreme as X
    models = [cm.config_c3(k) for k in range(nmodels)]
    settings = X.ForwardSettings()
    B.run(models[:2], settings, group_size=1, workers=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = B.run(models, settings, group_size=4, workers=1)
    torch.cuda.synchronize()
    log(f"profile: {nmodels} uneven models, {r.groups} groups, {r.solves} solves on one thread, "
        f"{time.perf_counter() - t0:.3f} s (Green tensors and chi0 {r.green_seconds:.3f} s, solves {r.solve_seconds:.3f} s), "
        f"converged {r.converged}, max true residual {r.max_true_residual:.2e}")


def hbm_probe(nz, reps=10):
    """HBM rates of the streaming contraction's copy path without arithmetic (xtr_probe_stream): blocks
    read into LDS with non-temporal LDS-DMA loads through a ring of 32 KB units, one 512-thread
    workgroup per CU, and KB written after each block with 16-byte lanes. Three mixes: read only; the
    contraction's (per folded block at nz = 64: 264 KB of Green tensor + 24.6 KB of spectrum read,
    24.6 KB written, modelled as 288 KB read / 24 KB written; at nz = 32: 78.8 / 12.3 KB as 192 / 30);
    and 1:1 (the FFT passes read and write about as much as each other). About 4.7 GB read per launch."""
    import torch
    from extreme_amd import xtreme as X
    read_kb, write_kb = (288, 24) if nz == 64 else (192, 30)
    blocks = (4700 << 20) // (read_kb << 10)
    src = torch.ones(blocks * read_kb * 128, dtype=torch.float64, device="cuda")
    dst = torch.empty(blocks * read_kb * 128, dtype=torch.float64, device="cuda")
    out = {}
    for name, wkb in (("read_only", 0), ("contraction_mix", write_kb), ("copy_1to1", read_kb)):
        for _ in range(2):
            X.probe_stream(src, dst, blocks, read_kb, wkb)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            X.probe_stream(src, dst, blocks, read_kb, wkb)
        b.record()
        b.synchronize()
        ms = a.elapsed_time(b) / reps
        nbytes = blocks * (read_kb + wkb) * 1024
        out[name] = {"read_kb_per_block": read_kb, "write_kb_per_block": wkb, "ms": ms,
                     "gbs": nbytes / (ms * 1e-3) / 1e9}
        log(f"hbm probe {name}: {out[name]['gbs']:.0f} GB/s")
    del src, dst
    torch.cuda.empty_cache()
    out["blocks"] = blocks
    return out


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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(problem, applies, mean_rhs_applies_per_freq):
    """Times the oracle's FP64 operator apply (the FFT convolution + packed-block contraction the
    reference runs per matvec, on the same model, frequency and Green tensor) on the host cores, on a
    bounded sample of `applies` applies, and scales to seconds per frequency with the GPU run's mean
    operator applies per frequency (both polarisations)."""
    import torch
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    m = problem.model
    t0 = time.perf_counter()
    G = problem.op.green_tensor()                 # reference layout, all 4 Nx Ny blocks, on the host
    fetch = time.perf_counter() - t0
    torch.cuda.empty_cache()
    oop = O.Operator(m, G)
    oop.prepare("A", m.zeta(problem.omega).reshape(-1), m.zeta0(problem.omega))
    x = problem.chi0[0].cpu().numpy()
    oop.apply(x)  # first touch of the buffers
    t0 = time.perf_counter()
    for _ in range(applies):
        oop.apply(x)
    per_apply = (time.perf_counter() - t0) / applies
    del oop, G
    return {"value": per_apply * mean_rhs_applies_per_freq, "unit": "s/freq", "cores": threads,
            "kind": "port", "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{applies} FP64 operator applies of the oracle (C restatement, OpenMP, self-written FFT) "
                      f"at {m.nx}x{m.ny}x{m.nz} with this run's Green tensor at {problem.frequency:.4g} Hz, "
                      f"{per_apply:.2f} s each, times the GPU run's {mean_rhs_applies_per_freq:.0f} operator applies "
                      "per frequency (both polarisations): an extrapolation; Krylov vector work not included",
            "green_fetch_s": fetch}


if __name__ == "__main__":
    main()
