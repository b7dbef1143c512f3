#!/usr/bin/env python3
"""Gradient evaluations per second of the Helmholtz solve-and-gradient engine.

Workload (the metric case): 512x512 interior + 24-cell PML (560x560 nodes),
dx = 0.25 mm, f = 0.8 MHz, 256-element ring at r = 56 mm, every element
transmits and all receive (256 RHS per solve).  Synthetic data: the model is
1500 m/s + 40 m/s of Gaussian-filtered seeded noise (sigma 3 mm); the
observations are the engine's refined forward solve on that model.  The
timed evaluations each run on their own model, c0 (1 + 1e-3 k b) with c0 =
1500 m/s and b a smooth bump (k = 1..K), as consecutive evaluations of an NCG
loop do; the same K evaluations on c0 alone are timed after them
("same_model"), and the accuracy legs use c0.  Each factorization seeds its leading
(PML) chain steps with the previous factorization's pages of the same key, as the
evaluations of an NCG loop do; the fully profiled evaluation that gives the phase
split and the roofline follows the warm-up.

One step = one wust_oracle(nargout=2) call: device assembly from c -> block
factor -> forward solves of all 256 TX -> FP64 source fit and loss -> adjoint
solves -> FP64 gradient.  Inputs (c, Y, mask) are resident in HBM before the
timed region.  "refined" = hand-written complex-FP32 MFMA factor and sweeps
with FP64 residual refinement; "plain" = the reference's GPU algorithm on
rocSOLVER/rocBLAS (one getrf+getrs per Schur block, one cgemm per column),
unrefined; "c64" = the engine's kernels without refinement.  The refinement
stops at a relative FP64 residual of 1e-10 per right-hand side (--tol).

After the timed GPU steps the same full step runs on the host in FP64: the
sparse-LU oracle (scipy SuperLU, one thread; median of --cpu-repeats runs) is
the CPU baseline, an FP64 block-LU on every granted host thread its multicore
variant, and the GPU modes' field, misfit and gradient errors are measured
against the oracle over all 256 transmitters.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "grad evals/s at 512²×256 TX: refined c64 vs plain c64; rel err vs FP64 LU"
MFMA_FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload`, from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary (already scaled by the
    copy-kernel calibration), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def host_info():
    """CPU model, logical CPUs visible to this process, and the thread count the CPU legs may use (the
    OMP_NUM_THREADS share the box grants, else every visible CPU)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS")
    threads = max(1, min(visible, int(share))) if share and share.isdigit() else visible
    return dict(cpu_model=model, nproc=os.cpu_count(), visible_cpus=visible, threads=threads)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed evaluations first (the first, on c0 after the observations' true model, falls back "
                         "from seeds that no longer fit; the profiled evaluation follows the warm-up)")
    ap.add_argument("--tol", type=float, default=1e-10, help="refinement stop (relative FP64 residual)")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--c64-steps", type=int, default=3, help="timed grad evals of the unrefined c64 engine mode")
    ap.add_argument("--cpu-repeats", type=int, default=3,
                    help="full 256-TX FP64 sparse-LU evaluations timed for the CPU baseline (median)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the FP64 oracle accuracy / CPU baseline leg")
    ap.add_argument("--workload", default="ring560",
                    choices=["ring560", "continuation8", "breast1088", "dtft", "cfg1"],
                    help="ring560: the metric case (default); continuation8: eight frequency stages on their own "
                         "grids, one grad eval each; breast1088: a full 36-step reconstruction at 1088^2 x 512 TX; "
                         "dtft: RF -> frequency DTFT of 256 TX x 256 RX x 4096 float32 samples, 12 frequencies; "
                         "cfg1: five NCG iterations at 160^2 x 64 TX on the GPU and on the FP64 oracle")
    ap.add_argument("--concurrent", action="store_true",
                    help="continuation8: run the stages on separate streams from host threads")
    return ap.parse_args()


def synthetic_obs(G, case, dev, tol):
    """Refined forward solves on the true model at every frequency of the case, gathered at the receivers."""
    import torch
    ny, nx, ntx = case.ny, case.nx, case.n_tx
    Y = torch.zeros((ntx, case.n_rx, len(case.frequencies_hz)), dtype=torch.complex128, device=dev)
    tx0 = torch.as_tensor(case.tx_index - 1, device=dev)
    rx0 = torch.as_tensor(case.rx_index - 1, device=dev)
    opts = dict(backend="gpu", stencil_bounds=case.cfg["stencil_bounds"], tol=tol, device=dev)
    for fi, f in enumerate(case.frequencies_hz):
        log(f"observations: forward solve {fi + 1}/{len(case.frequencies_hz)} at {f / 1e6:.3f} MHz")
        truth = G.HelmholtzSolver(case.x_m, case.y_m, torch.as_tensor(case.c_true, device=dev), None, f, -1,
                                  case.cfg["pml_strength"], case.cfg["pml_m"], opts)
        src = torch.zeros((ntx, nx, ny), dtype=torch.complex128, device=dev)
        src.view(ntx, -1)[torch.arange(ntx, device=dev), tx0] = 1
        u, _ = truth.solve(src.permute(2, 1, 0), False)
        Y[:, :, fi] = u.permute(2, 1, 0).reshape(ntx, -1)[:, rx0]
        truth.close()
        del u, src
    return dict(x_m=case.x_m, y_m=case.y_m, shape_yx=(ny, nx), n_tx=ntx, tx_index=case.tx_index,
                rx_index=case.rx_index, frequencies_hz=list(case.frequencies_hz), Y=Y,
                mask=torch.ones((ntx, case.n_rx, len(case.frequencies_hz)), dtype=torch.bool, device=dev))


def run_continuation(args):
    """Eight frequency stages on their own grids (201^2 ... 669^2 interiors + 24-cell PML, 256 TX each);
    one pass = one refined grad eval per stage at homogeneous 1500 m/s."""
    import threading

    import torch

    from waveforminversionust_amd import helmholtz as G
    from waveforminversionust_amd import workloads as W
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    cases = W.continuation_cases()
    stages = []
    for case in cases:
        obs = synthetic_obs(G, case, dev, args.tol)
        c0 = torch.full((case.ny, case.nx), 1500.0, dtype=torch.float64, device=dev)
        stages.append((case, obs, c0, dict(case.cfg, mode="refined", tol=args.tol, device=dev)))
    log("stages: " + ", ".join(f"{c.ny}^2 @ {c.frequencies_hz[0] / 1e6:.1f} MHz" for c in cases))
    streams = [torch.cuda.Stream(device=dev) for _ in stages] if args.concurrent else None
    per_stage = [0.0] * len(stages)

    def one(i):
        case, obs, c0, cfg = stages[i]
        t = time.perf_counter()
        if streams:
            with torch.cuda.stream(streams[i]):
                G.wust_oracle(c0, obs, 0, cfg)
        else:
            G.wust_oracle(c0, obs, 0, cfg)
        per_stage[i] += time.perf_counter() - t

    def one_pass():
        if streams:
            th = [threading.Thread(target=one, args=(i,)) for i in range(len(stages))]
            for t in th:
                t.start()
            for t in th:
                t.join()
        else:
            for i in range(len(stages)):
                one(i)

    for _ in range(args.warmup):
        one_pass()
    torch.cuda.synchronize()
    per_stage = [0.0] * len(stages)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_pass()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    line = {
        "metric": "continuation passes/s (8 stages x 256 TX, one refined grad eval per stage)",
        "value": round(args.steps / elapsed, 4), "unit": "passes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 (fp32 MFMA) + f64 refinement", "data": "synthetic",
        "config": {"workload": "continuation8", "concurrent_streams": bool(args.concurrent),
                   "stages": [dict(grid=c.ny, frequency_hz=c.frequencies_hz[0], dx_m=c.dx,
                                   ms=round(1e3 * per_stage[i] / args.steps, 2)) for i, c in enumerate(cases)],
                   "refine_tol": args.tol},
    }
    print(json.dumps(line), flush=True)


def run_dtft(args):
    """RF -> frequency DTFT: 256 TX x 256 RX x 4096 float32 samples (40 MHz), 12 frequencies 0.3-1.2 MHz."""
    import numpy as np
    import torch

    from waveforminversionust_amd import observations as OB
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n_time, n_rx, n_tx, dt = 4096, 256, 256, 25e-9
    freqs = np.linspace(0.3e6, 1.2e6, 12)
    t = 1e-6 + np.arange(n_time) * dt
    g = torch.Generator(device=dev).manual_seed(5)
    rf = torch.randn((n_tx, n_rx, n_time), generator=g, device=dev, dtype=torch.float32).permute(2, 1, 0)
    active = np.ones((n_tx, n_rx), bool)
    for _ in range(args.warmup):
        OB.rf_dtft(rf, t, freqs, active)
    torch.cuda.synchronize()
    kms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P, ms = OB.rf_dtft(rf, t, freqs, active, return_kernel_ms=True)
        kms += ms
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ncols = n_rx * n_tx
    rf_bytes = n_time * ncols * 4
    out_bytes = ncols * len(freqs) * 8
    flops = 2.0 * 2 * len(freqs) * n_time * ncols
    k_s = kms / 1e3 / args.steps
    gbs = (rf_bytes + out_bytes) / k_s / 1e9
    tfs = flops / k_s / 1e12
    cpu = None
    if not args.no_cpu:
        from threadpoolctl import threadpool_limits

        from oracle import observations_fp64 as O
        sub = 8
        x = rf[:, :, :sub].cpu().numpy()
        with threadpool_limits(1):
            c0 = time.perf_counter()
            want = O.rf_dtft(x, t, freqs, active[:sub], single=False)
            c_s = time.perf_counter() - c0
        got = P[:sub].cpu().numpy().astype(np.complex128)
        cpu = dict(value=round(sub / c_s, 3), unit="TX/s", cores=1, kind="port",
                   sample=f"FP64 numpy oracle (1 thread) on {sub} of {n_tx} TX ({n_rx} RX x {n_time} samples): "
                          f"{c_s:.3f} s; GPU rel err vs it {np.linalg.norm(got - want) / np.linalg.norm(want):.2e}")
    line = {
        "metric": "RF->frequency DTFT transmitters/s (256 RX x 4096 float32 samples, 12 frequencies)",
        "value": round(n_tx * args.steps / elapsed, 1), "unit": "TX/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (MFMA) from f32 RF, c64 result", "data": "synthetic",
        "config": {"workload": "dtft_256tx_256rx_4096t_12f", "n_tx": n_tx, "n_rx": n_rx, "n_time": n_time,
                   "n_freq": len(freqs)},
        # FP64 MFMA time (4 nf Nt N flops at 78.6 TFLOP/s, 0.16 ms) exceeds the HBM time of the RF stream
        # (0.13 ms at 8 TB/s), so the kernel is priced against the FP64 MFMA peak
        "roofline": {"kernel": "dtft_kernel", "bound": "mfma", "achieved": round(tfs, 2), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(tfs / 78.6, 4), "traffic": None,
                     "avg_launch_us": round(k_s * 1e6, 2), "algorithmic_flops_per_launch": flops,
                     "algorithmic_bytes_per_launch": rf_bytes + out_bytes, "hbm_gbs": round(gbs, 1),
                     "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_cfg1(args):
    """The reference's CPU-runnable case: 128^2 interior + 16-cell PML, 64-element ring, two discs,
    0.5 MHz, 40 dB noise; five NCG iterations from 1500 m/s (mask r < 26 mm, bounds 1400-1600 m/s)."""
    import numpy as np
    import torch

    from waveforminversionust_amd import helmholtz as G
    from waveforminversionust_amd import reconstruct as Rc
    from waveforminversionust_amd import workloads as W
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    case = W.cpu_reference_case()
    obs = synthetic_obs(G, case, dev, 1e-12)
    rng = np.random.default_rng(40)
    Y = obs["Y"].cpu().numpy()
    for t in range(case.n_tx):                        # complex Gaussian noise at 40 dB SNR per transmitter
        sig = np.sqrt(np.mean(np.abs(Y[t, :, 0]) ** 2) / 10 ** 4 / 2)
        Y[t, :, 0] += sig * (rng.standard_normal(case.n_rx) + 1j * rng.standard_normal(case.n_rx))
    obs = dict(obs, Y=Y, mask=np.ones(Y.shape, bool))
    X, Yg = np.meshgrid(case.x_m, case.y_m)
    mask = X ** 2 + Yg ** 2 < 26e-3 ** 2
    cfg = dict(case.cfg, step_damping=0.25, max_update_mps=12.0, bounds_mps=[1400.0, 1600.0], filter_cutoff=0.0,
               filter_order=4, update_mask=mask, frequency_schedule_hz=[0.5e6] * 5, tol=args.tol)
    init = 1500 * np.ones((case.ny, case.nx))
    for _ in range(args.warmup):
        Rc.wust_reconstruct(obs, init, dict(cfg, device=dev))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = Rc.wust_reconstruct(obs, init, dict(cfg, device=dev))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    c = res["c_mps"].cpu().numpy()
    err = lambda m: float(np.linalg.norm((m - case.c_true)[mask]) / np.linalg.norm(case.c_true[mask]))
    cpu = None
    if not args.no_cpu:
        from threadpoolctl import threadpool_limits

        from oracle import helmholtz_fp64 as H
        from oracle import ncg_fp64 as O
        with threadpool_limits(1):
            c0 = time.perf_counter()
            co, _ = O.wust_reconstruct(obs, init, dict(cfg, schedule=[0] * 5), H.wust_oracle,
                                       H.directional_curvature)
            c_s = time.perf_counter() - c0
        cpu = dict(value=round(c_s, 2), unit="s", cores=1, kind="port",
                   sample=f"FP64 oracle loop (scipy SuperLU, 1 thread), the same 5 NCG iterations: {c_s:.2f} s; "
                          f"final model rel err {err(co):.4%}, max |c_gpu - c_cpu| {np.max(np.abs(c - co)):.2e} m/s")
    line = {
        "metric": "cfg1 reconstruction wall time (160^2, 64 TX, 5 NCG iterations at 0.5 MHz)",
        "value": round(elapsed, 4), "unit": "s", "n_gpus": 1, "steps": 5, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / 5, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (fp32 MFMA) + f64 refinement", "data": "synthetic",
        "config": {"workload": case.name, "n_tx": case.n_tx, "refine_tol": args.tol, "snr_db": 40},
        "model_rel_err_initial_final": [round(err(init), 6), round(err(c), 6)],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_breast(args):
    """Full reconstruction: 12 frequencies 0.3-1.2 MHz x 3 NCG iterations on 1088^2 nodes, 512 TX."""
    import numpy as np
    import torch

    from waveforminversionust_amd import helmholtz as G
    from waveforminversionust_amd import reconstruct as Rc
    from waveforminversionust_amd import workloads as W
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    case = W.breast_case()
    t0 = time.perf_counter()
    obs = synthetic_obs(G, case, dev, args.tol)
    log(f"observations: {len(case.frequencies_hz)} forward solves in {time.perf_counter() - t0:.1f} s")
    X, Y = np.meshgrid(case.x_m, case.y_m)
    mask = X ** 2 + Y ** 2 < 55e-3 ** 2
    cfg = dict(case.cfg, step_damping=0.25, max_update_mps=12.0, bounds_mps=[1400.0, 1650.0], filter_cutoff=0.5,
               filter_order=4, update_mask=mask,
               frequency_schedule_hz=[f for f in case.frequencies_hz for _ in range(3)], tol=args.tol, device=dev)
    init = torch.full((case.ny, case.nx), 1500.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = Rc.wust_reconstruct(obs, init, cfg, progress=log)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    rec = res["records"]
    c = res["c_mps"].cpu().numpy()
    err0 = float(np.linalg.norm((1500.0 - case.c_true)[mask]) / np.linalg.norm(case.c_true[mask]))
    err = float(np.linalg.norm((c - case.c_true)[mask]) / np.linalg.norm(case.c_true[mask]))
    line = {
        "metric": "full reconstruction wall time (1088^2, 512 TX, 12 frequencies x 3 NCG steps)",
        "value": round(elapsed, 2), "unit": "s", "n_gpus": 1, "steps": len(rec), "warmup": 0,
        "ms_per_step": round(1e3 * elapsed / max(len(rec), 1), 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (fp32 MFMA) + f64 refinement", "data": "synthetic",
        "config": {"workload": case.name, "grid": f"{case.ny}x{case.nx} (1024^2 interior + 32-cell PML)",
                   "dx_m": case.dx, "n_tx": case.n_tx, "source_batch_size": case.cfg["source_batch_size"],
                   "refine_tol": args.tol},
        "per_step_ms": {"factor": round(1e3 * float(np.mean([r["factor_seconds"] for r in rec])), 1),
                        "forward": round(1e3 * float(np.mean([r["forward_seconds"] for r in rec])), 1),
                        "linearized": round(1e3 * float(np.mean([r["linearized_seconds"] for r in rec])), 1),
                        "wall": round(1e3 * float(np.mean([r["wall_seconds"] for r in rec])), 1)},
        "loss_first_last": [rec[0]["loss_before_update"], rec[-1]["loss_before_update"]],
        "model_rel_err_initial_final": [round(err0, 6), round(err, 6)],
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: a multi-GPU run would "
              "replicate the factor on every GPU, shard the transmitters and all-reduce the gradient, "
              "which is not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "continuation8":
        return run_continuation(args)
    if args.workload == "breast1088":
        return run_breast(args)
    if args.workload == "dtft":
        return run_dtft(args)
    if args.workload == "cfg1":
        return run_cfg1(args)
    import numpy as np
    import torch

    from waveforminversionust_amd import _native as N
    from waveforminversionust_amd import helmholtz as G
    from waveforminversionust_amd import workloads as W

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    lib = N.load()
    case = W.metric_case()
    f = case.frequencies_hz[0]
    ny, nx, ntx = case.ny, case.nx, case.n_tx
    log(f"workload {case.name}: {ny}x{nx} nodes, {ntx} TX, f={f/1e6} MHz")

    # ---- synthetic observations (untimed) ----
    opts = dict(backend="gpu", stencil_bounds=case.cfg["stencil_bounds"], tol=args.tol, device=dev)
    truth = G.HelmholtzSolver(case.x_m, case.y_m, torch.as_tensor(case.c_true, device=dev), None, f, -1,
                              case.cfg["pml_strength"], case.cfg["pml_m"], opts)
    src = torch.zeros((ntx, nx, ny), dtype=torch.complex128, device=dev)
    tx0 = torch.as_tensor(case.tx_index - 1, device=dev)
    src.view(ntx, -1)[torch.arange(ntx, device=dev), tx0] = 1
    u, _ = truth.solve(src.permute(2, 1, 0), False)
    rx0 = torch.as_tensor(case.rx_index - 1, device=dev)
    Y = u.permute(2, 1, 0).reshape(ntx, -1)[:, rx0].contiguous()   # [ntx, nrx]
    truth.close()
    del u, src
    obs = dict(x_m=case.x_m, y_m=case.y_m, shape_yx=(ny, nx), n_tx=ntx, tx_index=case.tx_index,
               rx_index=case.rx_index, frequencies_hz=[f], Y=Y[:, :, None],
               mask=torch.ones((ntx, case.n_rx, 1), dtype=torch.bool, device=dev))
    c0 = torch.full((ny, nx), 1500.0, dtype=torch.float64, device=dev)
    cfg = dict(case.cfg, mode="refined", tol=args.tol, device=dev)

    def step(mode="refined", c=None):
        return G.wust_oracle(c0 if c is None else c, obs, 0, dict(cfg, mode=mode))

    # each timed evaluation runs on its own model, as the evaluations of an NCG loop do:
    # c0 (1 + 1e-3 k b), k = 1..K, b a smooth bump, so every factorization meets the Newton-Schulz
    # round profile learned from the previous model rather than from the same operator
    yy = torch.linspace(-1, 1, ny, dtype=torch.float64, device=dev)[:, None]
    xx = torch.linspace(-1, 1, nx, dtype=torch.float64, device=dev)[None, :]
    bump = torch.exp(-((xx - 0.2) ** 2 + (yy + 0.1) ** 2) / 0.18)
    models = [c0 * (1 + 1e-3 * (k + 1) * bump) for k in range(args.steps)]

    # ---- warm-up, then one fully profiled step for the phase breakdown ----
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # the sweep GEMM is the largest kernel class of a step (about half of it at
    # 560^2 x 256 TX); its flops per launch are fixed, so it carries the roofline
    dominant = "sweep_gemm"
    dom_idx = N.PROFILE_CLASSES.index(dominant)
    lib.wust_profile_reset()
    lib.wust_profile_enable((1 << len(N.PROFILE_CLASSES)) - 1)
    step()
    torch.cuda.synchronize()
    phases = {}
    for ci, name in enumerate(N.PROFILE_CLASSES):
        cnt, ms = N.C.c_int64(), N.C.c_double()
        lib.wust_profile_read(ci, N.C.byref(cnt), N.C.byref(ms))
        if cnt.value:
            phases[name] = dict(launches=cnt.value, ms=round(ms.value, 3))
    step_cnt, step_ms, step_units = N.C.c_int64(), N.C.c_double(), N.C.c_int64()
    lib.wust_profile_read(dom_idx, N.C.byref(step_cnt), N.C.byref(step_ms))
    lib.wust_profile_units(dom_idx, N.C.byref(step_units))
    largest = max(phases, key=lambda k: phases[k]["ms"]) if phases else dominant
    if largest != dominant:
        log(f"note: largest kernel class of the profiled step is {largest}, roofline stays on {dominant}")
    lib.wust_profile_enable(0)
    lib.wust_profile_reset()

    # ---- timed region: K refined gradient evaluations; one launch in SAMPLE of
    # the dominant kernel class is timed by its dispatch packet throughout (an
    # event-timed launch's completion signal idles the device for several
    # microseconds before the next launch, so timing all of them would slow
    # the region it measures) ----
    SAMPLE = 32
    lib.wust_profile_stride(SAMPLE)
    lib.wust_profile_enable(1 << dom_idx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(c=models[i])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cnt, ms, units = N.C.c_int64(), N.C.c_double(), N.C.c_int64()
    lib.wust_profile_read(dom_idx, N.C.byref(cnt), N.C.byref(ms))
    lib.wust_profile_units(dom_idx, N.C.byref(units))
    lib.wust_profile_enable(0)
    lib.wust_profile_stride(1)
    value = args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    log(f"refined: {ms_per_step:.1f} ms/grad eval, dominant {dominant}: {cnt.value} sampled launches, "
        f"{ms.value:.1f} ms")
    # the same K evaluations on one model (each factorization on the profile learned from that very
    # operator: the best case), which also gives the accuracy leg its refined result at c0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss_ref, g_ref = step()
    torch.cuda.synchronize()
    same_elapsed = time.perf_counter() - t0
    same_model = dict(value=round(args.steps / same_elapsed, 4),
                      ms_per_step=round(1e3 * same_elapsed / args.steps, 2))
    log(f"refined, one model: {same_model['ms_per_step']} ms/grad eval")

    # roofline of the dominant kernel (algorithmic flops; n = Ny-2 interior rows per block)
    n, m, R = ny - 2, nx - 2, ntx
    # each launch = up to two chain steps, each an n x n times n x R complex product (8 flops / MAC);
    # the library counts the chain steps its profiled launches executed (a solve has 2m-1 in m+1
    # launches, fewer when the forward elimination skips columns where the right-hand side is zero)
    # `achieved` is priced on the fully profiled evaluation just before the timed region: every launch
    # of the class timed by its dispatch packet (the per-launch durations rocprofv3's kernel trace
    # reports); the timed region itself times one launch in SAMPLE ("sampled"), whose isolated
    # completion signal adds ~1.5-2 us to the launch it times
    total_flops = 8.0 * n * n * R * step_units.value
    flops_per_launch = total_flops / max(step_cnt.value, 1)
    per_unit = (f"8*n^2*R flops per chain step (n={n}, R={R}); {step_units.value} chain steps executed in the "
                f"{step_cnt.value} launches of one profiled evaluation (a full solve has 2m-1 = {2 * m - 1} chain "
                f"steps in m+1 = {m + 1} launches)")
    avg_launch_s = step_ms.value / 1e3 / max(step_cnt.value, 1)
    achieved = (total_flops / (step_ms.value / 1e3)) / 1e12 if step_ms.value > 0 else 0.0
    sampled_flops = 8.0 * n * n * R * units.value
    sampled = dict(launches=cnt.value, sample_stride=SAMPLE,
                   avg_launch_us=round(ms.value / 1e3 / max(cnt.value, 1) * 1e6, 2),
                   achieved=round((sampled_flops / (ms.value / 1e3)) / 1e12, 2) if ms.value > 0 else 0.0)
    roofline = dict(kernel=dominant, bound="mfma", achieved=round(achieved, 2), peak=MFMA_FP32_PEAK_TFLOPS,
                    unit="TFLOP/s", frac=round(achieved / MFMA_FP32_PEAK_TFLOPS, 4),
                    traffic=measured_traffic(dominant, case.name),
                    timed_launches=step_cnt.value, avg_launch_us=round(avg_launch_s * 1e6, 2),
                    algorithmic_flops_per_launch=round(flops_per_launch / 1e9, 4), per_unit=per_unit,
                    sampled_in_timed_region=sampled)

    # ---- the engine's own unrefined mode (same kernels, no FP64 refinement) ----
    c64 = None
    if args.c64_steps > 0:
        step("c64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.c64_steps):
            loss_c64, g_c64 = step("c64")
        torch.cuda.synchronize()
        ce = time.perf_counter() - t0
        # refined_time_over_c64: what the FP64 refinement (one correction solve per direction, residual
        # passes) costs on top of the unrefined kernels
        c64 = dict(value=round(args.c64_steps / ce, 4), ms_per_step=round(1e3 * ce / args.c64_steps, 2),
                   steps=args.c64_steps,
                   refined_time_over_c64=round((same_elapsed / args.steps) / (ce / args.c64_steps), 3))
        log(f"c64: {c64['ms_per_step']} ms/grad eval (refined / c64 time {c64['refined_time_over_c64']})")

    # ---- plain comparator ----
    plain = None
    if not args.no_plain:
        step("plain")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            loss_plain, g_plain = step("plain")
        torch.cuda.synchronize()
        pe = time.perf_counter() - t0
        plain = dict(value=round(args.plain_steps / pe, 4), ms_per_step=round(1e3 * pe / args.plain_steps, 1),
                     steps=args.plain_steps,
                     refined_over_plain=round(value / (args.plain_steps / pe), 2),
                     grad_rel_diff_vs_refined=float(torch.linalg.norm(g_plain - g_ref) / torch.linalg.norm(g_ref)))
        log(f"plain: {plain['ms_per_step']} ms/grad eval -> refined/plain = {plain['refined_over_plain']}x")

    # ---- FP64 CPU evaluations of the whole step: accuracy + CPU baseline ----
    accuracy, cpu = None, None
    if not args.no_cpu:
        from threadpoolctl import threadpool_limits

        from oracle import blocklu_fp64 as BL
        from oracle import helmholtz_fp64 as H
        host = host_info()
        Yh = Y.cpu().numpy()
        obs_h = dict(x_m=case.x_m, y_m=case.y_m, shape_yx=(ny, nx), n_tx=ntx, tx_index=case.tx_index,
                     rx_index=case.rx_index, frequencies_hz=[f], Y=Yh[:, :, None],
                     mask=np.ones((ntx, case.n_rx, 1), bool))
        c0h = np.full((ny, nx), 1500.0)
        cfg_h = dict(case.cfg, source_batch_size=ntx)
        # (i) the sparse-LU oracle (scipy SuperLU, the analogue of MATLAB's decomposition(H,'lu')) on one
        # thread: the full 256-TX gradient evaluation, median of --cpu-repeats runs; the first run also
        # keeps the forward fields for the field error
        runs = []
        with threadpool_limits(1):
            for r in range(max(1, args.cpu_repeats)):
                log(f"CPU baseline: FP64 sparse-LU grad eval {r + 1}/{max(1, args.cpu_repeats)} (1 thread)")
                t0 = time.perf_counter()
                if r == 0:
                    lo, go, so = H.wust_oracle(c0h, obs_h, 0, cfg_h, 3)
                else:
                    H.wust_oracle(c0h, obs_h, 0, cfg_h, 2)
                runs.append(time.perf_counter() - t0)
        t_med = float(np.median(runs))
        # (ii) FP64 block-tridiagonal LU (the reference's GPU algorithm in FP64) on every host thread the
        # box grants this process
        with threadpool_limits(host["threads"]):
            log(f"CPU baseline: FP64 block-LU grad eval ({host['threads']} threads)")
            t0 = time.perf_counter()
            lb, gb = BL.wust_oracle_blocklu(c0h, obs_h, 0, cfg_h)
            t_blk = time.perf_counter() - t0
        cpu = dict(value=round(1.0 / t_med, 6), unit="grad evals/s", cores=1, kind="port",
                   sample=(f"the full step on the FP64 oracle: 560x560 grid, all 256 TX and RX, scipy SuperLU on 1 "
                           f"thread (assembly, LU of H and H^H, 2x256 solves, fit, gradient), median of "
                           f"{len(runs)} runs: {t_med:.2f} s per grad eval"),
                   runs_s=[round(t, 2) for t in runs],
                   multicore=dict(value=round(1.0 / t_blk, 6), unit="grad evals/s", cores=host["threads"],
                                  kind="port", seconds=round(t_blk, 2),
                                  sample=(f"FP64 block-tridiagonal LU over grid columns (dense per-block inverses and "
                                          f"multi-RHS products on the host BLAS, {host['threads']} threads), the "
                                          f"same full step: {t_blk:.2f} s"),
                                  grad_rel_diff_vs_sparse_lu=float(np.linalg.norm(gb - go) / np.linalg.norm(go))),
                   host=host)
        log(f"CPU baseline: sparse LU {t_med:.2f} s (runs {runs}), block LU {t_blk:.2f} s on {host['threads']} threads")
        # accuracy of each GPU mode against the FP64 oracle over the same full step (all 256 TX)
        uo = np.concatenate(so.virtual, axis=2) / so.solver.virtual_source_multiplier[:, :, None]
        inner = (slice(1, -1), slice(1, -1))
        results = {"refined": (loss_ref, g_ref)}
        if c64 is not None:
            results["c64"] = (loss_c64, g_c64)
        if plain is not None:
            results["plain"] = (loss_plain, g_plain)
        accuracy = {}
        for mode, (lg, gg) in results.items():
            gg = gg.cpu().numpy()
            sm = G.HelmholtzSolver(case.x_m, case.y_m, c0, None, f, -1, case.cfg["pml_strength"], case.cfg["pml_m"],
                                   dict(opts, mode=mode))
            srcs = torch.zeros((ntx, nx, ny), dtype=torch.complex128, device=dev)
            srcs.view(ntx, -1)[torch.arange(ntx, device=dev), tx0] = 1
            ug = sm.solve(srcs.permute(2, 1, 0), False)[0].cpu().numpy()
            sm.close()
            del srcs
            accuracy[mode] = dict(
                field_rel_err=float(np.linalg.norm(ug[inner] - uo[inner]) / np.linalg.norm(uo[inner])),
                misfit_rel_err=float(abs(lg - lo) / abs(lo)),
                grad_rel_err=float(np.linalg.norm(gg - go) / np.linalg.norm(go)))
        accuracy["sample"] = "all 256 TX and RX of the timed step, full 560x560 grid (fields: interior nodes)"
        log(f"accuracy vs FP64 (all 256 TX): {accuracy}")

    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 MFMA) + f64 refinement",
        "data": "synthetic",
        "config": {"workload": case.name, "grid": f"{ny}x{nx} (512^2 interior + 24-cell PML)", "dx_m": case.dx,
                   "frequency_hz": f, "n_tx": ntx, "n_rx": case.n_rx, "mode": "refined", "refine_tol": args.tol,
                   "models": ("timed evaluation k on c0 (1 + 1e-3 k b), c0 = 1500 m/s, b a smooth bump "
                              "(a new operator per evaluation, as in an NCG loop)")},
        "same_model": same_model,
        "c64": c64,
        "plain": plain,
        "accuracy_vs_fp64": accuracy,
        "phases_ms_one_step": phases,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
