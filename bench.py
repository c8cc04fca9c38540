"""Throughput benchmark of the MI355X Fenton-Karma forward-Euler path.

One bench "step" = one full forward solve of the single-large-tissue
workload: 1024 x 1024 grid, PARAMSET_3, scarred diffusivity, S1 (WEST stripe
at step 0) + S2 (NORTH half field at a seeded step in [28k, 36k)), 1e5 Euler
steps at dt = dx = 0.01, inputs resident in HBM. Metric: giga cell-updates per
second (GCUPS = H * W * Euler steps / s / 1e9).

Prints ONE JSON line (rank 0). Progress goes to stderr.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--sim-steps S] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak, GB/s
# What `traffic` counts: rocprofv3 FETCH_SIZE + WRITE_SIZE, the L2-to-fabric bytes,
# which include what the 256 MB MALL serves (a 1024^2 grid's fields fit there).
TRAFFIC_NOTE = "FETCH_SIZE + WRITE_SIZE (L2-to-fabric; includes MALL hits)"
# What `achieved` counts: the algorithmic bytes of the Euler steps a launch
# performs, so for a two-step launch it is an effective bandwidth (the launch
# moves the state once for both steps; `traffic` is what it moved).
ACHIEVED_BASIS = "28 B per cell-update x the launch's cell-updates (effective for two-step launches)"
BYTES_PER_CELL_UPDATE = 28  # read v, w, u, D + write v, w, u (fp32)

# Step-kernel variant -> (kernel the variant launches, Euler steps per launch).
KERNELS = {"stream8": ("k_stream<8, true>", 1), "stream16": ("k_stream<16, true>", 1),
           "stream32": ("k_stream<32, true>", 1), "tile16": ("k_tile<16, true>", 1),
           "tile32": ("k_tile<32, true>", 1), "tb64x2": ("k_tb<64, 2, 1024>", 2),
           "tb64x4": ("k_tb<64, 4, 1024>", 4), "tb32x2": ("k_tb<32, 2, 256>", 2),
           "tb32x4": ("k_tb<32, 4, 512>", 4), "streamc8": ("k_stream_c<8>", 1),
           "streamc16": ("k_stream_c<16>", 1), "stream4": ("k_stream<4, true>", 1),
           "streamc4": ("k_stream_c<4>", 1), "stream2": ("k_stream<2, true>", 1),
           "streamc2": ("k_stream_c<2>", 1), "pair2": ("k_stream2<2, true>", 1),
           "pair4": ("k_stream2<4, true>", 1), "streamt8": ("k_stream_t<8>", 2),
           "streamt12": ("k_stream_t<12>", 2), "streamt16": ("k_stream_t<16>", 2),
           "streamt24": ("k_stream_t<24>", 2)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, H, W):
    """HBM bytes per launch of `kernel` at H x W from the committed rocprofv3
    --pmc summary (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel of known
    size); None when no summary covers this kernel and size."""
    return measured_traffic_key("%s %dx%d" % (kernel, H, W))


def measured_traffic_key(what):
    """The `traffic_bytes_per_launch[what] = N` line of a summary in profiles/."""
    import glob
    import re

    key = "traffic_bytes_per_launch[%s]" % what
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    m = re.search(r"=\s*(\d+)", line)
                    if m:
                        return int(m.group(1))
    return None


def cpu_affinity():
    """Cores this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def cpu_cores():
    """Threads of the CPU baseline: every core of the affinity mask, bounded
    by OMP_NUM_THREADS where the host sets it (the CPU share of one GPU on a
    shared node)."""
    n = cpu_affinity()
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(sim, target_s=15.0):
    """fp32 C restatement of the reference op graph (oracle/fk_oracle.c) on
    the host cores, over a bounded sample of the same workload."""
    import numpy as np

    from oracle import c_oracle, fk_oracle

    stim = [fk_oracle.Stimulus(s.protocol, s.field) for s in sim.stimuli]
    st = fk_oracle.init(sim.shape, np.float32)
    cores = cpu_cores()
    H, W = sim.shape
    n0 = 10
    t0 = time.perf_counter()
    st = c_oracle.forward_euler(st, 0, n0, sim.params, sim.D, stim, sim.dt, sim.dx,
                                np.float32, cores)
    dt0 = time.perf_counter() - t0
    n1 = int(max(10, min(50000, n0 * target_s / max(dt0, 1e-6))))
    t0 = time.perf_counter()
    c_oracle.forward_euler(st, n0, n0 + n1, sim.params, sim.D, stim, sim.dt, sim.dx,
                           np.float32, cores)
    el = time.perf_counter() - t0
    return {"value": H * W * n1 / el / 1e9, "unit": "GCUPS", "cores": cores, "kind": "port",
            "sample": "%dx%d grid, Euler steps %d..%d of the same inputs, fp32 C "
                      "restatement of the reference op graph, %d OpenMP threads, %.1f s"
                      % (H, W, n0, n0 + n1, cores, el),
            "cpu_model": cpu_model(), "cores_in_affinity": cpu_affinity(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_model():
    """Model name of the host CPU as lscpu reports it (/proc/cpuinfo when lscpu
    is absent), with the socket, core and thread counts."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    name = info.get("Model name")
    if not name:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        name = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    extra = ", ".join("%s %s" % (k.lower(), info[k]) for k in
                      ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                       "CPU max MHz") if info.get(k))
    return "%s (%s)" % (name, extra) if name and extra else name


def bench_batched(n_sims, n_steps, seed, per_sim_steps=64):
    """256 independent simulations, sides from {64,128,256,384,512}, seeded
    parameter sets, scar maps and S1/S2 timings: one batched graph-replayed
    run of n_steps against per-simulation launches (one _forward_euler call
    per simulation) over per_sim_steps steps."""
    import torch

    from cardiax_amd import solve, workloads
    from cardiax_amd.batch import Batch, BatchSim

    dev = torch.device("cuda", 0)
    sims = workloads.c3_inputs(seed, n=n_sims, steps=n_steps)
    bs = []
    for s in sims:
        stims = [(x.protocol, torch.from_numpy(x.field).to(dev)) for x in s.stimuli]
        bs.append(BatchSim(solve.init(s.shape), s.params, torch.from_numpy(s.D).to(dev),
                           stims, float(s.t0), s.dt, s.dx))
    cells = sum(s.D.size for s in sims)
    with Batch(bs) as b:
        tuned = b.autotune(32)
        log("batch autotune: %s" % json.dumps(tuned))
        b.run(64)  # warm-up: records the graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run(n_steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        n_bad = sum(int(not bool(torch.isfinite(x.u).all().item())) for x in b.states)
        umax = max(float(x.u.abs().max().item()) for x in b.states)
    batched = cells * n_steps / el / 1e9
    # the tuned work list: one or two Euler steps per launch
    per_launch = 2 if tuned["two_step"] else 1
    kernel = ("k_batch_step_t<%d>" if tuned["two_step"] else "k_batch_step<%d>") % tuned["rows"]
    # per-simulation launches: the same simulations, one call each
    for x in bs[:4]:
        solve._forward_euler(x.state, 0, 4, x.params, x.diffusivity, x.stimuli, x.dt, x.dx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for x in bs:
        solve._forward_euler(x.state, x.t, x.t + per_sim_steps, x.params, x.diffusivity,
                             x.stimuli, x.dt, x.dx)
    torch.cuda.synchronize()
    el1 = time.perf_counter() - t0
    per_sim = cells * per_sim_steps / el1 / 1e9
    return {"workload": "%d simulations, sides %s, %.1fM cells, %d Euler steps"
                        % (n_sims, list(workloads.C3_SIDES), cells / 1e6, n_steps),
            "gcups": round(batched, 3), "seconds": round(el, 3),
            "per_sim_launch_gcups": round(per_sim, 3),
            "per_sim_launch_steps": per_sim_steps,
            "speedup_vs_per_sim": round(batched / per_sim, 2), "autotune": tuned,
            "u_finite": n_bad == 0,
            "n_nonfinite_sims": n_bad, "u_absmax": round(umax, 4),
            "roofline": dict(leg_roofline(BYTES_PER_CELL_UPDATE * cells * per_launch,
                                          el / (n_steps / per_launch), kernel,
                                          "%d simulations, %.1f M cells"
                                          % (n_sims, cells / 1e6)),
                             achieved_basis=ACHIEVED_BASIS)}


def traffic_rate(traffic, seconds_per_launch):
    """Measured bytes per launch over this run's time per launch, GB/s (None
    without a measurement): the rate beside the algorithmic one."""
    return None if traffic is None else round(traffic / seconds_per_launch / 1e9, 1)


def leg_roofline(bytes_per_launch, seconds_per_launch, kernel, key):
    """HBM roofline of a leg's dominant kernel: algorithmic bytes per launch
    over the measured time per launch, with the traffic a committed
    rocprofv3 --pmc summary measured for it (None when there is none)."""
    achieved = bytes_per_launch / seconds_per_launch / 1e9
    traffic = measured_traffic_key("%s %s" % (kernel, key))
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "traffic_gbs": traffic_rate(traffic, seconds_per_launch),
            "kernel": kernel,
            "per_launch_us": round(seconds_per_launch * 1e6, 3),
            "bytes_per_launch": int(bytes_per_launch), "traffic_counts": TRAFFIC_NOTE}


def bench_adjoint(seed, size=256, n_steps=2000, frame_every=50, reps=2):
    """Gradient of the frame MSE with respect to D (size^2) and tau_d, tau_r,
    tau_si by the hand-written adjoint with sqrt(N) checkpoints, against the
    plain forward of the same rollout."""
    import numpy as np
    import torch

    from cardiax_amd import adjoint, solve, workloads
    from cardiax_amd.params import PARAMSET_3 as prm

    dev = torch.device("cuda", 0)
    shape = (size, size)
    rng = np.random.default_rng(seed)
    Dstar = torch.from_numpy(workloads.scar_map(shape, rng)).to(dev)
    sim = workloads.c1_inputs(seed, shape=shape, steps=n_steps)
    stims = [(x.protocol, torch.from_numpy(x.field).to(dev)) for x in sim.stimuli]
    st = solve.init(shape)
    cps = list(range(0, n_steps + 1, frame_every))
    target = torch.stack([x.u for x in solve.forward(st, cps, prm, Dstar, stims, 0.01, 0.01)])
    D = torch.full(shape, 1e-3, device=dev, requires_grad=True)
    taus = [torch.tensor(float(getattr(prm, k)), device=dev, requires_grad=True)
            for k in adjoint.GRAD_FIELDS]

    def fwd_only():
        with torch.no_grad():
            return adjoint.rollout(D, *taus, st, prm, stims, 0, n_steps, frame_every, 0.01, 0.01)

    def fwd_bwd():
        D.grad = None
        fr = adjoint.rollout(D, *taus, st, prm, stims, 0, n_steps, frame_every, 0.01, 0.01)
        ((fr - target) ** 2).mean().backward()

    fwd_only(), fwd_bwd()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fwd_only()
    torch.cuda.synchronize()
    tf = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        fwd_bwd()
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / reps
    cells = size * size * n_steps
    per_step = adjoint_step_seconds(st, prm, Dstar, stims, shape, t_mid=n_steps // 2)
    fwd_adj = ADJ_FWD_BWD_BYTES * cells / tb / 1e9
    return {"workload": "%dx%d, %d Euler steps, u frames every %d, MSE vs a hidden scar D-map, "
                        "grad wrt D and tau_d/tau_r/tau_si, sqrt(N) checkpoints"
                        % (size, size, n_steps, frame_every),
            "forward_s": round(tf, 4), "fwd_adjoint_s": round(tb, 4),
            "cost_ratio": round(tb / tf, 2),
            "forward_gcups": round(cells / tf / 1e9, 3),
            "fwd_adjoint_gcups": round(cells / tb / 1e9, 3),
            "fwd_adjoint_hbm": {"bytes_per_cell_step": ADJ_FWD_BWD_BYTES,
                                "achieved": round(fwd_adj, 1),
                                "frac": round(fwd_adj / HBM_PEAK_GBS, 4)},
            "roofline": dict(leg_roofline(ADJ_BWD_BYTES * size * size, per_step,
                                          "k_adj_main<true>", "%dx%d" % shape),
                             design_bytes_per_launch=ADJ_BWD_DESIGN_BYTES * size * size),
            "grad_finite": bool(torch.isfinite(D.grad).all().item())}


# Algorithmic bytes of the adjoint per cell-step: the reverse step reads the
# cotangent (12), the state (12) and D (4), writes the cotangent (12) and
# read-modify-writes dL/dD (8); forward + backward adds the forward step (28)
# and its recompute from the checkpoints (28).
ADJ_BWD_BYTES = 48
ADJ_FWD_BWD_BYTES = 104
# What the fused reverse kernel moves by construction per cell: it also reads
# the derived fields of its step (28), lam_v/lam_w, v, w, u, D, dL/dD and the
# previous u (32 more in all) and writes the next step's derived fields (28).
ADJ_BWD_DESIGN_BYTES = 104


def adjoint_step_seconds(st, prm, D, stims, shape, t_mid, n_short=10, n_long=40):
    """Time of one reverse step (k_adj_main<true>) on the stream it runs on:
    fk_adjoint_segment over n_long and n_short steps of a trajectory from the
    middle of the rollout, HIP events around each; the difference cancels the
    segment's fixed launches."""
    import ctypes

    import torch

    from cardiax_amd import _lib, solve

    dev = torch.device("cuda", 0)
    H, W = shape
    L = _lib.lib()
    x = solve._forward_euler(st, 0, t_mid, prm, D, stims, 0.01, 0.01)
    stim = solve._Stimuli(stims, dev, shape)
    p = _lib.make_params(prm)
    Dd = torch.as_tensor(D, device=dev)
    s = torch.cuda.current_stream()
    states = torch.empty((n_long, 3, H, W), device=dev)
    for f in range(3):
        states[0, f].copy_(x[f])
    _lib.check(L.fk_forward_trajectory(ctypes.byref(p), Dd.data_ptr(), x.v.data_ptr(),
                                       x.w.data_ptr(), x.u.data_ptr(), H, W, stim.pointer(),
                                       stim.n, float(t_mid), n_long - 1, 0.01, 0.01,
                                       states[1].data_ptr(), s.cuda_stream))
    nb = L.fk_adjoint_workspace_bytes(H, W)
    work = torch.empty(nb // 8 + 1, dtype=torch.float64, device=dev)
    g = torch.Generator(device="cpu").manual_seed(1)
    lam = [torch.randn(shape, generator=g).to(dev) * 1e-3 for _ in range(3)]
    gD = torch.zeros(shape, device=dev)
    _lib.check(L.fk_adjoint_prepare(Dd.data_ptr(), H, W, work.data_ptr(), nb, s.cuda_stream))

    def seg(n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        _lib.check(L.fk_adjoint_segment(
            ctypes.byref(p), Dd.data_ptr(), H, W, stim.pointer(), stim.n, float(t_mid), n,
            0.01, 0.01, states.data_ptr(), None, lam[0].data_ptr(), lam[1].data_ptr(),
            lam[2].data_ptr(), gD.data_ptr(), work.data_ptr(), nb, s.cuda_stream))
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    seg(n_long)  # warm-up
    t_long = min(seg(n_long) for _ in range(3))
    t_short = min(seg(n_short) for _ in range(3))
    return (t_long - t_short) / (n_long - n_short)


def bench_inverse(seed, n=32, iters=100, n_steps=1000, frame_every=50, per_instance_iters=2):
    """Batched inverse problem: n instances of 128^2..512^2 each recover a
    hidden scar D-map by `iters` Adam iterations of forward + adjoint, all
    instances through one batched rollout plan (one launch per forward step
    and per reverse step for all of them). For comparison, a few iterations
    of the per-instance path (one rollout per instance, 8 streams), which
    gives the same bits."""
    import numpy as np
    import torch

    from cardiax_amd import inverse
    from cardiax_amd import workloads as wl

    inst = inverse.make_instances(seed, n=n, n_steps=n_steps, frame_every=frame_every)
    inverse.recover(inst[:2], iters=1)  # warm-up (kernels, allocator)
    cells = sum(int(x.D_true.numel()) for x in inst)
    with inverse._plan(inst, wl.DT, wl.DX) as plan:  # forward only
        Ds = [torch.full_like(x.D_true, inverse.D_MAX) for x in inst]
        plan.forward(Ds)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.forward(Ds)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        plan_gb = plan.device_bytes / 1e9
    res = inverse.recover(inst, iters=iters, d_err_at=(50,) if iters > 50 else ())
    per = inverse.recover(inst, iters=per_instance_iters, batched=False)
    loss = res["loss"]
    d0, d1 = np.array(res["d_err_initial"]), np.array(res["d_err_final"])
    return {"workload": "%d instances, sides %s, %d Adam iterations of forward + adjoint, "
                        "%d Euler steps each, frames every %d, float64 loss"
                        % (n, sorted({int(x.D_true.shape[0]) for x in inst}), iters, n_steps,
                           frame_every),
            "seconds": round(res["seconds"], 3),
            "fwd_adjoint_gcups": round(cells * n_steps * iters / res["seconds"] / 1e9, 3),
            "per_instance_fwd_adjoint_gcups": round(
                cells * n_steps * per_instance_iters / per["seconds"] / 1e9, 3),
            "forward_only_seconds": round(tf, 4),
            "forward_only_gcups": round(cells * n_steps / tf / 1e9, 3),
            "cost_ratio": round(res["seconds"] / iters / tf, 2),
            "plan_device_gb": round(plan_gb, 3),
            # summed over instances (a few see no wave reach their scars, so their
            # loss starts near 1e-17 and a per-instance ratio says nothing)
            "loss_sum_final_over_initial": round(float(loss[:, -1].sum() / loss[:, 0].sum()), 6),
            "loss_ratio_median": round(float(np.median(loss[:, -1] / loss[:, 0])), 6),
            "d_err_initial": round(float(np.mean(d0)), 4),
            "d_err_final": round(float(np.mean(d1)), 4),
            # the configured 50 iterations, on the way to `iters`
            "d_err_after_50_iterations": (round(float(np.mean(res["d_err_at"][50])), 4)
                                          if 50 in res["d_err_at"] else None),
            "d_err_per_instance": [[int(x.D_true.shape[0]), round(float(a), 4), round(float(b), 4)]
                                   for x, a, b in zip(inst, d0, d1)],
            "instances_improved": int((d1 < d0).sum()),
            # the same over the cells a target frame shows excited (the cells the
            # frames constrain)
            "reached_frac_mean": round(float(np.mean(res["reached_frac"])), 4),
            "d_err_reached_initial": round(float(np.mean(res["d_err_reached_initial"])), 4),
            "d_err_reached_final": round(float(np.mean(res["d_err_reached_final"])), 4)}


def bench_heun(seed, size=1024, n_steps=2000):
    """TimeIntegrator.HEUN (cardiax/solve.py:103-111) on the single-tissue
    workload: two streaming launches per step. Algorithmic bytes per Heun
    cell-step: stage 1 reads v, w, u, D and writes k1 (28 B); stage 2 reads
    x0, k1 and D, rebuilds x1 = x0 + dt k1 in registers and writes the state
    (40 B): 68 B."""
    import torch

    from cardiax_amd import solve, workloads

    sim = workloads.c2_inputs(seed, shape=(size, size), steps=n_steps)
    dev = torch.device("cuda", 0)
    stims = [(s.protocol, torch.from_numpy(s.field).to(dev)) for s in sim.stimuli]
    D = torch.from_numpy(sim.D).to(dev)
    st0 = solve.init((size, size))
    solve._forward_heun(st0, 0, 200, sim.params, D, stims, sim.dt, sim.dx)  # warm-up
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    out = solve._forward_heun(st0, 0, n_steps, sim.params, D, stims, sim.dt, sim.dx)
    e1.record(stream)
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3
    gcups = size * size * n_steps / sec / 1e9
    # HBM bytes per Heun step from the committed PMC summary of both stages (None if absent)
    R = 2 if size * size <= 1536 * 1024 else 4
    t1 = measured_traffic("k_stream_heun<%d, 1>" % R, size, size)
    t2 = measured_traffic("k_stream_heun<%d, 2>" % R, size, size)
    traffic = t1 + t2 if t1 is not None and t2 is not None else None
    return {"workload": "%dx%d grid, %d Heun steps (two launches each)" % (size, size, n_steps),
            "gcups": round(gcups, 3), "us_per_step": round(sec / n_steps * 1e6, 3),
            "hbm_frac": round(gcups * 68 / HBM_PEAK_GBS, 4),
            "traffic_bytes_per_step": traffic,
            "u_finite": bool(torch.isfinite(out.u).all().item())}


def bench_slab(seed, n_steps, size=1024, n_parts=8, depths=(1, 4, 8), tune=True):
    """One GPU's share of the size x size tissue split into n_parts row slabs:
    size / n_parts owned rows plus 4k ghost rows on each side, the depth a
    halo exchange every k steps needs (the stencil reaches 4 rows per step).
    The slab's rows come from the middle of the single-tissue workload. GCUPS
    count owned cells only; the ghost rows are the recomputation that k steps
    between exchanges cost. No exchange is made: there is one GPU."""
    import numpy as np
    import torch

    from cardiax_amd import autotune, solve, workloads

    sim = workloads.c2_inputs(seed, shape=(size, size), steps=n_steps)
    dev = torch.device("cuda", 0)
    own = size // n_parts
    r_own = (size - own) // 2
    res = {"workload": "%d of %d row slabs of a %dx%d grid (%d owned rows), %d Euler steps"
                       % (1, n_parts, size, size, own, n_steps), "slabs": []}
    for k in depths:
        r0, r1 = r_own - 4 * k, r_own + own + 4 * k
        shape = (r1 - r0, size)
        D = torch.from_numpy(np.ascontiguousarray(sim.D[r0:r1])).to(dev)
        stims = [(s.protocol, torch.from_numpy(np.ascontiguousarray(s.field[r0:r1])).to(dev))
                 for s in sim.stimuli]
        variant = autotune.tune(shape, n_steps=300)["best"] if tune else autotune.resolved(shape)
        st0 = solve.init(shape)
        solve._forward_euler(st0, 0, 200, sim.params, D, stims, sim.dt, sim.dx)  # warm-up
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        solve._forward_euler(st0, 0, n_steps, sim.params, D, stims, sim.dt, sim.dx)
        e1.record(stream)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / n_steps
        res["slabs"].append({"steps_per_exchange": k, "rows": shape[0], "variant": variant,
                             "us_per_step": round(us, 3),
                             "owned_gcups": round(own * size / us / 1e3, 3),
                             "exchange_bytes_per_neighbour": 4 * k * size * 4})
    return res


def bench_sequence(seed, n_steps, size=1200, every=500, reshape=256):
    """deepx/generate.py:135-210 at the reference's production shape (1200^2
    grid, a checkpoint every 500 steps, frames resized to 256^2,
    experiments/generate_fd_data_256.py:7-12), over n_steps steps: the whole
    generator loop (forward, on-GPU resize, copy of every frame to pinned host
    memory; no file) against the same checkpointed forward alone."""
    import numpy as np
    import torch

    from cardiax_amd import generate, io, solve, workloads
    from cardiax_amd.stimulus import Protocol, Stimulus

    sim = workloads.c2_inputs(seed, shape=(size, size), steps=n_steps)
    dev = torch.device("cuda", 0)
    stims = [Stimulus(Protocol(*s.protocol), torch.from_numpy(s.field).to(dev))
             for s in sim.stimuli]
    D = torch.from_numpy(sim.D).to(dev)
    cps = list(range(0, n_steps, every))
    from cardiax_amd import autotune

    tuned = autotune.tune((size, size), n_steps=300)["best"]  # setup, as for the headline
    generate.sequence(0, 3 * every, every, sim.dt, sim.dx, sim.params, D, stims, None,
                      reshape=(reshape, reshape), plot_while=False)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    frames = generate.sequence(0, n_steps, every, sim.dt, sim.dx, sim.params, D, stims, None,
                               reshape=(reshape, reshape), plot_while=False)
    t_seq = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solve.forward(solve.init((size, size)), cps, sim.params, D, stims, sim.dt, sim.dx)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    # the resize kernel alone, one frame of 3 fields
    x = torch.stack(list(st[-1]))
    o = torch.empty((3, reshape, reshape), device=dev)
    io.imresize(x, (reshape, reshape), out=o)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(50):
        io.imresize(x, (reshape, reshape), out=o)
    e1.record()
    torch.cuda.synchronize()
    cells = size * size * (cps[-1] - cps[0])
    return {"workload": "%dx%d grid, %d Euler steps, a frame every %d steps resized to %dx%d "
                        "and copied to pinned host memory (%d frames)"
                        % (size, size, cps[-1], every, reshape, reshape, len(frames)),
            "seconds": round(t_seq, 4), "gcups": round(cells / t_seq / 1e9, 3),
            "forward_only_seconds": round(t_fwd, 4),
            "forward_only_gcups": round(cells / t_fwd / 1e9, 3),
            "overhead_frac": round(t_seq / t_fwd - 1.0, 4),
            "resize_us_per_frame": round(e0.elapsed_time(e1) * 1e3 / 50, 2), "variant": tuned,
            "frames_finite": bool(np.isfinite(frames).all())}


FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64, TFLOP/s


def bench_scar(seed, size=1200, reps=5):
    """deepx/generate.py:79-83 (random_diffusivity: the spline scar of
    deepx/utils_scars.py:246-426) at the reference's default 1200^2 image:
    the whole call (host contours + device rasterisation, blur and rescale)
    and each device kernel on its own, timed with HIP events on the current
    stream. The blur dominates: 2 kh kw FP64 flops per pixel (kh = kw =
    int(0.1 size)), against the FP64 peak. The CPU leg is the oracle
    pipeline (brute-force crossing numbers, scipy's FFT convolution, FP64
    rescale) on one core."""
    import numpy as np
    import torch

    from cardiax_amd import generate, scars

    shape = (size, size)
    generate.random_diffusivity(seed, shape)  # warm-up (library load, kernels)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        D = generate.random_diffusivity(seed + r, shape)
    torch.cuda.synchronize()
    t_call = (time.perf_counter() - t0) / reps
    g1, g2, ksize, sigma = scars._blur_draws(seed, shape)
    t0 = time.perf_counter()
    cs = scars.blob_contours(g2, shape, scars.DEF_PARAMS, scars.spline_centroids(g1))
    t_host = time.perf_counter() - t0
    k = scars.gauss_kernel(ksize, sigma)
    mask = scars.rasterize(cs, shape)
    blurred = scars.convolve_same(mask, k)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    stream = torch.cuda.current_stream()
    n = 20
    ev[0].record(stream)
    for _ in range(n):
        scars.convolve_same(mask, k)
    ev[1].record(stream)
    for _ in range(n):
        scars.scar_to_diffusivity(blurred)
    ev[2].record(stream)
    torch.cuda.synchronize()
    blur_s = ev[0].elapsed_time(ev[1]) / 1e3 / n
    rescale_s = ev[1].elapsed_time(ev[2]) / 1e3 / n
    t0 = time.perf_counter()
    for _ in range(n):
        scars.rasterize(cs, shape, out=mask)  # synchronises per call
    raster_s = (time.perf_counter() - t0) / n
    flops = 2.0 * ksize * ksize * size * size
    achieved = flops / blur_s / 1e12
    # CPU: the oracle pipeline on the same contours
    from oracle import scar_oracle as so

    t0 = time.perf_counter()
    m = so.union_mask(cs, shape)
    t_mask = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = so.convolve_same_fft(m, so.gauss_kernel(ksize, sigma)).astype(np.float32)
    so.scar_to_diffusivity(b)
    t_fft = time.perf_counter() - t0
    return {"workload": "random_diffusivity %dx%d: %d blobs, %dx%d Gaussian blur (sigma %.1f)"
                        % (size, size, len(cs), ksize, ksize, sigma),
            "ms_per_map": round(t_call * 1e3, 3), "host_contours_ms": round(t_host * 1e3, 3),
            "raster_ms_incl_sync": round(raster_s * 1e3, 3),
            "blur_ms": round(blur_s * 1e3, 3), "rescale_us": round(rescale_s * 1e6, 2),
            "roofline": {"bound": "fp64", "achieved": round(achieved, 2),
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                         "kernel": "k_convolve_same", "flops_per_launch": int(flops)},
            "D_range": [float(D.min()), float(D.max())],
            "cpu_oracle_ms": {"raster": round(t_mask * 1e3, 1),
                              "fft_blur_and_rescale": round(t_fft * 1e3, 1), "cores": 1}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sim-steps", type=int, default=100_000)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip timing the plain one-thread-per-cell fp32 kernel")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run torch's copy kernel over MIB MiB 20 times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    ap.add_argument("--batch-sims", type=int, default=256,
                    help="simulations of the batched workload (0 = skip it)")
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--inverse-iters", type=int, default=100,
                    help="Adam iterations of the batched inverse problem (0 = skip it)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (the size-based default kernel runs)")
    ap.add_argument("--variant", default=None,
                    help="run the single-tissue solve with this step kernel (no tuning), "
                         "e.g. stream2 or streamt8; the other legs keep their own choice")
    ap.add_argument("--sequence-steps", type=int, default=20_000,
                    help="Euler steps of the dataset-generator leg (0 skips it)")
    ap.add_argument("--scar-size", type=int, default=1200,
                    help="image side of the random D-map generator leg (0 skips it)")
    ap.add_argument("--slab-steps", type=int, default=4000,
                    help="Euler steps of the one-slab-of-eight leg (0 skips it)")
    args = ap.parse_args()
    # stdout carries the one JSON line; the library's progress prints go to stderr
    json_out = sys.stdout
    sys.stdout = sys.stderr

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from cardiax_amd import solve, workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        n = args.copy_calibration << 18
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n}),
              file=json_out, flush=True)
        return
    H = W = args.size
    sim = workloads.c2_inputs(args.seed, shape=(H, W), steps=args.sim_steps)
    stims = [(s.protocol, torch.from_numpy(s.field).to(dev)) for s in sim.stimuli]
    D = torch.from_numpy(sim.D).to(dev)
    st0 = solve.init((H, W))
    torch.cuda.synchronize()

    from cardiax_amd import autotune

    tuned = None
    forced = None
    if args.variant is not None:
        from cardiax_amd import _lib

        if args.variant not in _lib.VARIANTS:
            sys.exit("unknown variant %r (one of %s)" % (args.variant, sorted(_lib.VARIANTS)))
        forced = _lib.lib().fk_set_forward_variant(_lib.VARIANTS[args.variant])
    elif not args.no_tune:  # setup, outside the timed region
        # on the workload's own state after S2 (the spiral regime that fills
        # most of the run), with its parameters, D-map and stimuli
        t_mid = min(sim.t1 - 300, sim.meta["s2_start"] + (sim.t1 - sim.meta["s2_start"]) // 4)
        mid = solve._forward_euler(st0, sim.t0, t_mid, sim.params, D, stims, sim.dt, sim.dx)
        tuned = autotune.tune_on(mid, t_mid, sim.params, D, stims, sim.dt, sim.dx, n_steps=300)
        tuned["state_at_step"] = t_mid
        del mid
        log("autotune %dx%d: %s" % (H, W, json.dumps(tuned)))
    variant = autotune.resolved((H, W))

    def run():
        return solve._forward_euler(st0, sim.t0, sim.t1, sim.params, D, stims, sim.dt, sim.dx)

    for i in range(args.warmup):
        out = run()
        torch.cuda.synchronize()
        log("warmup %d/%d done" % (i + 1, args.warmup))

    # Timed region: K full solves; HIP events on the stream the kernels use.
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        out = run()
        log("step %d/%d enqueued" % (i + 1, args.steps))
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)
    finite = bool(torch.isfinite(out.u).all().item())
    umax = float(out.u.abs().max().item())
    cell_updates = H * W * sim.t1 * args.steps
    gcups = cell_updates / wall / 1e9

    # Dominant kernel: the tuned variant's, one launch per `depth` Euler steps.
    kernel, depth = KERNELS[variant]
    launches = -(-sim.t1 // depth) * args.steps
    per_launch_s = (ev_ms / 1e3) / launches
    achieved = BYTES_PER_CELL_UPDATE * H * W * depth / per_launch_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(kernel, H, W),
                "traffic_gbs": traffic_rate(measured_traffic(kernel, H, W), per_launch_s),
                "kernel": kernel,
                "per_launch_us": round(per_launch_s * 1e6, 3),
                "bytes_per_launch": BYTES_PER_CELL_UPDATE * H * W * depth,
                "achieved_basis": ACHIEVED_BASIS, "traffic_counts": TRAFFIC_NOTE}

    extra = {"event_gcups": round(cell_updates / (ev_ms / 1e3) / 1e9, 3),
             "variant": variant, "autotune": tuned,
             "u_finite": finite, "u_absmax": round(umax, 4)}
    if not args.no_plain:
        n = 2000
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        solve._forward_euler_plain(st0, 0, n, sim.params, D, stims, sim.dt, sim.dx)
        e1.record(stream)
        torch.cuda.synchronize()
        plain = H * W * n / (e0.elapsed_time(e1) / 1e3) / 1e9
        extra["plain_fp32_kernel_gcups"] = round(plain, 3)
        extra["speedup_vs_plain_fp32"] = round(extra["event_gcups"] / plain, 3)

    if forced is not None:  # the other legs run their own kernels again
        _lib.lib().fk_set_forward_variant(forced)

    if args.batch_sims > 0:
        log("batched workload ...")
        extra["batched"] = bench_batched(args.batch_sims, args.batch_steps, args.seed)
        log("batched: %s" % json.dumps(extra["batched"]))

    if not args.no_adjoint:
        log("adjoint workload ...")
        extra["adjoint"] = bench_adjoint(args.seed)
        log("adjoint: %s" % json.dumps(extra["adjoint"]))

    if args.inverse_iters > 0:
        log("inverse problem ...")
        extra["inverse"] = bench_inverse(args.seed, iters=args.inverse_iters)
        log("inverse: %s" % json.dumps(extra["inverse"]))

    log("heun ...")
    extra["heun"] = bench_heun(args.seed)
    log("heun: %s" % json.dumps(extra["heun"]))

    if args.slab_steps > 0:
        log("row slab ...")
        extra["slab"] = bench_slab(args.seed, args.slab_steps, size=H, tune=not args.no_tune)
        log("slab: %s" % json.dumps(extra["slab"]))

    if args.sequence_steps > 0:
        log("dataset generator ...")
        extra["sequence"] = bench_sequence(args.seed, args.sequence_steps)
        log("sequence: %s" % json.dumps(extra["sequence"]))

    if args.scar_size > 0:
        log("scar D-map generator ...")
        extra["scar"] = bench_scar(args.seed, args.scar_size)
        log("scar: %s" % json.dumps(extra["scar"]))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(sim)
        log("cpu baseline: %.4f GCUPS on %d cores" % (cpu["value"], cpu["cores"]))

    rec = {
        "metric": "GCUPS + % HBM roofline (fwd, batched, fwd+adjoint) vs plain fp32 at 1 GPU",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scar D-map and S1/S2 stimuli)",
        "config": {"workload": "single large tissue: %dx%d grid, %d Euler steps per bench "
                               "step, PARAMSET_3, scarred D, S1 WEST + S2 NORTH at step %d"
                               % (H, W, sim.t1, sim.meta["s2_start"]),
                   "scars": "1-5 seeded elliptical scars, axes 20-200 cells, D scaled by "
                            "U[0.1, 0.2] with tapers of at least 12 cells (the reference "
                            "scheme diverges on sharper, deeper scars, FP64 included)",
                   "H": H, "W": W, "euler_steps": sim.t1, "dt": sim.dt, "dx": sim.dx,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(rec), file=json_out, flush=True)


if __name__ == "__main__":
    main()
