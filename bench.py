"""Benchmark: synaptic events/s of the network step on one MI355X.

Workload of `value` (default): the single large block -- N = 5,000,000 LIF
neurons, in-degree 1,000 (5e9 synapses, int32 index + fp32 weight),
per-synapse delays U{1..10} steps, per-neuron Poisson noise 5-20 Hz,
dt = 0.1 ms, BOLD per 1,000-neuron voxel once per ms; one step = one pass of
the fused neuron kernel + the event-driven delivery kernels (+ the BOLD
kernel every 10th step).  Inputs are resident in HBM before timing starts.

python bench.py [--gpus 1] [--steps 10000] [--warmup 200]

Order of the main line: build; exactly --warmup steps; delivery autotuning
(its steps reported as `tuning_steps`); BOLD attached; the step-batch graphs
captured (`capture_steps`); then exactly --steps timed steps, all replayed
from captured graphs.  After the timed region, `--kernel-window` more steps
run as direct launches with HIP events around each kernel phase, on the
block's own stream; they give the per-kernel times of `roofline` and
`kernels` and are not part of `value`.

Beside the main line's fields: `plain` (the 10,000-neuron plain block,
in-degree 100, 8,000 steps after 100 warm-up steps and delivery tuning,
which also tries the persistent launch -- the workload the CPU comparison
is quoted on), `cpu_baseline` (the oracle's FP32 restatement of
the reference algorithm -- a full pass over every synapse per step, as
torch.sparse.mm in models/block.py:76-81 -- with OpenMP on the host cores,
on those same plain-block inputs), `da` (DA window seconds: the
64-member batched ensemble with the simulator; surrogate-assisted against
simulator-only windows on 256 whole-brain members; the surrogate's
matrix-core GEMM throughput)
and `regions` (the 90-region, 2e7-neuron share of a whole-brain model, timed
like the main line).  --da-windows 0, --region-steps 0, --plain-steps 0 and
--no-cpu skip them.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix-core peak
B_SEGMENT = 16  # rowptr begin (8) + delay-segment ends (4) + spike-list entry (4) per delivered (row, delay)
B_SYNAPSE = 8   # packed int32 target + fp32 weight per synaptic event
BOLD_SPS = 10   # the large block's BOLD integrator runs every 10 steps (1 ms)


def neuron_bytes(launch, per_neuron_noise=True, pair=False):
    """Bytes per neuron per step that neuron_step_kernel reads and writes.

    Dynamic state: V 4, Jp 16, acc 8 (16 pair), ref_end 4 read; V 4, Jp 16,
    flags 1 written (+ the acc clear with direct delivery).  Constants: g 16,
    pop 4, noise p 4, and either a 1-byte parameter-set id or 56 B of
    per-neuron V_u, decay, g_L, V_L, dt/C, V_th, V_reset, I_ext."""
    acc = 16 if pair else 8
    b = 4 + 16 + acc + 4 + 4 + 16 + 1 + 16 + 4 + (4 if per_neuron_noise else 0)
    b += 1 if launch.get("neuron_psets", 0) > 0 else 56
    if launch.get("deliver_mode", 0) == 0:
        b += acc
    return b


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def pmc_traffic(kernel_prefix, workload, events_per_launch):
    """HBM bytes per launch from a committed rocprofv3 PMC summary (a fixed
    part plus bytes per synaptic event, scaled to this run's events), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel_prefix}_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    fixed = re.search(r"traffic_fixed_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    per = re.search(r"traffic_bytes_per_event:\s*([0-9.eE+]+)", txt)
    if not (fixed and per):
        return None
    return float(fixed.group(1)) + float(per.group(1)) * events_per_launch


def run_counted(nb, steps, **kw):
    """Run `steps` steps through captured graphs and return (wall seconds,
    events, spikes); the GPU is idle before and after the timed span."""
    import torch

    ev = np.zeros(steps, np.uint64)
    sp = np.zeros(steps, np.uint32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nb.run(steps, graph=True, events=ev, spikes=sp, **kw)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, ev, sp


def capture(nb, steps, **kw):
    """Run once (untimed) every step-batch shape that a run of `steps` steps
    replays, so the timed run replays graphs that are already captured.
    Returns the number of steps this advanced the network."""
    kmax = 1000
    done = 0
    for k in sorted({min(steps, kmax), steps % kmax if steps > kmax else 0} - {0}):
        nb.run(k, graph=True, **kw)
        done += k
    return done


# ----------------------------------------------------------------- CPU
def cpu_plain(steps, warmup):
    """FP32 oracle, full-synapse pass per step (pull form, the torch.sparse.mm
    algorithm), OpenMP over the host cores, on the plain block's inputs: the
    same synapses (read back from the GPU generator), the same parameters and
    the same Philox noise stream."""
    from digital_twin_brain_amd import networks
    from oracle import oracle as orc

    nb, props = networks.build_plain()
    csr = dict(rowptr=nb.get("rowptr"), post=nb.get("post"), kind=nb.get("kind"), delay=nb.get("delay"),
               w0=nb.get("w0"), w1=None)
    nb.close()
    outdeg = np.diff(csr["rowptr"])
    ob = orc.OracleBlock(props, csr, 0.1, real="f32", pull=True)
    for _ in range(warmup):
        ob.run(noise_mode=1, seed=3)
    ev = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        ob.run(noise_mode=1, seed=3)
        ev += int(outdeg[ob.get("spike").astype(bool)].sum())
    dt = time.perf_counter() - t0
    cores = orc.omp_threads("f32")
    # cores: the OpenMP threads used; nproc: the CPUs this process may run on (what `nproc` prints)
    return dict(value=ev / dt, unit="synaptic events/s", cores=cores, nproc=len(os.sched_getaffinity(0)),
                kind="port", cpu=cpu_model(),
                sample=(f"the whole plain-block workload: oracle FP32 pull (a full pass over all "
                        f"{int(outdeg.sum())} synapses per step, the torch.sparse.mm algorithm of "
                        f"models/block.py:76-81) on the same 10,000-neuron network, parameters and noise stream as "
                        f"the GPU line, {steps} steps timed after {warmup} warm-up steps, {dt:.2f} s, {ev} events, "
                        f"{cores} OpenMP threads"),
                seconds=dt, steps=steps, events=ev)


def cpu_large_sample(n_sample, in_degree, steps_warm, steps_timed):
    """The same FP32 pull oracle on the large block's workload scaled down to
    n_sample neurons (same in-degree, delays and parameters)."""
    from digital_twin_brain_amd import networks
    from oracle import oracle as orc

    nb, props = networks.build_large(n=n_sample, in_degree=in_degree)
    csr = dict(rowptr=nb.get("rowptr"), post=nb.get("post"), kind=nb.get("kind"), delay=nb.get("delay"),
               w0=nb.get("w0"), w1=None)
    nb.close()
    outdeg = np.diff(csr["rowptr"])
    ob = orc.OracleBlock(props, csr, 0.1, real="f32", pull=True)
    for _ in range(steps_warm):
        ob.run(noise_mode=1, seed=3)
    ev = 0
    t0 = time.perf_counter()
    for _ in range(steps_timed):
        ob.run(noise_mode=1, seed=3)
        ev += int(outdeg[ob.get("spike").astype(bool)].sum())
    dt = time.perf_counter() - t0
    return dict(value=ev / dt, unit="synaptic events/s", cores=orc.omp_threads("f32"),
                sample=(f"large-block workload scaled to N={n_sample} (in-degree {in_degree}, delays 1-10, same "
                        f"parameters): {steps_timed} steps timed after {steps_warm} warm-up steps, {dt:.1f} s, "
                        f"{ev} events"))


# ------------------------------------------------------------ GPU lines
def plain_line(steps, warmup, log):
    """The plain block: N = 10,000 (80% excitatory), g_ui ~ U[0.001, 0.05],
    in-degree 100, weights U[0,1], delay 1, noise 10 Hz, dt = 0.1 ms."""
    from digital_twin_brain_amd import networks

    from digital_twin_brain_amd.autotune import tune_delivery

    nb, props = networks.build_plain()
    nb.run(warmup, noise_mode=1, graph=True)
    s0 = nb.steps
    best, table = tune_delivery(nb, dict(noise_mode=1), steps=20, use_cache=False)
    tuning = nb.steps - s0
    log(f"plain block: delivery {best}, persistent {table.get('persistent step us')} us/step against "
        f"{table.get('best shape step us')} us/step")
    cap = capture(nb, steps, noise_mode=1)
    wall, ev, sp = run_counted(nb, steps, noise_mode=1)
    out = {"metric": "synaptic events/s", "value": float(ev.sum()) / wall, "unit": "events/s", "steps": steps,
           "warmup": warmup, "tuning_steps": tuning, "capture_steps": cap, "ms_per_step": wall / steps * 1e3,
           "events_per_step": float(ev.mean()), "spikes_per_step": float(sp.mean()),
           "config": {"workload": "plain block", "neurons": int(nb.n), "in_degree": 100, "synapses": int(nb.nnz),
                      "g_ui": "U[0.001, 0.05]", "weights": "U[0, 1] compact fp32", "delay_steps": 1, "dt_ms": 0.1,
                      "noise_hz": 10},
           "launch": nb.get_launch()}
    log(f"plain block: {out['value']:.3e} events/s, {out['ms_per_step'] * 1e3:.2f} us/step, "
        f"{out['events_per_step']:.0f} events/step")
    # the same network behind the drop-in `block` interface (models/block.py), untuned:
    # run() hands back each step's statistics (a host round trip per step); run_steps()
    # replays a captured graph and returns them all at the end
    import torch

    from digital_twin_brain_amd.block import block

    rp, post, kind, w = nb.get("rowptr"), nb.get("post"), nb.get("kind"), nb.get("w0")
    pre = np.repeat(np.arange(nb.n), np.diff(rp))
    ch = kind.astype(np.int64) * 2
    idx = np.concatenate([np.stack([ch, post, pre]), np.stack([ch + 1, post, pre])], axis=1)
    W = torch.sparse_coo_tensor(torch.from_numpy(idx), torch.from_numpy(np.concatenate([w, w])), (4, nb.n, nb.n))
    nb.close()
    B = block(torch.from_numpy(props), W, delta_t=0.1, seed=3)
    for _ in range(20):
        B.run(noise_rate=1e-3)
    t0 = time.perf_counter()
    for _ in range(300):
        B.run(noise_rate=1e-3)
    per_run = (time.perf_counter() - t0) / 300 * 1e6
    B.run_steps(steps, noise_rate=1e-3)  # capture
    t0 = time.perf_counter()
    B.run_steps(steps, noise_rate=1e-3)
    per_batched = (time.perf_counter() - t0) / steps * 1e6
    out["dropin_block"] = {"run_us_per_step": per_run, "run_steps_us_per_step": per_batched,
                           "note": "models/block.py interface, default launch, noise 10 Hz as a scalar rate"}
    log(f"plain block behind the drop-in interface: run() {per_run:.1f} us/step, run_steps() "
        f"{per_batched:.1f} us/step")
    B._nb.close()
    return out


def build_region_share():
    """90 regions, one GPU's share of a whole-brain model: 2e7 neurons, 1e10 synapses."""
    from digital_twin_brain_amd import networks

    nb, props, spec, _ = networks.build_regions()
    config = {"workload": "90 regions, one GPU's share of a whole-brain model", "neurons": int(nb.n),
              "synapses": int(nb.nnz), "in_degree": "400 intra + 100 inter (Exp(1) connectome)",
              "region_sizes": "lognormal sigma 0.8", "delays": "1-40 steps", "dt_ms": 0.1,
              "external": "Poisson, 20% of input", "bold": "per region, 1 ms", "weights": "compact fp32"}
    return nb, spec, config


def attach_region_bold(nb, spec):
    vop = (np.arange(nb.S) // 2).astype(np.int32)  # one BOLD voxel per region
    nb.attach_bold(steps_per_sample=BOLD_SPS, delta_t=1e-3, voxel_of_pop=vop,
                   voxel_size=np.array([s[0] for s in spec], np.float64), nvox=len(spec))


def region_share_line(steps, warmup, log):
    """The region workload timed like the main one (graph-replayed batches,
    autotuned delivery), reported beside it."""
    import torch
    from digital_twin_brain_amd.autotune import tune_delivery

    t0 = time.perf_counter()
    nb, spec, config = build_region_share()
    nb.run(warmup, noise_mode=1, graph=True)
    s0 = nb.steps
    tune_delivery(nb, dict(noise_mode=1), steps=20, use_cache=False)
    nb.run((-nb.steps) % BOLD_SPS, noise_mode=1, graph=False)
    tuning = nb.steps - s0
    attach_region_bold(nb, spec)
    cap = capture(nb, steps, noise_mode=1)
    wall, ev, _ = run_counted(nb, steps, noise_mode=1)
    out = {"metric": "synaptic events/s", "value": float(ev.sum()) / wall, "unit": "events/s", "steps": steps,
           "warmup": warmup, "tuning_steps": tuning, "capture_steps": cap, "ms_per_step": wall / steps * 1e3,
           "events_per_step": float(ev.mean()), "config": config, "launch": nb.get_launch()}
    log(f"region share: {out['value']:.3e} events/s, {out['ms_per_step']:.3f} ms/step "
        f"({time.perf_counter() - t0:.1f} s with build and tuning)")
    nb.close()
    del nb
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", choices=("large", "regions", "plain"), default="large",
                    help="large: one 5e6-neuron block; regions: 90 regions, 2e7 neurons, connectome, delays 1-40; "
                         "plain: only the 10,000-neuron plain block line (for profiling)")
    ap.add_argument("--n", type=int, default=5_000_000)
    ap.add_argument("--in-degree", type=int, default=1000)
    ap.add_argument("--kernel-window", type=int, default=200,
                    help="direct-launch steps with per-kernel HIP events, run after the timed region")
    ap.add_argument("--plain-steps", type=int, default=8000, help="timed steps of the plain block (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=8000, help="timed CPU steps of the plain block")
    ap.add_argument("--cpu-sample", type=int, default=100_000, help="neurons of the large-block CPU sample (0: skip)")
    ap.add_argument("--cpu-sample-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--region-steps", type=int, default=10000,
                    help="steps of the region workload reported beside the main one (0: skip)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--da-windows", type=int, default=20)
    ap.add_argument("--train-windows", type=int, default=196,
                    help="surrogate training windows (x256 members: 196 gives 50,176 training windows)")
    ap.add_argument("--member-neurons", type=int, default=10_000, help="neurons per whole-brain DA member")
    ap.add_argument("--no-graph", action="store_true", help="launch the timed steps directly (for counter runs)")
    ap.add_argument("--launch", default="", metavar="MODE,GRID,SEGS[,EPT[,VARIANT]]",
                    help="fixed delivery launch shape instead of the autotuner (e.g. 1,1024,128,32,0)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the single-GPU build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if a.steps < BOLD_SPS or a.steps % BOLD_SPS:
        print(f"--steps must be a positive multiple of {BOLD_SPS}: the BOLD integrator of the timed workload runs "
              f"every {BOLD_SPS} steps (1 ms)", file=sys.stderr)
        sys.exit(2)
    if a.warmup < 0 or a.kernel_window < 0:
        print("--warmup and --kernel-window must be >= 0", file=sys.stderr)
        sys.exit(2)

    import torch

    from digital_twin_brain_amd import networks

    torch.cuda.init()
    if a.copy_calibration:
        n = a.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                          "launches": 5}), flush=True)
        return
    if a.workload == "plain":
        out = plain_line(a.plain_steps, 100, log)
        out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
                    "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return
    graph = not a.no_graph
    t0 = time.perf_counter()
    spec = None
    if a.workload == "regions":
        nb, spec, config = build_region_share()
    else:
        nb, props = networks.build_large(n=a.n, in_degree=a.in_degree)
        config = {"workload": "single large block", "neurons": a.n, "in_degree": a.in_degree,
                  "synapses": a.n * a.in_degree, "delays": "1-10 steps", "dt_ms": 0.1,
                  "noise_hz": "5-20 per neuron", "bold": "per 1000-neuron voxel, 1 ms", "weights": "compact fp32"}
    log(f"built {a.workload} network n={nb.n} nnz={nb.nnz} S={nb.S} D={nb.max_delay} in "
        f"{time.perf_counter() - t0:.1f} s")
    # exactly --warmup steps (before BOLD is attached, so any count works)
    if a.warmup:
        nb.run(a.warmup, noise_mode=1, graph=graph)
    log(f"warm-up {a.warmup} steps done")
    s0 = nb.steps
    if a.launch:
        v = [int(x) for x in a.launch.split(",")]
        mode, grid, segs = v[:3]
        ept = v[3] if len(v) > 3 else 32
        var = v[4] if len(v) > 4 else 0
        nb.set_launch(deliver_mode=mode, deliver_grid=grid, stage_segs=segs, stage_ept=ept, stage_variant=var)
        log(f"fixed delivery launch: mode={mode} grid={grid} segs={segs} ept={ept} variant={var}")
    elif not a.no_tune:
        from digital_twin_brain_amd.autotune import tune_delivery

        best, table = tune_delivery(nb, dict(noise_mode=1), steps=20, use_cache=False)
        log(f"autotuned delivery launch: {best} from {json.dumps(table)}")
    # the BOLD integrator starts at a whole millisecond
    nb.run((-nb.steps) % BOLD_SPS, noise_mode=1, graph=False)
    tuning_steps = nb.steps - s0
    if a.workload == "regions":
        attach_region_bold(nb, spec)
    else:
        nb.attach_bold(steps_per_sample=BOLD_SPS, delta_t=1e-3)
    S = nb.S
    K = a.steps
    cap = capture(nb, K, noise_mode=1) if graph else 0
    log(f"tuning {tuning_steps} steps, graph capture {cap} steps")
    events = np.zeros(K, np.uint64)
    nspk = np.zeros(K, np.uint32)
    bold = np.zeros((K // BOLD_SPS, nb.nvox), np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nb.run(K, noise_mode=1, graph=graph, events=events, spikes=nspk, bold=bold)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tot_ev = int(events.sum())
    log(f"timed {K} steps in {wall:.3f} s: {tot_ev} events, {int(nspk.sum())} spikes, "
        f"mean rate {nspk.sum() / K / nb.n / 1e-4:.2f} Hz delivered")
    # per-kernel times: an untimed window of direct launches after the timed region
    M = max(BOLD_SPS, a.kernel_window - a.kernel_window % BOLD_SPS)
    ev_w = np.zeros(M, np.uint64)
    sp_w = np.zeros(M, np.uint32)
    kms = np.zeros((M, 2), np.float32)
    nb.run(M, noise_mode=1, graph=False, events=ev_w, spikes=sp_w, kernel_ms=kms)
    t_neu = float(kms[:, 0].mean()) * 1e-3
    t_del = float(kms[:, 1].mean()) * 1e-3
    launch = nb.get_launch()
    bpn = neuron_bytes(launch)
    b_neu = bpn * nb.n
    # delivered (row, delay) segments: every spike of the last D steps contributes one
    sp_all = np.concatenate([nspk, sp_w]).astype(np.float64)
    seg = np.convolve(sp_all, np.ones(nb.max_delay))[:sp_all.shape[0]][K:]
    b_del = float((B_SEGMENT * seg + B_SYNAPSE * ev_w.astype(np.float64)).mean())
    del_name = ("delivery (deliver_stage_kernel + deliver_accum_kernel)" if launch.get("deliver_mode", 0) == 1
                else "delivery (deliver_kernel)")
    if t_neu >= t_del:
        dom, bytes_l, tl, prof = "neuron_step_kernel", b_neu, t_neu, "neuron_step_kernel"
    else:
        dom, bytes_l, tl, prof = del_name, b_del, t_del, "delivery"
    achieved = bytes_l / tl / 1e9
    kernels = {"window_steps": M,
               "neuron_step_kernel": {"avg_us": t_neu * 1e6, "bytes_per_launch": b_neu, "bytes_per_neuron": bpn,
                                      "GBps": b_neu / t_neu / 1e9},
               del_name: {"avg_us": t_del * 1e6, "bytes_per_launch": b_del, "GBps": b_del / t_del / 1e9,
                          "events_per_launch": float(ev_w.mean()),
                          "atomic_rmw_bytes_per_launch": 8.0 * float(ev_w.mean())}}
    log(f"kernels: {json.dumps(kernels)}")
    nb.close()
    del nb
    torch.cuda.empty_cache()

    plain = None
    if a.plain_steps > 0 and a.workload == "large":
        plain = plain_line(a.plain_steps, 100, log)

    da = None
    if a.da_windows > 0 and a.workload == "large":
        from digital_twin_brain_amd.da_runner import run_ensemble_da
        from digital_twin_brain_amd.region_da import run_region_surrogate_da

        # the batched ensemble: 64 uneven members, simulator-only windows
        t0 = time.perf_counter()
        c4, da4, _ = run_ensemble_da(E=64, voxels=90, windows=a.da_windows, window_ms=800.0, dt=0.1, log=log)
        da4.block_model.nb.close()
        del da4
        torch.cuda.empty_cache()
        c4["total_s"] = time.perf_counter() - t0
        # surrogate-assisted against simulator-only windows on 256 whole-brain members
        t0 = time.perf_counter()
        c5, rec = run_region_surrogate_da(E=256, R=90, member_n=a.member_neurons, train_windows=a.train_windows,
                                          windows=a.da_windows, window_ms=800.0, dt=0.1, subset=16, every=5,
                                          log=log)
        rec["da"].block_model.nb.close()
        rec["judge"].nb.close()
        da = {"metric": "DA window seconds (surrogate vs full sim)", "unit": "s", "higher_is_better": False,
              "full_sim_window_s": c5["sim_window_s_mean"], "surrogate_window_s": c5["sur_window_s_mean"],
              "gain": c5["gain"], "nrmse_median": c5["nrmse_median"], "nrmse_max": c5["nrmse_max"],
              "config": (f"256 whole-brain members (90 regions, {a.member_neurons} neurons, in-degree 100, delays "
                         f"1-40) + truth, 4 g_ui per region, 800 ms windows, EnKF FP64, surrogate trained on "
                         f"{c5['train_samples']} simulated windows, judge: 16 members every 5th window"),
              "detail": c5, "total_s": time.perf_counter() - t0,
              "batched_ensemble": {"config": "64 uneven members (2k-200k neurons, in-degree 100-500) + truth, 90 "
                                             "voxels, 800 ms windows, EnKF FP64, simulator only",
                                   "window_s": c4["window_s_mean"], "detail": c4}}
        del rec
        # the surrogate's matrix-core product, live: one layer of the MLP and a large square product
        from digital_twin_brain_amd.surrogate import gemm_kernel_name, gemm_throughput

        g_small = gemm_throughput(256, 512, 512, reps=50)
        g_big = gemm_throughput(8192, 8192, 8192, reps=10)
        da["surrogate_gemm"] = {"bound": "mfma", "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                                "layer_256x512x512_kernel": gemm_kernel_name(256, 512, 512),
                                "layer_256x512x512_us": g_small["us"],
                                "square_8192_kernel": gemm_kernel_name(8192, 8192, 8192),
                                "square_8192_tflops": g_big["tflops"],
                                "square_8192_frac": g_big["tflops"] / MFMA_BF16_PEAK_TFLOPS}
        log(f"surrogate GEMM: 256x512x512 {g_small['us']:.1f} us, 8192^3 {g_big['tflops']:.0f} TFLOP/s")
        torch.cuda.empty_cache()

    regions = None
    if a.region_steps > 0 and a.workload == "large":
        regions = region_share_line(a.region_steps - a.region_steps % BOLD_SPS, 200, log)

    cpu = None
    if not a.no_cpu and a.workload == "large":
        t0 = time.perf_counter()
        cpu = cpu_plain(a.cpu_steps, 100)
        log(f"cpu baseline (plain block) {cpu['value']:.3e} events/s ({time.perf_counter() - t0:.1f} s)")
        if plain is not None:
            plain["vs_cpu"] = plain["value"] / cpu["value"]
        if a.cpu_sample > 0:
            t0 = time.perf_counter()
            cpu["large_block_sample"] = cpu_large_sample(a.cpu_sample, a.in_degree, 200, a.cpu_sample_steps)
            log(f"cpu baseline (large-block sample) {cpu['large_block_sample']['value']:.3e} events/s "
                f"({time.perf_counter() - t0:.1f} s)")

    value = tot_ev / wall
    traffic = pmc_traffic(prof, a.workload, float(ev_w.mean())) if prof == "delivery" else None
    out = {
        "metric": "synaptic events/s", "value": value, "unit": "events/s", "n_gpus": 1, "steps": K,
        "warmup": a.warmup, "tuning_steps": tuning_steps, "capture_steps": cap, "ms_per_step": wall / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "kernels": kernels,
        "launch": launch,
        "plain": plain,
        "da": da,
        "regions": regions,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
