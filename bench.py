#!/usr/bin/env python3
"""Primary photons/s of the MI355X engine on the voxel-CT workload.

Workload (one GPU): the 512x512x256 body phantom of 0.8 mm voxels (air, lung,
adipose, soft tissue, cortical bone; 40 seeded ellipsoids), a point cone-beam
source at SID 600 mm with a 120 kVp Kramers spectrum (2.5 mm Al), a flat
256x256 detector of 1 mm GOS pixels at SDD 1000 mm, Compton + photoelectric +
Rayleigh, FP64 dose + uncertainty tallies.  Set up through the GGEMS API
(ggems_amd.configs.voxel_ct), then timed on the engine with every input
resident in HBM: one step is the workload's run of 1e9 primaries
(--photons-per-step), one launch laid out in batches of 1,048,576 particle
slots as GGEMS lays them out.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).

--workload mesh_ct measures the meshed-phantom scene (two icosphere meshes
with a curved 888x64 detector, 80 keV) and --workload acquisition the CT
acquisition in one launch (360 projections around each of the 128^3, 256^3
and 384^3 ellipsoid phantoms, budgets scaled by --budget-scale); both print
their own JSON line.
"""

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "primary photons/s on 1 MI355X (voxel CT); scatter error vs 1e9-photon MC"
HBM_PEAK_GBS = 8000.0
# VALU issue peak of one MI355X: the published FP32 vector rate, 157.3 TFLOP/s
# = 256 CUs x 4 SIMDs x 2.4 GHz x 64 FLOP per cycle, is one wave64 FMA
# (128 FLOP) per SIMD every 2 cycles
VALU_PEAK_INSTR_PER_S = 256 * 4 * 2.4e9 / 2


def algorithmic_bytes_per_primary(stats):
    """B = S*1 (uint8 label gather) + I*40 (FP64 edep + edep^2 and int32 hit
    read-modify-writes) + D*8 (int32 histogram RMW) + 48 (source / final state)."""
    p = max(stats["primaries"], 1)
    s = stats["voxel_steps"] / p
    i = stats["interactions"] / p
    d = stats["detector_interactions"] / p
    return s * 1.0 + i * 40.0 + d * 8.0 + 48.0, s, i, d


def cpu_model():
    """the host CPU's model name (/proc/cpuinfo), or None"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(workdir, n_per_step, seed, target_s, threads, min_primaries=10_000_000):
    """The CPU oracle on a bounded sample of particle slots of the same
    workload: consecutive chunks of slots until at least target_s seconds
    and min_primaries primaries."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    from scene import voxel_ct_scene

    sc = voxel_ct_scene(workdir)
    prepared = sc.prepare_oracle()  # tallies allocated and faulted in outside the timed region
    # consecutive chunks of particle slots until about target_s seconds of CPU work
    begin, chunk, done, elapsed = 0, 2048, 0, 0.0
    while (elapsed < target_s or done < min_primaries) and begin + chunk <= 1048576:
        t0 = time.perf_counter()
        r = sc.run_oracle(n_per_step, seed, threads=threads, slots=(begin, begin + chunk), prepared=prepared)
        elapsed += time.perf_counter() - t0
        done += r["stats"]["primaries"]
        begin += chunk
        rate = done / elapsed
        per_slot = r["stats"]["primaries"] / chunk
        want = max((target_s - elapsed) * rate, min_primaries - done) / per_slot
        chunk = int(min(max(2048, want / 2 if want > 4096 else want + 1), 1048576 - begin))
        if chunk <= 0:
            break
    return {
        "value": done / elapsed,
        "unit": "photons/s",
        "cores": threads,
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": "CPU oracle (oracle/, C restatement of the reference kernels), particle slots 0..%d of a "
                  "%d-primary voxel-CT step: %d primaries in %.1f s" % (begin, n_per_step, done, elapsed),
    }


def profile_record(sampling):
    """The committed PMC summary of the transport kernel, if it was measured on this engine source and
    sampling mode."""
    path = os.path.join(REPO, "profiles", "transport_kernel_voxel_ct.json")
    try:
        with open(path) as f:
            rec = json.load(f)
        from ggems_amd.profiling import source_digest
        ok = rec.get("source_sha256") == source_digest(REPO) and rec.get("sampling", "reference") == sampling
        return rec if ok else None
    except (OSError, ValueError, KeyError):
        return None


def bench_mesh_ct(args):
    """photons/s through the two icosphere meshes (no enclosing box: GGEMS
    would let the box claim every photon) and the curved detector"""
    import torch

    from ggems_amd import configs
    from ggems_amd.ggems import GGEMS
    wd = os.path.join(REPO, "ggems_amd", "data", "_gen", "mesh_ct")
    keep = configs.mesh_ct(wd, n_particles=args.photons_per_step, water_box=False, dosimetry=True)
    g = GGEMS()
    g.initialize(args.seed)
    eng = g.engine()
    for _ in range(args.warmup):
        eng.run(args.photons_per_step)
    eng.reset_tallies()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.run(args.photons_per_step)
        kernel_ms.append(eng.last_kernel_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = eng.stats()
    total = args.steps * args.photons_per_step
    out = {"metric": "primary photons/s on 1 MI355X (mesh CT)", "value": total / wall, "unit": "photons/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "two icosphere meshes (r 50 mm, 81,920 triangles, soft tissue; r 15 mm, bone), "
                                  "80 keV cone beam at SID 600 mm, curved 888x64 GOS detector at SDD 1040 mm, "
                                  "dose map on the outer mesh",
                      "photons_per_step": args.photons_per_step,
                      "mesh_steps_per_primary": st["mesh_steps"] / total,
                      "mesh_interactions_per_primary": st["mesh_interactions"] / total},
           "kernel_ms": sum(kernel_ms) / len(kernel_ms), "kernel_variant": "general"}
    g.delete()
    del keep
    print(json.dumps(out), flush=True)


def bench_acquisition(args):
    """The batched CT acquisition: 360 one-degree projections around each of
    the 128^3, 256^3 and 384^3 ellipsoid phantoms, 1,080 projections of
    uneven budgets in ONE captured launch (each projection sees its own
    phantom), launch shape chosen by the autotuner; value = primaries / the
    launch's transport time"""
    import numpy as np

    from ggems_amd import configs
    from ggems_amd.ggems import GGEMS
    lo, hi = int(1_000_000 * args.budget_scale), int(30_000_000 * args.budget_scale)
    sides = (128, 256, 384)
    wd = os.path.join(REPO, "ggems_amd", "data", "_gen", "acquisition")
    keep = configs.acquisition(wd, sides=sides, seeds=(1, 2, 3))
    rot, budgets, only = [], [], []
    for k in range(len(sides)):
        r, b = configs.acquisition_plan(n_angles=args.angles, budget_range=(lo, hi), seed=5 + k)
        rot.append(r)
        budgets.append(b)
        only += [k] * args.angles
    rot, budgets = np.concatenate(rot), np.concatenate(budgets)
    g = GGEMS()
    g.initialize(args.seed)
    eng = g.engine()
    eng.set_sampling(args.sampling)
    g.autotune(4_000_000)
    g.run_acquisition(rot, budgets, "deg", only)
    grid = list(eng.last_grid())
    st = g.engine_stats()
    ms = g.run_time_ms()
    n = int(budgets.sum())
    assert st["primaries"] == n, st
    per = [{"side": s, "primaries": int(budgets[k * args.angles:(k + 1) * args.angles].sum())}
           for k, s in enumerate(sides)]
    g.delete()
    del keep
    out = {"metric": "primary photons/s on 1 MI355X (batched CT acquisition in one launch)", "value": n / ms * 1e3,
           "unit": "photons/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "ms_per_step": ms,
           "config": {"workload": "%d projections (1 degree) around each of three seeded ellipsoid phantoms of "
                                  "128^3, 256^3, 384^3 1 mm voxels in one captured launch, budgets uniform in "
                                  "[%d, %d], flat 512x384 GOS panel of 0.78 mm at SDD 1000 mm, 120 kVp"
                                  % (args.angles, lo, hi),
                      "sampling": args.sampling, "projections": int(rot.shape[0]), "primaries": n,
                      "voxel_steps_per_primary": st["voxel_steps"] / n, "grid": grid, "phantoms": per}}
    print(json.dumps(out), flush=True)


BF16_PEAK_TFLOPS = 2500.0  # dense MFMA, MI355X


def bench_surrogate(args):
    """scatter surrogate: training views from the engine, U-Net training on
    the MFMA convolutions, held-out error against the low-count Monte Carlo"""
    import numpy as np
    import torch

    from ggems_amd import surrogate as S
    wd = os.path.join(REPO, "ggems_amd", "data", "_gen", "surrogate")

    def log(msg):
        print("[surrogate] " + msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    rate = S.open_field_rate(wd, args.n_high, seed=3)
    tr = S.dataset_chunks(wd, range(1, 1 + args.train_phantoms), args.train_angles, 100, args.n_low, args.n_high,
                          rate, log=log)
    te = S.dataset_chunks(wd, range(100001, 100001 + args.test_phantoms), args.test_angles, 100000, args.n_low,
                          args.n_eval, rate, log=log)
    n_train, n_test = tr["x"].shape[0], te["x"].shape[0]
    t_data = time.perf_counter() - t0
    net, t_train, losses = S.train(tr["x"], tr["y"], steps=args.train_steps, batch=args.train_batch, seed=1,
                                   log=log)
    S.predict(net, te["x"][:8])  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pred = S.predict(net, te["x"])
    t_infer = (time.perf_counter() - t0) / n_test
    err = S.scatter_errors(pred, te["y"], te["var"], te["low"])
    base = S.scatter_errors(S.drr_baseline(tr["x"], tr["y"])(te["x"]), te["y"], te["var"], te["low"])
    # wall time per view: the surrogate's inputs (low-count MC, ray trace) and
    # its inference, against plain MC with enough primaries to match its error
    t_sur = float(np.mean(te["t_low"] + te["t_drr"])) + t_infer
    mc_rate = args.n_eval / float(np.mean(te["t_high"]))
    t_mc = err["mc_primaries_factor"] * args.n_low / mc_rate
    flops_step = 3 * args.train_batch * S.conv_flops(net, 256, 256)
    # the MFMA convolution on its own: the widest decoder layer, batch 8
    x = torch.randn(8, 64, 64, 192, device="cuda").bfloat16()
    w = torch.randn(64, 9 * 192, device="cuda").bfloat16()
    for _ in range(3):
        S.conv3x3(x, w)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        S.conv3x3(x, w)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    conv_tflops = 2 * 8 * 64 * 64 * 64 * 9 * 192 / (ms * 1e-3) / 1e12
    out = {"metric": "scatter-surrogate RMS error on held-out views (relative to the mean scatter)",
           "value": err["pred"], "unit": "relative RMS", "n_gpus": 1, "higher_is_better": False,
           "dtype": "bf16", "data": "synthetic",
           "config": {"workload": "U-Net (4 levels, 16-32-64-128, 3x3 conv, BF16 MFMA, FP32 accumulate) on 256x256 "
                                  "views of seeded 96x96x64 body phantoms (2 mm), 120 kVp, generated by batched "
                                  "acquisitions (one launch per group of phantoms x gantry angles); inputs: "
                                  "%d-primary MC projection (counts and scatter counts) and the ray-traced "
                                  "transmission; training targets: "
                                  "%d-primary MC scatter; held-out targets: %d-primary MC scatter"
                                  % (args.n_low, args.n_high, args.n_eval),
                      "train_views": n_train, "test_views": n_test, "train_steps": args.train_steps,
                      "train_batch": args.train_batch},
           "low_count_mc_relative_rms": err["low"],
           "target_noise_relative_rms": err["target_noise"],
           "surrogate_vs_truth_relative_rms": err["pred_vs_truth"],
           "low_count_mc_vs_truth_relative_rms": err["low_vs_truth"],
           "drr_scatter_model_vs_truth_relative_rms": base["pred_vs_truth"],
           "mc_primaries_for_equal_error": err["mc_primaries_factor"] * args.n_low,
           "seconds_per_view": {"surrogate": t_sur, "low_count_mc": float(np.mean(te["t_low"])),
                                "ray_trace": float(np.mean(te["t_drr"])), "inference": t_infer,
                                "plain_mc_at_equal_error": t_mc},
           "wall_time_ratio_at_equal_error": t_mc / t_sur,
           "train_seconds": t_train, "data_seconds": t_data,
           "train_conv_tflops": flops_step * args.train_steps / t_train / 1e12,
           "losses": losses[::max(1, len(losses) // 8)],
           "roofline": {"bound": "mfma", "achieved": conv_tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": conv_tflops / BF16_PEAK_TFLOPS, "kernel": "conv_fwd_kernel",
                        "shape": "8x64x64, 192 -> 64 channels", "kernel_ms": ms}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons-per-step", type=int, default=None,
                    help="primaries per step (default: the workload's run, 1e9 for voxel_ct, 1e8 for mesh_ct)")
    ap.add_argument("--seed", type=int, default=777)
    ap.add_argument("--workdir", default=os.path.join(REPO, "ggems_amd", "data", "_gen", "voxel_ct"))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    # the host's share of CPUs for one GPU where OMP_NUM_THREADS states it; otherwise every core
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--round-steps", type=int, default=0,
                    help="voxel steps a wave makes at least between regroupings (0: engine default)")
    ap.add_argument("--interact-batch", type=int, default=0,
                    help="resolve interactions inside the voxel loop once this many lanes wait (0: engine default)")
    ap.add_argument("--cold-passes", type=int, default=0,
                    help="navigation passes per round for stopped photons (0: engine default)")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--general-kernel", action="store_true", help="run the general kernel variant")
    ap.add_argument("--sampling", choices=("reference", "tau"), default="tau",
                    help="voxel-step sampling: one exponential number of mean free paths per flight (tau, "
                         "statistically equivalent to GGEMS's, the default) or GGEMS's three exponentials per "
                         "step (reference, GGEMS's histories)")
    ap.add_argument("--no-compare-sampling", dest="compare_sampling", action="store_false",
                    help="skip timing the other sampling mode after the headline")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a plain device copy of this size (calibrates profiler byte counters)")
    ap.add_argument("--workload", choices=("voxel_ct", "mesh_ct", "acquisition", "surrogate"), default="voxel_ct")
    ap.add_argument("--n-low", type=int, default=1_000_000, help="surrogate: primaries of the input projection")
    ap.add_argument("--n-high", type=int, default=100_000_000, help="surrogate: primaries of the target")
    ap.add_argument("--n-eval", type=int, default=1_000_000_000, help="surrogate: primaries of held-out targets")
    ap.add_argument("--train-phantoms", type=int, default=8, help="surrogate: seeded training phantoms")
    ap.add_argument("--train-angles", type=int, default=24, help="surrogate: gantry angles per training phantom")
    ap.add_argument("--test-phantoms", type=int, default=4, help="surrogate: held-out phantoms")
    ap.add_argument("--test-angles", type=int, default=3, help="surrogate: gantry angles per held-out phantom")
    ap.add_argument("--train-steps", type=int, default=8000)
    ap.add_argument("--train-batch", type=int, default=16, help="surrogate: views per Adam step")
    ap.add_argument("--angles", type=int, default=360, help="acquisition: projections per phantom")
    ap.add_argument("--budget-scale", type=float, default=1.0,
                    help="acquisition: budgets uniform in [1e6, 3e7] times this")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU (--gpus 1); the photon budget of a run shards across GPUs by "
              "particle slots, with one sum-reduction of the tallies at the end", file=sys.stderr)
        sys.exit(2)

    if args.photons_per_step is None:
        args.photons_per_step = 100_000_000 if args.workload == "mesh_ct" else 1_000_000_000
    if args.workload == "mesh_ct":
        return bench_mesh_ct(args)
    if args.workload == "acquisition":
        return bench_acquisition(args)
    if args.workload == "surrogate":
        return bench_surrogate(args)

    import torch

    from ggems_amd import configs
    from ggems_amd.ggems import GGEMS

    if args.copy_calibration_mb:
        n = args.copy_calibration_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    keep = configs.voxel_ct(args.workdir, n_particles=args.photons_per_step, seed=20240, dosimetry=True)
    g = GGEMS()
    g.initialize(args.seed)
    eng = g.engine()
    eng.set_launch(args.round_steps, args.interact_batch, args.blocks_per_cu, args.general_kernel,
                   args.cold_passes)
    eng.set_sampling(args.sampling)
    for _ in range(args.warmup):
        eng.run(args.photons_per_step)
    eng.reset_tallies()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.run(args.photons_per_step)
        kernel_ms.append(eng.last_kernel_ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stats = eng.stats()
    total = args.steps * args.photons_per_step
    assert stats["primaries"] == total, stats
    value = total / wall
    other = None
    if args.compare_sampling:
        # the other sampling mode on the same scene, after the timed region
        mode = "reference" if args.sampling == "tau" else "tau"
        eng.set_sampling(mode)
        eng.run(args.photons_per_step)
        eng.reset_tallies()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        other_ms = []
        for _ in range(args.steps):
            eng.run(args.photons_per_step)
            other_ms.append(eng.last_kernel_ms())
        torch.cuda.synchronize()
        other_wall = time.perf_counter() - t1
        ost = eng.stats()
        other = {"sampling": mode, "value": total / other_wall, "ms_per_step": other_wall / args.steps * 1e3,
                 "kernel_ms": sum(other_ms) / len(other_ms), "voxel_steps_per_primary": ost["voxel_steps"] / total,
                 "interactions_per_primary": ost["interactions"] / total}
    bpp, s, i, d = algorithmic_bytes_per_primary(stats)
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) * 1e-3
    achieved = bpp * args.photons_per_step / avg_kernel_s / 1e9
    prof = profile_record(args.sampling)
    traffic = prof["traffic_bytes_per_primary"] * args.photons_per_step if prof and prof.get(
        "traffic_bytes_per_primary") else None
    steps_per_launch = stats["voxel_steps"] / args.steps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "voxel CT: 512x512x256 body phantom (0.8 mm, 5 materials, 40 seeded ellipsoids), "
                        "120 kVp Kramers + 2.5 mm Al point cone beam at SID 600 mm, flat 256x256 GOS detector "
                        "(1 mm) at SDD 1000 mm, Compton+PE+Rayleigh, FP64 dose+uncertainty",
            "photons_per_step": args.photons_per_step,
            "sampling": args.sampling,
            "voxel_steps_per_primary": s,
            "interactions_per_primary": i,
            "detector_interactions_per_primary": d,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "transport_kernel",
            "kernel_variant": "general" if eng.last_grid()[2] else "specialised",
            "grid": list(eng.last_grid()[:2]),
            "kernel_ms": avg_kernel_s * 1e3,
            "bytes_per_primary": bpp,
        },
    }
    if other:
        out["other_sampling"] = other
    if prof and prof.get("valu_instr_per_step"):
        # issue-rate roofline of the same kernel (VALU wave instructions per voxel
        # step from the committed SQ counters of this source)
        valu = prof["valu_instr_per_step"] * steps_per_launch / avg_kernel_s
        out["valu_roofline"] = {
            "bound": "valu",
            "achieved": valu,
            "peak": VALU_PEAK_INSTR_PER_S,
            "unit": "wave instructions/s",
            "frac": valu / VALU_PEAK_INSTR_PER_S,
            "instr_per_voxel_step": prof["valu_instr_per_step"],
        }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workdir, args.photons_per_step, args.seed, args.cpu_seconds,
                                           args.cpu_threads)
    g.delete()
    del keep
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
