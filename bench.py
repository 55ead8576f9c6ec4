#!/usr/bin/env python3
"""Throughput of the MS-OSM many-DM grid on one MI355X.

Workload (the metric's configuration): complex int8 baseband, 2^27 samples at 64 MS/s, band
1218-1282 MHz (fc 1250 MHz, BW 64 MHz), 512 DM trials uniformly over 0-100 pc/cm^3, FFT length
2^16 (reference rule, M = 2^15, delaycount 54), 32 blocks per push, every trial folded to 1024
bins at P = 5.0 ms inside the dedispersion kernel. One step = the whole 2^27-sample record through
all 512 trials, input resident in HBM.

Prints one JSON line: DM-trial*samples/s, the dedispersion kernel's roofline position (HIP events
on the plan's stream) and the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "acceleratedms-osm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E

# Uneven grid (--workload uneven): complex64 baseband, 2^26 samples at 32 MS/s, band 784-816 MHz,
# 1024 DM trials over 0-1000 pc/cm^3, each at its own FFT length (the reference's minimum order for
# that DM, 2^14..2^18 here), complex64 series out; one launch over all lengths vs one per DM.
UNEVEN = dict(workload="msosm_uneven_c64", samples=1 << 26, fs_hz=32e6, fc_hz=800e6, bw_hz=32e6, num_dms=1024,
              dm_lo=0.0, dm_hi=1000.0, push_blocks=1, output="complex64 series", input="complex64, normal")

WORKLOAD = dict(workload="msosm_dm_grid_int8", samples=1 << 27, fs_hz=64e6, fc_hz=1250e6, bw_hz=64e6,
                num_dms=512, dm_lo=0.0, dm_hi=100.0, fft_len=1 << 16, batch=32, nbins=1024, period_s=5.0e-3,
                input="int8 I/Q, uniform in [-127,127]")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(design, n_dm=24, n_samples=1 << 22):
    """FP64 oracle (single thread) on a bounded sample: the first n_dm trials of the same grid over
    n_samples samples, dedispersed and folded exactly as the GPU path."""
    import numpy as np

    from oracle import oracle as orc

    rng = np.random.default_rng(1)
    q = rng.integers(-127, 128, size=2 * n_samples, dtype=np.int8)
    x = orc.int8_to_complex(q)
    idx = np.linspace(0, len(design.dms) - 1, n_dm).astype(int)
    delays = design.delays[idx]
    H = design.H[idx].astype(np.complex128)
    t0 = time.perf_counter()
    out = orc.msosm(x, design.M, delays, H)
    for d in range(n_dm):
        orc.fold(np.abs(out[d]) ** 2, WORKLOAD["nbins"], 0.0, 1.0 / WORKLOAD["fs_hz"], WORKLOAD["period_s"])
    dt = time.perf_counter() - t0
    return {"value": n_dm * n_samples / dt, "unit": "DM-trial*samples/s", "cores": 1, "kind": "port",
            "sample": f"{n_dm} of the 512 DM trials x {n_samples} samples (2^{n_samples.bit_length() - 1}), "
                      f"FFT length 2^16, FP64 oracle incl. fold, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pushes", type=int, default=0, help="pushes per step (default: the whole record)")
    ap.add_argument("--workload", choices=["dm_grid", "uneven"], default="dm_grid")
    ap.add_argument("--plain-pushes", type=int, default=8,
                    help="uneven workload: pushes timed in the one-launch-per-DM form")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (reference kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copied 3 x 1 GiB")
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              f"path (DM sharding, no exchange) is not built.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven":
        bench_uneven(args)
        return

    import numpy as np
    import torch

    import msosm_amd as ms

    W = WORKLOAD
    f0 = W["fc_hz"] - W["bw_hz"] / 2
    dms = np.linspace(W["dm_lo"], W["dm_hi"], W["num_dms"]).astype(np.float32)
    t = time.perf_counter()
    design = ms.Design(W["bw_hz"], f0, dms)
    log(f"design: M={design.M} L={2 * design.M} delaycount={design.delaycount} ({time.perf_counter() - t:.1f} s)")
    assert 2 * design.M == W["fft_len"]

    N, B, M, D = W["samples"], W["batch"], design.M, W["num_dms"]
    per = B * M
    pushes = N // per
    if args.pushes:
        pushes = min(pushes, args.pushes)
    torch.cuda.init()
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(1234)
    data = torch.randint(-127, 128, (N * 2,), dtype=torch.int8, device=dev, generator=gen)
    torch.cuda.synchronize()

    plan = ms.Plan.from_design(design, B, in_format=ms.IN_I8, out_flags=ms.OUT_NONE, device=0)
    fold = ms.Fold(W["period_s"], W["fs_hz"], per, W["nbins"], num_series=D, device=0)
    plan.attach_fold(fold)
    info = plan.info()
    log(f"plan: FBS={info.FBS} workgroups={info.workgroups} device bytes={info.device_bytes / 2**30:.2f} GiB")
    base = data.data_ptr()

    def step():
        for c in range(pushes):
            plan.push(base + c * per * 2)

    for _ in range(args.warmup):
        step()
    plan.synchronize()
    torch.cuda.synchronize()
    plan.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    plan.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    inv_ms, inv_n = plan.profile_read(1)
    fwd_ms, fwd_n = plan.profile_read(0)
    plan.profile(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = D * pushes * per / ((t1 - t0) / args.steps)

    # roofline of the dominant kernel: algorithmic bytes per launch = D*B*M DM-samples x 16 B
    # (the L = 2M gathered complex64 bins read per M output samples; the fold writes ~nothing)
    inv_avg_s = inv_ms / inv_n / 1e3
    alg_bytes = D * B * M * 16.0
    achieved = alg_bytes / inv_avg_s
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "inv_kernel<256,256>",
            "avg_launch_ms": inv_avg_s * 1e3, "launches": inv_n, "alg_bytes_per_launch": alg_bytes}
    tpath = os.path.join(ROOT, "profiles", "traffic_msosm_dm_grid_int8.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            roof["traffic"] = json.load(fh).get("hbm_bytes_per_launch")
    log(f"fwd kernel avg {fwd_ms / max(fwd_n, 1):.3f} ms x {fwd_n}, inv kernel avg {inv_avg_s * 1e3:.3f} ms x {inv_n}")

    prof, sums, cnts = fold.read()
    assert int(cnts.sum()) > 0

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(design)
        log(f"cpu baseline: {cpu['value']:.3e} DM-trial*samples/s ({cpu['sample']})")

    out = {
        "metric": "DM-trial*samples/s", "value": value, "unit": "DM-trial*samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": W, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    fold.close()


def bench_uneven(args):
    """Uneven DM grid: the one-launch form over all length classes against one launch per DM."""
    import numpy as np
    import torch

    import msosm_amd as ms

    W = UNEVEN
    f0 = W["fc_hz"] - W["bw_hz"] / 2
    dms = np.linspace(W["dm_lo"], W["dm_hi"], W["num_dms"]).astype(np.float32)
    t = time.perf_counter()
    design = ms.UnevenDesign(W["bw_hz"], f0, dms)
    multi = ms.Multi(design, push_blocks=W["push_blocks"], in_format=ms.IN_C64, out_flags=ms.OUT_C64, device=0)
    lengths = {int(2 * m): int((design.M == m).sum()) for m in np.unique(design.M)}
    log(f"design + plans: {multi.num_classes} length classes {lengths} ({time.perf_counter() - t:.1f} s)")
    N, D, per = W["samples"], W["num_dms"], multi.push_samples
    pushes = N // per
    if args.pushes:
        pushes = min(pushes, args.pushes)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(4321)
    data = torch.randn(N, dtype=torch.complex64, device=dev, generator=gen)
    torch.cuda.synchronize()
    base = data.data_ptr()

    def run(n, first=0):
        for c in range(n):
            multi.push(base + ((first + c) % (N // per)) * per * 8)
        multi.synchronize()

    res = {}
    for name, mode, n in (("one_launch", ms.LAUNCH_PERSISTENT, pushes),
                          ("per_dm", ms.LAUNCH_PER_DM, min(args.plain_pushes, pushes))):
        multi.set_launch_mode(mode)
        run(2)
        multi.profile(True)
        t0 = time.perf_counter()
        steps = args.steps if name == "one_launch" else 1
        for _ in range(steps):
            run(n)
        t1 = time.perf_counter()
        kms, kn = multi.profile_read(1)
        multi.profile(False)
        res[name] = dict(rate=D * n * per * steps / (t1 - t0), ms_per_push=(t1 - t0) * 1e3 / (n * steps),
                         kernel_ms=kms, launches=kn, pushes=n * steps)
        log(f"{name}: {res[name]['rate']:.3e} DM-trial*samples/s, {res[name]['ms_per_push']:.3f} ms/push, "
            f"{kn} dedispersion launches, {kms:.1f} ms in them")
    one = res["one_launch"]
    alg = D * per * 24.0  # 16 B gathered spectrum + 8 B complex64 output per DM-sample
    avg_s = one["kernel_ms"] / one["launches"] / 1e3
    roof = {"bound": "hbm", "achieved": alg / avg_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": alg / avg_s / HBM_PEAK, "traffic": None, "kernel": "inv_multi_kernel",
            "avg_launch_ms": avg_s * 1e3, "launches": one["launches"], "alg_bytes_per_launch": alg}
    out = {
        "metric": "DM-trial*samples/s", "value": one["rate"], "unit": "DM-trial*samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 2, "ms_per_step": one["ms_per_push"] * pushes, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": dict(W, length_classes=lengths), "roofline": roof, "cpu_baseline": None,
        "per_dm_launch_form": {"value": res["per_dm"]["rate"], "pushes": res["per_dm"]["pushes"],
                               "speedup_of_one_launch": one["rate"] / res["per_dm"]["rate"]},
    }
    print(json.dumps(out), flush=True)
    multi.close()


if __name__ == "__main__":
    main()
