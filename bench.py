#!/usr/bin/env python3
"""Throughput of the MS-OSM many-DM grid on one MI355X.

Workload (the metric's configuration): complex int8 baseband, 2^27 samples at 64 MS/s, band
1218-1282 MHz (fc 1250 MHz, BW 64 MHz), 512 DM trials uniformly over 0-100 pc/cm^3, FFT length
2^16 (reference rule, M = 2^15, delaycount 54), 32 blocks per push, every trial folded to 1024
bins at P = 5.0 ms inside the dedispersion kernel. One step = the whole 2^27-sample record through
all 512 trials, input resident in HBM.

Prints one JSON line: DM-trial*samples/s, the dedispersion kernel's roofline position (HIP events
on the plan's stream), the CPU port timed on a bounded sample of the same workload, the relative
error of each workload against the FP64 oracle (sampled trials, outside the timed region), a check
of the timed launch's folds against the sequential launch, and a nested object for the uneven
1024-trial grid (one launch over all FFT lengths against one launch per trial).
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
ORDERS = ["xcd_dm_major", "linear", "xcd_block_major", "xcd_tile4", "xcd_tile8", "xcd_tile16", "xcd_tile2"]
SPLIT = 8  # schedule flag: role-split launch (MSOSM_SCHEDULE_SPLIT)
ACCUM = 16  # schedule flag: accumulate launch (MSOSM_SCHEDULE_ACCUM)
FORMS = {"split": SPLIT, "accum": ACCUM}


def order_name(order: int) -> str:
    return ORDERS[order & 7] + ("+accum" if order & ACCUM else "+split" if order & SPLIT else "")


def order_code(name: str) -> int:
    base, _, form = name.partition("+")
    return ORDERS.index(base) | FORMS.get(form, 0)

# Uneven grid (--workload uneven): complex64 baseband, 2^26 samples at 32 MS/s, band 784-816 MHz,
# 1024 DM trials over 0-1000 pc/cm^3, each at its own FFT length (the reference's minimum order for
# that DM, 2^14..2^18 here), complex64 series out; one launch over all lengths vs one per DM. A push is
# 32 blocks of the longest length's M = 2^17 (2^22 samples), the reference speed harness's batch
# (test_msosm_speed.cpp:119-120), as is the one-launch-per-DM form it is compared with; the sub-unit
# spectra of one push take 60 GB of device memory.
UNEVEN = dict(workload="msosm_uneven_c64", samples=1 << 26, fs_hz=32e6, fc_hz=800e6, bw_hz=32e6, num_dms=1024,
              dm_lo=0.0, dm_hi=1000.0, push_blocks=32, output="complex64 series", input="complex64, normal")

# DM refinement (--workload refine): complex64 baseband, 2^24 samples at 16 MS/s, band 1392-1408 MHz,
# a dispersed pulsar (DM 123.456, P = 33 ms, Gaussian pulse of 3 % FWHM, amplitude 0.5, AWGN sigma 1),
# 512-bin folds; 64 start DMs within +-0.5 refined by DM gradients vs a 1e-4 grid over +-0.5.
REFINE = dict(workload="msosm_dm_refine", samples=1 << 24, fs_hz=16e6, fc_hz=1400e6, bw_hz=16e6, true_dm=123.456,
              period_s=33.0e-3, duty=0.03, amp=0.5, sigma=1.0, nbins=512, num_starts=64, start_halfwidth=0.5,
              grid_step=1e-4, batch=32)

WORKLOAD = dict(workload="msosm_dm_grid_int8", samples=1 << 27, fs_hz=64e6, fc_hz=1250e6, bw_hz=64e6,
                num_dms=512, dm_lo=0.0, dm_hi=100.0, fft_len=1 << 16, batch=32, nbins=1024, period_s=5.0e-3,
                input="int8 I/Q, uniform in [-127,127]")


def alg_bytes_per_dm_sample(in_bytes: float, out_bytes: float, num_dms: int, fft_len: int, samples: int) -> float:
    """Algorithmic HBM bytes per DM-trial*sample: S + O + F/D + 16 L/N with S = 16 (the L gathered
    complex64 bins of a block per M outputs), O the output bytes per sample, F = in_bytes + 16 (input
    read and its spectrum written once, shared by the D trials) and 16 L/N the chirp read once per trial
    over the record."""
    return 16.0 + out_bytes + (in_bytes + 16.0) / num_dms + 16.0 * fft_len / samples


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Host threads for the CPU baseline: the process's CPU share, capped by OMP_NUM_THREADS if set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def cpu_baseline(design, n_samples=3 << 24):
    """FP32 CPU port of the same MS-OSM + fold (oracle/cpu_msosm.c: FFTW3-API transforms from Intel MKL,
    OpenMP over DM trials on every host thread) on a bounded sample: all 512 trials of the grid over the
    first n_samples samples of a uniform int8 input, dedispersed and folded to the same 1024 bins."""
    import numpy as np

    from oracle import oracle as orc

    rng = np.random.default_rng(1)
    q = rng.integers(-127, 128, size=2 * n_samples, dtype=np.int8)
    x = orc.int8_to_complex(q).astype(np.complex64)
    threads = host_threads()
    W = WORKLOAD
    t0 = time.perf_counter()
    orc.cpu_msosm_fold(x, design.M, design.delays, design.H, W["nbins"], 0.0, 1.0 / W["fs_hz"], W["period_s"],
                       nthreads=threads, call_size=W["batch"] * design.M)
    dt = time.perf_counter() - t0
    n_dm = len(design.dms)
    tz = (n_samples & -n_samples).bit_length() - 1  # n_samples = odd x 2^tz
    size = f"2^{tz}" if n_samples == 1 << tz else f"{n_samples >> tz} x 2^{tz}"
    return {"value": n_dm * n_samples / dt, "unit": "DM-trial*samples/s", "cores": threads, "kind": "port",
            "sample": f"all {n_dm} DM trials x {n_samples} samples ({size}), FFT length "
                      f"2^{design.M.bit_length()}, FP32 MS-OSM + fold in {dt:.1f} s on {threads} threads "
                      f"(OpenMP over trials); transforms: {orc.cpu_backend()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pushes", type=int, default=0, help="pushes per step (default: the whole record)")
    ap.add_argument("--workload", choices=["dm_grid", "uneven", "refine", "save"], default="dm_grid")
    ap.add_argument("--order", choices=ORDERS + [o + "+" + f for o in ORDERS for f in FORMS], default=None,
                    help="fixed unit order instead of the autotuner (profiling runs)")
    ap.add_argument("--refine-steps", type=int, default=20)
    ap.add_argument("--push-blocks", type=int, default=0,
                    help="uneven workload: blocks of the longest length per push (default 32)")
    ap.add_argument("--graph", action="store_true", help="uneven workload: run each one-launch push as a hipGraph")
    ap.add_argument("--plain-pushes", type=int, default=4,
                    help="uneven workload: pushes timed in the one-launch-per-DM form (at most a quarter record)")
    ap.add_argument("--no-uneven", action="store_true",
                    help="default workload: leave out the nested uneven-grid measurement")
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
    if args.workload == "refine":
        bench_refine(args)
        return
    if args.workload == "save":
        bench_save(args, pushes=args.pushes or 6)
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
    if args.order:
        plan.set_schedule(0, order_code(args.order))
        wg, order = plan.schedule()
        schedule = {"workgroups": wg, "order": order_name(order), "autotuned": False}
    else:
        # tile schedule: measured on the first pushes of the record (this resets the plan and the fold);
        # five timed pushes per candidate, since the best unit orders lie within 1 % of each other
        wg, order, tms = plan.autotune(base, 6)
        schedule = {"workgroups": wg, "order": order_name(order), "autotuned": True}
        log(f"autotune: {schedule} ({tms:.3f} ms per dedispersion launch)")

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

    # roofline of the dominant kernel: algorithmic bytes per launch = the per-DM-sample figure
    # S + O + F/D + 16 L/N (O = 0: the fold writes ~nothing per sample) times the D*B*M DM-samples of
    # one launch
    inv_avg_s = inv_ms / inv_n / 1e3
    per_dm_sample = alg_bytes_per_dm_sample(2, 0, D, 2 * M, N)
    alg_bytes = D * B * M * per_dm_sample
    achieved = alg_bytes / inv_avg_s
    kname, kdesc = dedispersion_kernel(order, 2 * M)
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None, "kernel": kdesc,
            "avg_launch_ms": inv_avg_s * 1e3, "launches": inv_n, "alg_bytes_per_launch": alg_bytes,
            "alg_bytes_per_dm_sample": per_dm_sample}
    roof.update(read_traffic(kname, "msosm_dm_grid_int8", order_name(order)))
    log(f"fwd kernel avg {fwd_ms / max(fwd_n, 1):.3f} ms x {fwd_n}, inv kernel avg {inv_avg_s * 1e3:.3f} ms x {inv_n}")

    prof, sums, cnts = fold.read()
    assert int(cnts.sum()) > 0
    plan.close()
    fold.close()

    # outside the timed region: the timed folds of sampled trials against the sequential launch of
    # the same design over the same pushes (identical counts; sums equal up to FP64 atomic order)
    check = timed_output_check(ms, design, data, pushes, args.warmup + args.steps, sums, cnts, order)
    log(f"timed-output check: {check}")
    rel_err = {"dm_grid_int8": rel_err_dm_grid(ms, design, data, order)}
    log(f"rel err dm grid: {rel_err['dm_grid_int8']}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(design)
        log(f"cpu baseline: {cpu['value']:.3e} DM-trial*samples/s ({cpu['sample']})")
    del data
    torch.cuda.empty_cache()

    save = None if args.no_uneven else bench_save(args, nested=True)
    nested = None
    if not args.no_uneven:
        nested = bench_uneven(args, nested=True)
        rel_err["uneven_c64"] = nested.pop("rel_err")

    out = {
        "metric": "DM-trial*samples/s", "value": value, "unit": "DM-trial*samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": dict(W, schedule=schedule), "roofline": roof, "cpu_baseline": cpu,
        "rel_err_vs_fp64": rel_err, "timed_output_check": check, "uneven_grid": nested, "save_mode": save,
    }
    print(json.dumps(out), flush=True)


def dedispersion_kernel(order: int, fft_len: int):
    """(file stem, description) of the kernel a one-length dedispersion launch runs."""
    if order & ACCUM and fft_len in (1 << 14, 1 << 15, 1 << 16):
        return "acc_kernel", f"acc_kernel<{fft_len >> 14}>"
    if order & SPLIT:
        n = int(round(fft_len ** 0.5))
        return "inv_rs_kernel", f"inv_rs_kernel<{n}>"
    lg = fft_len.bit_length() - 1
    n1, n2 = 1 << ((lg + 1) // 2), 1 << (lg // 2)
    return "inv_kernel", f"inv_kernel<{n1},{n2}>"


def read_traffic(kname: str, workload: str, schedule: str) -> dict:
    """L2-to-fabric bytes per launch of a kernel from its committed rocprofv3 counter summary (TCC
    FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel's readings), if one exists for this schedule."""
    tpath = os.path.join(ROOT, "profiles", f"traffic_{kname}_{workload}.json")
    if not os.path.exists(tpath):
        return {}
    with open(tpath) as fh:
        t = json.load(fh)
    if t.get("schedule") not in (None, schedule):
        return {"traffic_note": f"counter profile is of schedule {t.get('schedule')}, not {schedule}"}
    return {"traffic": t.get("l2_fabric_bytes_per_launch"),
            "traffic_counts": "L2-to-fabric bytes per launch (TCC FETCH_SIZE + WRITE_SIZE, copy-calibrated; "
                              "the MALL may serve part before HBM)",
            "traffic_source": "profiles/" + t.get("source", "")}


def timed_output_check(ms, design, data, pushes: int, records: int, sums, cnts, order: int, ntrials: int = 8) -> dict:
    """Fold the same pushes through a plan of sampled trials in the linear unit order, in the timed launch's
    arithmetic (the accumulate form, or the two-pass form that the role split shares with the sequential
    launch): the series bits do not depend on the schedule, so the counts must be identical and the sums
    equal up to the order of their FP64 atomic additions."""
    import numpy as np

    W = WORKLOAD
    pick = np.linspace(1, len(design.dms) - 1, ntrials).astype(int)
    sub = ms.Design(design.bw, design.f0, design.dms[pick], M=design.M)
    assert np.array_equal(sub.delays, design.delays[pick])
    per = W["batch"] * design.M
    plan = ms.Plan.from_design(sub, W["batch"], in_format=ms.IN_I8, out_flags=ms.OUT_NONE, device=0)
    fold = ms.Fold(W["period_s"], W["fs_hz"], per, W["nbins"], num_series=ntrials, device=0)
    plan.attach_fold(fold)
    ref_order = ms.ORDER_LINEAR | (order & ACCUM)
    plan.set_schedule(0, ref_order)
    base = data.data_ptr()
    for _ in range(records):
        for c in range(pushes):
            plan.push(base + c * per * 2)
    _, s2, c2 = fold.read()
    plan.close()
    fold.close()
    counts_equal = bool(np.array_equal(cnts[pick], c2))
    rel = float(np.max(np.abs(sums[pick] - s2) / np.maximum(np.abs(s2), 1e-300)))
    if not counts_equal or rel > 1e-9:
        raise AssertionError(f"timed launch disagrees with the sequential launch: counts equal {counts_equal}, "
                             f"max sum rel diff {rel:.3e}")
    return {"trials": [float(d) for d in design.dms[pick]], "records": records,
            "reference": "linear unit order, " + ("accumulate form" if order & ACCUM else "sequential two-pass form"),
            "fold_counts_equal": counts_equal, "max_fold_sum_rel_diff": rel}


def rel_err_series(got, ref) -> tuple[float, float]:
    import numpy as np

    l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    mx = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    return l2, mx


def rel_err_dm_grid(ms, design, data, order: int, ntrials: int = 4, calls: int = 2) -> dict:
    """Series of sampled trials of the grid (same M, design, int8 input) against the FP64 oracle over the
    record's first calls pushes, from zero history in both, in the timed launch's form (accumulate,
    role split or sequential: the unit order does not change the series bits)."""
    import numpy as np

    from oracle import oracle as orc

    W = WORKLOAD
    pick = np.linspace(1, len(design.dms) - 1, ntrials).astype(int)
    sub = ms.Design(design.bw, design.f0, design.dms[pick], M=design.M)
    per = W["batch"] * design.M
    n = per * calls
    q = data[: 2 * n].cpu().numpy()
    plan = ms.Plan.from_design(sub, W["batch"], in_format=ms.IN_I8, out_flags=ms.OUT_C64, device=0)
    plan.set_schedule(0, order)
    outs = []
    for c in range(calls):
        plan.push(q[2 * c * per: 2 * (c + 1) * per].reshape(-1, 2))
        outs.append(plan.output_c64())
    plan.close()
    got = np.concatenate(outs, axis=1)
    ref = orc.msosm(orc.int8_to_complex(q), design.M, design.delays[pick], design.H[pick].astype(np.complex128))
    errs = [rel_err_series(got[j], ref[j]) for j in range(ntrials)]
    return {"rel_l2_max": max(e[0] for e in errs), "max_abs_rel_max": max(e[1] for e in errs),
            "trials": [float(d) for d in design.dms[pick]], "samples": n, "bar": 1e-5,
            "schedule": order_name(order)}


def bench_uneven(args, nested: bool = False):
    """Uneven DM grid: the one-launch form over all length classes against one launch per DM, the latter
    at the reference speed harness's 32 blocks per call (test_msosm_speed.cpp:119-120) and at the best
    of a few batch sizes."""
    import numpy as np
    import torch

    import msosm_amd as ms

    W = dict(UNEVEN)
    if args.push_blocks:
        W["push_blocks"] = args.push_blocks
    f0 = W["fc_hz"] - W["bw_hz"] / 2
    dms = np.linspace(W["dm_lo"], W["dm_hi"], W["num_dms"]).astype(np.float32)
    t = time.perf_counter()
    design = ms.UnevenDesign(W["bw_hz"], f0, dms)
    lengths = {int(2 * m): int((design.M == m).sum()) for m in np.unique(design.M)}
    N, D = W["samples"], W["num_dms"]
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(4321)
    data = torch.randn(N, dtype=torch.complex64, device=dev, generator=gen)
    torch.cuda.synchronize()
    base = data.data_ptr()

    def timed(multi, mode, n_pushes, steps, graph=False):
        multi.set_launch_mode(mode)
        multi.set_graph(graph)
        per = multi.push_samples
        total = N // per

        def run(n, first=0):
            for c in range(n):
                multi.push(base + ((first + c) % total) * per * 8)
            multi.synchronize()

        run(2)
        multi.profile(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            run(n_pushes)
        t1 = time.perf_counter()
        kms, kn = multi.profile_read(1)
        multi.profile(False)
        return dict(rate=D * n_pushes * per * steps / (t1 - t0), ms_per_push=(t1 - t0) * 1e3 / (n_pushes * steps),
                    kernel_ms=kms, launches=kn, pushes=n_pushes * steps, push_blocks=per // int(design.M.max()))

    multi = ms.Multi(design, push_blocks=W["push_blocks"], in_format=ms.IN_C64, out_flags=ms.OUT_C64, device=0)
    log(f"uneven: {multi.num_classes} length classes {lengths} ({time.perf_counter() - t:.1f} s)")
    per = multi.push_samples
    pushes = N // per
    if args.pushes and not nested:
        pushes = min(pushes, args.pushes)
    # each length class tunes its schedule; then one launch over all classes or one tuned launch per class
    t = time.perf_counter()
    mode, tms = multi.autotune(base, 3)
    mode_name = {ms.LAUNCH_PERSISTENT: "one launch over all classes", ms.LAUNCH_PER_CLASS: "one launch per class"}[mode]
    log(f"uneven autotune: {mode_name}, {tms:.3f} ms per push ({time.perf_counter() - t:.1f} s)")
    # outside the timed region, on the tuned grid (autotune has reset it): the first push against the oracle
    multi.set_launch_mode(mode)
    rel = rel_err_uneven(ms, design, multi, data) if nested else None
    one = timed(multi, mode, pushes, args.steps, graph=args.graph and not nested and mode == ms.LAUNCH_PERSISTENT)
    log(f"uneven tuned: {one['rate']:.3e} DM-trial*samples/s, {one['ms_per_push']:.3f} ms/push")
    multi.close()
    # the plain form: one launch per DM, at the harness's 32 blocks per call and at other batches
    per_dm = {}
    for pb in (4, 32, 64):
        m2 = ms.Multi(design, push_blocks=pb, in_format=ms.IN_C64, out_flags=ms.OUT_C64, device=0)
        n_p = max(1, min(args.plain_pushes, (N // m2.push_samples) // 4))
        per_dm[pb] = timed(m2, ms.LAUNCH_PER_DM, n_p, 1)
        log(f"uneven per-DM launches at {pb} blocks: {per_dm[pb]['rate']:.3e} DM-trial*samples/s")
        m2.close()
    best_pb = max(per_dm, key=lambda k: per_dm[k]["rate"])
    # the reference's per-DM structure with the vendor FFT library, at the harness's 32 blocks per call
    lib_fft = per_dm_library_fft(design, data, 32)
    log(f"uneven per-DM form with library FFTs at 32 blocks: {lib_fft['value']:.3e} DM-trial*samples/s "
        f"({lib_fft['trials_timed']} trials timed), rel-L2 {max(lib_fft['rel_l2_per_length'].values()):.2e}")
    per_sample = alg_bytes_per_dm_sample(8, 8, D, int(2 * design.M.max()), N)
    alg = D * per * per_sample
    # dedispersion time per push: one launch, or one launch per class
    avg_s = one["kernel_ms"] / one["pushes"] / 1e3
    roof = {"bound": "hbm", "achieved": alg / avg_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": alg / avg_s / HBM_PEAK, "traffic": None,
            "kernel": "inv_multi_kernel" if mode == ms.LAUNCH_PERSISTENT else "per-class launches (tuned)",
            "avg_launch_ms": avg_s * 1e3, "launches": one["launches"], "launches_per_push": one["launches"] / one["pushes"],
            "alg_bytes_per_launch": alg, "alg_bytes_per_dm_sample": per_sample}
    if not (args.graph and not nested) and mode == ms.LAUNCH_PERSISTENT:
        roof.update(read_traffic("inv_multi_kernel", "msosm_uneven_c64", "one launch per push"))
    elif mode == ms.LAUNCH_PER_CLASS:
        roof.update(read_traffic("per_class", "msosm_uneven_c64", "one launch per class"))
    del data
    torch.cuda.empty_cache()
    plain = {"harness_batch_32": {"value": per_dm[32]["rate"], "pushes": per_dm[32]["pushes"],
                                  "speedup_of_one_launch": one["rate"] / per_dm[32]["rate"]},
             "best_batch": {"push_blocks": best_pb, "value": per_dm[best_pb]["rate"],
                            "pushes": per_dm[best_pb]["pushes"],
                            "speedup_of_one_launch": one["rate"] / per_dm[best_pb]["rate"]},
             "tried": {str(k): v["rate"] for k, v in per_dm.items()},
             "library_fft_per_dm": dict(lib_fft, speedup_of_one_launch=one["rate"] / lib_fft["value"])}
    out = {
        "metric": "DM-trial*samples/s", "value": one["rate"], "unit": "DM-trial*samples/s",
        "ms_per_step": one["ms_per_push"] * pushes, "steps": args.steps,
        "config": dict(W, length_classes=lengths, graph=bool(args.graph and not nested),
                       schedule={"launch_mode": mode_name, "autotuned": True, "ms_per_push_tuning": tms}),
        "roofline": roof,
        "per_dm_launch_form": plain,
    }
    if nested:
        out["rel_err"] = rel
        return out
    out.update({"n_gpus": 1, "warmup": 2, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic", "cpu_baseline": None})
    print(json.dumps(out), flush=True)


def per_dm_library_fft(design, data, push_blocks: int, pushes: int = 2, stride: int = 8) -> dict:
    """The reference's GPU form on this GPU, one DM trial at a time (msosm_gpu_batch.cpp:193-201): per trial
    and push, every bin's spectrum gathered from the ring of forward transforms by the bin's delay in
    blocks, times the chirp, the push's blocks inverse-transformed as one batched FFT of the vendor library
    (torch.fft on hipFFT/rocFFT), and each block's first half discarded into the trial's series. The
    forward transforms are computed once per push and length class and shared by its trials (the
    reference's instances each compute their own). Timed over `pushes` pushes after a warm-up push, on
    every `stride`-th trial of each length class; the first push (zero history) of one trial per class is
    compared with the FP64 oracle."""
    import numpy as np
    import torch

    from oracle import oracle as orc

    dev = data.device
    Mmax = int(design.M.max())
    per = push_blocks * Mmax
    n_push = pushes + 1
    x = data[: n_push * per]
    classes = []
    for M in np.unique(design.M):
        M = int(M)
        L, B = 2 * M, per // M
        sel = np.nonzero(design.M == M)[0][::stride]
        dmax = int(max(int(design.delays[i].max()) for i in sel))
        # ring of forward transforms: row dmax + j is block j's segment [x_(j-1), x_j] (zero before the
        # record); rows below dmax stay zero (blocks before the record)
        ring = torch.zeros(dmax + n_push * B, L, dtype=torch.complex64, device=dev)
        xpad = torch.cat([torch.zeros(M, dtype=torch.complex64, device=dev), x])
        segs = xpad.unfold(0, L, M)
        rel = [torch.from_numpy(np.arange(B)[:, None] + dmax - design.delays[i][None, :].astype(np.int64)).to(dev)
               for i in sel]
        H = [torch.from_numpy(np.ascontiguousarray(design.H[i]).astype(np.complex64)).to(dev) for i in sel]
        out = torch.empty(len(sel), per, dtype=torch.complex64, device=dev)
        classes.append(dict(M=M, L=L, B=B, sel=sel, dmax=dmax, ring=ring, segs=segs, rel=rel, H=H, out=out))

    def push(p, keep=None):
        for c in classes:
            M, L, B = c["M"], c["L"], c["B"]
            c["ring"][c["dmax"] + p * B: c["dmax"] + (p + 1) * B] = torch.fft.fft(c["segs"][p * B:(p + 1) * B], dim=1)
            for j in range(len(c["sel"])):
                Y = torch.gather(c["ring"], 0, c["rel"][j] + p * B) * c["H"][j]
                y = torch.fft.ifft(Y, dim=1)
                c["out"][j].view(B, M).copy_(y[:, M:])
            if keep is not None:
                keep[M] = c["out"][-1].cpu().numpy().copy()

    first = {}
    push(0, first)  # warm-up, and the checked push
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in range(1, n_push):
        push(p)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    xs = x[:per].cpu().numpy().astype(np.complex128)
    errs = {}
    for c in classes:
        i = int(c["sel"][-1])
        ref = orc.msosm(xs, c["M"], design.delays[i][None], design.H[i][None].astype(np.complex128))[0]
        errs[str(c["L"])] = rel_err_series(first[c["M"]], ref)[0]
    trials = sum(len(c["sel"]) for c in classes)
    del classes
    torch.cuda.empty_cache()
    return {"value": trials * per * pushes / dt, "trials_timed": trials, "push_blocks": push_blocks,
            "pushes": pushes, "rel_l2_per_length": errs}


def rel_err_uneven(ms, design, multi, data) -> dict:
    """One trial of each length class against the FP64 oracle over the record's first push (zero history),
    through the given grid in its current (autotuned, freshly reset) state, so in the timed launch forms."""
    import numpy as np

    from oracle import oracle as orc

    per = multi.push_samples
    x = data[:per].cpu().numpy()
    multi.push(x)
    multi.synchronize()
    res = {}
    for M in np.unique(design.M):
        d = int(np.nonzero(design.M == M)[0][-1])
        got = multi.output(d)
        ref = orc.msosm(x.astype(np.complex128), int(M), design.delays[d][None], design.H[d][None].astype(np.complex128))[0]
        l2, mx = rel_err_series(got, ref)
        res[str(int(2 * M))] = {"dm": float(design.dms[d]), "rel_l2": l2, "max_abs_rel": mx}
    return {"per_length": res, "rel_l2_max": max(v["rel_l2"] for v in res.values()), "samples": per, "bar": 1e-5}


def bench_save(args, pushes: int = 6, repeats: int = 3, nested: bool = False):
    """Save mode end to end, as the reference's speed harness times it (test_msosm_speed.cpp:143-208): the
    grid's 512 trials with uint16-pair output, each push's int8 input copied from pinned host memory and
    its 512 x 32 x 32768 uint16 pairs (2 GiB) copied back into one pinned host buffer, `repeats` times over
    `pushes` pushes, mean and standard deviation of the wall time with the copies included. Pipelined:
    input and output copies on the plan's copy streams, one synchronize per repeat (the harness's loop);
    serialised: the same with a synchronize after every push."""
    import ctypes as C

    import numpy as np
    import torch

    import msosm_amd as ms

    W = WORKLOAD
    f0 = W["fc_hz"] - W["bw_hz"] / 2
    design = ms.Design(W["bw_hz"], f0, np.linspace(W["dm_lo"], W["dm_hi"], W["num_dms"]).astype(np.float32))
    B, M, D = W["batch"], design.M, W["num_dms"]
    per = B * M
    lib = ms.lib()

    def pinned(nbytes):
        p = C.c_void_p()
        ms._check(lib.msosm_host_alloc(C.c_ulonglong(nbytes), C.byref(p)))
        return p

    rng = np.random.default_rng(7)
    q = rng.integers(-127, 128, size=2 * per * pushes, dtype=np.int8)
    pin_in = pinned(q.nbytes)
    C.memmove(pin_in, q.ctypes.data, q.nbytes)
    out_bytes = D * per * 4
    pin_out = pinned(out_bytes)
    plan = ms.Plan.from_design(design, B, in_format=ms.IN_I8, out_flags=ms.OUT_U16, device=0)
    plan.set_schedule(0, order_code("xcd_tile16+accum"))
    res = {}
    for name, sync_each in (("pipelined", False), ("serialised", True)):
        times = []
        for rep in range(repeats + 1):  # the first repeat warms up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for c in range(pushes):
                ms._check(lib.msosm_plan_push(plan.h, C.c_void_p(pin_in.value + c * per * 2), ms.INPUT_HOST_ASYNC))
                ms._check(lib.msosm_plan_get_output(plan.h, ms.OUT_U16, pin_out))
                if sync_each:
                    plan.synchronize()
            plan.synchronize()
            if rep:
                times.append((time.perf_counter() - t0) * 1e3)
        t = np.array(times)
        res[name] = {"mean_ms": float(t.mean()), "std_ms": float(t.std()),
                     "rate": D * per * pushes / (t.mean() / 1e3), "d2h_GB_per_s": out_bytes * pushes / (t.mean() / 1e3) / 1e9}
        log(f"save mode {name}: {t.mean():.1f} +- {t.std():.1f} ms for {pushes} pushes "
            f"({res[name]['rate']:.3e} DM-trial*samples/s, {res[name]['d2h_GB_per_s']:.1f} GB/s out)")
    plan.close()
    lib.msosm_host_free(pin_in)
    lib.msosm_host_free(pin_out)
    out = {"workload": "msosm_dm_grid_int8 save mode (uint16-pair output, host in and out)", "pushes": pushes,
           "repeats": repeats, "output_bytes_per_push": out_bytes, "unit": "DM-trial*samples/s",
           "schedule": "xcd_tile16+accum", **res,
           "overlap_gain": res["serialised"]["mean_ms"] / res["pipelined"]["mean_ms"]}
    if nested:
        return out
    print(json.dumps(out), flush=True)


def synth_pulsar(n, fs, f0, dm, period, duty, amp, sigma, seed, device):
    """Dispersed pulsar baseband on the GPU: complex unit-variance noise modulated by a Gaussian pulse
    (FWHM = duty * period, peak amplitude amp), dispersed with the direct chirp e^{+i phi} over the
    whole record (FP64), plus complex AWGN of standard deviation sigma."""
    import math

    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    t = torch.arange(n, device=device, dtype=torch.float64) / fs
    ph = torch.remainder(t / period, 1.0) - 0.5
    sig = duty / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    a = amp * torch.exp(-0.5 * (ph / sig) ** 2)
    x = a * torch.randn(n, dtype=torch.complex128, device=device, generator=g)
    k = torch.arange(n, device=device, dtype=torch.float64)
    f = torch.where(k < n // 2, fs / 2 + k * fs / n, k * fs / n - fs / 2)
    w, w0 = 2 * math.pi * f, 2 * math.pi * f0
    phase = 4 * math.pi ** 2 * 4.15e15 * dm * w * w / ((w + w0) * w0 * w0)
    x = torch.fft.ifft(torch.fft.fft(x) * torch.polar(torch.ones_like(phase), phase))
    x = x + sigma * torch.randn(n, dtype=torch.complex128, device=device, generator=g)
    return x.to(torch.complex64)


def bench_refine(args):
    """Gradient DM refinement (forward fold + DM-derivative pass per step) against a fine DM grid."""
    import numpy as np
    import torch

    from msosm_amd import autograd as ag

    W = REFINE
    f0 = W["fc_hz"] - W["bw_hz"] / 2
    dev = torch.device("cuda:0")
    N = W["samples"]
    rec = synth_pulsar(N, W["fs_hz"], f0, W["true_dm"], W["period_s"], W["duty"], W["amp"], W["sigma"], 2024, dev)
    torch.cuda.synchronize()
    noise_var = ag.profile_noise_var(rec, N, W["nbins"])
    rng = np.random.default_rng(20261203)
    starts = W["true_dm"] + rng.uniform(-W["start_halfwidth"], W["start_halfwidth"], W["num_starts"])
    lo, hi = W["true_dm"] - W["start_halfwidth"], W["true_dm"] + W["start_halfwidth"]
    # every trial is designed at its own DM on each pass; M and the ring depth cover the whole window
    eng = ag.ProfileEngine(W["bw_hz"], f0, starts.astype(np.float32), W["batch"], W["nbins"], W["period_s"],
                           dm_max=float(np.float32(hi)))
    log(f"refine: M={eng.M} delaycount={eng.delaycount} noise_var={noise_var:.3e}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dm_ref, loss_ref, steps = ag.refine_dm(eng, rec, starts, noise_var, steps=args.refine_steps, bounds=(lo, hi))
    t_ref = time.perf_counter() - t0
    best = int(np.argmin(loss_ref))
    # the grid: every DM at 1e-4 over +-0.5 around the truth, in engine batches of 1024 trials
    grid = W["true_dm"] + np.arange(-W["start_halfwidth"], W["start_halfwidth"] + 1e-12, W["grid_step"])
    t0 = time.perf_counter()
    gl = []
    for s0 in range(0, grid.size, 1024):
        dms = grid[s0:s0 + 1024].astype(np.float32)
        ge = ag.ProfileEngine(W["bw_hz"], f0, dms, W["batch"], W["nbins"], W["period_s"],
                              dm_max=float(np.float32(hi)))
        assert ge.M == eng.M
        prof, _, _ = ge.run(rec.data_ptr(), N, False)
        gl.append(ag.snr_loss(torch.from_numpy(prof), noise_var).numpy())
        del ge
    t_grid = time.perf_counter() - t0
    gl = np.concatenate(gl)
    gbest = int(np.argmin(gl))
    # trend or noise: a line through the means of 20 consecutive grid stretches (the loss is smooth on
    # 1e-4 scales, so single grid points are not independent), and the pulse smear over the window
    nblk = 20
    blk = gl[: (gl.size // nblk) * nblk].reshape(nblk, -1)
    xm = grid[: blk.size].reshape(nblk, -1).mean(axis=1)
    ym = blk.mean(axis=1)
    A = np.vstack([xm - xm.mean(), np.ones(nblk)]).T
    coef, res, _, _ = np.linalg.lstsq(A, ym, rcond=None)
    sigma = float(np.sqrt(res[0] / (nblk - 2))) if res.size else 0.0
    slope_se = sigma / float(np.sqrt(((xm - xm.mean()) ** 2).sum()))
    kdm = 4.15e15
    smear = kdm * W["start_halfwidth"] * (f0 ** -2 - (f0 + W["bw_hz"]) ** -2)
    # the two-stage search: a grid at 1e-3 (the DM scale on which the loss decorrelates here), then
    # gradient refinement from its 4 deepest basins
    t0 = time.perf_counter()
    srch = ag.search_dm(rec, N, W["bw_hz"], f0, lo, hi, noise_var, W["nbins"], W["period_s"], W["batch"], 1e-3,
                        polish=4, steps=args.refine_steps)
    t_srch = time.perf_counter() - t0
    search = {"best_dm": srch["best_dm"], "best_loss": srch["best_loss"], "seconds": t_srch,
              "dm_record_passes": srch["passes"], "coarse_step": 1e-3, "coarse_best_dm": srch["coarse_best_dm"],
              "coarse_best_loss": srch["coarse_best_loss"], "polished_starts": srch["starts"],
              "abs_dm_difference_to_grid_best": abs(srch["best_dm"] - float(grid[gbest])),
              "loss_le_grid_best": bool(srch["best_loss"] <= gl[gbest])}
    log(f"search: {search}")
    landscape = {"grid_loss_min": float(gl.min()), "grid_loss_max": float(gl.max()),
                 "grid_loss_at_true_dm": float(gl[int(np.argmin(np.abs(grid - W["true_dm"])))]),
                 "trend_slope_per_dm": float(coef[0]), "trend_slope_stderr": slope_se,
                 "stretch_mean_std": float(ym.std()),
                 "smear_at_window_edge_us": smear * 1e6, "pulse_fwhm_us": W["duty"] * W["period_s"] * 1e6}
    out = {
        "metric": "DM refinement", "workload": W["workload"], "config": W,
        "refined_best_dm": float(dm_ref[best]), "refined_best_loss": float(loss_ref[best]),
        "grid_best_dm": float(grid[gbest]), "grid_best_loss": float(gl[gbest]),
        "abs_dm_difference": abs(float(dm_ref[best]) - float(grid[gbest])),
        "refine_seconds": t_ref, "refine_steps": steps, "refine_dm_passes": W["num_starts"] * steps * 3,
        "grid_seconds": t_grid, "grid_trials": int(grid.size),
        "refined_dm_spread": [float(np.min(dm_ref)), float(np.max(dm_ref))],
        "refined_better_than_grid": bool(loss_ref[best] <= gl[gbest]), "landscape": landscape,
        "coarse_grid_then_gradient": search,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
