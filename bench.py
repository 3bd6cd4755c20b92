"""Benchmark of the MI355X NLM path: one JSON line on stdout.

Default workload (the headline): 3D NLM of a 256^3 FP32 ellipsoid phantom
(64 ellipsoids, semi-axes 5-40 voxels, sigma 0.05), patch 5^3 (r_p = 2),
search 11^3 (r_s = 5), h = 0.05.  One step = one forward pass over the whole
volume through the C-ABI; the input is resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
    python bench.py --pmc-pass [--workload NAME]   (program for rocprofv3 --pmc)

Roofline: the kernels are FP32-VALU bound.  Algorithmic work per
(voxel, offset) pair is 13 flop (diff, square, 3 separable running-sum
updates of 2 flop, scale, exp, and 3 for the two accumulations), so one
forward is N * (2 r_s + 1)^3 * 13 flop; achieved = that / kernel time,
peak = 157.3 TFLOP/s (MI355X FP32 vector).  HBM traffic per launch comes
from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs kept in profiles/.

The default line also carries, as objects of their own: the batched
persistent launch of 64 uneven volumes against per-volume launches
("batched"), fwd+bwd against fwd on 8 x 128^3 ("grad"), the reference's own
default configuration on its 8-bit data type (512 x 512 x 128, NS 17 /
SS 13: "reference_default") and its run.sh sweep ("sweep").

cpu_baseline: the CPU FP32 path of the oracle (separable, OpenMP), timed on
this host's cores on a bounded sample of the same volume (a slab of z-planes,
or for large patches one x-y window of a plane); the sample's outputs are also
compared with the GPU's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-gnlm_amd"))

PEAK_FP32_TFLOPS = 157.3
FLOP_PER_PAIR_3D = 13
FLOP_PER_PAIR_2D = 11

WORKLOADS = {
    "nlm3d_256_p5_s11": dict(dims=(256, 256, 256), n_ell=64, ax=(5.0, 40.0), sigma=0.05, rp=2, rs=5, h=0.05,
                             seed=20261206),
    "nlm3d_96x96x48_p3_s7": dict(dims=(96, 96, 48), n_ell=12, ax=(4.0, 20.0), sigma=0.05, rp=1, rs=3, h=0.05,
                                 seed=20261204),
    # the reference's default configuration (3DNLM/config.h:4-9: 512 x 512 x 128,
    # SS 13, NS 17; 8-bit input, h = PARAMETER = 10 on 0-255, 3DNLM/main.cu:13)
    "nlm3d_512x512x128_p17_s13_u8": dict(dims=(512, 512, 128), n_ell=64, ax=(5.0, 40.0), sigma=0.05, rp=8, rs=6,
                                         h=10.0, seed=20261210, u8=True),
    "nlm2d_2048_p5_s11": dict(dims2=(2048, 2048), sigma=0.08, rp=2, rs=5, h=0.08, seed=20261205),
    # the reference's own 2D geometry ("2D NLM/11x11/main.cu": patch 11, search
    # 21, uint8 input, h = PARAMETER = 10 on 0-255), the one with a published
    # figure: 50.90 ms = 82.4 Mpixels/s at 2048^2 on an sm_86 GPU
    # ("2D NLM/11x11/resource-11.txt:807")
    "nlm2d_2048_p11_s21_u8": dict(dims2=(2048, 2048), sigma=0.08, rp=5, rs=10, h=10.0, seed=20261205, u8=True,
                                  published=82.4),
}


def _traffic_from_profiles(name, shape=None):
    """HBM bytes per launch measured by rocprofv3 PMC runs, if recorded: one
    figure per workload, or one per launch shape of its tile family (the
    shapes give the same bits but differ in register spills, so in traffic)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f).get(name)
    except (OSError, ValueError):
        return None
    if isinstance(t, dict):
        return t.get(str(shape))
    return t


def cpu_sample(nx, ny, nz, rp, rs, planes=0, budget_s=4.0):
    """The bounded CPU-baseline sample of a 3D workload.

    Cost of the separable CPU path on 16 threads: about 7e-12 s x (x y voxels)
    x (planes + 2 r_p) x offsets x (2 r_p + 1)^2.4 (256 planes of the headline
    volume, 256^2 s11 p5, take 9-13 s; 2 planes of 512^2 s13 p17 took 81 s).
    Takes as many centred z-planes as budget_s (per timed run; the baseline
    is the median of three) allows, or `planes` if > 0.
    When even one plane of the full x-y extent would take longer, it takes one
    plane of a centred x-y window instead; the CPU path then runs on the
    window grown by R = r_p + r_s in every axis (clipped to the volume: a
    clipped face is the volume's own, where reflect-101 applies as usual), so
    its interior sees exactly the samples the full volume gives it.

    Returns (z0, nplanes, y0, wy, x0, wx) of the sample and the crop
    (cz0, cz1, cy0, cy1, cx0, cx1) the CPU path runs on."""
    R = rp + rs
    per = 7e-12 * (2 * rs + 1) ** 3 * (2 * rp + 1) ** 2.4
    nplanes = int(budget_s / (per * nx * ny)) - 2 * rp
    wx, wy = nx, ny
    if nplanes < 1:
        nplanes = 1
        side = int((budget_s / (per * (1 + 2 * rp))) ** 0.5)
        w = max(32, (side - 2 * R) // 16 * 16)
        wx, wy = (w if w + 2 * R <= nx else nx), (w if w + 2 * R <= ny else ny)
    nplanes = max(1, min(planes if planes > 0 else nplanes, nz))
    z0, y0, x0 = nz // 2 - nplanes // 2, (ny - wy) // 2, (nx - wx) // 2
    if (wx, wy) == (nx, ny):
        crop = (0, nz, 0, ny, 0, nx)
    else:
        crop = (max(0, z0 - R), min(nz, z0 + nplanes + R), max(0, y0 - R), min(ny, y0 + wy + R), max(0, x0 - R),
                min(nx, x0 + wx + R))
    return (z0, nplanes, y0, wy, x0, wx), crop


def _threads():
    """Threads the CPU leg uses: OMP_NUM_THREADS if set, else the affinity count."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    if n > 0:
        return n
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_info():
    """(model name, physical cores) of this host, from /proc/cpuinfo: distinct
    (physical id, core id) pairs among the CPUs this process may run on."""
    model, cores, cur = None, set(), {}
    try:
        allowed = os.sched_getaffinity(0)
    except AttributeError:
        allowed = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f.read().splitlines() + [""]:
                if not line.strip():
                    if cur and (allowed is None or int(cur.get("processor", -1)) in allowed):
                        cores.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if k.strip() == "model name" and model is None:
                    model = v.strip()
    except OSError:
        pass
    return model, (len(cores) or None)


def _median_time(fn, runs=3):
    """Median wall time of `runs` calls of fn (the last call's result kept)."""
    ts, out = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts, out


def batch_summary(r):
    """The batched-workload object carried on the headline line."""
    return {"workload": "nlm3d_batch64", "ratio": r["value"], "plain_ms": r["plain_ms"],
            "batched_ms": r["batched_ms"], "ratio_vs_unsplit": r["ratio_vs_unsplit"],
            "plain_unsplit_ms": r["plain_unsplit_ms"], "bit_identical": r["bit_identical"], "steps": r["steps"],
            "warmup": r["warmup"], "tiles": r["config"]["tiles"], "workgroups": r["config"]["workgroups"],
            "target": ">= 1.5x per-volume launches, identical outputs"}


def grad_summary(r):
    """The differentiable-workload object carried on the headline line."""
    return {"workload": "nlm3d_grad_8x128", "ratio": r["value"], "fwd_ms": r["fwd_ms"],
            "fwd_bwd_ms": r["fwd_bwd_ms"], "steps": r["steps"], "warmup": r["warmup"], "fit_h": r["fit_h"],
            "target": "fwd+bwd <= 3x fwd"}


def ref_default_summary(steps=3, warmup=1):
    """The reference's own default configuration (3DNLM/config.h:4-9: 512 x
    512 x 128, SS 13, NS 17, PARAMETER 10) on its data type, 8-bit samples
    resident on the device, through gnlm3d_forward_u8 -- the object the
    headline line carries for it (kernel time by HIP events, no CPU work;
    parity at this size is a GPU test)."""
    import torch
    import gnlm
    wl = WORKLOADS["nlm3d_512x512x128_p17_s13_u8"]
    nx, ny, nz = wl["dims"]
    rp, rs, h = wl["rp"], wl["rs"], wl["h"]
    dev = torch.device("cuda:0")
    u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=wl["n_ell"], axis_lo=wl["ax"][0], axis_hi=wl["ax"][1],
                          background=0.05, sigma=wl["sigma"], seed=wl["seed"])
    u8 = torch.from_numpy(np.clip(np.rint(u * 255.0), 0, 255).astype(np.uint8)).to(dev)
    out = torch.empty(u8.shape, dtype=torch.float32, device=dev)
    for _ in range(warmup):
        gnlm.nlm3d_u8(u8, rp, rs, h, out=out)
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        gnlm.nlm3d_u8(u8, rp, rs, h, out=out)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nvox = nx * ny * nz
    tflops = nvox * (2 * rs + 1) ** 3 * FLOP_PER_PAIR_3D / (ms * 1e-3) / 1e12
    integer = gnlm.forward_u8_route(nx, ny, nz, rp, rs)[1] == gnlm.KERNEL_U8_INTEGER
    return {"workload": "nlm3d_512x512x128_p17_s13_u8", "ms": round(ms, 3),
            "Mvoxels_s": round(nvox / ms / 1e3, 1),
            "kernel": "nlm3d_fwd_u8 (integer patch distances)" if integer else "FP32 kernels on a float copy",
            "achieved_TFLOPs": round(tflops, 3), "frac": round(tflops / PEAK_FP32_TFLOPS, 4),
            "steps": steps, "warmup": warmup, "traffic": _traffic_from_profiles("nlm3d_512x512x128_p17_s13_u8")}


def sweep_summary(r):
    """The run.sh sweep object carried on the headline line."""
    gp = [p["Gpairs_s"] for p in r["points"]]
    return {"workload": "nlm3d_sweep_512", "points": len(r["points"]), "total_ms": r["ms_per_step"],
            "Mvoxels_s": r["value"], "Gpairs_s_min": min(gp), "Gpairs_s_max": max(gp),
            "steps": r["steps"], "warmup": r["warmup"]}


def bench_batch(args):
    """64 volumes of uneven size (dims from {32, 48, ..., 192} per axis),
    ellipsoid phantoms with their own sigma ~ U[0.02, 0.10] and h = sigma,
    r_p in {1, 2}, r_s in {3, 5}: one graph-replayed persistent launch vs 64
    per-volume launches; outputs must be bit-identical.  The per-volume
    launches are timed twice: as the library runs them (small volumes take
    the dz split, gnlm_set_dz_split) and with the split off (one workgroup per
    tile, as the batched launch runs every tile)."""
    import torch
    import gnlm
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(20261208)
    dims_set = np.arange(32, 193, 16)
    vols, outs_a, outs_b, outs_c, pairs = [], [], [], [], 0
    for i in range(64):
        nx, ny, nz = (int(v) for v in rng.choice(dims_set, 3))
        sigma = float(rng.uniform(0.02, 0.10))
        rp, rs = int(rng.choice([1, 2])), int(rng.choice([3, 5]))
        u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=12, axis_lo=4, axis_hi=20, background=0.05, sigma=sigma,
                              seed=20261208 * 100 + i)
        t = torch.from_numpy(u).to(dev)
        vols.append((t, rp, rs, sigma))
        outs_a.append(torch.empty_like(t))
        outs_b.append(torch.empty_like(t))
        outs_c.append(torch.empty_like(t))
        pairs += nx * ny * nz * (2 * rs + 1) ** 3
    plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs_b)])
    ntiles, nblocks = plan.info()

    def plain():
        for (u, rp, rs, h), o in zip(vols, outs_a):
            gnlm.nlm3d(u, rp, rs, h, out=o)

    def plain_unsplit():
        for (u, rp, rs, h), o in zip(vols, outs_c):
            gnlm.nlm3d(u, rp, rs, h, out=o)

    def batched():
        plan.run()

    res = {}
    try:
        for name, fn in (("plain", plain), ("batched", batched), ("plain_unsplit", plain_unsplit)):
            if name == "plain_unsplit":
                # re-tune the launch shapes without the split
                gnlm.set_dz_split(False)
                gnlm.reset_tuning()
            for _ in range(max(args.warmup, 1)):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn()
            torch.cuda.synchronize()
            res[name] = (time.perf_counter() - t0) / args.steps
    finally:
        gnlm.set_dz_split(True)
        gnlm.reset_tuning()
    identical = all(torch.equal(a, b) and torch.equal(a, c) for a, b, c in zip(outs_a, outs_b, outs_c))
    nvox = sum(u.numel() for u, *_ in vols)
    return {
        "metric": "batched persistent speedup vs 64 per-volume launches",
        "value": round(res["plain"] / res["batched"], 4),
        "unit": "x",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(res["batched"] * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "nlm3d_batch64", "volumes": 64, "voxels": nvox, "pairs": pairs, "tiles": ntiles,
                   "workgroups": nblocks},
        "plain_ms": round(res["plain"] * 1e3, 4), "batched_ms": round(res["batched"] * 1e3, 4),
        "plain_unsplit_ms": round(res["plain_unsplit"] * 1e3, 4),
        "ratio_vs_unsplit": round(res["plain_unsplit"] / res["batched"], 4),
        "batched_Mvox_s": round(nvox / res["batched"] / 1e6, 2), "bit_identical": identical,
    }


def bench_grad(args):
    """8 volumes of 128^3 (ellipsoid phantoms, sigma 0.05), patch 3^3, search
    9^3, per-volume learnable h starting at 0.1: forward vs forward+adjoint of
    the MSE loss against the clean phantoms, then 50 gradient steps on h."""
    import torch
    import gnlm
    dev = torch.device("cuda:0")
    rp, rs = 1, 4
    us, cs = [], []
    for i in range(8):
        u, c = gnlm.phantom3d(128, 128, 128, n_ellipsoids=24, axis_lo=5, axis_hi=30, background=0.05, sigma=0.05,
                              seed=20261207 + i)
        us.append(torch.from_numpy(u).to(dev))
        cs.append(torch.from_numpy(c).to(dev))
    n = us[0].numel()

    def fwd(h):
        return [gnlm.nlm3d(u, rp, rs, float(hv), want_Z=True) for u, hv in zip(us, h)]

    def fwd_bwd(h):
        gh = []
        loss = 0.0
        for (u, c, hv) in zip(us, cs, h):
            y, Z = gnlm.nlm3d(u, rp, rs, float(hv), want_Z=True)
            r = y - c
            g = (2.0 / n) * r
            _, dh = gnlm.nlm3d_backward(u, y, Z, g, rp, rs, float(hv))
            gh.append(dh)
            loss = loss + (r * r).mean()
        return loss, gh

    h0 = [0.1] * 8
    times = {}
    for name, fn in (("fwd", lambda: fwd(h0)), ("fwd_bwd", lambda: fwd_bwd(h0))):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / args.steps
    # 50 Adam steps on log h per volume
    logh = torch.full((8,), float(np.log(0.1)), device=dev)
    m = torch.zeros(8, device=dev)
    v = torch.zeros(8, device=dev)
    lr, b1, b2 = 0.05, 0.9, 0.999
    traj = []
    t0 = time.perf_counter()
    for step in range(1, 51):
        h = torch.exp(logh)
        loss, gh = fwd_bwd(h.tolist())
        g = torch.stack(gh) * h  # d/dlogh
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        logh = logh - lr * (m / (1 - b1 ** step)) / (torch.sqrt(v / (1 - b2 ** step)) + 1e-12)
        if step in (1, 10, 25, 50):
            traj.append({"step": step, "loss": float(loss), "h_mean": float(h.mean())})
    fit_s = time.perf_counter() - t0
    return {
        "metric": "fwd+bwd / fwd time ratio, 8x128^3 p3 s9",
        "value": round(times["fwd_bwd"] / times["fwd"], 4),
        "unit": "x",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(times["fwd_bwd"] * 1e3, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "nlm3d_grad_8x128", "volumes": 8, "dims": [128, 128, 128], "r_p": rp, "r_s": rs,
                   "h0": 0.1},
        "fwd_ms": round(times["fwd"] * 1e3, 4), "fwd_bwd_ms": round(times["fwd_bwd"] * 1e3, 4),
        "fit_h": {"steps": 50, "seconds": round(fit_s, 3), "trajectory": traj,
                  "h_final": [round(float(x), 5) for x in torch.exp(logh).tolist()]},
    }


def bench_sweep(args):
    """The reference's parameter sweep (3DNLM/run.sh:25-34): a 512 x 512 8-bit
    image stacked `depth` times (3DNLM/nlm.cpp:53-59), SS in {11, 13}, NS in
    {13, 15, 17}, depth in {16, 32, 64, 128}, h = PARAMETER = 10 on 0-255.
    The image is synthetic (a 512^2 phantom quantised to 8 bits) and stays
    8-bit on the device, as the reference's does (gnlm3d_forward_u8).  Each
    point is timed with HIP events over `steps` forwards after `warmup`."""
    import torch
    import gnlm
    dev = torch.device("cuda:0")
    img, _ = gnlm.phantom2d(512, 512, sigma=0.05, seed=20261209)
    img = np.clip(np.rint(img * 255.0), 0, 255).astype(np.uint8)
    steps, warmup = max(1, min(args.steps, 3)), max(0, min(args.warmup, 1))
    points = []
    t_all = time.perf_counter()
    for depth in (16, 32, 64, 128):
        vol = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(img, (depth, 512, 512)))).to(dev)
        out = torch.empty(vol.shape, dtype=torch.float32, device=dev)
        for ss in (11, 13):
            for ns in (13, 15, 17):
                rs, rp = ss // 2, ns // 2
                for _ in range(warmup):
                    gnlm.nlm3d_u8(vol, rp, rs, 10.0, out=out)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                torch.cuda.synchronize()
                e0.record()
                for _ in range(steps):
                    gnlm.nlm3d_u8(vol, rp, rs, 10.0, out=out)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / steps
                nvox = 512 * 512 * depth
                integer = gnlm.forward_u8_route(512, 512, depth, rp, rs)[1] == gnlm.KERNEL_U8_INTEGER
                points.append({"SS": ss, "NS": ns, "depth": depth,
                               "kernel": "8-bit integer distances" if integer else
                               f"FP32 variant {gnlm.tuned_variant(rp, rs)} on a float copy",
                               "ms": round(ms, 3), "Mvoxels_s": round(nvox / ms / 1e3, 1),
                               "Gpairs_s": round(nvox * ss ** 3 / ms / 1e6, 1)})
                print(f"sweep SS={ss} NS={ns} depth={depth}: {ms:.2f} ms", file=sys.stderr, flush=True)
    total = sum(p["ms"] for p in points)
    return {
        "metric": "Mvoxels/s 3D NLM 512^2 x depth, SS 11-13, NS 13-17 (run.sh sweep), all points",
        "value": round(sum(512 * 512 * p["depth"] for p in points) / total / 1e3, 1),
        "unit": "Mvoxels/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(total, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "nlm3d_sweep_512", "h": 10.0, "input": "512^2 8-bit phantom stacked along z"},
        "points": points, "wall_s": round(time.perf_counter() - t_all, 2),
    }


def pmc_pass(args):
    """One launch of the workload's kernels, then a device copy of a known
    size (64 Mi floats = 256 MiB read, 256 MiB written) whose FETCH_SIZE and
    WRITE_SIZE readings calibrate the byte counters.  Prints one line."""
    import torch
    import gnlm
    dev = torch.device("cuda:0")
    if args.workload == "nlm3d_grad_8x128":
        u, c = gnlm.phantom3d(128, 128, 128, n_ellipsoids=24, axis_lo=5, axis_hi=30, background=0.05, sigma=0.05,
                              seed=20261207)
        ut, ct = torch.from_numpy(u).to(dev), torch.from_numpy(c).to(dev)
        y, Z = gnlm.nlm3d(ut, 1, 4, 0.1, want_Z=True)
        gnlm.nlm3d_backward(ut, y, Z, (2.0 / ut.numel()) * (y - ct), 1, 4, 0.1)
    elif args.workload == "nlm3d_batch64":
        args.steps, args.warmup = 1, 0
        bench_batch(args)
    else:
        wl = WORKLOADS[args.workload]
        if "dims" in wl:
            nx, ny, nz = wl["dims"]
            u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=wl["n_ell"], axis_lo=wl["ax"][0], axis_hi=wl["ax"][1],
                                  background=0.05, sigma=wl["sigma"], seed=wl["seed"])
            if wl.get("u8"):
                u8 = np.clip(np.rint(u * 255.0), 0, 255).astype(np.uint8)
                gnlm.nlm3d_u8(torch.from_numpy(u8).to(dev), wl["rp"], wl["rs"], wl["h"])
            else:
                gnlm.nlm3d(torch.from_numpy(u).to(dev), wl["rp"], wl["rs"], wl["h"])
        else:
            u, _ = gnlm.phantom2d(*wl["dims2"], sigma=wl["sigma"], seed=wl["seed"])
            if wl.get("u8"):
                u8 = np.clip(np.rint(u * 255.0), 0, 255).astype(np.uint8)
                gnlm.nlm2d_u8(torch.from_numpy(u8).to(dev), wl["rp"], wl["rs"], wl["h"])
            else:
                gnlm.nlm2d(torch.from_numpy(u).to(dev), wl["rp"], wl["rs"], wl["h"])
    a = torch.empty(64 * 1024 * 1024, device=dev)
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"pmc pass done: {args.workload} + 256 MiB copy", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="nlm3d_256_p5_s11",
                    choices=sorted(WORKLOADS) + ["nlm3d_batch64", "nlm3d_grad_8x128", "nlm3d_sweep_512"])
    ap.add_argument("--cpu-planes", type=int, default=0,
                    help="z-planes in the CPU baseline sample (0: as many as about 15 s of CPU work on 16 cores "
                         "allows -- all 256 planes of the headline volume)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: skip the batched, differentiable, reference-default and sweep "
                         "workloads the default run appends to the line (for profiler runs)")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="one launch of the workload plus a 256 MiB calibration copy (for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the volume path would shard into z-slabs across GPUs; only one GPU's share "
              "is implemented and measured)", file=sys.stderr)
        sys.exit(2)

    import torch
    import gnlm

    if args.pmc_pass:
        pmc_pass(args)
        return
    if args.workload == "nlm3d_batch64":
        print(json.dumps(bench_batch(args)))
        return
    if args.workload == "nlm3d_grad_8x128":
        print(json.dumps(bench_grad(args)))
        return
    if args.workload == "nlm3d_sweep_512":
        print(json.dumps(bench_sweep(args)))
        return
    wl = WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    rp, rs, h = wl["rp"], wl["rs"], wl["h"]
    if "dims" in wl:
        nx, ny, nz = wl["dims"]
        u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=wl["n_ell"], axis_lo=wl["ax"][0], axis_hi=wl["ax"][1],
                              background=0.05, sigma=wl["sigma"], seed=wl["seed"])
        pairs = (2 * rs + 1) ** 3
        fpp = FLOP_PER_PAIR_3D
        run = lambda t, o: gnlm.nlm3d(t, rp, rs, h, out=o)  # noqa: E731
        unit_vox = "Mvoxels/s"
    else:
        nx, ny = wl["dims2"]
        nz = 1
        u, _ = gnlm.phantom2d(nx, ny, sigma=wl["sigma"], seed=wl["seed"])
        pairs = (2 * rs + 1) ** 2
        fpp = FLOP_PER_PAIR_2D
        run = lambda t, o: gnlm.nlm2d(t, rp, rs, h)  # noqa: E731
        unit_vox = "Mpixels/s"
    nvox = nx * ny * nz
    if wl.get("u8"):
        # 8-bit input resident in HBM, through the 8-bit entry points
        # (gnlm3d_forward_u8: integer patch distances; gnlm2d_forward_u8: the
        # tiled kernel reading the bytes, exact sliding-window patch sums)
        u = np.clip(np.rint(u * 255.0), 0, 255).astype(np.float32)
        u8 = torch.from_numpy(u.astype(np.uint8)).to(dev)

        def run(t, o):
            if nz > 1:
                return gnlm.nlm3d_u8(u8, rp, rs, h, out=o)
            return gnlm.nlm2d_u8(u8, rp, rs, h, out=o)
    ud = torch.from_numpy(u).to(dev)
    out = torch.empty_like(ud)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        y = run(ud, out)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        y = run(ud, out)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_kernel = e0.elapsed_time(e1) / args.steps
    ms_step = wall * 1e3 / args.steps
    value = nvox * args.steps / wall / 1e6

    flop = float(nvox) * pairs * fpp
    achieved = flop / (ms_kernel * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                "kernel_ms": round(ms_kernel, 4), "flop_per_launch": flop}
    if nz > 1 and wl.get("u8") and gnlm.forward_u8_route(nx, ny, nz, rp, rs)[1] == gnlm.KERNEL_U8_INTEGER:
        roofline["kernel"] = "nlm3d_fwd_u8 (8-bit samples, integer patch distances)"
        roofline["traffic"] = _traffic_from_profiles(args.workload)
    elif nz > 1:
        # the tile variant and the launch shape the first call's tuning chose
        shape, nshapes = gnlm.forward_shape((nz, ny, nx), rp, rs)
        roofline["variant"] = gnlm.tuned_variant(rp, rs)
        roofline["launch_shape"] = {"chosen": shape, "candidates": nshapes}
        roofline["traffic"] = _traffic_from_profiles(args.workload, shape)
    else:
        if wl.get("u8"):
            roofline["kernel"] = "nlm2d_fwd_tiled, 8-bit build (bytes staged in LDS, sliding-window patch sums)"
        roofline["traffic"] = _traffic_from_profiles(args.workload)

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle
        from nlm_metrics import fp32_units, phantom_parity, underflow_bound
        yc = y.cpu().numpy()
        threads = _threads()
        model, phys = _cpu_info()
        if nz > 1:
            (z0, nplanes, y0, wy, x0, wx), (cz0, cz1, cy0, cy1, cx0, cx1) = cpu_sample(nx, ny, nz, rp, rs,
                                                                                      args.cpu_planes)
            R = rp + rs
            crop = np.ascontiguousarray(u[cz0:cz1, cy0:cy1, cx0:cx1])
            win = (slice(None), slice(y0 - cy0, y0 - cy0 + wy), slice(x0 - cx0, x0 - cx0 + wx))
            tcpu, tall, full32 = _median_time(lambda: oracle.nlm3d_separable_f32(
                crop, rp, rs, h, z0=z0 - cz0, z1=z0 - cz0 + nplanes, nthreads=threads))
            ref = full32[win]
            got = yc[z0:z0 + nplanes, y0:y0 + wy, x0:x0 + wx]
            svox = nplanes * wx * wy
            sample = (f"z-planes [{z0},{z0 + nplanes}) of the same volume ({svox} voxels)" if (wx, wy) == (nx, ny)
                      else f"the {wx} x {wy} window at ({x0}, {y0}) of z-plane {z0} ({svox} voxels; the CPU "
                      f"path runs on the window grown by {R})")
        else:
            tcpu, tall, ref = _median_time(lambda: oracle.nlm2d_separable_f32(u, rp, rs, h, nthreads=threads))
            got = yc
            sample = "the whole image"
            svox = nvox
        cpu = {"value": round(svox / tcpu / 1e6, 3), "unit": unit_vox, "cores": threads, "kind": "port",
               "sample": sample, "seconds": round(tcpu, 3), "runs_s": [round(t, 3) for t in tall],
               "statistic": "median of 3 timed runs",
               "cpu_model": model, "physical_cores": phys,
               "path": "the oracle's separable FP32 path (OpenMP), the CPU restatement of the reference's math"}
        # the brute-force definition on a bounded sample of the 96x96x48
        # configuration (patch 3^3, search 7^3), reported for that config only
        if nz > 1:
            ub, _ = gnlm.phantom3d(96, 96, 48, n_ellipsoids=12, axis_lo=4.0, axis_hi=20.0, background=0.05,
                                   sigma=0.05, seed=20261204)
            zb0, zb1 = 18, 30
            crop_b = np.ascontiguousarray(ub[zb0 - 4:zb1 + 4])
            tb, _, _ = _median_time(lambda: oracle.nlm3d_brute(crop_b, 1, 3, 0.05), runs=1)
            cpu["brute_force_config1"] = {
                "value": round(96 * 96 * (zb1 - zb0 + 8) / tb / 1e6, 4), "unit": "Mvoxels/s",
                "sample": f"planes [{zb0 - 4},{zb1 + 4}) of the 96x96x48 phantom (r_p 1, r_s 3, h 0.05), "
                          "the FP32 brute-force definition on OpenMP threads", "seconds": round(tb, 3)}
        d = np.abs(got.astype(np.float64) - ref)
        floor = 1e-3 * float(ref.max() - ref.min())
        parity = {"max_abs_vs_cpu_fp32": float(d.max()),
                  "max_rel_vs_cpu_fp32": float((d / np.maximum(np.abs(ref), 1e-6)).max()),
                  "max_rel_floor_1e-3_range": float((d / np.maximum(np.abs(ref), floor)).max())}
        # the FP64 re-run of the same definition, with the normaliser Z and the
        # magnitude sum A: 3D, the separable FP64 path on (at most) 16 planes
        # of the sample; 2D, the separable FP64 path on the whole image
        if nz > 1:
            n64 = min(16, nplanes)
            ref64, Z64, A64, B64 = oracle.nlm3d_separable_f64_za(crop, rp, rs, h, z0=z0 - cz0, z1=z0 - cz0 + n64,
                                                                 nthreads=threads, with_b=True)
            ref64, Z64, A64, B64 = ref64[win], Z64[win], A64[win], B64[win]
            got64, ref32_64 = got[:n64], ref[:n64]
            what = (f"{n64} planes from z = {z0}" if (wx, wy) == (nx, ny)
                    else f"the same {wx} x {wy} window of z-plane {z0}")
        else:
            ref64, Z64, A64, B64 = oracle.nlm2d_separable_f64(u, rp, rs, h, nthreads=threads, with_b=True)
            got64, ref32_64 = yc, ref
            what = "the whole image"
        d64 = np.abs(got64.astype(np.float64) - ref64)
        parity["max_rel_vs_fp64"] = float((d64 / np.maximum(np.abs(ref64), 1e-6)).max())
        parity["max_abs_vs_fp64"] = float(d64.max())
        parity["fp64_sample"] = what
        # FP32 weights below 2^-126 may flush to 0 (8-bit images with flat
        # black regions, where y is ~1e-58 in FP64)
        uf = underflow_bound(pairs, float(np.abs(u).max()), ref64, Z64)
        rep = phantom_parity(got64, ref32_64, ref64, Z64, A64, underflow=uf, B64=B64)
        parity["bar"] = ("unfloored rel <= 1e-5 vs CPU FP32 and FP64 where |y| >= 0.5% of the range; "
                         "|gpu - fp64| <= 32 * 2^-24 * (A + B)/Z + pairs * 2^-126 * (max|u| + |y|) / Z everywhere, "
                         "A = sum w |u|, B = sum w a (|u| + |y|), w = exp(-a)")
        parity.update({k: rep[k] for k in ("tau", "voxels_below_tau", "max_rel_vs_fp32_above_tau",
                                           "max_rel_vs_fp64_above_tau", "max_err_in_fp32_cond_units",
                                           "voxels_unfloored_rel_over_tol", "ok")})
        dc = np.abs(ref32_64.astype(np.float64) - ref64)
        parity["cpu_fp32_max_rel_vs_fp64"] = float((dc / np.maximum(np.abs(ref64), 1e-6)).max())
        parity["cpu_fp32_max_abs_vs_fp64"] = float(dc.max())
        parity["cpu_fp32_err_in_fp32_cond_units"] = float(fp32_units(dc, A64, Z64, uf, B64).max())
        # the A-only measure (no weight-error term), GPU and CPU FP32
        parity["err_in_fp32_units_of_A_over_Z"] = float(fp32_units(d64, A64, Z64, uf).max())
        parity["cpu_fp32_err_in_fp32_units_of_A_over_Z"] = float(fp32_units(dc, A64, Z64, uf).max())

    line = {
        "metric": ((f"Mvoxels/s 3D NLM {nx}^3" if nx == ny == nz else f"Mvoxels/s 3D NLM {nx}x{ny}x{nz}") +
                   f" p{2 * rp + 1} s{2 * rs + 1} FP32" if nz > 1 else
                   (f"Mpixels/s 2D NLM {nx}^2 p{2 * rp + 1} s{2 * rs + 1}" + (" uint8" if wl.get("u8") else ""))),
        "value": round(value, 3),
        "unit": unit_vox,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / wl["published"], 4) if wl.get("published") else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "dims": [nx, ny, nz], "r_p": rp, "r_s": rs, "h": h,
                   "seed": wl["seed"], "input_hash": gnlm.hash_f32(u)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    if args.workload == "nlm3d_256_p5_s11" and not args.no_extras:
        # the two other targets, on the same line: the batched persistent
        # launch (64 uneven volumes) and fwd+bwd vs fwd (8 x 128^3)
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = 3, 1
        line["batched"] = batch_summary(bench_batch(sub))
        line["grad"] = grad_summary(bench_grad(sub))
        # the reference's own default configuration and its run.sh sweep, on
        # its 8-bit data type
        line["reference_default"] = ref_default_summary()
        sub.steps, sub.warmup = 1, 1
        line["sweep"] = sweep_summary(bench_sweep(sub))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
