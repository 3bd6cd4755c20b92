"""Benchmark of the MI355X NLM path: one JSON line on stdout.

Default workload (the headline): 3D NLM of a 256^3 FP32 ellipsoid phantom
(64 ellipsoids, semi-axes 5-40 voxels, sigma 0.05), patch 5^3 (r_p = 2),
search 11^3 (r_s = 5), h = 0.05.  One step = one forward pass over the whole
volume through the C-ABI; the input is resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Roofline: the kernels are FP32-VALU bound.  Algorithmic work per
(voxel, offset) pair is 13 flop (diff, square, 3 separable running-sum
updates of 2 flop, scale, exp, and 3 for the two accumulations), so one
forward is N * (2 r_s + 1)^3 * 13 flop; achieved = that / kernel time,
peak = 157.3 TFLOP/s (MI355X FP32 vector).  HBM traffic per launch comes
from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs kept in profiles/.

cpu_baseline: the CPU FP32 path of the oracle (separable, OpenMP), timed on
this host's cores on a bounded sample of the same volume (a slab of z-planes);
the sample's outputs are also compared with the GPU's.
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
    "nlm2d_2048_p5_s11": dict(dims2=(2048, 2048), sigma=0.08, rp=2, rs=5, h=0.08, seed=20261205),
}


def _traffic_from_profiles(name):
    """HBM bytes per launch measured by rocprofv3 PMC runs, if recorded."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="nlm3d_256_p5_s11", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-planes", type=int, default=8, help="z-planes in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the volume path shards into z-slabs across GPUs by design, but only the one-GPU "
              "share is built and measured)", file=sys.stderr)
        sys.exit(2)

    import torch
    import gnlm

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
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": _traffic_from_profiles(args.workload),
                "kernel_ms": round(ms_kernel, 4), "flop_per_launch": flop}

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        yc = y.cpu().numpy()
        threads = os.cpu_count() or 1
        try:
            threads = len(os.sched_getaffinity(0))
        except AttributeError:
            pass
        if nz > 1:
            nplanes = max(1, min(args.cpu_planes, nz))
            z0 = nz // 2 - nplanes // 2
            tc = time.perf_counter()
            ref = oracle.nlm3d_separable_f32(u, rp, rs, h, z0=z0, z1=z0 + nplanes, nthreads=threads)
            tcpu = time.perf_counter() - tc
            got = yc[z0:z0 + nplanes]
            sample = f"z-planes [{z0},{z0 + nplanes}) of the same volume ({nplanes * nx * ny} voxels)"
            svox = nplanes * nx * ny
        else:
            tc = time.perf_counter()
            ref = oracle.nlm2d_separable_f32(u, rp, rs, h, nthreads=threads)
            tcpu = time.perf_counter() - tc
            got = yc
            sample = "the whole image"
            svox = nvox
        cpu = {"value": round(svox / tcpu / 1e6, 3), "unit": unit_vox, "cores": threads, "kind": "port",
               "sample": sample, "seconds": round(tcpu, 3)}
        d = np.abs(got.astype(np.float64) - ref)
        parity = {"max_abs_vs_cpu_fp32": float(d.max()),
                  "max_rel_vs_cpu_fp32": float((d / np.maximum(np.abs(ref), 1e-6)).max())}

    line = {
        "metric": "Mvoxels/s 3D NLM 256^3 p5 s11 FP32" if nz > 1 else "Mpixels/s 2D NLM",
        "value": round(value, 3),
        "unit": unit_vox,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "dims": [nx, ny, nz], "r_p": rp, "r_s": rs, "h": h,
                   "seed": wl["seed"], "input_hash": gnlm.hash_f32(u)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
