"""Benchmark: TGV2-L1 deshade of a 256^3 volume (1000 iterations + DCT Poisson
bias solve + epilogue) on one MI355X, reported as voxel-iterations/s.

A step is one full deshade of the synthetic 256^3 volume with the input
already resident in HBM (the C ABI's device-memory mode).  The roofline
object is for the dominant kernel, timed with HIP events on the library's own
stream over the timed region.  The CPU baseline is the oracle (the reference
algorithm restated in C++/OpenMP, FP32 mirror order) on a bounded sample of
the same workload on the host cores.

    python bench.py [--steps K] [--warmup W] [--storage f32|bf16]
                    [--dct reference|f32|bf16x3] [--cpu-iters N] [--no-cpu-baseline]
                    [--no-split]
    python bench.py --copy-calibration     # plain 1 GiB device copies (PMC scale)

Unless --no-split, the same workload also runs in the split form (BF16 field
storage + BF16x3 MFMA DCT) and the line carries its throughput, its speed-up
over the plain form and the norm-wise relative error of its u and shading
against the plain form's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# element accesses per voxel for one launch of each sweep (f is always FP32)
#   dual    read u_bar, v_bar x3, p x3, q x6 ; write p x3, q x6              22
#   primal  read p x3, q x6, u, f, v x3 ; write u, u_bar, v x3, v_bar x3      22
#   fused   read u, f, v x3, p x3, q x6 ; write u, v x3, p x3, q x6          27
SWEEPS = ("dual_sweep", "primal_sweep", "fused_sweep")


def sweep_bytes(kind, storage_bytes, n):
    if kind == "dual_sweep":
        return 22 * storage_bytes * n
    if kind == "primal_sweep":
        return (21 * storage_bytes + 4) * n
    return (26 * storage_bytes + 4) * n


def copy_calibration(nbytes=1 << 30, reps=3):
    """Plain device-to-device copies of a known size: under rocprofv3 --pmc the
    FETCH_SIZE / WRITE_SIZE of these dispatches against nbytes give the
    counter's scale for the sweep kernels' traffic figures."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def run_pipeline(ctx, native, f_dev, shape, p, steps, warmup, out):
    """Warm-up, then `steps` timed deshades; returns (seconds, profile)."""
    u_dev, l_dev, r_dev = out

    def step():
        ctx.deshade_pipeline_device(f_dev, shape, p, u_dev, l_dev, r_dev)

    for _ in range(warmup):
        step()
        print("warmup step done", file=sys.stderr, flush=True)
    ctx.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        print(f"timed step {k + 1}/{steps} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile(False)
    return elapsed, prof


def rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--iterations", type=int, default=1000)
    ap.add_argument("--storage", choices=["f32", "bf16"], default="f32")
    ap.add_argument("--dct", choices=["reference", "f32", "bf16x3"], default="reference")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the TGV sweeps would shard into "
              "z-slabs exchanging one-plane halos, which is not built). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    from ambient_image_processor_amd import native, phantoms

    n_side = args.size
    shape = (n_side, n_side, n_side)
    f, _, _ = phantoms.deshade_phantom(shape, seed=2, kind="spheres", noise=0.03, bias_order=3,
                                       bias_sigma=0.05)
    n = f.size
    tau = float(1.0 / np.sqrt(20.0))
    storage = native.STORAGE_BF16 if args.storage == "bf16" else native.STORAGE_F32
    dct = {"reference": native.DCT_REFERENCE, "f32": native.DCT_F32,
           "bf16x3": native.DCT_BF16X3}[args.dct]
    p = native.default_params(lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau, theta=1.0,
                              iterations=args.iterations, data_term=native.DATA_L1,
                              storage=storage, dct_precision=dct,
                              flags=native.FLAG_SET_NEGATIVE_ZERO)

    ctx = native.Context(0)
    f_dev = ctx.alloc(n * 4)
    out = (ctx.alloc(n * 4), ctx.alloc(n * 4), ctx.alloc(n * 4))
    u_dev, l_dev, r_dev = out
    ctx.upload(f_dev, f)

    elapsed, prof = run_pipeline(ctx, native, f_dev, shape, p, args.steps, args.warmup, out)

    voxel_iters = float(n) * args.iterations * args.steps
    value = voxel_iters / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps

    sb = 2 if storage == native.STORAGE_BF16 else 4
    kernels = {}
    for name in SWEEPS:
        ms, cnt = prof[name]
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        b = sweep_bytes(name, sb, n)
        kernels[name] = {"avg_us": 1000.0 * avg_ms, "launches": cnt,
                         "bytes_per_launch": b, "gbs": b / (avg_ms * 1e-3) / 1e9}
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    achieved = kernels[dom]["gbs"]
    sweep_ms = sum(prof[k][0] for k in SWEEPS)
    sweep_b = sum(sweep_bytes(k, sb, n) * prof[k][1] for k in SWEEPS)
    it_ms = sweep_ms / (args.iterations * args.steps)
    iter_bytes = sweep_b / (args.iterations * args.steps)

    split = None
    if not args.no_split and storage == native.STORAGE_F32 and dct == native.DCT_REFERENCE:
        u_plain, l_plain = ctx.download(u_dev, shape), ctx.download(l_dev, shape)
        ps = native.default_params(lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau,
                                   theta=1.0, iterations=args.iterations,
                                   data_term=native.DATA_L1, storage=native.STORAGE_BF16,
                                   dct_precision=native.DCT_BF16X3,
                                   flags=native.FLAG_SET_NEGATIVE_ZERO)
        s_el, _ = run_pipeline(ctx, native, f_dev, shape, ps, args.steps, args.warmup, out)
        s_val = voxel_iters / s_el
        split = {"form": "bf16 field storage + bf16x3 MFMA DCT", "value": s_val,
                 "ms_per_step": 1000.0 * s_el / args.steps, "speedup_vs_plain": s_val / value,
                 "u_rel_err_vs_plain": rel_err(ctx.download(u_dev, shape), u_plain),
                 "shading_rel_err_vs_plain": rel_err(ctx.download(l_dev, shape), l_plain),
                 "err_norm": "max|a-b| / max|b|"}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        oracle.build()
        st = oracle.TGVState(f, np.float32)
        st.iterate(1, lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau, data_term="l1")
        t1 = time.perf_counter()
        st.iterate(args.cpu_iters, lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau,
                   data_term="l1")
        cpu_s = time.perf_counter() - t1
        cpu = {"value": float(n) * args.cpu_iters / cpu_s, "unit": "voxel-iters/s",
               "cores": oracle.num_threads(), "kind": "port",
               "sample": f"{args.cpu_iters} TGV2-L1 iterations of the same {n_side}^3 volume "
                         f"(FP32 reference op order, OpenMP), {cpu_s:.1f} s"}

    line = {
        "metric": "voxel-iters/s TGV-DCT deshade 256\u00b3 + %HBM roofline; split vs plain at 1e-4 err",
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if storage == native.STORAGE_F32 else "bf16-storage/f32-arith",
        "data": "synthetic (32 seeded spheres + DCT bias + N(0,0.03^2) noise)",
        "config": {"workload": f"tgv2_l1_deshade_{n_side}^3", "shape": list(shape),
                   "iterations": args.iterations, "lambda": 0.5, "alpha0": 2.0, "alpha1": 1.0,
                   "tau_sigma": "1/sqrt(20)", "storage": args.storage, "dct": args.dct,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": kernels[dom]["bytes_per_launch"]},
        "iteration": {"ms": it_ms, "bytes": iter_bytes,
                      "gbs": iter_bytes / (it_ms * 1e-3) / 1e9,
                      "voxel_iters_per_s": n / (it_ms * 1e-3)},
        "kernels": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in kernels.items()},
        "other_kernels_ms": {k: round(prof[k][0], 3) for k in ("dct", "poisson", "divergence",
                                                                "epilogue", "init")},
        "split": split,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.free(f_dev); ctx.free(u_dev); ctx.free(l_dev); ctx.free(r_dev)
    ctx.close()


if __name__ == "__main__":
    main()
