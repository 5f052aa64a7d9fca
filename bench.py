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
                    [--no-split] [--no-host]
    python bench.py --workload poisson_512      # DCT Poisson solve 512x512x256
    python bench.py --workload bf16_384         # BF16 field storage, L2, 384^3, 800 it
    python bench.py --workload adjoint_4x128    # unrolled TGV2-L2 + adjoint, 4 x 128^3
    python bench.py --workload tgvk3_256        # TGV-k (k = 3) deshade 256^3
    python bench.py --workload downsampled_512  # 512^3 resized to 256^3, masked TGV-k, back
    python bench.py --workload masked_256       # masked TGV2-L1 deshade 256^3, spherical mask
    python bench.py --workload planar_4096      # 2-D TGV2-L1 deshade of a 4096^2 image
    python bench.py --copy-calibration          # plain 1 GiB device copies (PMC scale)

Unless --no-split, the default workload also runs in two split forms -- FP32
fields with the BF16x3 MFMA DCT, and BF16 field storage with the BF16x3 DCT --
and the line carries, for each, its throughput, its speed-up over the plain
form (FP32 fields, reference-numerics DCT) and the max- and RMS-norm relative
errors of its u and shading against the plain form's.  Unless --no-energy, one
more deshade runs with the FP64 energy / gap reduced inside the fused sweep
every 100 iterations (`energy`: the final gap and brackets, the energy launch's
time against the plain fused launch).  Unless --no-host, one more deshade runs
from host buffers (PCIe transfers included) and is reported as `host_buffers`.
Unless --no-configs, the other benchmark configurations run briefly under the
same clock (`configs`): the 512x512x256 DCT Poisson solve in the three DCT
precisions (3 solves each), the 384^3 TGV2-L2 deshade with FP32 and with BF16
fields (800 iterations each), and the unrolled TGV2-L2 denoise + adjoint of
4 x 128^3.  Where tests/golden/longrun.npz (the oracle's full-size runs) is
present, the line also states whether the plain form equals the FP32 oracle bit
for bit (per-plane sums of u and the shading) and the FP32 oracle's distance to
FP64 at this configuration.
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
FP32_PEAK_TFLOPS = 157.3       # dense FP32 (vector / matrix)
BF16_PEAK_TFLOPS = 2500.0      # dense BF16 MFMA
# Algorithmic bytes (the per-unit figure of the reference algorithm, f FP32):
# one voxel-iteration is 44 element accesses (176 B FP32, 90 B BF16 fields):
#   dual half    read u_bar, v_bar x3, p x3, q x6 ; write p x3, q x6          22
#   primal half  read p x3, q x6, u, f, v x3 ; write u, u_bar, v x3, v_bar x3  22
# A dual or primal sweep launch processes half a voxel-iteration per voxel; a
# fused launch (primal half of k + dual half of k + 1) one.  The fused kernel
# itself needs only 27 accesses per voxel (u, f, v x3, p x3, q x6 read; u, v x3,
# p x3, q x6 written): u_bar and v_bar never reach HBM.
SWEEPS = ("dual_sweep", "primal_sweep", "fused_sweep")
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# the default workload (256^3, FP32 fields, default plan), FETCH_SIZE scaled by
# 2.0 and WRITE_SIZE by 1.0 (what the counters read for a plain 1 GiB copy);
# summaries in profiles/.  Keyed by (kernel, storage bytes, edge length); the
# TGV-k entries by (kernel, levels, edge length) (k = 3, FP32 fields).
MEASURED_TRAFFIC = {("fused_sweep", 4, 256): 1857559048,
                    ("dual_sweep", 4, 256): 1532122086,
                    ("primal_sweep", 4, 256): 1532759954,
                    ("tgvk_dual", 1, 256): 2908697085,
                    ("tgvk_primal", 1, 256): 3563855669,
                    ("tgvk_fused", 1, 256): 3479840059}


def tgvk_bytes(kind, levels, n):
    """Algorithmic bytes of one TGV-k launch (FP32 fields), L = k - 2 levels:
      dual    read u_bar, v_bar x3, p x3, q x6, per level w_bar x6, r x6;
              write p x3, q x6, per level r x6                    (22 + 18 L) x 4
      primal  read u, f, p x3, q x6, v x3, per level r x6, w x6;
              write u, u_bar, v x3, v_bar x3, per level w x6, w_bar x6  (22 + 24 L) x 4
      fused   (primal half of one iteration + dual half of the next; u_bar,
              v_bar, w_bar stay on chip) read u, f, v x3, p x3, q x6, per level
              w x6, r x6; write u, v x3, p x3, q x6, per level w x6, r x6
                                                                  (27 + 24 L) x 4"""
    per = {"tgvk_dual": 22 + 18 * levels, "tgvk_primal": 22 + 24 * levels,
           "tgvk_fused": 27 + 24 * levels}[kind]
    return 4 * per * n


def tgvk_kernels(prof, levels, n):
    """Per-kernel averages of the TGV-k launches that ran, and the one with the
    most device time.  For the fused sweep `two_launch_equivalent_gbs` is the
    two launches' bytes over its time (one fused launch does one iteration's
    work)."""
    kernels = {}
    for name in ("tgvk_dual", "tgvk_primal", "tgvk_fused"):
        ms, cnt = prof[name]
        if cnt == 0:
            continue
        b = tgvk_bytes(name, levels, n)
        kernels[name] = {"avg_us": 1000.0 * ms / cnt, "launches": cnt, "bytes_per_launch": b,
                         "gbs": b / (ms / cnt * 1e-3) / 1e9}
        if name == "tgvk_fused":
            b2 = tgvk_bytes("tgvk_dual", levels, n) + tgvk_bytes("tgvk_primal", levels, n)
            kernels[name]["two_launch_equivalent_gbs"] = b2 / (ms / cnt * 1e-3) / 1e9
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    return kernels, dom


def sweep_bytes(kind, storage_bytes, n):
    """Algorithmic bytes per launch: every field the kernel's own formulation
    reads or writes, moved once (f is FP32):
      dual   read u_bar, v_bar x3, p x3, q x6; write p x3, q x6         22 fields
      primal read u, f, p x3, q x6, v x3; write u, u_bar, v x3, v_bar x3 21 + f
      fused  (primal half of one iteration + dual half of the next; u_bar and
             v_bar stay on chip) read u, f, v x3, p x3, q x6; write u, v x3,
             p x3, q x6                                                26 + f
    (108 B per voxel-iteration fused at FP32; 176 B for the two sweeps.)  The
    fused 2-D sweep (depth 1) takes v_z, p_z, q_zz, q_xz, q_yz as 0 without
    moving them: reads u, f, v x2, p x2, q x3, writes u, v x2, p x2, q x3 =
    16 fields + f (68 B per pixel-iteration at FP32)."""
    if kind == "fused_sweep_2d":
        return (16 * storage_bytes + 4) * n
    if kind == "dual_sweep":
        return 22 * storage_bytes * n
    if kind == "primal_sweep":
        return (21 * storage_bytes + 4) * n
    return (26 * storage_bytes + 4) * n


def two_sweep_bytes(storage_bytes, n):
    """Bytes of one iteration in the two-sweep form (dual + primal): the 176 B
    per voxel-iteration at FP32 that the fused sweep does in 108 B."""
    return sweep_bytes("dual_sweep", storage_bytes, n) + sweep_bytes("primal_sweep", storage_bytes, n)


PLAN_OVERRIDES = []   # (name, value) from --plan, set on every context of the run


def new_context():
    """A library context on GPU 0 with the --plan launch-plan overrides (tuning;
    results do not depend on them)."""
    from ambient_image_processor_amd import native
    ctx = native.Context(0)
    for name, value in PLAN_OVERRIDES:
        ctx.set_override(name, value)
    return ctx


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def rel_err(a, b):
    """max |a - b| / max |b|"""
    return float(np.max(np.abs(a.astype(np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


def rms_err(a, b):
    """||a - b||_2 / ||b||_2"""
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def forward_difference(a, axis):
    """D+ along axis 0 = x (fastest), 1 = y, 2 = z: a[i+1] - a[i], 0 at the last index."""
    ax = 2 - axis
    out = np.zeros_like(a)
    lo = [slice(None)] * 3
    hi = [slice(None)] * 3
    lo[ax], hi[ax] = slice(0, -1), slice(1, None)
    out[tuple(lo)] = a[tuple(hi)] - a[tuple(lo)]
    return out


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


PROFILE_EVERY = 8   # HIP events around one launch in 8 of each kernel class


def timed(ctx, step, steps, warmup, every=PROFILE_EVERY):
    """Warm-up, then `steps` timed calls of step(); returns (seconds, profile).
    Inside the timed region HIP events on the library's stream time one launch
    in `every` of each kernel class (the events' own cost, about 2% of a step
    when every launch is bracketed, stays out of the others); profile[class] =
    (average of the timed launches x all launches in ms, all launches);
    profile["fused_tiles"] = (tiles swept, tiles in the grids) of the 3-D fused
    sweep (a masked run skips the tiles that own no active voxel)."""
    for _ in range(warmup):
        step()
        print("warmup step done", file=sys.stderr, flush=True)
    ctx.synchronize()
    ctx.profile(True, every=every)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for k in range(steps):
        step()
        print(f"timed step {k + 1}/{steps} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    measured = ctx.profile_read()
    counts = ctx.profile_counts()
    tiles = ctx.profile_fused_tiles()
    ctx.profile(False)
    prof = {k: ((ms / n * counts[k]) if n else 0.0, counts[k]) for k, (ms, n) in measured.items()}
    prof["fused_tiles"] = tiles
    return elapsed, prof


def sweep_kernels(prof, sb, n, planar=False):
    kernels = {}
    for name in SWEEPS:
        ms, cnt = prof[name]
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        b = sweep_bytes("fused_sweep_2d" if planar and name == "fused_sweep" else name, sb, n)
        kernels[name] = {"avg_us": 1000.0 * avg_ms, "launches": cnt,
                         "bytes_per_launch": b, "gbs": b / (avg_ms * 1e-3) / 1e9}
        if name == "fused_sweep":   # one fused launch does one iteration's work
            kernels[name]["two_sweep_equivalent_gbs"] = (two_sweep_bytes(sb, n) /
                                                         (avg_ms * 1e-3) / 1e9)
    return kernels


def rounded(d):
    return {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()}
            for k, v in d.items()}


def deshade_256(args):
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

    def params(storage_, dct_):
        return native.default_params(lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau,
                                     theta=1.0, iterations=args.iterations,
                                     data_term=native.DATA_L1, storage=storage_,
                                     dct_precision=dct_, flags=native.FLAG_SET_NEGATIVE_ZERO)

    p = params(storage, dct)
    ctx = new_context()
    f_dev = ctx.alloc(n * 4)
    u_dev, l_dev, r_dev = ctx.alloc(n * 4), ctx.alloc(n * 4), ctx.alloc(n * 4)
    ctx.upload(f_dev, f)

    def step_with(pp):
        return lambda: ctx.deshade_pipeline_device(f_dev, shape, pp, u_dev, l_dev, r_dev)

    elapsed, prof = timed(ctx, step_with(p), args.steps, args.warmup)
    voxel_iters = float(n) * args.iterations * args.steps
    value = voxel_iters / elapsed
    ms_per_step = 1000.0 * elapsed / args.steps

    sb = 2 if storage == native.STORAGE_BF16 else 4
    kernels = sweep_kernels(prof, sb, n)
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    achieved = kernels[dom]["gbs"]
    sweep_ms = sum(prof[k][0] for k in SWEEPS)
    sweep_b = sum(sweep_bytes(k, sb, n) * prof[k][1] for k in SWEEPS)
    it_ms = sweep_ms / (args.iterations * args.steps)
    iter_bytes = sweep_b / (args.iterations * args.steps)
    it_rate = n / (it_ms * 1e-3)

    split = None
    if not args.no_split and storage == native.STORAGE_F32 and dct == native.DCT_REFERENCE:
        u_plain, l_plain = ctx.download(u_dev, shape), ctx.download(l_dev, shape)
        split = {"err_norms": "rel: max|a-b|/max|b|; rms: ||a-b||_2/||b||_2 (vs the plain form)"}
        for key, form, st_, dct_ in (
                ("dct", "fp32 fields + bf16x3 MFMA DCT", native.STORAGE_F32, native.DCT_BF16X3),
                ("storage", "bf16 fields + bf16x3 MFMA DCT", native.STORAGE_BF16,
                 native.DCT_BF16X3)):
            s_el, s_prof = timed(ctx, step_with(params(st_, dct_)), args.steps, args.warmup)
            s_val = voxel_iters / s_el
            u_s, l_s = ctx.download(u_dev, shape), ctx.download(l_dev, shape)
            s_kern = sweep_kernels(s_prof, 2 if st_ == native.STORAGE_BF16 else 4, n)
            split[key] = {"form": form, "value": s_val, "ms_per_step": 1000.0 * s_el / args.steps,
                          "speedup_vs_plain": s_val / value,
                          "u_rel_err": rel_err(u_s, u_plain), "u_rms_err": rms_err(u_s, u_plain),
                          "shading_rel_err": rel_err(l_s, l_plain),
                          "shading_rms_err": rms_err(l_s, l_plain),
                          "fused_sweep": rounded({"k": s_kern["fused_sweep"]})["k"]
                          if "fused_sweep" in s_kern else None}
            if "fused_sweep" in s_kern:
                split[key]["fused_sweep"]["hbm_frac"] = round(
                    s_kern["fused_sweep"]["gbs"] / HBM_PEAK_GBS, 4)

    # in-sweep FP64 energy / gap every 100 iterations (the fused sweep's energy
    # form) on one more deshade: its launch time against the plain fused sweep
    energy = None
    if not args.no_energy and storage == native.STORAGE_F32 and dct == native.DCT_REFERENCE:
        every = 100
        ctx.set_energy_interval(every)
        e_el, e_prof = timed(ctx, step_with(p), 1, 0, every=1)
        trace = ctx.energy_trace()
        ctx.set_energy_interval(0)
        ms_e, n_e = e_prof["fused_sweep_energy"]
        ms_f, n_f = e_prof["fused_sweep"]
        fused_us = 1000.0 * ms_f / max(n_f, 1)
        en_us = 1000.0 * ms_e / max(n_e, 1)
        energy = {"every": every, "launches": n_e, "energy_launch_us": en_us,
                  "fused_launch_us": fused_us,
                  "energy_launch_over_fused": en_us / fused_us if n_e else None,
                  "loop_overhead": (ms_e + ms_f) / (fused_us * 1e-3 * (n_e + n_f)) - 1.0
                  if n_e else None,
                  "ms_per_step": 1000.0 * e_el,
                  "final": {"iteration": int(trace[-1, 0]), "E": trace[-1, 1],
                            "gap": trace[-1, 5], "brackets": [trace[-1, 6], trace[-1, 7],
                                                               trace[-1, 8]],
                            "l1_infeasibility": trace[-1, 9], "v_residual": trace[-1, 10]},
                  "gap_trace": [[int(r[0]), r[5]] for r in trace]}
        print("energy trace done", file=sys.stderr, flush=True)

    host = None
    if not args.no_host:
        t0 = time.perf_counter()
        ctx.deshade_pipeline(f, p)
        h_s = time.perf_counter() - t0
        host = {"value": float(n) * args.iterations / h_s, "ms_per_step": 1000.0 * h_s,
                "note": "one deshade from host buffers: H2D of f, D2H of u, shading, deshaded"}

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
               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
               "sample": f"{args.cpu_iters} TGV2-L1 iterations of the same {n_side}^3 volume "
                         f"(FP32 reference op order, OpenMP), {cpu_s:.1f} s"}

    # the other benchmark configurations, short runs under the same clock
    configs = None
    if not args.no_configs and n_side == 256:
        import types
        sub = types.SimpleNamespace(steps=3, warmup=1, iterations=1000)
        configs = {}
        for name, fn in (("poisson_512x512x256", poisson_512), ("bf16_storage_384", bf16_384),
                         ("adjoint_4x128", adjoint_4x128)):
            t0 = time.perf_counter()
            r_ = fn(sub, emit=False)
            r_ = {k: v for k, v in r_.items()
                  if k not in ("n_gpus", "higher_is_better", "scaling", "vs_baseline")}
            r_["wall_s"] = time.perf_counter() - t0
            configs[name] = r_
            print(f"{name} done", file=sys.stderr, flush=True)

    gold = longrun_golden()
    if split is not None and gold is not None and "c2_fp64_u_max" in gold:
        split["fp32_oracle_vs_fp64"] = {
            "note": "the plain form's oracle (FP32 mirror order) against the FP64 oracle at "
                    "this configuration, offline (tests/golden/longrun.npz)",
            "u_max": float(gold["c2_fp64_u_max"]), "u_rms": float(gold["c2_fp64_u_rms"]),
            "shading_max": float(gold["c2_fp64_shading_max"]),
            "shading_rms": float(gold["c2_fp64_shading_rms"])}
        if args.iterations == 1000:
            split["plain_equals_fp32_oracle"] = bool(
                np.array_equal(plane_sums(u_plain), gold["c2_u_sums"]) and
                np.array_equal(plane_sums(l_plain), gold["c2_shading_sums"]))

    traffic = MEASURED_TRAFFIC.get((dom, sb, n_side))
    avg_s = kernels[dom]["avg_us"] * 1e-6
    line = {
        "metric": "voxel-iters/s TGV-DCT deshade 256³ + %HBM roofline; split vs plain at 1e-4 err",
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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": kernels[dom]["bytes_per_launch"],
                     "two_sweep_equivalent_gbs": kernels[dom].get("two_sweep_equivalent_gbs"),
                     "measured_hbm_gbs": (traffic / avg_s / 1e9) if traffic else None},
        "iteration": {"ms": it_ms, "bytes": iter_bytes,
                      "gbs": iter_bytes / (it_ms * 1e-3) / 1e9, "voxel_iters_per_s": it_rate},
        "kernels": rounded(kernels),
        "other_kernels_ms_per_step": {k: round(prof[k][0] / args.steps, 3)
                                      for k in ("dct", "poisson", "divergence", "epilogue", "init")},
        "split": split,
        "energy": energy,
        "host_buffers": host,
        "cpu_baseline": cpu,
        "configs": configs,
    }
    print(json.dumps(line))
    for ptr in (f_dev, u_dev, l_dev, r_dev):
        ctx.free(ptr)
    ctx.close()


def _tgv2_variant(args, f, mask, flags, workload, metric, data, cpu_note):
    """One TGV2-L1 deshade pipeline per step on the given volume (masked and/or
    2-D per `flags`): throughput, the dominant sweep's roofline, per-kernel
    times and the oracle's loop on a bounded sample as the CPU baseline."""
    from ambient_image_processor_amd import native
    shape = f.shape
    n = f.size
    tau = float(1.0 / np.sqrt(20.0))
    p = native.default_params(lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau, theta=1.0,
                              iterations=args.iterations, data_term=native.DATA_L1, flags=flags)
    ctx = new_context()
    bufs = [ctx.alloc(n * 4) for _ in range(4 if mask is None else 5)]
    f_dev, u_dev, l_dev, r_dev = bufs[:4]
    m_dev = bufs[4] if mask is not None else None
    ctx.upload(f_dev, f)
    if mask is not None:
        ctx.upload(m_dev, mask)
    elapsed, prof = timed(ctx, lambda: ctx.deshade_pipeline_device(
        f_dev, shape, p, u_dev, l_dev, r_dev, mask_dev=m_dev), args.steps, args.warmup)
    value = float(n) * args.iterations * args.steps / elapsed
    in_mask = None
    if mask is not None and (flags & native.FLAG_MASKED_ITERATION):
        act = int(np.count_nonzero(mask >= 1))
        in_mask = {"active_voxels": act, "fraction": act / n,
                   "voxel_iters_per_s": float(act) * args.iterations * args.steps / elapsed}
    kernels = sweep_kernels(prof, 4, n, planar=bool(flags & native.FLAG_PLANAR))
    swept, total = prof["fused_tiles"]
    if "fused_sweep" in kernels and total and swept < total:
        # a masked run: the fused launches move the bytes of the tiles they sweep
        # (tiles of equal size here: the plans divide 256^3 evenly)
        fs = kernels["fused_sweep"]
        fs["tiles_swept_fraction"] = swept / total
        fs["whole_volume_bytes_per_launch"] = fs["bytes_per_launch"]
        fs["bytes_per_launch"] = int(round(fs["bytes_per_launch"] * swept / total))
        fs["gbs"] = fs["bytes_per_launch"] / (fs["avg_us"] * 1e-6) / 1e9
        fs.pop("two_sweep_equivalent_gbs", None)
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        oracle.build()
        planar = bool(flags & native.FLAG_PLANAR)
        m_it = mask if (flags & native.FLAG_MASKED_ITERATION) else None
        st = oracle.TGVState(f, np.float32)
        kw = dict(lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau, data_term="l1",
                  mask=m_it, planar=planar)
        st.iterate(1, **kw)
        t1 = time.perf_counter()
        st.iterate(args.cpu_iters, **kw)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": float(n) * args.cpu_iters / cpu_s, "unit": "voxel-iters/s",
               "cores": oracle.num_threads(), "kind": "port",
               "sample": f"{args.cpu_iters} {cpu_note} (FP32 reference op order, OpenMP), "
                         f"{cpu_s:.1f} s"}
    print(json.dumps({
        "metric": metric, "value": value, "unit": "voxel-iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": data,
        "config": {"workload": workload, "shape": list(shape), "iterations": args.iterations,
                   "lambda": 0.5, "alpha0": 2.0, "alpha1": 1.0, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kernels[dom]["gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": kernels[dom]["bytes_per_launch"]},
        "in_mask": in_mask,
        "kernels": rounded(kernels),
        "other_kernels_ms_per_step": {k: round(prof[k][0] / args.steps, 3)
                                      for k in ("dct", "poisson", "divergence", "epilogue", "init")},
        "cpu_baseline": cpu}))
    for ptr in bufs:
        ctx.free(ptr)
    ctx.close()


def masked_256(args):
    """Masked TGV2-L1 deshade (tgv_deshade_masked/, TGVDeshadeMaskedProcessor) of
    the 256^3 volume with a spherical mask (mask 1 inside a ball of radius
    --mask-radius x 256, by default 0.495, about half of the voxels; 0.5
    outside): the masked iteration (updates only where mask >= 1), the Poisson
    solve, the shading times the mask, the clamp and the background added back.
    The masked sweeps also read one active byte per voxel (not counted in the
    algorithmic bytes); fused-sweep tiles that own no active voxel skip, so the
    fused sweep's bytes are those of the tiles it sweeps (the first launch of a
    call sweeps all of them), and the line also gives the rate over the in-mask
    voxels."""
    from ambient_image_processor_amd import native, phantoms
    n_side = args.size
    shape = (n_side, n_side, n_side)
    f, _, _ = phantoms.deshade_phantom(shape, seed=2, kind="spheres", noise=0.03, bias_order=3,
                                       bias_sigma=0.05)
    c = (n_side - 1) / 2.0
    z, y, x = np.ogrid[0:n_side, 0:n_side, 0:n_side]
    ball = ((z - c) ** 2 + (y - c) ** 2 + (x - c) ** 2) < (args.mask_radius * n_side) ** 2
    mask = np.where(ball, 1.0, 0.5).astype(np.float32)
    flags = (native.FLAG_MASKED_ITERATION | native.FLAG_SET_NEGATIVE_ZERO
             | native.FLAG_ADD_BACKGROUND_BACK)
    _tgv2_variant(args, f, mask, flags, f"tgv2_l1_masked_deshade_{n_side}^3",
                  f"voxel-iters/s masked TGV2-L1 deshade {n_side}^3 + %HBM roofline",
                  "synthetic (32 seeded spheres + DCT bias + N(0,0.03^2) noise, spherical mask)",
                  f"masked TGV2-L1 iterations of the same {n_side}^3 volume")


def planar_4096(args):
    """2-D TGV2-L1 deshade (tgv2_l1_deshade_2d, the reference's depth-1 loop) of
    a 4096 x 4096 image, as many pixels as the 256^3 volume has voxels."""
    from ambient_image_processor_amd import native, phantoms
    side = 4096
    f, _, _ = phantoms.deshade_phantom((1, side, side), seed=2, kind="spheres", noise=0.03,
                                       bias_order=3, bias_sigma=0.05)
    _tgv2_variant(args, f, None, native.FLAG_PLANAR | native.FLAG_SET_NEGATIVE_ZERO,
                  f"tgv2_l1_deshade_2d_{side}^2",
                  f"pixel-iters/s 2-D TGV2-L1 deshade {side}^2 + %HBM roofline",
                  "synthetic (seeded discs + DCT bias + N(0,0.03^2) noise), depth 1",
                  f"2-D TGV2-L1 iterations of the same {side}^2 image")


def poisson_512(args, emit=True):
    """DCT Poisson integration of v = D+ phi at 512 x 512 x 256, phi a sum of
    64 cosine modes: B-divergence, forward DCT with the eigenvalue divide, and
    inverse DCT; the exact answer is phi - mean(phi) = phi."""
    from ambient_image_processor_amd import native, phantoms
    shape = (256, 512, 512)
    n = int(np.prod(shape))
    _, phi = phantoms.poisson_field(shape, seed=3)
    v = [forward_difference(phi, a).astype(np.float32) for a in range(3)]
    ctx = new_context()
    vd = [ctx.alloc(n * 4) for _ in range(3)]
    for d, a in zip(vd, v):
        ctx.upload(d, a)
    out = ctx.alloc(n * 4)
    flops = 4.0 * n * sum(shape)
    res = {}
    for name, prec in (("reference", native.DCT_REFERENCE), ("f32", native.DCT_F32),
                       ("bf16x3", native.DCT_BF16X3)):
        step = lambda: ctx.integrate_gradients_poisson_device(*vd, shape, out, prec)
        el, prof = timed(ctx, step, args.steps, args.warmup)
        l = ctx.download(out, shape)
        ms, cnt = prof["dct"]
        pass_ms = ms / max(cnt, 1)
        dct_ms = ms / args.steps
        mf_peak = BF16_PEAK_TFLOPS if prec == native.DCT_BF16X3 else FP32_PEAK_TFLOPS
        issued = 3 * flops if prec == native.DCT_BF16X3 else flops
        res[name] = {"ms_per_solve": 1000.0 * el / args.steps, "dct_ms_per_solve": dct_ms,
                     "dct_pass_us": 1000.0 * pass_ms,
                     "mfma_frac": (issued / (dct_ms * 1e-3) / 1e12 / mf_peak
                                   if prec != native.DCT_REFERENCE else None),
                     "hbm_frac": 6 * 2 * n * 4 / (dct_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "max_rel_err_vs_closed_form": rel_err(l, phi),
                     "rms_rel_err_vs_closed_form": rms_err(l, phi)}
        print(f"{name} done", file=sys.stderr, flush=True)
    best = res["bf16x3"]
    line = {
        "metric": "DCT Poisson solves/s 512x512x256 (B-div + DCT-II + eigen divide + DCT-III)",
        "value": 1000.0 / best["ms_per_solve"], "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": best["ms_per_solve"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3/f32-acc", "data": "synthetic (64 seeded cosine modes, closed form)",
        "config": {"workload": "poisson_512x512x256", "shape": list(shape),
                   "flops_per_solve_fp32_equiv": flops},
        "variants": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv)
                         for kk, vv in v.items()} for k, v in res.items()}}
    if emit:
        print(json.dumps(line))
    for d in vd + [out]:
        ctx.free(d)
    ctx.close()
    return line


def longrun_golden():
    """The oracle's full-size results (tests/golden/longrun.npz, written by
    tests/golden/gen_longrun.py), or None."""
    path = os.path.join(ROOT, "tests", "golden", "longrun.npz")
    if not os.path.exists(path):
        return None
    with np.load(path) as z:
        return {k: z[k] for k in z.files}


def plane_sums(a):
    return np.sum(np.asarray(a, dtype=np.float32), axis=(1, 2), dtype=np.float64)


def bf16_384(args, emit=True):
    """TGV2-L2 deshade at 384^3, 800 iterations, BF16 field storage against the
    FP32 path (errors of u and of the shading l), and, where the oracle's
    full-size results are committed, whether the BF16 run equals the oracle's
    BF16 mode bit for bit (per-plane sums) together with that mode's offline
    errors against the FP64 oracle."""
    from ambient_image_processor_amd import native, phantoms
    w = phantoms.workload("deshade_384")
    f = w["f"]
    shape = f.shape
    n = f.size
    iters = args.iterations if args.iterations != 1000 else w["iterations"]
    ctx = new_context()
    f_dev = ctx.alloc(n * 4)
    u_dev, l_dev, r_dev = ctx.alloc(n * 4), ctx.alloc(n * 4), ctx.alloc(n * 4)
    ctx.upload(f_dev, f)
    out = {}
    for name, storage in (("f32", native.STORAGE_F32), ("bf16", native.STORAGE_BF16)):
        p = native.default_params(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"],
                                  tau=w["tau"], sigma=w["tau"], theta=1.0, iterations=iters,
                                  data_term=native.DATA_L2, storage=storage,
                                  dct_precision=native.DCT_F32, flags=0)
        el, prof = timed(ctx, lambda: ctx.deshade_pipeline_device(f_dev, shape, p, u_dev, l_dev,
                                                                 r_dev), 1, 0)
        sb = 2 if storage == native.STORAGE_BF16 else 4
        out[name] = {"s": el, "voxel_iters_per_s": n * iters / el,
                     "kernels": rounded(sweep_kernels(prof, sb, n)),
                     "u": ctx.download(u_dev, shape), "l": ctx.download(l_dev, shape)}
    res = {k: {kk: vv for kk, vv in v.items() if kk not in ("u", "l")} for k, v in out.items()}
    b, r = out["bf16"], out["f32"]
    gold = longrun_golden()
    vs64 = None
    if gold is not None and "c4_u_sums" in gold and iters == 800:
        # the shading of this run comes from the F32 MFMA DCT; the oracle's from the
        # reference-numerics DCT, so only u is compared bit for bit
        same = bool(np.array_equal(plane_sums(b["u"]), gold["c4_u_sums"]))
        vs64 = {"u_equals_oracle_bf16_mode": same,
                "note": "oracle BF16 mode and FP32 run against the FP64 oracle at this "
                        "configuration (offline, tests/golden/longrun.npz)",
                "bf16_u_max": float(gold["c4_bf16_u_max"]),
                "bf16_u_rms": float(gold["c4_bf16_u_rms"]),
                "bf16_shading_max": float(gold["c4_bf16_shading_max"]),
                "bf16_shading_rms": float(gold["c4_bf16_shading_rms"]),
                "f32_u_max": float(gold["c4_f32_u_max"]), "f32_u_rms": float(gold["c4_f32_u_rms"]),
                "f32_shading_max": float(gold["c4_f32_shading_max"]),
                "f32_shading_rms": float(gold["c4_f32_shading_rms"])}
    line = {
        "metric": "voxel-iters/s TGV2-L2 deshade 384^3 800 it, BF16 field storage",
        "value": b["voxel_iters_per_s"], "unit": "voxel-iters/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": 1000.0 * b["s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16-storage/f32-arith",
        "data": "synthetic (spheres, bias max 0.2, noise 0.02)",
        "config": {"workload": "tgv2_l2_deshade_384^3_bf16", "iterations": iters},
        "speedup_vs_f32": b["voxel_iters_per_s"] / r["voxel_iters_per_s"],
        "u_rel_err_vs_f32": rel_err(b["u"], r["u"]), "u_rms_err_vs_f32": rms_err(b["u"], r["u"]),
        "shading_rel_err_vs_f32": rel_err(b["l"], r["l"]),
        "shading_rms_err_vs_f32": rms_err(b["l"], r["l"]),
        "vs_fp64": vs64,
        "runs": res}
    if emit:
        print(json.dumps(line))
    for ptr in (f_dev, u_dev, l_dev, r_dev):
        ctx.free(ptr)
    ctx.close()
    return line


def adjoint_4x128(args, emit=True):
    """Unrolled TGV2-L2 denoise (200 iterations) and its hand-written adjoint
    (checkpoints every 20) on 4 volumes of 128^3; MSE loss to the clean volumes."""
    from ambient_image_processor_amd import native, phantoms
    w = phantoms.workload("denoise_128x4")
    vols, cleans = w["f"], w["clean"]
    shape = vols.shape[1:]
    n = int(np.prod(shape))
    iters = w["iterations"]
    p = native.default_params(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"],
                              tau=w["tau"], sigma=w["tau"], theta=1.0, iterations=iters,
                              data_term=native.DATA_L2)
    ctx = new_context()
    f_dev, g_dev, u_dev, gf_dev = (ctx.alloc(n * 4) for _ in range(4))
    t_fwd = t_adj = 0.0
    import ctypes
    # warm-up: one forward + adjoint of the first volume (buffers, tables, code objects)
    ctx.upload(f_dev, vols[0])
    ctx.upload(g_dev, np.zeros(shape, np.float32))
    ckpt = getattr(args, "checkpoint_every", 0) or w["checkpoint_every"]
    ctx.tgv2_denoise_adjoint_device(f_dev, shape, p, g_dev, u_dev, gf_dev, ckpt)
    ctx.synchronize()
    for b in range(vols.shape[0]):
        ctx.upload(f_dev, vols[b])
        ctx.synchronize()
        t0 = time.perf_counter()   # the forward alone, device-resident like the adjoint
        native._check(native.lib().amb_tgv2_denoise(
            ctx._h, native.MEM_DEVICE, ctypes.c_void_p(f_dev), *native._dims(shape),
            ctypes.byref(p), ctypes.c_void_p(u_dev), native.CALLBACK(), None))
        ctx.synchronize()
        t_fwd += time.perf_counter() - t0
        u = ctx.download(u_dev, shape)
        ctx.upload(g_dev, ((2.0 / (vols.shape[0] * n)) * (u - cleans[b])).astype(np.float32))
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.tgv2_denoise_adjoint_device(f_dev, shape, p, g_dev, u_dev, gf_dev, ckpt)
        t_adj += time.perf_counter() - t0
        print(f"volume {b} done", file=sys.stderr, flush=True)
    vi = float(n) * iters * vols.shape[0]
    line = {
        "metric": "voxel-iters/s unrolled TGV2-L2 denoise forward + adjoint, 4 x 128^3, 200 it",
        "value": vi / t_adj, "unit": "voxel-iters/s (forward + reverse)", "n_gpus": 1,
        "steps": vols.shape[0], "warmup": 0, "ms_per_step": 1000.0 * t_adj / vols.shape[0],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (12 seeded boxes + N(0,0.05^2) per volume)",
        "config": {"workload": "tgv2_l2_adjoint_4x128^3", "iterations": iters,
                   "checkpoint_every": ckpt},
        "forward_only_voxel_iters_per_s": vi / t_fwd,
        "adjoint_over_forward_time": t_adj / t_fwd}
    if emit:
        print(json.dumps(line))
    for ptr in (f_dev, g_dev, u_dev, gf_dev):
        ctx.free(ptr)
    ctx.close()
    return line


def tgvk3_256(args):
    """TGV-k deshade (k = 3, the widget's default weights 1, 4, 7;
    TGVKDeshadeWidget.ui / TGVKDeshadeProcessor.cpp:272-281) of the 256^3
    volume: the TGV-k loop, the Poisson solve and the epilogue."""
    from ambient_image_processor_amd import native, phantoms
    n_side = args.size
    shape = (n_side, n_side, n_side)
    f, _, _ = phantoms.deshade_phantom(shape, seed=2, kind="spheres", noise=0.03, bias_order=3,
                                       bias_sigma=0.05)
    n = f.size
    alpha = (1.0, 4.0, 7.0)
    levels = len(alpha) - 2
    tau = float(1.0 / np.sqrt(20.0))
    p = native.default_params(lam=0.5, tau=tau, sigma=tau, theta=1.0, iterations=args.iterations,
                              data_term=native.DATA_L1, flags=native.FLAG_SET_NEGATIVE_ZERO)
    ctx = new_context()
    f_dev = ctx.alloc(n * 4)
    u_dev, l_dev, r_dev = ctx.alloc(n * 4), ctx.alloc(n * 4), ctx.alloc(n * 4)
    ctx.upload(f_dev, f)
    elapsed, prof = timed(ctx, lambda: ctx.tgvk_deshade_pipeline_device(
        f_dev, shape, p, alpha, u_dev, l_dev, r_dev), args.steps, args.warmup)
    value = float(n) * args.iterations * args.steps / elapsed
    kernels, dom = tgvk_kernels(prof, levels, n)
    achieved = kernels[dom]["gbs"]
    traffic = MEASURED_TRAFFIC.get((dom, levels, n_side))
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle
        oracle.build()
        st = oracle.TGVKState(f, len(alpha), np.float32)
        st.iterate(1, lam=0.5, alpha=alpha, tau=tau, sigma=tau)
        t1 = time.perf_counter()
        st.iterate(args.cpu_iters // 2, lam=0.5, alpha=alpha, tau=tau, sigma=tau)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": float(n) * (args.cpu_iters // 2) / cpu_s, "unit": "voxel-iters/s",
               "cores": oracle.num_threads(), "kind": "port",
               "sample": f"{args.cpu_iters // 2} TGV-k (k=3) iterations of the same {n_side}^3 "
                         f"volume (FP32 reference op order, OpenMP), {cpu_s:.1f} s"}
    print(json.dumps({
        "metric": "voxel-iters/s TGV-k (k=3) deshade 256^3 + %HBM roofline",
        "value": value, "unit": "voxel-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (32 seeded spheres + DCT bias + N(0,0.03^2) noise)",
        "config": {"workload": f"tgvk3_l1_deshade_{n_side}^3", "shape": list(shape),
                   "iterations": args.iterations, "lambda": 0.5, "alpha": list(alpha),
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": kernels[dom]["bytes_per_launch"],
                     "measured_hbm_gbs": (traffic / (kernels[dom]["avg_us"] * 1e-6) / 1e9)
                     if traffic else None},
        "kernels": rounded(kernels),
        "other_kernels_ms_per_step": {k: round(prof[k][0] / args.steps, 3)
                                      for k in ("dct", "poisson", "divergence", "epilogue", "init")},
        "cpu_baseline": cpu}))
    for ptr in (f_dev, u_dev, l_dev, r_dev):
        ctx.free(ptr)
    ctx.close()


def downsampled_512(args):
    """The downsampled TGV-k deshade as tgvk_deshade_downsampled_main.cpp runs it
    (factor 0.5, order 3, alpha[i] = i + 1, negatives to zero, background added
    back) on a 512^3 volume with a mask: the masked input and the mask resized
    to 256^3 with cubic B-splines, the masked TGV-k deshade there, u and l
    resized back, the tail.  The resampling kernels are reported separately."""
    from ambient_image_processor_amd import native, phantoms
    shape = (512, 512, 512)
    f, _, _ = phantoms.deshade_phantom(shape, seed=2, kind="spheres", noise=0.03, bias_order=3,
                                       bias_sigma=0.05)
    z, y, x = np.ogrid[0:512, 0:512, 0:512]
    mask = (((z - 256.0) ** 2 + (y - 256.0) ** 2 + (x - 256.0) ** 2) < 240.0 ** 2)
    mask = mask.astype(np.float32)
    n = f.size
    iters = args.iterations
    tau = float(1.0 / np.sqrt(20.0))
    alpha = (1.0, 2.0, 3.0)
    p = native.default_params(lam=0.5, tau=tau, sigma=tau, theta=1.0, iterations=iters,
                              flags=native.FLAG_MASKED_ITERATION | native.FLAG_SET_NEGATIVE_ZERO
                              | native.FLAG_ADD_BACKGROUND_BACK)
    ctx = new_context()
    bufs = [ctx.alloc(n * 4) for _ in range(5)]
    f_dev, m_dev, u_dev, l_dev, r_dev = bufs
    ctx.upload(f_dev, f)
    ctx.upload(m_dev, mask)
    del f, mask
    elapsed, prof = timed(ctx, lambda: ctx.tgvk_deshade_downsampled_device(
        f_dev, shape, p, 0.5, alpha, m_dev, u_dev, l_dev, r_dev), args.steps, args.warmup)
    small = native.resize_extents(shape, 0.5)
    ns = int(np.prod(small))
    value = float(ns) * iters * args.steps / elapsed
    resize_ms, resize_n = prof["resize"]
    kernels, dom = tgvk_kernels(prof, 1, ns)
    print(json.dumps({
        "metric": "voxel-iters/s (resized volume) downsampled TGV-k deshade 512^3 -> 256^3",
        "value": value, "unit": "voxel-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (32 seeded spheres + DCT bias + N(0,0.03^2) noise, spherical mask)",
        "config": {"workload": "tgvk3_downsampled_512^3_to_256^3", "factor": 0.5,
                   "iterations": iters, "alpha": list(alpha), "lambda": 0.5,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kernels[dom]["gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": kernels[dom]["bytes_per_launch"]},
        "resize": {"ms_per_step": resize_ms / args.steps, "calls_per_step": resize_n / args.steps,
                   "note": "4 resamplings per deshade (f*mask and mask down, u and l up), "
                           "timed as 2 scopes"},
        "kernels": rounded(kernels),
        "other_kernels_ms_per_step": {k: round(prof[k][0] / args.steps, 3)
                                      for k in ("dct", "poisson", "divergence", "epilogue", "init")},
        "cpu_baseline": None}))
    for ptr in bufs:
        ctx.free(ptr)
    ctx.close()


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
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--mask-radius", type=float, default=0.495,
                    help="masked_256: ball radius as a fraction of the edge")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--workload", choices=["deshade_256", "poisson_512", "bf16_384",
                                           "adjoint_4x128", "tgvk3_256", "downsampled_512",
                                           "masked_256", "planar_4096"],
                    default="deshade_256")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--checkpoint-every", type=int, default=0,
                    help="adjoint workload: checkpoint interval (default: the configuration's 20; "
                         "1 keeps every state)")
    ap.add_argument("--plan", action="append", default=[], metavar="NAME=VALUE",
                    help="launch-plan override (amb_context_set_plan_override), repeatable")
    args = ap.parse_args()
    for item in args.plan:
        name, _, value = item.partition("=")
        PLAN_OVERRIDES.append((name.strip(), int(value)))

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only (the TGV sweeps would shard into z-slabs "
              "exchanging one-plane halos, which is not implemented). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    {"deshade_256": deshade_256, "poisson_512": poisson_512, "bf16_384": bf16_384,
     "adjoint_4x128": adjoint_4x128, "tgvk3_256": tgvk3_256,
     "downsampled_512": downsampled_512, "masked_256": masked_256,
     "planar_4096": planar_4096}[args.workload](args)


if __name__ == "__main__":
    main()
