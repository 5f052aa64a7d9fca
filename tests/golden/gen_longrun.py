"""Writes tests/golden/longrun.npz: the oracle's results of the long benchmark
runs at their full sizes, so that GPU tests can check them bit for bit without
the volumes (a GPU result equal to the oracle's has the same per-plane sums),
and the FP64 oracle's distance to them.

  c2_*  TGV2-L1 deshade of the 256^3 sphere phantom (workload deshade_256:
        seed 2, 1000 iterations, lambda 0.5, alpha0 2, alpha1 1,
        tau = sigma = 1/sqrt(20)), FP32 mirror order, reference-numerics DCT
        Poisson solve and the CLI epilogue (clamp):
          {u,vx,vy,vz,shading,deshaded}_sums  per z plane, the FP64 sum of the
                                              float32 values (numpy's order)
          u_z128, shading_z128                two whole planes
          fp64_{u,shading}_{max,rms}          the FP64 run's u and shading
                                              against FP32: max|a-b|/max|b|,
                                              ||a-b||/||b|| (b = FP64)
  c4_*  TGV2-L2 deshade at 384^3 (workload deshade_384: seed 4, bias max 0.2,
        noise 0.02, 800 iterations, lambda 1), BF16 field storage (the
        oracle's BF16 mode), FP32 Poisson solve in reference numerics:
          {u,vx,vy,vz,shading}_sums, u_z192
          bf16_{u,shading}_{max,rms}, f32_{u,shading}_{max,rms}   the BF16-storage
                                              and the FP32 runs against FP64
  c5_*  the unrolled TGV2-L2 denoise of workload denoise_128x4 (4 x 128^3,
        200 iterations, alpha0 2, alpha1 1, lambda 1): the FP64 loss
        L = (1/4) sum_b mean((u_b - clean_b)^2) and its central differences in
        alpha0, alpha1 and lambda (shared by the 4 volumes) at eps 1e-3 and 1e-5:
          loss, eps, fd (len(eps) x 3: alpha0, alpha1, lambda)

Run:  python tests/golden/gen_longrun.py [c2] [c4] [c5]   (all by default; the
oracle's OpenMP loop on the host's cores: c2 about 40 minutes, c5 30, c4 two
and a half hours on 8 cores).  Each
stage's arrays are cached under --cache (default /tmp/amb_longrun_cache), and
the npz is rewritten with every stage present in the cache.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, "longrun.npz")

from ambient_image_processor_amd import phantoms  # noqa: E402
from oracle import oracle  # noqa: E402


def say(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)


def plane_sums(a):
    return np.sum(np.asarray(a, dtype=np.float32), axis=(1, 2), dtype=np.float64)


def err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return (float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)),
            float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))


def run_loop(f, iters, prm, data_term, dtype, bf16=False, chunk=50, tag=""):
    st = oracle.TGVState(f, dtype, bf16=bf16)
    done = 0
    while done < iters:
        k = min(chunk, iters - done)
        st.iterate(k, data_term=data_term, **prm)
        done += k
        say(f"{tag} {done}/{iters}")
    return st


def stage_c2(cache):
    w = phantoms.workload("deshade_256")
    f = w["f"]
    prm = dict(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"], tau=w["tau"],
               sigma=w["tau"], theta=1.0)
    out = {}
    st = run_loop(f, 1000, prm, "l1", np.float32, tag="c2 fp32")
    l = oracle.integrate_poisson(*st.v, dtype=np.float32)
    r, l2 = oracle.deshade_epilogue(st.u, l, None, False, True, False)
    u32, s32 = st.u.copy(), l2
    for name, a in (("u", st.u), ("vx", st.v[0]), ("vy", st.v[1]), ("vz", st.v[2]),
                    ("shading", l2), ("deshaded", r)):
        out[f"c2_{name}_sums"] = plane_sums(a)
    out["c2_u_z128"] = st.u[128].copy()
    out["c2_shading_z128"] = l2[128].copy()
    del st
    say("c2 fp32 done")
    st = run_loop(f, 1000, prm, "l1", np.float64, tag="c2 fp64")
    l64 = oracle.integrate_poisson(*st.v, dtype=np.float64)
    (out["c2_fp64_u_max"], out["c2_fp64_u_rms"]) = err(u32, st.u)
    (out["c2_fp64_shading_max"], out["c2_fp64_shading_rms"]) = err(s32, l64)
    say("c2 fp64 done", out["c2_fp64_u_max"], out["c2_fp64_shading_max"])
    np.savez(os.path.join(cache, "c2.npz"), **out)


def stage_c4(cache):
    w = phantoms.workload("deshade_384")
    f = w["f"]
    prm = dict(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"], tau=w["tau"],
               sigma=w["tau"], theta=1.0)
    iters = w["iterations"]
    out = {}
    st = run_loop(f, iters, prm, "l2", np.float32, bf16=True, tag="c4 bf16")
    lb = oracle.integrate_poisson(*st.v, dtype=np.float32)
    for name, a in (("u", st.u), ("vx", st.v[0]), ("vy", st.v[1]), ("vz", st.v[2]),
                    ("shading", lb)):
        out[f"c4_{name}_sums"] = plane_sums(a)
    out["c4_u_z192"] = st.u[192].copy()
    ub = st.u.copy()
    del st
    say("c4 bf16 done")
    st = run_loop(f, iters, prm, "l2", np.float32, tag="c4 fp32")
    u32 = st.u.copy()
    l32 = oracle.integrate_poisson(*st.v, dtype=np.float32)
    del st
    say("c4 fp32 done")
    st = run_loop(f, iters, prm, "l2", np.float64, tag="c4 fp64")
    l64 = oracle.integrate_poisson(*st.v, dtype=np.float64)
    (out["c4_bf16_u_max"], out["c4_bf16_u_rms"]) = err(ub, st.u)
    (out["c4_bf16_shading_max"], out["c4_bf16_shading_rms"]) = err(lb, l64)
    (out["c4_f32_u_max"], out["c4_f32_u_rms"]) = err(u32, st.u)
    (out["c4_f32_shading_max"], out["c4_f32_shading_rms"]) = err(l32, l64)
    say("c4 fp64 done", {k: v for k, v in out.items() if k.endswith(("_max", "_rms"))})
    np.savez(os.path.join(cache, "c4.npz"), **out)


def stage_c5(cache):
    w = phantoms.workload("denoise_128x4")
    vols, cleans = w["f"].astype(np.float64), w["clean"].astype(np.float64)
    base = dict(alpha0=w["alpha0"], alpha1=w["alpha1"], lam=w["lam"])
    it = w["iterations"]

    def loss(**kw):
        prm = dict(base, **kw)
        tot = 0.0
        for b in range(vols.shape[0]):
            st = oracle.tgv2(vols[b], it, prm["lam"], prm["alpha0"], prm["alpha1"], w["tau"],
                             w["tau"], 1.0, "l2", np.float64)
            tot += float(np.mean((st.u - cleans[b]) ** 2))
        return tot / vols.shape[0]

    out = {"c5_loss": loss()}
    eps = np.array([1e-3, 1e-5])
    fd = np.zeros((len(eps), 3))
    for i, e in enumerate(eps):
        for j, name in enumerate(("alpha0", "alpha1", "lam")):
            fd[i, j] = (loss(**{name: base[name] + e}) - loss(**{name: base[name] - e})) / (2 * e)
        say("c5 eps", e, fd[i])
    out["c5_eps"] = eps
    out["c5_fd"] = fd
    np.savez(os.path.join(cache, "c5.npz"), **out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("stages", nargs="*", default=["c2", "c4", "c5"])
    ap.add_argument("--cache", default="/tmp/amb_longrun_cache")
    ap.add_argument("--merge-only", action="store_true",
                    help="rewrite the npz from the cached stages without running any")
    args = ap.parse_args()
    os.makedirs(args.cache, exist_ok=True)
    oracle.build()
    for s in ([] if args.merge_only else args.stages):
        {"c2": stage_c2, "c4": stage_c4, "c5": stage_c5}[s](args.cache)
    merged = {}
    for s in ("c2", "c4", "c5"):
        p = os.path.join(args.cache, s + ".npz")
        if os.path.exists(p):
            with np.load(p) as z:
                merged.update({k: z[k] for k in z.files})
    np.savez_compressed(OUT, **merged)
    say("wrote", OUT, sorted(merged))


if __name__ == "__main__":
    main()
