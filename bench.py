#!/usr/bin/env python3
"""Benchmark of the transform3D path on one MI355X.

Workload: 512^3 real grid, FP32, field sin(x)sin(2y)sin(3z) + N(0, 1e-3)
noise (synthetic, generated on the GPU from a fixed seed), input memory
order (2,1,0), spectrum memory order (1,2,0).  One step = one R2C forward
3D transform plus one C2R backward 3D transform (R2CFFT_S x, CFFT_FORWARD_S
y and z; then CFFT_BACKWARD_S, CFFT_BACKWARD_S, C2RFFT_S), called through
the library's C ABI on device-resident arrays.

value = B_alg / t_step with B_alg = 2 (V_r + 5 V_c): one read and one write
of the array per 1D stage of the reference's 3-stage algorithm, in each
direction (V_r = 4 N^3 bytes, V_c = 8 (N/2+1) N^2 bytes).

Prints one JSON line on stdout.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "p3dfft.3_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "512^3 R2C FP32 fwd+bwd GB/s (% HBM peak); Poisson solves/s + p99 at err ≤1e-5"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def b_alg(n, e):
    vr = e * n ** 3
    vc = 2 * e * (n // 2 + 1) * n * n
    return 2 * (vr + 5 * vc)


def cpu_baseline(n_sample=256):
    """The FP64 oracle (single-threaded C restatement of the reference's
    FFTW path) on a bounded sample: one R2C+C2R round trip at n_sample^3 with
    the same memory orders.  Rate = B_alg(n_sample, FP32) / t."""
    import numpy as np
    import oracle_binding as orc
    x = np.random.default_rng(2).uniform(-1.0, 1.0, (n_sample,) * 3)
    t0 = time.perf_counter()
    spec = orc.exec3d([5, 9, 9], [n_sample] * 3, [2, 1, 0], [1, 2, 0], x)
    back = orc.exec3d([7, 11, 11], [n_sample // 2 + 1, n_sample, n_sample], [1, 2, 0], [2, 1, 0], spec)
    t = time.perf_counter() - t0
    err = float(np.abs(back / n_sample ** 3 - x).max())
    return {"value": round(b_alg(n_sample, 4) / t / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle/ FP64 C restatement, one R2C+C2R round trip at {n_sample}^3 "
                      f"(mem order (2,1,0)->(1,2,0)), single thread, {t:.1f} s, roundtrip max err {err:.1e}; "
                      f"rate = B_alg({n_sample}^3, FP32) / t"}


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic_pass_kernel_512cubed_fp32.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, run a plain 256 MiB device copy three times (the "
                         "known-size kernel that scales HBM byte counters under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the 1x1 (single-GPU) transform path only; "
                                   "use --gpus 1"}), flush=True)
        sys.exit(2)

    import torch
    import p3dfft3 as P

    n = args.n
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    P.setup()
    pg = P.init_proc_grid((1, 1, 1))
    mo1, mo2 = (2, 1, 0), (1, 2, 0)
    g_real = P.init_data_grid((n, n, n), -1, pg, (0, 1, 2), mo1)
    g_spec = P.init_data_grid((n // 2 + 1, n, n), 0, pg, (0, 1, 2), mo2)
    tf = P.init_3Dtype((P.type_id("R2CFFT_S"), P.type_id("CFFT_FORWARD_S"), P.type_id("CFFT_FORWARD_S")))
    tb = P.init_3Dtype((P.type_id("C2RFFT_S"), P.type_id("CFFT_BACKWARD_S"), P.type_id("CFFT_BACKWARD_S")))
    pf = P.plan_3Dtrans(g_real, g_spec, tf)
    pb = P.plan_3Dtrans(g_spec, g_real, tb)
    log("forward plan :", P.describe(pf))
    log("backward plan:", P.describe(pb))

    # input in memory order (2,1,0): storage index [x][y][z], z fastest
    gen = torch.Generator(device="cuda").manual_seed(2)
    x = torch.arange(n, device="cuda", dtype=torch.float64) * (2 * torch.pi / n)
    f = (torch.sin(x)[:, None, None] * torch.sin(2 * x)[None, :, None] * torch.sin(3 * x)[None, None, :])
    f = (f + 1e-3 * torch.randn((n, n, n), generator=gen, device="cuda", dtype=torch.float64)).float()
    f = f.contiguous()
    # spectrum in memory order (1,2,0): storage dims (z, x, y) -> shape [y][x][z]
    spec = torch.empty((n, n // 2 + 1, n), dtype=torch.complex64, device="cuda")
    back = torch.empty_like(f)
    st = P.stream()
    torch.cuda.synchronize()

    def step():
        P.exec_async(pf, f.data_ptr(), spec.data_ptr(), -1, 0, st, prec=4)
        P.exec_async(pb, spec.data_ptr(), back.data_ptr(), -1, 0, st, prec=4)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    P.profile(pf, True)
    P.profile(pb, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    nf, avg_f, tot_f = P.profile_read(pf)
    nb, avg_b, tot_b = P.profile_read(pb)
    P.profile(pf, False)
    P.profile(pb, False)

    err = (torch.linalg.vector_norm((back.double() / n ** 3 - f.double())) /
           torch.linalg.vector_norm(f.double())).item()
    t_step = t / args.steps
    balg = b_alg(n, 4)
    value = balg / t_step / 1e9
    launches = nf * len(avg_f) + nb * len(avg_b)
    avg_launch_ms = (tot_f + tot_b) / max(launches, 1)
    per_launch_bytes = balg / (len(avg_f) + len(avg_b))
    achieved = per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9
    log(f"step {t_step * 1e3:.3f} ms  {value:.1f} GB/s  roundtrip rel L2 {err:.2e}")
    log("forward pass avg ms:", [round(v, 4) for v in avg_f], "backward:", [round(v, 4) for v in avg_b])
    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 3 x 268435456-byte device copies done")
    traffic = load_traffic()
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 256^3 sample) ...")
        cpu = cpu_baseline()
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "roundtrip_rel_l2": err,
        "config": {"workload": f"{n}^3 R2C+C2R FP32 round trip, in mem order (2,1,0), spectrum (1,2,0), "
                               "sin(x)sin(2y)sin(3z)+N(0,1e-3)",
                   "grid": [n, n, n], "types": "R2CFFT_S,CFFT_FORWARD_S,CFFT_FORWARD_S / back",
                   "b_alg_bytes": balg, "parallelism": "1 GPU, 1x1 processor grid"},
        "roofline": {"kernel": "p3e::pass_kernel_pow2<float,M> (M=512 C2C, M=256 R2C/C2R; all 6 stage launches)", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "bytes_per_launch_alg": per_launch_bytes,
                     "avg_launch_ms": round(avg_launch_ms, 5)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
