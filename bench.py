"""Benchmark of the VCA solver-core hot path on one MI355X.

Workload (default): the 12-site 3x4 square Hubbard cluster, U=8, half-filled
sector Nup=Ndw=6 (853,776 states), FP64 -- the plain FP64 GPU
configuration.  A "step" is one H x v application on that sector with the
vector resident in HBM.  Reported beside it in the same JSON line:
  * the MIXED form (complex64 H x v) on the same sector;
  * the ground-state Lanczos (FP64, and MIXED with FP64 refinement) with
    its E0 error against FP64;
  * Omega evals/s: (k, omega) log-det evaluations per second of the Omega
    integrand on the 2x2 BHZ cluster (n=16 in two spin blocks of 8), 128x128
    k-grid, 8192 Matsubara frequencies at beta=200, FP64 and MIXED;
  * the roofline of the H x v kernel (algorithmic bytes 2 x 16 B per state
    per application over its event-timed duration, against 8 TB/s);
  * the CPU baseline: the oracle's restatement of the reference H x v loops
    (C, -O2 -funroll-loops, one core), timed on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--full]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--full", action="store_true", help="also run the full vca_solve of the 3x4 cluster")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bath16", action="store_true",
                    help="skip the 8-site + 8-bath-level sector (165,636,900 states)")
    ap.add_argument("--omega-nk", type=int, default=128)
    ap.add_argument("--omega-nfreq", type=int, default=8192)
    ap.add_argument("--only", choices=["all", "hxv", "omega"], default="all",
                    help="hxv: only the H x v timing plus a 256 MiB copy kernel that "
                         "calibrates HBM byte counters (for rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from vca_amd import configs
    from vca_amd.api import VcaSolver

    dev = torch.device("cuda:0")
    p, hloc, hk, _, _ = configs.config_square_3x4(nk=2)
    s = VcaSolver(p)
    s.vca_set_hcluster(hloc)
    out = {}

    # ---------------------------------------------------------- H x v ----
    def time_hxv(prec):
        sec = s.sector(6, 6, prec)
        dt = torch.complex128 if prec == "fp64" else torch.complex64
        g = torch.Generator(device=dev).manual_seed(1)
        v = torch.randn(sec.dim, dtype=dt, device=dev, generator=g)
        hv = torch.empty_like(v)
        torch.cuda.synchronize()
        sec.hxv_device(v.data_ptr(), hv.data_ptr(), repeat=max(1, args.warmup))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ms = sec.hxv_device(v.data_ptr(), hv.data_ptr(), repeat=args.steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return sec, ms, wall

    sec64, ms64, wall64 = time_hxv("fp64")
    sec32, ms32, wall32 = time_hxv("mixed")
    dim = sec64.dim
    log(f"H x v FP64: {ms64 * 1e3:.1f} us/apply; MIXED: {ms32 * 1e3:.1f} us/apply; dim {dim}")
    if args.only == "hxv":
        # calibration kernel for the HBM byte counters: a copy of 256 MiB
        a = torch.ones(1 << 25, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"hxv_fp64_us": ms64 * 1e3, "hxv_mixed_us": ms32 * 1e3, "dim": dim,
                          "copy_bytes_each_way": a.numel() * 8}), flush=True)
        return

    # --------------------------------------------------- ground state ----
    t0 = time.perf_counter()
    e64, _, it64 = sec64.lanczos_gs(tol=1e-12, niter=512, want_vec=False)
    t_gs64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    emx, _, itmx = sec32.lanczos_gs(tol=1e-12, niter=512, want_vec=False)
    t_gsmx = time.perf_counter() - t0
    log(f"GS FP64 E0={e64:.14f} ({it64} it, {t_gs64:.2f}s); MIXED E0={emx:.14f} "
        f"(refine {itmx} it, {t_gsmx:.2f}s); dE0 rel {abs(emx - e64) / abs(e64):.2e}")

    # --------------------------------------------------------- Omega -----
    po, hlo, hko, _, _ = configs.config_bhz_2x2(nk=args.omega_nk, nfreq=args.omega_nfreq,
                                               precision="fp64")
    so = VcaSolver(po)
    t0 = time.perf_counter()
    om64 = so.vca_solve(hlo, hko)
    t_solve_bhz = time.perf_counter() - t0
    zs = 1j * (2 * np.arange(args.omega_nfreq) + 1) * np.pi / po.beta
    nk = hko.shape[6]
    so.omega_integrand(zs[:64], "fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f64 = so.omega_integrand(zs, "fp64")
    t_om64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    f32 = so.omega_integrand(zs, "mixed")
    t_om32 = time.perf_counter() - t0
    s64 = (3 - po.nspin) * 2.0 / po.beta * f64.sum()
    s32 = (3 - po.nspin) * 2.0 / po.beta * f32.sum()
    om_err = abs(s32 - s64) / abs(s64)
    evals = nk * args.omega_nfreq
    log(f"Omega BHZ: {evals} (k,w) evals FP64 {t_om64:.3f}s, MIXED {t_om32:.3f}s; "
        f"integral rel err mixed {om_err:.2e}; full solve {t_solve_bhz:.2f}s, Omega={om64:.10f}")

    # ------------------------------- 8 sites + 8 bath levels, 1.66e8 states
    bath16 = None
    if not args.no_bath16:
        pb, hlb, hkb, bhb, bvb = configs.config_bath_2x4(nk=2, precision="mixed")
        sb = VcaSolver(pb)
        sb.vca_set_bath(bhb, bvb)
        sb.vca_set_hcluster(hlb)
        out16 = {}
        for prec in ("mixed", "fp64"):
            sec = sb.sector(8, 8, prec)
            dt = torch.complex128 if prec == "fp64" else torch.complex64
            g = torch.Generator(device=dev).manual_seed(2)
            v = torch.randn(sec.dim, dtype=dt, device=dev, generator=g)
            hv = torch.empty_like(v)
            sec.hxv_device(v.data_ptr(), hv.data_ptr(), repeat=2)
            ms = sec.hxv_device(v.data_ptr(), hv.data_ptr(), repeat=5)
            del v, hv
            torch.cuda.empty_cache()
            out16[prec] = {"ms_per_hxv": ms, "gstates_per_s": sec.dim / (ms * 1e-3) / 1e9,
                           "hbm_gbs_algorithmic": 2 * (16 if prec == "fp64" else 8) * sec.dim
                           / (ms * 1e-3) / 1e9}
            log(f"bath16 {prec}: {ms:.3f} ms/apply, dim {sec.dim}")
            if prec == "mixed":
                t0 = time.perf_counter()
                e_mx, _, it_mx = sec.lanczos_gs(tol=1e-12, niter=512, want_vec=False)
                out16["gs_mixed"] = {"e0": e_mx, "seconds": time.perf_counter() - t0,
                                     "refine_iters": it_mx}
            else:
                t0 = time.perf_counter()
                e_64, _, it_64 = sec.lanczos_gs(tol=1e-12, niter=512, want_vec=False)
                out16["gs_fp64"] = {"e0": e_64, "seconds": time.perf_counter() - t0,
                                    "iters": it_64}
            sec.close()
        out16["e0_rel_err_mixed"] = abs(out16["gs_mixed"]["e0"] - out16["gs_fp64"]["e0"]) / abs(
            out16["gs_fp64"]["e0"])
        out16["dim"] = 12870 * 12870
        bath16 = out16
        log(f"bath16 GS: mixed {out16['gs_mixed']} fp64 {out16['gs_fp64']}")

    full = None
    if args.full:
        pf, hlf, hkf, _, _ = configs.config_square_3x4(nk=64)
        sf = VcaSolver(pf)
        t0 = time.perf_counter()
        omf = sf.vca_solve(hlf, hkf)
        full = dict(seconds=time.perf_counter() - t0, omega=omf, **sf.results())
        log(f"full 3x4 solve: {full['seconds']:.1f}s Omega={omf:.10f}")

    # ---------------------------------------------------- CPU baseline ---
    cpu = None
    if not args.no_cpu:
        from oracle import vca_ref as R
        from tests._models import to_oracle
        om = R.Model(to_oracle(p), hloc)
        v = np.random.default_rng(3).normal(size=dim).astype(np.complex128)
        n = 0
        t0 = time.perf_counter()
        while True:
            om.hxv(6, 6, v)
            n += 1
            if time.perf_counter() - t0 > 10.0 or n >= 20:
                break
        tcpu = (time.perf_counter() - t0) / n
        cpu = {"value": dim / tcpu / 1e9, "unit": "Gstates/s", "cores": 1, "kind": "port",
               "sample": f"{n} H x v applications on the 853,776-state sector (6,6), "
                         f"oracle C restatement of the reference loops, -O2 -funroll-loops"}
        log(f"CPU 1-core: {tcpu:.3f} s/apply")

    gst64 = dim / (ms64 * 1e-3) / 1e9
    gst32 = dim / (ms32 * 1e-3) / 1e9
    bytes64 = 2 * 16 * dim
    line = {
        "metric": "H×v Gstates/s + Ω evals/s, mixed vs FP64 on 1 GPU at ΔE0<1e-10",
        "value": gst64,
        "unit": "Gstates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random vectors; Hubbard model inputs generated in-repo)",
        "config": {"workload": "hubbard 3x4 square cluster U=8 half-filled sector (6,6), "
                               "853,776 states, H x v FP64", "sector_dim": dim},
        "mixed": {"value": gst32, "ms_per_step": ms32, "dtype": "f32 (complex64)"},
        "ground_state": {"e0_fp64": e64, "iters_fp64": it64, "seconds_fp64": t_gs64,
                         "e0_mixed": emx, "seconds_mixed": t_gsmx,
                         "e0_rel_err_mixed": abs(emx - e64) / abs(e64)},
        "omega": {"workload": f"BHZ 2x2 Norb=2 Nspin=2, {nk} k x {args.omega_nfreq} Matsubara",
                  "evals": evals, "evals_per_s_fp64": evals / t_om64,
                  "evals_per_s_mixed": evals / t_om32, "integral_rel_err_mixed": om_err},
        "roofline": {"bound": "hbm", "achieved": bytes64 / (ms64 * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
    }
    if full:
        line["full_solve"] = full
    if bath16:
        line["bath16"] = bath16
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
