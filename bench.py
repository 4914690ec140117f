"""Benchmark of the VCA solver-core hot path on one MI355X.

Headline workload: the 12-site 3x4 square Hubbard cluster, U=8, half-filled
sector Nup=Ndw=6 (853,776 states), FP64.  A "step" is one H x v application
on that sector with the vector resident in HBM; `value` is Gstates/s =
dim / (average kernel time of one application, HIP events on the sector's
stream).  Reported beside it in the same JSON line:
  * mixed: the complex64 H x v on the same sector;
  * batched: 64 vectors per launch on the same sector (the form the
    excitation chains run in), FP64 and MIXED;
  * ground_state: the Lanczos ground state, FP64 and MIXED (FP32 Lanczos +
    FP64 refinement), with the E0 error of MIXED against FP64;
  * omega: (k, omega) log-det evaluations per second of the Omega integrand
    on the 2x2 BHZ cluster (Norb=2, Nspin=2: two spin blocks of 8),
    128x128 k-grid x 8192 Matsubara frequencies at beta=200, FP64 and MIXED,
    the MIXED error of the integral, and the roofline of the log-det
    kernel (flops it performs per evaluation over the measured time);
  * full_solve: vca_solve of the 3x4 cluster end to end (ground state, the
    excitation chains of 300 steps on the N+-1 sectors, Omega on the 64x64
    k-grid), then G, Sigma and G0 on the default 5000-point Matsubara and
    real-axis grids (pole terms per second of the grid kernel, its FP64
    roofline, and the oracle's pole sums on one core beside it);
  * batched_2x5: vca_solve of the 10-site 2x5 cluster, every (Nup, Ndw)
    sector, ARPACK-equivalent eigenpairs per sector, all chains batched;
  * chain4: the 4-site chain end to end (every sector, Matsubara-sum Omega
    at beta=100 over 4096 frequencies, G on 512 real-axis points), on the
    GPU and by the oracle on one host core, with the differences;
  * bath16: the 8-site + 8-bath-level sector (165,636,900 states), H x v
    FP64 and MIXED on complex vectors (the tuned form: the transposed one on
    this sector) and on real vectors (the Hamiltonian is real), and the
    ground state both ways (it runs on real vectors);
  * roofline: of the FP64 H x v kernel, algorithmic bytes (one read of v and
    one write of Hv, 2 x 16 B per state) over the event-timed duration,
    against 8 TB/s; traffic = HBM bytes per launch measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/hxv_traffic.json);
  * cpu_baseline: the oracle's C restatement of the reference H x v loops
    (-O2 -funroll-loops) on one core on a bounded sample, plus all host
    cores with the reference's DW-slice split, and the Omega integrand
    (numpy) on a sample of frequencies.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
                    [--only hxv|omega|full|c3|c1|bath16]
                    [--no-full] [--no-c3] [--no-bath16] [--no-cpu]
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
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hxv_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def logdet_flops(nb: int, nblk: int) -> float:
    """Real flops of one (k, omega) evaluation of the log-det kernel: per
    spin block of size nb the contraction H_k G' (nb^3 complex MACs, 8 flops
    each) and the LU elimination (nb^3/3 complex MACs)."""
    return nblk * (8.0 * nb ** 3 + 8.0 * nb ** 3 / 3.0)


# real flops of one pole term w / (z - p) accumulated into G: z - p (1),
# |z - p|^2 (3), reciprocal (1), w * conj(z - p) (6), scale (2), add (2)
POLE_TERM_FLOPS = 15.0


def run_full(cpu=False):
    """vca_solve of the 3x4 cluster end to end, then G, Sigma, G0 on the
    default grids; with cpu, the oracle's pole sums on a sample of the same
    grid for the CPU figure."""
    from vca_amd import configs
    from vca_amd.api import VcaSolver

    pf, hlf, hkf, _, _ = configs.config_square_3x4(nk=64)
    sf = VcaSolver(pf)
    t0 = time.perf_counter()
    omf = sf.vca_solve(hlf, hkf)
    r = sf.results()
    full = {"workload": "3x4 square cluster U=8, sector (6,6) ground state, GF chains "
                        "on the N+-1 sectors (300 steps), Omega on the 64x64 k-grid",
            "seconds": time.perf_counter() - t0, "omega": omf,
            "lanczos_step_gstates_per_s": r["hxv_states"] / max(r["hxv_seconds"], 1e-30)
            / 1e9,
            **{k: r[k] for k in ("omegaprime", "omega_integral", "nstates", "nchains",
                                 "hxv_count", "t_diag", "t_gf", "t_omega",
                                 "omega_evals")}}
    log(f"full 3x4 solve: {full['seconds']:.1f}s Omega={omf:.10f} "
        f"(diag {r['t_diag']:.2f}s, GF {r['t_gf']:.2f}s, Omega {r['t_omega']:.2f}s, "
        f"{r['nchains']} chains)")
    # G, Sigma, G0 on the default grids (Lmats = Lreal = 5000): the pole sums
    # and the per-frequency inversions of build_gf_normal / build_sigma_normal
    grids = {}
    for axis in ("matsubara", "realaxis"):
        t1 = time.perf_counter()
        sf.gf_grid(axis)
        grids[axis + "_seconds"] = time.perf_counter() - t1
    r2 = sf.results()
    rate = r2["grid_pole_evals"] / max(r2["grid_seconds"], 1e-30)
    grids.update(npoles=r2["npoles"], pole_evals=r2["grid_pole_evals"],
                 pole_kernel_seconds=r2["grid_seconds"], pole_evals_per_s=rate,
                 roofline={"bound": "valu", "flops_per_pole_term": POLE_TERM_FLOPS,
                           "achieved": rate * POLE_TERM_FLOPS / 1e12, "peak": FP64_PEAK_TFLOPS,
                           "unit": "TFLOP/s",
                           "frac": rate * POLE_TERM_FLOPS / 1e12 / FP64_PEAK_TFLOPS})
    if cpu:
        # the same pole sums on one core (the oracle's numpy restatement of
        # the loops of VCA_GF_NORMAL.f90:1024-1031) over 64 Matsubara points
        import numpy as np
        from oracle import vca_ref as R
        from tests._models import to_oracle
        po = to_oracle(pf)
        gd = R.GFData(po, 0.0, 1.0)
        for il in range(1, pf.nlat + 1):
            for jl in range(1, pf.nlat + 1):
                nch = 2 if il == jl else 4
                gd.elems[(il, jl, 1, 1, 1)] = [[sf.gf_poles(il, jl, 1, 1, 1, 1, c)
                                                for c in range(1, nch + 1)]]
        wm = np.pi / pf.beta * (2 * np.arange(1, 65) - 1)
        t1 = time.perf_counter()
        R.gf_cluster_grid(gd, 1j * wm)
        tc = time.perf_counter() - t1
        grids["cpu_baseline"] = {
            "pole_evals_per_s": r2["npoles"] * len(wm) / tc, "cores": 1, "kind": "port",
            "sample": f"{r2['npoles']} poles x {len(wm)} Matsubara points, the oracle's "
                      "numpy pole sums (one core)"}
        log(f"grids CPU: {r2['npoles'] * len(wm) / tc / 1e9:.3f} G pole terms/s")
    full["grids"] = grids
    log(f"grids 3x4 ({r2['npoles']} poles, 2 x 5000 points): Matsubara "
        f"{grids['matsubara_seconds'] * 1e3:.1f} ms, real {grids['realaxis_seconds'] * 1e3:.1f} ms "
        f"(pole-sum kernels {r2['grid_seconds'] * 1e3:.2f} ms, "
        f"{grids['pole_evals_per_s'] / 1e9:.1f} G pole terms/s)")
    # the same solve in the mixed form (complex64 ground-state Lanczos with
    # FP64 refinement, complex64 chains, FP32 log-dets), with its errors
    # against the FP64 solve: E0, Omega, and G_ij on the 5000 Matsubara
    # points, each element relative to its own size
    import numpy as np
    g64 = sf.gf_grid("matsubara", ("g",))["g"]
    pf.precision = "mixed"
    sm = VcaSolver(pf)
    t0 = time.perf_counter()
    omm = sm.vca_solve(hlf, hkf)
    tm = time.perf_counter() - t0
    rm = sm.results()
    gm = sm.gf_grid("matsubara", ("g",))["g"]
    # each element against its own largest value over the grid (elements that
    # vanish by symmetry against 1e-8 of the largest element)
    scale = np.maximum(np.abs(g64).max(axis=-1, keepdims=True), 1e-8 * np.abs(g64).max())
    gerr = float(np.max(np.abs(gm - g64) / scale))
    full["mixed"] = {
        "seconds": tm, "omega": omm, "speedup_vs_fp64": full["seconds"] / tm,
        "t_diag": rm["t_diag"], "t_gf": rm["t_gf"], "t_omega": rm["t_omega"],
        "e0_rel_err": abs(rm["omegaprime"] - r["omegaprime"]) / abs(r["omegaprime"]),
        "omega_rel_err": abs(omm - omf) / abs(omf),
        "g_matsubara_max_rel_err": gerr,
        "lanczos_step_gstates_per_s": rm["hxv_states"] / max(rm["hxv_seconds"], 1e-30) / 1e9}
    log(f"full 3x4 solve mixed: {tm:.1f}s ({full['seconds'] / tm:.2f}x FP64; diag "
        f"{rm['t_diag']:.2f}s, GF {rm['t_gf']:.2f}s), dE0 {full['mixed']['e0_rel_err']:.1e}, "
        f"dOmega {full['mixed']['omega_rel_err']:.1e}, max dG_ij(iw)/|G_ij| {gerr:.1e}")
    return full


def run_c1(cpu=True):
    """The reference's CPU-runnable case: the 4-site chain (U=4, half
    filling, Nk=32), every sector, Omega by the Matsubara sum at beta=100 over
    4096 frequencies, G on 512 real-axis points with eta=0.05; on the GPU and
    (cpu) on the host by the oracle, with the differences."""
    import numpy as np
    from vca_amd import configs
    from vca_amd.api import VcaSolver

    p, hloc, hk, _, _ = configs.config_chain4(nk=32)
    p.omega_mode, p.omega_nfreq, p.beta = "matsubara", 4096, 100.0
    p.lreal, p.eps = 512, 0.05
    t0 = time.perf_counter()
    s = VcaSolver(p)
    om = s.vca_solve(hloc, hk)
    greal = s.gf_grid("realaxis", ("g",))["g"]
    tg = time.perf_counter() - t0
    out = {"workload": "4-site Hubbard chain U=4, half filling, Nk=32, every sector; Omega by "
                       "the Matsubara sum (beta=100, 4096 frequencies); G on 512 real-axis "
                       "points, eta=0.05",
           "gpu_seconds": tg, "omega": om}
    if cpu:
        from oracle import vca_ref as R
        from tests._models import to_oracle
        po = to_oracle(p)
        t0 = time.perf_counter()
        ref = R.vca_solve(po, hloc, hk, omega_mode="matsubara", nfreq=4096)
        g_cpu, _, _ = R.sigma_grid(ref["model"], ref["gf"], 1)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"seconds": tc, "cores": 1, "kind": "port",
                               "sample": "the whole case: the oracle's numpy/C restatement"}
        out["omega_rel_diff"] = abs(om - ref["sft_potential"]) / abs(ref["sft_potential"])
        out["g_real_max_diff"] = float(np.abs(greal - g_cpu).max())
    log(f"C1 chain4: GPU {tg:.2f}s Omega={om:.10f}" +
        (f"; CPU {out['cpu_baseline']['seconds']:.2f}s, dOmega {out['omega_rel_diff']:.1e}, "
         f"max dG(w) {out['g_real_max_diff']:.1e}" if cpu else ""))
    return out


def run_c3():
    """vca_solve of the 2x5 cluster over every sector."""
    from vca_amd import configs
    from vca_amd.api import VcaSolver

    pc, hlc, hkc, _, _ = configs.config_batched_2x5(nk=8)
    sc = VcaSolver(pc)
    t0 = time.perf_counter()
    omc = sc.vca_solve(hlc, hkc)
    r = sc.results()
    c3 = {"workload": "2x5 cluster U=6, Nspin=2, every (Nup,Ndw) sector (121, dims "
                      "1..63,504), up to 6 eigenpairs per sector (ARPACK-equivalent), "
                      "all chains batched per target sector",
          "seconds": time.perf_counter() - t0, "omega": omc,
          "sector_threads": int(os.environ.get("VCA_SECTOR_THREADS", "16")),
          "lanczos_step_gstates_per_s": r["hxv_states"] / max(r["hxv_seconds"], 1e-30) / 1e9,
          **{k: r[k] for k in ("omegaprime", "nstates", "nchains", "hxv_count",
                               "t_diag", "t_gf", "t_omega")}}
    log(f"2x5 every-sector solve: {c3['seconds']:.1f}s Omega={omc:.10f} "
        f"(diag {r['t_diag']:.2f}s, GF {r['t_gf']:.2f}s, {r['nchains']} chains)")
    return c3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-full", action="store_true", help="skip the end-to-end 3x4 solve")
    ap.add_argument("--no-c3", action="store_true", help="skip the every-sector 2x5 solve")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bath16", action="store_true",
                    help="skip the 8-site + 8-bath-level sector (165,636,900 states)")
    ap.add_argument("--omega-nk", type=int, default=128)
    ap.add_argument("--omega-nfreq", type=int, default=8192)
    ap.add_argument("--only", choices=["all", "hxv", "omega", "full", "c3", "c1", "bath16"],
                    default="all",
                    help="hxv: the H x v timing plus a 256 MiB copy that calibrates the HBM "
                         "byte counters; omega: the Omega integrand timing; full: the "
                         "end-to-end 3x4 solve; c3: the every-sector 2x5 solve; bath16: the "
                         "1.66e8-state bath sector (rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); how the path would "
              "shard over several GPUs is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from vca_amd import configs
    from vca_amd.api import DeviceArray, VcaSolver, device_synchronize

    def rand_vec(n, prec, seed):
        rng = np.random.default_rng(seed)
        dt = np.complex128 if prec == "fp64" else np.complex64
        return (rng.standard_normal(n, dtype=np.float32)
                + 1j * rng.standard_normal(n, dtype=np.float32)).astype(dt)

    # ------------------------------- 8 sites + 8 bath levels, 1.66e8 states
    def run_bath16():
        pb, hlb, hkb, bhb, bvb = configs.config_bath_2x4(nk=2, precision="mixed")
        sb = VcaSolver(pb)
        sb.vca_set_bath(bhb, bvb)
        sb.vca_set_hcluster(hlb)
        out16 = {}
        # FP64 first: the MIXED ground state builds an FP64 twin sector for
        # its refinement, whose launch shape is then already tuned
        for prec in ("fp64", "mixed"):
            sec = sb.sector(8, 8, prec)
            v = rand_vec(sec.dim, prec, 2)
            dv = DeviceArray(v.shape, v.dtype).upload(v)
            dh = DeviceArray(v.shape, v.dtype)
            del v
            sec.hxv_device(dv.ptr, dh.ptr, repeat=2)
            device_synchronize()
            ms = sec.hxv_device(dv.ptr, dh.ptr, repeat=5)
            device_synchronize()
            dv.free()
            dh.free()
            out16[prec] = {"ms_per_hxv": ms, "gstates_per_s": sec.dim / (ms * 1e-3) / 1e9,
                           "hbm_gbs_algorithmic": 2 * (16 if prec == "fp64" else 8) * sec.dim
                           / (ms * 1e-3) / 1e9}
            log(f"bath16 {prec}: {ms:.3f} ms/apply, dim {sec.dim}")
            # the real-vector product the ground state runs (real Hamiltonian)
            rdt = np.float64 if prec == "fp64" else np.float32
            rv = np.random.default_rng(3).standard_normal(sec.dim, dtype=np.float32).astype(rdt)
            drv = DeviceArray(rv.shape, rdt).upload(rv)
            drh = DeviceArray(rv.shape, rdt)
            del rv
            sec.hxv_device_real(drv.ptr, drh.ptr, repeat=2)
            ms_r = sec.hxv_device_real(drv.ptr, drh.ptr, repeat=5)
            device_synchronize()
            drv.free()
            drh.free()
            rsz = 8 if prec == "fp64" else 4
            out16[prec]["real_vector"] = {
                "ms_per_hxv": ms_r, "gstates_per_s": sec.dim / (ms_r * 1e-3) / 1e9,
                "hbm_gbs_algorithmic": 2 * rsz * sec.dim / (ms_r * 1e-3) / 1e9}
            log(f"bath16 {prec} real vectors: {ms_r:.3f} ms/apply")
            t0 = time.perf_counter()
            e_, _, it_ = sec.lanczos_gs(tol=1e-12, niter=512, want_vec=False)
            out16["gs_" + prec] = {"e0": e_, "seconds": time.perf_counter() - t0, "iters": it_}
            sec.close()
        out16["e0_rel_err_mixed"] = abs(out16["gs_mixed"]["e0"] - out16["gs_fp64"]["e0"]) / abs(
            out16["gs_fp64"]["e0"])
        out16["gs_speedup_mixed"] = out16["gs_fp64"]["seconds"] / out16["gs_mixed"]["seconds"]
        out16["dim"] = 12870 * 12870
        log(f"bath16 GS: mixed {out16['gs_mixed']} fp64 {out16['gs_fp64']}")
        return out16

    # --------------------------------------------------------- Omega -----
    def omega_section():
        po, hlo, hko, _, _ = configs.config_bhz_2x2(nk=args.omega_nk, nfreq=args.omega_nfreq,
                                                   precision="fp64")
        so = VcaSolver(po)
        t0 = time.perf_counter()
        om64 = so.vca_solve(hlo, hko)
        t_solve = time.perf_counter() - t0
        zs = 1j * (2 * np.arange(args.omega_nfreq) + 1) * np.pi / po.beta
        nk = hko.shape[6]
        so.omega_integrand(zs[:64], "fp64")
        so.omega_integrand(zs[:64], "mixed")
        r0 = so.results()
        t0 = time.perf_counter()
        f64 = so.omega_integrand(zs, "fp64")
        t64 = time.perf_counter() - t0
        r1 = so.results()
        t0 = time.perf_counter()
        f32 = so.omega_integrand(zs, "mixed")
        t32 = time.perf_counter() - t0
        r2 = so.results()
        k64 = r1["omega_seconds"] - r0["omega_seconds"]  # log-det kernel time (events)
        k32 = r2["omega_seconds"] - r1["omega_seconds"]
        s64 = (3 - po.nspin) * 2.0 / po.beta * f64.sum()
        s32 = (3 - po.nspin) * 2.0 / po.beta * f32.sum()
        err = abs(s32 - s64) / abs(s64)
        evals = nk * args.omega_nfreq
        fl = logdet_flops(8, 2)  # two spin blocks of 8 (Nlat*Norb = 8)
        log(f"Omega BHZ: {evals} (k,w) evals FP64 {t64:.3f}s (kernel {k64:.3f}s), MIXED "
            f"{t32:.3f}s (kernel {k32:.3f}s); integral rel err mixed {err:.2e}; "
            f"full solve {t_solve:.2f}s Omega={om64:.10f}")
        a64 = evals * fl / k64 / 1e12
        a32 = evals * fl / k32 / 1e12
        return {
            "workload": f"BHZ 2x2 Norb=2 Nspin=2 (2 spin blocks of 8), {nk} k x "
                        f"{args.omega_nfreq} Matsubara frequencies, beta=200",
            "evals": evals, "evals_per_s_fp64": evals / t64, "evals_per_s_mixed": evals / t32,
            "kernel_seconds_fp64": k64, "kernel_seconds_mixed": k32,
            "integral_rel_err_mixed": err, "omega_full_solve": om64,
            "full_solve_seconds": t_solve,
            "roofline": {"bound": "valu", "flops_per_eval": fl,
                         "fp64": {"achieved": a64, "peak": FP64_PEAK_TFLOPS,
                                  "unit": "TFLOP/s", "frac": a64 / FP64_PEAK_TFLOPS},
                         "mixed": {"achieved": a32, "peak": FP32_PEAK_TFLOPS,
                                   "unit": "TFLOP/s", "frac": a32 / FP32_PEAK_TFLOPS}},
        }, (po, hlo, hko, so)

    if args.only == "omega":
        om, _ = omega_section()
        print(json.dumps({"omega": om}), flush=True)
        return
    if args.only == "full":
        print(json.dumps({"full_solve": run_full(cpu=not args.no_cpu)}), flush=True)
        return
    if args.only == "c3":
        print(json.dumps({"batched_2x5": run_c3()}), flush=True)
        return
    if args.only == "c1":
        print(json.dumps({"chain4": run_c1(cpu=not args.no_cpu)}), flush=True)
        return
    if args.only == "bath16":
        print(json.dumps({"bath16": run_bath16()}), flush=True)
        return

    p, hloc, hk, _, _ = configs.config_square_3x4(nk=2)
    s = VcaSolver(p)
    s.vca_set_hcluster(hloc)

    # ---------------------------------------------------------- H x v ----
    def time_hxv(prec):
        sec = s.sector(6, 6, prec)
        v = rand_vec(sec.dim, prec, 1)
        dv = DeviceArray(v.shape, v.dtype).upload(v)
        dh = DeviceArray(v.shape, v.dtype)
        sec.hxv_device(dv.ptr, dh.ptr, repeat=max(1, args.warmup))
        device_synchronize()
        ms = sec.hxv_device(dv.ptr, dh.ptr, repeat=args.steps)  # event-timed on its stream
        device_synchronize()
        dv.free()
        dh.free()
        return sec, ms

    sec64, ms64 = time_hxv("fp64")
    sec32, ms32 = time_hxv("mixed")
    dim = sec64.dim

    def time_hxv_batch(sec, prec, nb=64, steps=20):
        """nb stacked vectors per launch (the form the GF chains run in)."""
        v = np.tile(rand_vec(sec.dim, prec, 5), nb)
        dv = DeviceArray(v.shape, v.dtype).upload(v)
        dh = DeviceArray(v.shape, v.dtype)
        del v
        sec.hxv_device_batch(nb, dv.ptr, dh.ptr, repeat=2)
        device_synchronize()
        ms = sec.hxv_device_batch(nb, dv.ptr, dh.ptr, repeat=steps)
        device_synchronize()
        dv.free()
        dh.free()
        return {"nb": nb, "ms_per_launch": ms, "gstates_per_s": nb * sec.dim / (ms * 1e-3) / 1e9,
                "hbm_gbs_algorithmic": nb * sec.dim * 2 * (16 if prec == "fp64" else 8)
                / (ms * 1e-3) / 1e9}

    log(f"H x v FP64: {ms64 * 1e3:.2f} us/apply; MIXED: {ms32 * 1e3:.2f} us/apply; dim {dim}")
    if args.only == "hxv":
        # calibration kernel for the HBM byte counters: a device copy of
        # 256 MiB (the runtime's copy kernel)
        nbytes = 1 << 28
        a = DeviceArray((nbytes // 8,), np.float64).upload(np.ones(nbytes // 8))
        b = DeviceArray((nbytes // 8,), np.float64)
        from vca_amd import _lib as L
        import ctypes
        L.check(L.load().vca_device_copy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr),
                                         nbytes, 2))
        device_synchronize()
        print(json.dumps({"hxv_fp64_us": ms64 * 1e3, "hxv_mixed_us": ms32 * 1e3, "dim": dim,
                          "copy_bytes_each_way": nbytes}), flush=True)
        return

    batched = {"fp64": time_hxv_batch(sec64, "fp64"), "mixed": time_hxv_batch(sec32, "mixed")}
    log(f"H x v batched (64 vectors): FP64 {batched['fp64']['gstates_per_s']:.1f} Gstates/s, "
        f"MIXED {batched['mixed']['gstates_per_s']:.1f} Gstates/s")

    # --------------------------------------------------- ground state ----
    t0 = time.perf_counter()
    e64, _, it64 = sec64.lanczos_gs(tol=1e-12, niter=512, want_vec=False)
    t_gs64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    emx, _, itmx = sec32.lanczos_gs(tol=1e-12, niter=512, want_vec=False)
    t_gsmx = time.perf_counter() - t0
    log(f"GS FP64 E0={e64:.14f} ({it64} it, {t_gs64:.2f}s); MIXED E0={emx:.14f} "
        f"(refine {itmx} it, {t_gsmx:.2f}s); dE0 rel {abs(emx - e64) / abs(e64):.2e}")

    omega, (po, hlo, hko, so) = omega_section()

    # --------------------------------------------- end-to-end solves -----
    def guarded(name, fn):
        """Optional sections never cost the headline line: a failure is
        logged and recorded instead."""
        try:
            return fn()
        except Exception as exc:
            log(f"{name} failed: {exc}")
            return {"error": str(exc)}

    full = None if args.no_full else guarded("full 3x4 solve",
                                             lambda: run_full(cpu=not args.no_cpu))
    c3 = None if args.no_c3 else guarded("2x5 every-sector solve", run_c3)
    c1 = guarded("4-site chain", lambda: run_c1(cpu=not args.no_cpu))
    bath16 = None if args.no_bath16 else guarded("bath16", run_bath16)

    # ---------------------------------------------------- CPU baseline ---
    cpu = None
    if not args.no_cpu:
        from oracle import vca_ref as R
        from tests._models import to_oracle
        om = R.Model(to_oracle(p), hloc)
        v = np.random.default_rng(3).normal(size=dim).astype(np.complex128)

        def timed(workers, budget):
            n, t0 = 0, time.perf_counter()
            while True:
                om.hxv(6, 6, v, workers=workers)
                n += 1
                if time.perf_counter() - t0 > budget or n >= 20:
                    break
            return n, (time.perf_counter() - t0) / n

        n1, t1 = timed(0, 10.0)
        ncores = len(os.sched_getaffinity(0))
        threads = int(os.environ.get("OMP_NUM_THREADS", ncores))
        threads = max(1, min(threads, ncores))
        nn, tn = timed(threads, 5.0)
        cpu_om = None
        try:
            # the per-k work of the Omega integrand on the CPU: for each
            # sampled frequency M_k = 1 + (Hloc - H_k) G'(z) and log|det M_k|
            # over the 128x128 k-grid (numpy batched matmul + slogdet), with
            # G'(z) from the solved BHZ cluster
            po_o = to_oracle(po)
            hkl = R.hk_to_lso(po_o, hko)
            hll = R.nnn2lso(hlo, po.nlat, po.nspin, po.norb)
            zsamp = 1j * (2 * np.arange(4) + 1) * np.pi / po.beta
            gz = [R.nnn2lso(so.vca_gf_cluster(z), po.nlat, po.nspin, po.norb) for z in zsamp]
            t0 = time.perf_counter()
            for g in gz:
                m = np.eye(po.nlso)[None] + (hll[None] - hkl) @ g[None]
                np.mean(np.linalg.slogdet(m)[1])
            tz = (time.perf_counter() - t0) / len(zsamp)
            cpu_om = {"evals_per_s": hkl.shape[0] / tz, "threads": "numpy default",
                      "sample": f"{len(zsamp)} frequencies x {hkl.shape[0]} k-points of the "
                                "BHZ Omega integrand (16x16 contraction + slogdet per k)"}
        except Exception as exc:  # the CPU Omega sample is informative only
            log(f"CPU Omega sample skipped: {exc}")
        cpu = {"value": dim / t1 / 1e9, "unit": "Gstates/s", "cores": 1, "kind": "port",
               "sample": f"{n1} H x v applications on the 853,776-state sector (6,6), oracle C "
                         f"restatement of the reference loops (-O2 -funroll-loops), one core",
               "all_cores": {"value": dim / tn / 1e9, "cores": threads,
                             "sample": f"{nn} applications, DW-slice split over {threads} "
                                       "OpenMP threads (the reference's MPI decomposition)"},
               "omega": cpu_om}
        log(f"CPU 1-core: {t1:.3f} s/apply; {threads} threads: {tn:.3f} s/apply")

    gst64 = dim / (ms64 * 1e-3) / 1e9
    gst32 = dim / (ms32 * 1e-3) / 1e9
    bytes64 = 2 * 16 * dim
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("traffic_bytes")
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
        "data": "synthetic (seeded random vectors; Hubbard/BHZ/bath model inputs generated "
                "in-repo)",
        "config": {"workload": "hubbard 3x4 square cluster U=8 half-filled sector (6,6), "
                               "853,776 states, H x v FP64", "sector_dim": dim},
        "mixed": {"value": gst32, "ms_per_step": ms32, "dtype": "f32 (complex64)"},
        "batched": batched,
        "ground_state": {"e0_fp64": e64, "iters_fp64": it64, "seconds_fp64": t_gs64,
                         "e0_mixed": emx, "seconds_mixed": t_gsmx,
                         "e0_rel_err_mixed": abs(emx - e64) / abs(e64)},
        "omega": omega,
        "roofline": {"bound": "hbm", "achieved": bytes64 / (ms64 * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if full:
        line["full_solve"] = full
    if c3:
        line["batched_2x5"] = c3
    if c1:
        line["chain4"] = c1
    if bath16:
        line["bath16"] = bath16
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
