"""Benchmark of the VCA solver-core hot path on one MI355X.

Headline workload: the 8-site 2x4 Hubbard cluster with 8 bath levels in
generic geometry (seeded bath_h, bath_v), U=4, half-filled sector
Nup=Ndw=8: 12,870 x 12,870 = 165,636,900 states, the largest single-GPU
workload.  A "step" is one FP64 H x v application on that sector (complex128
vectors, the reference operator's type), with the vectors resident in HBM;
`value` is Gstates/s = dim / (average time of one application, HIP events
on the sector's stream).  Reported beside it in the same JSON line:
  * mixed: the complex64 H x v on the same sector; real_vector: the
    products on dim reals (FP64 and FP32) that the ground state and the
    real-coefficient chains of this real Hamiltonian run on;
  * ground_state: the Lanczos ground state, FP64 and MIXED (FP32 Lanczos +
    FP64 refinement), with the E0 error of MIXED against FP64;
  * roofline: the FP64 H x v against HBM (algorithmic bytes: one read of v
    and one write of Hv, 2 x 16 B per state) and against the FP64 VALU peak
    (algorithmic flops: 4 per hop and 4 for the diagonal per state, real
    amplitudes on complex vectors); traffic = HBM bytes per application from
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/hxv_traffic.json),
    used only when the stored launch shape is the one this run picked;
  * square3x4: the 12-site 3x4 Hubbard sector (6,6), 853,776 states: H x v
    FP64 and MIXED, 64 vectors per launch (the form the chains run in), the
    ground state;
  * omega: (k, omega) log-det evaluations per second of the Omega integrand
    on the 2x2 BHZ cluster (Norb=2, Nspin=2: two spin blocks of 8),
    128x128 k-grid x 8192 Matsubara frequencies at beta=200, FP64 and MIXED,
    the MIXED error of the integral, and the roofline of the log-det kernel;
  * full_solve: vca_solve of the 3x4 cluster end to end (ground state, the
    excitation chains of 300 steps on the N+-1 sectors, Omega on the 64x64
    k-grid) FP64 and MIXED, then G, Sigma and G0 on the default grids;
  * batched_2x5: vca_solve of the 10-site 2x5 cluster, every (Nup, Ndw)
    sector, ARPACK-equivalent eigenpairs per sector (the 77 sectors above
    the dense threshold in one batched eigensolve), all 760 chains as one
    chain set (two launches per Lanczos step for every chain);
  * chain4: the 4-site chain end to end (every sector, Matsubara-sum Omega
    at beta=100 over 4096 frequencies, G on 512 real-axis points), on the
    GPU and by the oracle on one host core, with the differences;
  * cpu_baseline: the oracle's C restatement of the reference H x v loops
    (-O2 -funroll-loops) on the headline sector, timed on a bounded sample:
    one of 256 slices of the reference's DW decomposition on one core (the
    local and up parts of 51 output columns and the down part of 51 up
    rows, 1/256 of one application), extrapolated to Gstates/s; all host
    cores on as many slices of a 512-way split; the 3x4 sector on one core
    and all cores; the Omega integrand (numpy) on a sample of frequencies.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
                    [--only c4|hxv|omega|full|c3|c1|square3x4|c4_solve]
                    [--c4-form plain|fp64|mixed] [--save F.npz] [--gs-residual R]
                    [--compare A.npz B.npz]
                    [--no-full] [--no-c3] [--no-cpu]
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


def rand_vec(n, prec, seed):
    """Seeded complex vector (complex128 for fp64, complex64 for mixed)."""
    import numpy as np
    x = np.random.default_rng(seed).random(2 * n, dtype=np.float32)
    x -= 0.5
    v = x.view(np.complex64)
    return v.astype(np.complex128) if prec == "fp64" else v


def stored_traffic(key, variant):
    """HBM bytes per application stored for workload `key` by the profiler
    runs, with a note; None when absent or measured on another launch shape."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, "no stored counters"
    with open(TRAFFIC_FILE) as f:
        rec = json.load(f).get(key)
    if rec is None:
        return None, "no stored counters for this workload"
    if rec.get("variant") != variant:
        return None, (f"stale: counters of launch shape {rec.get('variant')}, this run picked "
                      f"{variant}")
    return rec["traffic_bytes"], f"rocprofv3 FETCH_SIZE/WRITE_SIZE of launch shape {variant}"


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
                      "1..63,504), up to 6 eigenpairs per sector (ARPACK-equivalent; the "
                      "small sectors in one batched eigensolve), the 760 chains as one chain "
                      "set (two launches per Lanczos step for all of them)",
          "seconds": time.perf_counter() - t0, "omega": omc,
          "sector_threads": int(os.environ.get("VCA_SECTOR_THREADS", "16")),
          "lanczos_step_gstates_per_s": r["hxv_states"] / max(r["hxv_seconds"], 1e-30) / 1e9,
          **{k: r[k] for k in ("omegaprime", "nstates", "nchains", "hxv_count",
                               "t_diag", "t_gf", "t_omega")}}
    log(f"2x5 every-sector solve: {c3['seconds']:.1f}s Omega={omc:.10f} "
        f"(diag {r['t_diag']:.2f}s, GF {r['t_gf']:.2f}s, {r['nchains']} chains)")
    return c3


C4_FORMS = {
    # the plain FP64 GPU form: complex128 vectors everywhere, one chain per
    # launch, every Lanczos step launched eagerly (no graphs)
    "plain": ("fp64", {"VCA_CHAIN_BATCH": "1", "VCA_NO_GRAPHS": "1", "VCA_REAL_CHAINS": "0",
                       "VCA_REAL_GS": "0", "VCA_DERIVED_CHANNELS": "0"}),
    # FP64 with the batching form: chains of a target sector in one launch,
    # 16-step graphs, real vectors where the operator and state are real,
    # the c_i - i c_j channels from the diagonal chains
    "fp64": ("fp64", {}),
    # the batched mixed-precision form: FP32 ground-state Lanczos refined in
    # FP64, FP32 chains, FP32 log-dets on the matrix cores
    "mixed": ("mixed", {}),
    # the same with FP64 chains (G' at FP64 accuracy)
    "mixed_fp64_chains": ("mixed", {"VCA_CHAIN_PRECISION": "fp64"}),
    # FP32 only in the ground-state Lanczos (refined in FP64): FP64 chains
    # and FP64 log-dets
    "mixed_gs": ("mixed", {"VCA_CHAIN_PRECISION": "fp64", "VCA_OMEGA_PRECISION": "fp64"}),
}


def run_c4_solve(form, save=None, gs_residual=None):
    """vca_solve of the 16-site bath cluster end to end: the Lanczos ground
    state of the 165,636,900-state sector (8,8), the 240 excitation chains of
    200 steps on the 147,232,800-state sectors (9,8) and (7,8), and Omega on
    the 16x16 k-grid, in one of C4_FORMS; with save, E0, Omega and G'_ij at 64
    Matsubara and 64 real-axis points go to an .npz for the comparison of
    two forms (each form needs minutes, so they run as separate commands).
    gs_residual sets VCA_GS_RESIDUAL: the FP64 ground-state Lanczos then also
    stops on the Ritz residual, not on |dE| alone."""
    import numpy as np
    from vca_amd import configs
    from vca_amd.api import VcaSolver

    prec, env = C4_FORMS[form]
    env = dict(env)
    if gs_residual is not None:
        env["VCA_GS_RESIDUAL"] = repr(float(gs_residual))
    os.environ.update(env)
    os.environ.setdefault("VCA_PROGRESS", "1")
    p, hloc, hk, bh, bv = configs.config_bath_2x4(nk=16, precision=prec)
    s = VcaSolver(p)
    log(f"C4 solve, form {form} ({prec}; {env or 'default knobs'})")
    t0 = time.perf_counter()
    om = s.vca_solve(hloc, hk, bh, bv)
    secs = time.perf_counter() - t0
    r = s.results()
    out = {"workload": "8-site 2x4 Hubbard cluster + 8 bath levels, U=4: ground state of "
                       "sector (8,8) (165,636,900 states), 240 excitation chains of 200 steps "
                       "on sectors (9,8) and (7,8) (147,232,800 states each), Omega on the "
                       "16x16 k-grid",
           "form": form, "precision": prec, "knobs": env, "seconds": secs, "sft": om,
           **{k: r[k] for k in ("omegaprime", "omega_integral", "e0_mixed", "nstates",
                                "nchains", "nchains_derived", "hxv_count", "t_diag", "t_gf",
                                "t_omega",
                                "omega_evals")},
           "hxv_gstates_per_s": r["hxv_states"] / max(r["hxv_seconds"], 1e-30) / 1e9}
    log(f"C4 solve {form}: {secs:.1f} s (diag {r['t_diag']:.1f}, chains {r['t_gf']:.1f}, Omega "
        f"{r['t_omega']:.2f}), {r['nchains']} chains, Omega'={r['omegaprime']:.12f}, "
        f"SFT={om:.12f}")
    if save:
        wm = np.pi / 100.0 * (2 * np.arange(64) + 1)
        wr = np.linspace(-6.0, 6.0, 64)
        gm = np.stack([s.vca_gf_cluster(1j * w)[:, :, 0, 0, 0, 0] for w in wm])
        gr = np.stack([s.vca_gf_cluster(w + 0.05j)[:, :, 0, 0, 0, 0] for w in wr])
        np.savez(save, gm=gm, gr=gr, wm=wm, wr=wr, omegaprime=r["omegaprime"],
                 omega_integral=r["omega_integral"], sft=om)
        out["saved"] = save
    return out


def compare_c4_solves(ref_path, path):
    """Differences of a C4 solve saved by run_c4_solve against another (the
    reference form): E0, Omega, and each G'_ij against its own largest value
    over the points."""
    import numpy as np
    a, b = np.load(ref_path), np.load(path)
    out = {"e0_rel_err": float(abs(b["omegaprime"] - a["omegaprime"]) / abs(a["omegaprime"])),
           "sft_rel_err": float(abs(b["sft"] - a["sft"]) / abs(a["sft"])),
           "omega_integral_rel_err": float(abs(b["omega_integral"] - a["omega_integral"]) /
                                           abs(a["omega_integral"]))}
    for key in ("gm", "gr"):
        scale = np.maximum(np.abs(a[key]).max(axis=0), 1e-8 * np.abs(a[key]).max())
        out[f"{key}_max_rel_err"] = float((np.abs(b[key] - a[key]) / scale).max())
    return out


def c4_section(args, cpu, gs=True):
    """The headline: H x v on the 165,636,900-state bath sector, FP64 and
    MIXED, complex and real vectors, the ground state both ways (gs), and
    the oracle's CPU sample (cpu)."""
    import numpy as np
    from vca_amd import configs
    from vca_amd.api import DeviceArray, VcaSolver, device_synchronize

    pb, hlb, hkb, bhb, bvb = configs.config_bath_2x4(nk=2, precision="fp64")
    sb = VcaSolver(pb)
    sb.vca_set_bath(bhb, bvb)
    sb.vca_set_hcluster(hlb)
    out = {"workload": "8-site 2x4 Hubbard cluster + 8 bath levels (seeded generic bath_h, "
                       "bath_v), U=4, half-filled sector (8,8): 12,870 x 12,870 = 165,636,900 "
                       "states; H x v on complex128 (FP64) vectors resident in HBM"}
    vhost = None
    secs = {}
    for prec in ("fp64", "mixed"):
        sec = sb.sector(8, 8, prec)
        secs[prec] = sec
        v = rand_vec(sec.dim, prec, 2)
        dv = DeviceArray(v.shape, v.dtype).upload(v)
        dh = DeviceArray(v.shape, v.dtype)
        if prec == "fp64":
            vhost = v
        del v
        sec.hxv_device(dv.ptr, dh.ptr, repeat=max(1, args.warmup))
        device_synchronize()
        ms = sec.hxv_device(dv.ptr, dh.ptr, repeat=args.steps)  # events on its stream
        device_synchronize()
        dv.free()
        dh.free()
        esz = 16 if prec == "fp64" else 8
        out[prec] = {"ms_per_hxv": ms, "gstates_per_s": sec.dim / (ms * 1e-3) / 1e9,
                     "hbm_gbs_algorithmic": 2 * esz * sec.dim / (ms * 1e-3) / 1e9,
                     "variant": sec.variant()[0]}
        log(f"C4 H x v {prec}: {ms:.3f} ms/apply ({out[prec]['gstates_per_s']:.2f} Gstates/s)")
        # the real-vector product (real Hamiltonian): the ground state's form
        rdt = np.float64 if prec == "fp64" else np.float32
        rv = np.random.default_rng(3).random(sec.dim, dtype=np.float32).astype(rdt)
        drv = DeviceArray(rv.shape, rdt).upload(rv)
        drh = DeviceArray(rv.shape, rdt)
        del rv
        sec.hxv_device_real(drv.ptr, drh.ptr, repeat=max(1, args.warmup))
        ms_r = sec.hxv_device_real(drv.ptr, drh.ptr, repeat=args.steps)
        device_synchronize()
        drv.free()
        drh.free()
        out[prec]["real_vector"] = {
            "ms_per_hxv": ms_r, "gstates_per_s": sec.dim / (ms_r * 1e-3) / 1e9,
            "hbm_gbs_algorithmic": 2 * (esz // 2) * sec.dim / (ms_r * 1e-3) / 1e9}
        log(f"C4 H x v {prec} real vectors: {ms_r:.3f} ms/apply")
    sec = secs["fp64"]
    dim = sec.dim
    # algorithmic flops per state: real amplitudes on complex vectors, 4 per
    # hop (two FMAs) and 4 for the diagonal
    nnz = sum(sec.spin_coo(s)[0].size for s in (1, 2))
    hops_per_state = nnz / sec.dimup  # DimUp = DimDw
    flops_state = 4.0 * (1.0 + hops_per_state)
    out["hops_per_state"] = hops_per_state
    out["flops_per_state"] = flops_state
    for prec in ("fp64", "mixed") if gs else ():
        t0 = time.perf_counter()
        e_, _, it_ = secs[prec].lanczos_gs(tol=1e-12, niter=512, want_vec=False)
        out["gs_" + prec] = {"e0": e_, "seconds": time.perf_counter() - t0, "iters": it_}
    if gs:
        out["e0_rel_err_mixed"] = abs(out["gs_mixed"]["e0"] - out["gs_fp64"]["e0"]) / abs(
            out["gs_fp64"]["e0"])
        out["gs_speedup_mixed"] = out["gs_fp64"]["seconds"] / out["gs_mixed"]["seconds"]
        log(f"C4 GS: FP64 {out['gs_fp64']} MIXED {out['gs_mixed']} "
            f"dE0 {out['e0_rel_err_mixed']:.1e}")
    for s_ in secs.values():
        s_.close()
    ms64 = out["fp64"]["ms_per_hxv"]
    achieved = 32.0 * dim / (ms64 * 1e-3) / 1e9
    tflops = flops_state * dim / (ms64 * 1e-3) / 1e12
    traffic, tnote = stored_traffic("bath16_fp64", out["fp64"]["variant"])
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": tnote,
            "per": "one H x v application (all its launches)",
            "valu": {"achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP64_PEAK_TFLOPS, "flops_per_state": flops_state}}
    cpu_rec = None
    if cpu:
        from oracle import vca_ref as R
        from tests._models import oracle_model
        om = oracle_model(pb, hlb, bhb, bvb)
        hv = np.zeros(dim, np.complex128)
        t0 = time.perf_counter()
        om.hxv_slices(8, 8, vhost, hv, 256, 0, 1, 1)
        t1 = time.perf_counter() - t0
        ncores = len(os.sched_getaffinity(0))
        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", ncores)), ncores))
        t0 = time.perf_counter()
        om.hxv_slices(8, 8, vhost, hv, 512, 0, threads, threads)
        tn = time.perf_counter() - t0
        cpu_rec = {"value": dim / 256 / t1 / 1e9, "unit": "Gstates/s", "cores": 1,
                   "kind": "port",
                   "sample": "1/256 of one H x v on the 165,636,900-state sector: one slice of "
                             "the reference's DW decomposition (local and up parts of 51 output "
                             "columns, down part of 51 up rows), oracle C restatement of the "
                             f"reference loops (-O2 -funroll-loops), one core, {t1:.1f} s; "
                             "extrapolated to a whole application",
                   "all_cores": {"value": dim * threads / 512 / tn / 1e9, "cores": threads,
                                 "sample": f"{threads} slices of a 512-way DW split on "
                                           f"{threads} OpenMP threads, {tn:.1f} s"}}
        log(f"C4 CPU: 1 core {t1:.1f} s per 1/256 application; {threads} threads {tn:.1f} s "
            f"per {threads}/512")
    return out, roof, cpu_rec


def square3x4_section(args, cpu):
    """The 3x4 sector (6,6): single and batched products, the ground state."""
    import numpy as np
    from vca_amd import configs
    from vca_amd.api import DeviceArray, VcaSolver, device_synchronize

    p, hloc, hk, _, _ = configs.config_square_3x4(nk=2)
    s = VcaSolver(p)
    s.vca_set_hcluster(hloc)
    out = {"workload": "12-site 3x4 square cluster U=8, half-filled sector (6,6), 853,776 "
                       "states"}
    secs = {}
    for prec in ("fp64", "mixed"):
        sec = s.sector(6, 6, prec)
        secs[prec] = sec
        v = rand_vec(sec.dim, prec, 1)
        dv = DeviceArray(v.shape, v.dtype).upload(v)
        dh = DeviceArray(v.shape, v.dtype)
        sec.hxv_device(dv.ptr, dh.ptr, repeat=max(1, args.warmup))
        device_synchronize()
        ms = sec.hxv_device(dv.ptr, dh.ptr, repeat=max(args.steps, 200))
        device_synchronize()
        dv.free()
        dh.free()
        esz = 16 if prec == "fp64" else 8
        out[prec] = {"ms_per_hxv": ms, "gstates_per_s": sec.dim / (ms * 1e-3) / 1e9,
                     "hbm_gbs_algorithmic": 2 * esz * sec.dim / (ms * 1e-3) / 1e9,
                     "variant": sec.variant()[0]}
        # 64 stacked vectors per launch: the form the chains run in
        nb = 64
        vb = np.tile(rand_vec(sec.dim, prec, 5), nb)
        dvb = DeviceArray(vb.shape, vb.dtype).upload(vb)
        dhb = DeviceArray(vb.shape, vb.dtype)
        del vb
        sec.hxv_device_batch(nb, dvb.ptr, dhb.ptr, repeat=2)
        msb = sec.hxv_device_batch(nb, dvb.ptr, dhb.ptr, repeat=20)
        device_synchronize()
        dvb.free()
        dhb.free()
        out[prec]["batched"] = {"nb": nb, "ms_per_launch": msb,
                                "gstates_per_s": nb * sec.dim / (msb * 1e-3) / 1e9,
                                "hbm_gbs_algorithmic": nb * sec.dim * 2 * esz / (msb * 1e-3)
                                / 1e9}
        log(f"3x4 H x v {prec}: {ms * 1e3:.2f} us/apply, batched 64: "
            f"{out[prec]['batched']['gstates_per_s']:.1f} Gstates/s")
    dim = secs["fp64"].dim
    for prec in ("fp64", "mixed"):
        t0 = time.perf_counter()
        e_, _, it_ = secs[prec].lanczos_gs(tol=1e-12, niter=512, want_vec=False)
        out["gs_" + prec] = {"e0": e_, "seconds": time.perf_counter() - t0, "iters": it_}
    out["e0_rel_err_mixed"] = abs(out["gs_mixed"]["e0"] - out["gs_fp64"]["e0"]) / abs(
        out["gs_fp64"]["e0"])
    traffic, tnote = stored_traffic("square3x4_fp64", out["fp64"]["variant"])
    a = 32.0 * dim / (out["fp64"]["ms_per_hxv"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": a / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": tnote}
    if cpu:
        from tests._models import oracle_model
        om = oracle_model(p, hloc)
        vv = np.random.default_rng(3).normal(size=dim).astype(np.complex128)

        def timed(workers, budget):
            n, t0 = 0, time.perf_counter()
            while True:
                om.hxv(6, 6, vv, workers=workers)
                n += 1
                if time.perf_counter() - t0 > budget or n >= 20:
                    break
            return n, (time.perf_counter() - t0) / n

        n1, t1 = timed(0, 3.0)
        ncores = len(os.sched_getaffinity(0))
        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", ncores)), ncores))
        nn, tn = timed(threads, 3.0)
        out["cpu_baseline"] = {
            "value": dim / t1 / 1e9, "unit": "Gstates/s", "cores": 1, "kind": "port",
            "sample": f"{n1} H x v applications, oracle C loops, one core",
            "all_cores": {"value": dim / tn / 1e9, "cores": threads,
                          "sample": f"{nn} applications, DW-slice split over {threads} "
                                    "OpenMP threads"}}
    for s_ in secs.values():
        s_.close()
    return out


def omega_section(args, cpu):
    """Omega integrand evaluations per second on the 2x2 BHZ cluster."""
    import numpy as np
    from vca_amd import configs
    from vca_amd.api import VcaSolver

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
    # the same FP32 log-dets with the contraction on the vector ALUs
    os.environ["VCA_OMEGA_MFMA"] = "0"
    try:
        so.omega_integrand(zs[:64], "mixed")
        r3 = so.results()
        t0 = time.perf_counter()
        f32v = so.omega_integrand(zs, "mixed")
        t32v = time.perf_counter() - t0
        k32v = so.results()["omega_seconds"] - r3["omega_seconds"]
    finally:
        os.environ.pop("VCA_OMEGA_MFMA", None)
    errv = abs((3 - po.nspin) * 2.0 / po.beta * f32v.sum() - (3 - po.nspin) * 2.0 / po.beta
               * f64.sum()) / abs((3 - po.nspin) * 2.0 / po.beta * f64.sum())
    s64 = (3 - po.nspin) * 2.0 / po.beta * f64.sum()
    s32 = (3 - po.nspin) * 2.0 / po.beta * f32.sum()
    err = abs(s32 - s64) / abs(s64)
    evals = nk * args.omega_nfreq
    fl = logdet_flops(8, 2)  # two spin blocks of 8 (Nlat*Norb = 8)
    log(f"Omega BHZ: {evals} (k,w) evals FP64 {t64:.3f}s (kernel {k64:.3f}s), MIXED "
        f"{t32:.3f}s (kernel {k32:.3f}s, vector-ALU contraction {k32v:.3f}s); integral rel err "
        f"mixed {err:.2e} (vector ALU {errv:.2e}); full solve {t_solve:.2f}s Omega={om64:.10f}")
    a64 = evals * fl / k64 / 1e12
    a32 = evals * fl / k32 / 1e12
    out = {
        "workload": f"BHZ 2x2 Norb=2 Nspin=2 (2 spin blocks of 8), {nk} k x "
                    f"{args.omega_nfreq} Matsubara frequencies, beta=200",
        "evals": evals, "evals_per_s_fp64": evals / t64, "evals_per_s_mixed": evals / t32,
        "kernel_seconds_fp64": k64, "kernel_seconds_mixed": k32,
        "integral_rel_err_mixed": err,
        "mixed_vector_alu": {"evals_per_s": evals / t32v, "kernel_seconds": k32v,
                             "integral_rel_err": errv,
                             "note": "FP32 log-dets with the contraction on the vector ALUs "
                                     "(k_logdet<float, 8>) instead of the matrix cores "
                                     "(k_logdet_lane8<8, 8>)"},
        "omega_full_solve": om64,
        "full_solve_seconds": t_solve,
        "roofline": {"flops_per_eval": fl,
                     "fp64": {"bound": "valu", "achieved": a64, "peak": FP64_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": a64 / FP64_PEAK_TFLOPS},
                     "mixed": {"bound": "mfma", "achieved": a32, "peak": FP32_PEAK_TFLOPS,
                               "unit": "TFLOP/s", "frac": a32 / FP32_PEAK_TFLOPS,
                               "note": "contraction (8 nb^3 per block) on the FP32 matrix "
                                       "cores, LU (8 nb^3 / 3) on the vector ALUs; peak: "
                                       "dense FP32 MFMA"}},
    }
    if cpu:
        try:
            from oracle import vca_ref as R
            from tests._models import to_oracle
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
            out["cpu_baseline"] = {
                "evals_per_s": hkl.shape[0] / tz, "threads": "numpy default", "kind": "port",
                "sample": f"{len(zsamp)} frequencies x {hkl.shape[0]} k-points of the BHZ "
                          "Omega integrand (16x16 contraction + slogdet per k)"}
        except Exception as exc:  # the CPU Omega sample is informative only
            log(f"CPU Omega sample skipped: {exc}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-full", action="store_true", help="skip the end-to-end 3x4 solve")
    ap.add_argument("--no-c3", action="store_true", help="skip the every-sector 2x5 solve")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--omega-nk", type=int, default=128)
    ap.add_argument("--omega-nfreq", type=int, default=8192)
    ap.add_argument("--c4-form", choices=sorted(C4_FORMS), default="mixed",
                    help="the form of the c4_solve run")
    ap.add_argument("--save", default=None, help="c4_solve: .npz of E0, Omega and G'")
    ap.add_argument("--gs-residual", type=float, default=None,
                    help="c4_solve: Ritz-residual bound of the FP64 ground state (VCA_GS_RESIDUAL)")
    ap.add_argument("--compare", nargs=2, metavar=("REF", "OTHER"), default=None,
                    help="the differences of two saved c4_solve runs (no GPU)")
    ap.add_argument("--only", choices=["all", "c4", "hxv", "omega", "full", "c3", "c1",
                                       "square3x4", "c4_solve"],
                    default="all",
                    help="c4: the headline H x v and ground state on the 165,636,900-state "
                         "sector; hxv: the 3x4 and 165,636,900-state H x v plus a 256 MiB "
                         "copy that calibrates the HBM byte counters; omega: the Omega "
                         "integrand; full: the end-to-end 3x4 solve; c3: the every-sector "
                         "2x5 solve; c1: the 4-site chain; square3x4: the 3x4 sector; "
                         "c4_solve: the 16-site bath cluster end to end in --c4-form")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); how the path would "
              "shard over several GPUs is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)
    cpu = not args.no_cpu

    if args.compare:
        print(json.dumps(compare_c4_solves(*args.compare)), flush=True)
        return
    if args.only == "c4_solve":
        print(json.dumps({"c4_solve": run_c4_solve(args.c4_form, args.save, args.gs_residual)}), flush=True)
        return
    if args.only == "omega":
        print(json.dumps({"omega": omega_section(args, cpu)}), flush=True)
        return
    if args.only == "full":
        print(json.dumps({"full_solve": run_full(cpu=cpu)}), flush=True)
        return
    if args.only == "c3":
        print(json.dumps({"batched_2x5": run_c3()}), flush=True)
        return
    if args.only == "c1":
        print(json.dumps({"chain4": run_c1(cpu=cpu)}), flush=True)
        return
    if args.only == "square3x4":
        print(json.dumps({"square3x4": square3x4_section(args, cpu)}), flush=True)
        return
    if args.only == "hxv":
        # the H x v products only, then a device copy of 256 MiB that
        # calibrates the HBM byte counters (the runtime's copy kernel)
        import ctypes
        import numpy as np
        from vca_amd import _lib as L
        from vca_amd.api import DeviceArray, device_synchronize
        sq = square3x4_section(args, False)
        c4, _, _ = c4_section(args, False, gs=False)
        nbytes = 1 << 28
        a = DeviceArray((nbytes // 8,), np.float64).upload(np.ones(nbytes // 8))
        b = DeviceArray((nbytes // 8,), np.float64)
        L.check(L.load().vca_device_copy(ctypes.c_void_p(b.ptr), ctypes.c_void_p(a.ptr),
                                         nbytes, 2))
        device_synchronize()
        print(json.dumps({"square3x4_fp64_variant": sq["fp64"]["variant"],
                          "bath16_fp64_variant": c4["fp64"]["variant"],
                          "copy_bytes_each_way": nbytes}), flush=True)
        return

    c4, roof, cpu_c4 = c4_section(args, cpu)
    if args.only == "c4":
        print(json.dumps({"c4": c4, "roofline": roof, "cpu_baseline": cpu_c4}), flush=True)
        return

    def guarded(name, fn):
        """Optional sections never cost the headline line: a failure is
        logged and recorded instead."""
        try:
            return fn()
        except Exception as exc:
            log(f"{name} failed: {exc}")
            return {"error": str(exc)}

    sq = guarded("3x4 sector", lambda: square3x4_section(args, cpu))
    omega = guarded("Omega", lambda: omega_section(args, cpu))
    full = None if args.no_full else guarded("full 3x4 solve", lambda: run_full(cpu=cpu))
    c3 = None if args.no_c3 else guarded("2x5 every-sector solve", run_c3)
    c1 = guarded("4-site chain", lambda: run_c1(cpu=cpu))

    ms64 = c4["fp64"]["ms_per_hxv"]
    line = {
        "metric": "H×v Gstates/s + Ω evals/s, mixed vs FP64 on 1 GPU at ΔE0<1e-10",
        "value": c4["fp64"]["gstates_per_s"],
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
        "config": {"workload": "8-site 2x4 Hubbard cluster + 8 bath levels, U=4, half-filled "
                               "sector (8,8), 165,636,900 states, H x v FP64 (complex128)",
                   "sector_dim": 165636900},
        "mixed": {"value": c4["mixed"]["gstates_per_s"], "ms_per_step": c4["mixed"]["ms_per_hxv"],
                  "dtype": "f32 (complex64)"},
        "real_vector": {"fp64": c4["fp64"]["real_vector"], "fp32": c4["mixed"]["real_vector"]},
        "ground_state": {"e0_fp64": c4["gs_fp64"]["e0"], "iters_fp64": c4["gs_fp64"]["iters"],
                         "seconds_fp64": c4["gs_fp64"]["seconds"],
                         "e0_mixed": c4["gs_mixed"]["e0"],
                         "seconds_mixed": c4["gs_mixed"]["seconds"],
                         "e0_rel_err_mixed": c4["e0_rel_err_mixed"],
                         "speedup_mixed": c4["gs_speedup_mixed"]},
        "omega": omega,
        "roofline": roof,
        "cpu_baseline": cpu_c4,
        "bath16": c4,
        "square3x4": sq,
    }
    if full:
        line["full_solve"] = full
    if c3:
        line["batched_2x5"] = c3
    if c1:
        line["chain4"] = c1
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
