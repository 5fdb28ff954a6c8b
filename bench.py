"""
Benchmark: two-site TDVP sweep time on the XXZ chain (L=100, D=1024), mixed
precision versus the all-FP64 GPU path, with the state fidelity and energy
error between them.

Workload (fixed-D mode of the two-site TDVP workload): XXZ chain L=100, J=1,
Delta=0.8, h=-0.1, quantum numbers zeroed; random complex128 MPS from a fixed
seed with bonds min(D, 2^i, 2^(L-i)), left-orthonormalised; dt=-0.05j,
numiter_lanczos=8, tol_split=1e-10, bond cap D. A step is one TDVP time step
(one left-to-right plus right-to-left sweep) after the setup (right
orthonormalisation and environment blocks), with everything resident in HBM.

    python bench.py [--steps K] [--warmup W] [--L 100] [--D 1024] [--mixed fp32|bf16x6|bf16x3]

Prints one JSON line. `value` is the mixed-precision sweep time (s/sweep,
lower is better; by default the contractions run as three-term split-bf16
products with FP32 accuracy); the FP64 sweep time, the speed-up, |1 - fidelity| and the
relative energy error are extra fields. `roofline` is measured live, over the
timed sweeps, for the kernel with the largest share of the timed kernel time
among those with a roofline (HIP events attached to every launch on the
stream it runs on); `cpu_baseline` times the oracle (NumPy restatement of the
reference) on a bounded sample of the same workload.
"""

import argparse
import copy
import json
import os
import sys
import time

import numpy as np

T_START = time.perf_counter()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("2-site TDVP sweep time at D=1024: mixed vs FP64 GPU at |1−fidelity|≤1e-8")
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "bf16x6": 2500.0}
MIXED_DTYPE = {"bf16x6": "bf16x6 (bf16 MFMA on three-term splits of f32 operands, f32 accumulation)",
               "bf16x3": "bf16x3 (bf16 MFMA on two-term splits of f32 operands, f32 accumulation)",
               "fp32": "f32"}


def heff_flops(Dl, d, Dr, Ml, Mr):
    """Algorithmic flops of one effective-Hamiltonian application (8 per complex MAC):
    A.R, W mixing, L^T.(...)."""
    return 8.0 * (Dl * d * Dr * Mr * Dr + Ml * d * d * Mr * Dl * Dr + Dl * Ml * Dl * d * Dr)


def build_state(L, D, seed):
    import pytenet_amd as ptn
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [min(D, 2**min(i, 40), 2**min(L - i, 40)) for i in range(L + 1)]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random",
                  rng=np.random.default_rng(seed))
    psi.to_device()
    psi.orthonormalize(mode="left")
    return H, psi


def prewarm():
    """Load every kernel once on a tiny chain (not a timed or warmup step)."""
    import pytenet_amd as ptn
    H, psi = build_state(6, 8, 0)
    for prec in ("fp64", "fp32", "bf16x3", "bf16x6"):
        p = copy.deepcopy(psi)
        ptn.tdvp_twosite(H, p, -0.05j, 1, numiter_lanczos=4, tol_split=1e-10, precision=prec,
                         max_bond=8)


def run_leg(H, psi0, prec, warmup, steps, D, numiter, svd_precond32=None, decide=None):
    """One precision leg of the headline: setup, `warmup` untimed sweeps, then
    timed sweeps with the live launch timing of the complex products switched
    on. With `decide`, the number of timed sweeps is decide(t) for t = the
    duration of the first sweep run (warmup or timed), capped by `steps`.
    Returns (psi, seconds per timed sweep, timed sweeps, per-kernel timings)."""
    import torch
    import pytenet_amd as ptn
    from pytenet_amd import backend
    psi = copy.deepcopy(psi0)
    run = ptn.TwoSiteTDVP(H, psi, -0.05j, numiter, 1e-10, precision=prec, max_bond=D,
                          svd_precond32=svd_precond32)
    run.setup()
    torch.cuda.synchronize()
    first = None
    for k in range(warmup):
        t0 = time.perf_counter()
        run.sweep()
        torch.cuda.synchronize()
        first = first or time.perf_counter() - t0
        print(f"[bench] {prec}: warmup sweep {k + 1}/{warmup} {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    if decide is not None and first is not None:
        steps = max(1, min(steps, decide(first)))
    backend.timing_begin()
    t0 = time.perf_counter()
    k = 0
    while k < steps:
        t1 = time.perf_counter()
        run.sweep()
        k += 1
        if decide is not None and first is None:
            torch.cuda.synchronize()
            first = time.perf_counter() - t1
            steps = max(1, min(steps, decide(first)))
        print(f"[bench] {prec}: sweep {k}/{steps} done, bonds max {max(psi.bond_dims)}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    timings = backend.timing_end()
    run.finish()
    del run
    return psi, dt, steps, timings


def _kernel_bound(name):
    """(bound, peak, unit) of a timed kernel instance: the complex products
    and the Jacobi Gram kernel are MFMA-bound (8 flops per complex
    multiply-add, the peak of their arithmetic), the Jacobi apply kernel
    streams the rotated columns (HBM); the rotation solve is latency-bound and
    has no roofline."""
    if name.startswith("zgemm_bf16x3") or name.startswith("zgemm_bf16x6"):
        return "mfma", PEAK_TFLOPS["bf16x3"], "TFLOP/s"
    if (name.startswith("zgemm_kernel<float") or name.startswith("zgemm_grouped_kernel<float")
            or name == "k_jac_gram<float>"):
        return "mfma", PEAK_TFLOPS["fp32"], "TFLOP/s"
    if (name.startswith("zgemm_kernel<double") or name.startswith("zgemm_grouped_kernel<double")
            or name == "k_jac_gram<double>"):
        return "mfma", PEAK_TFLOPS["fp64"], "TFLOP/s"
    if name.startswith("k_jac_apply") or name.startswith("k_mix_sectors"):
        return "hbm", HBM_PEAK_GBS, "GB/s"
    return None, None, None


def kernel_roofline(timings):
    """Roofline entry of the dominant timed kernel instance of a leg (largest
    summed device time among the kernels with a roofline): algorithmic flops
    (or bytes) of its launches / their summed device time."""
    rows = [t for t in timings if _kernel_bound(t["name"])[0]]
    if not rows:
        return None
    top = rows[0]
    bound, peak, unit = _kernel_bound(top["name"])
    work = top["flop"] / 1e12 if bound == "mfma" else top["bytes"] / 1e9
    ach = work / (top["total_ms"] * 1e-3)
    total = sum(t["total_ms"] for t in timings)
    return {"kernel": top["name"], "bound": bound, "achieved": round(ach, 2), "peak": peak,
            "unit": unit, "frac": round(ach / peak, 4), "launches": top["launches"],
            "avg_us": round(top["avg_ms"] * 1e3, 2),
            "flop_per_launch": top["flop"] / top["launches"],
            "bytes_per_launch": top["bytes"] / top["launches"],
            "share_of_timed_kernel_time": round(top["total_ms"] / total, 4),
            "source": "HIP events around every launch on its stream, over the timed sweeps"}


def top_kernels(timings, n=8, rates=False):
    """The n timed kernel instances with the largest summed device time; with
    `rates`, each with its achieved rate and fraction of its roofline."""
    total = sum(t["total_ms"] for t in timings) or 1.0
    out = []
    for t in timings[:n]:
        row = {"name": t["name"], "launches": t["launches"], "total_ms": round(t["total_ms"], 1),
               "share": round(t["total_ms"] / total, 4)}
        bound, peak, unit = _kernel_bound(t["name"])
        if rates and bound:
            work = t["flop"] / 1e12 if bound == "mfma" else t["bytes"] / 1e9
            ach = work / (t["total_ms"] * 1e-3)
            row.update({"avg_us": round(t["avg_ms"] * 1e3, 2), "achieved": round(ach, 2),
                        "unit": unit, "frac": round(ach / peak, 4)})
        out.append(row)
    return out


def heff_roofline(D, prec, reps=10):
    """Average duration of the two-site effective-Hamiltonian application at
    D x 4 x D, M=5 (HIP events on torch's current stream, where the library
    launches), and its algorithmic TFLOP/s."""
    import torch
    from pytenet_amd import backend
    dev = backend.device()
    g = torch.Generator(device=dev).manual_seed(0)
    a = torch.randn(D, 4, D, dtype=torch.complex128, device=dev, generator=g)
    w = torch.randn(5, 4, 4, 5, dtype=torch.complex128, device=dev, generator=g)
    l = torch.randn(D, 5, D, dtype=torch.complex128, device=dev, generator=g)
    r = torch.randn(D, 5, D, dtype=torch.complex128, device=dev, generator=g)
    backend.heff_apply(a, w, l, r, prec)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        backend.heff_apply(a, w, l, r, prec)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return ms, heff_flops(D, 4, D, 5, 5) / (ms * 1e-3) / 1e12


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "svd_split_2048_kernels_and_hbm_traffic.txt")


def profiled_traffic(kernel):
    """HBM bytes per launch of a product kernel instance from the committed
    FETCH_SIZE / WRITE_SIZE counter summary (calibrated against a plain
    copy); None when the summary has no line for it."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch ") and kernel in line:
            return float(line.rsplit(":", 1)[1])
    return None


def svd_time(D, precond32, reps=2):
    """Seconds per 2D x 2D SVD of a seeded random complex matrix, and the
    (FP64, FP32) Jacobi sweep counts."""
    import torch
    from pytenet_amd import backend
    dev = backend.device()
    g = torch.Generator(device=dev).manual_seed(1)
    a = torch.randn(2 * D, 2 * D, dtype=torch.complex128, device=dev, generator=g)
    backend.svd(a, precond32=precond32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        _, _, _, inf = backend.svd(a, return_info=True, precond32=precond32)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, int(inf[0]), int(inf[3])


def cpu_baseline(D, L, numiter, budget_s=30.0):
    """Oracle (NumPy/LAPACK restatement of the reference) on a bounded sample:
    one two-site local update at a saturated bond (merge, Krylov exponential
    with numiter two-site matvecs, SVD split, left environment step) plus one
    one-site back-step, on random environments of the workload's shapes;
    extrapolated to a sweep of 2(L-2)+1 such updates."""
    import oracle as orc
    from pytenet_amd.hamiltonian import heisenberg_xxz_1d_mpo
    try:
        from threadpoolctl import threadpool_info
        cores = max(int(i.get("num_threads", 1)) for i in threadpool_info()) if threadpool_info() else os.cpu_count()
    except Exception:
        cores = os.cpu_count()
    rng = np.random.default_rng(3)
    op = heisenberg_xxz_1d_mpo(10, 1.0, 0.8, -0.1)
    w0, w1 = op.a[4], op.a[5]
    h2 = orc.mpo_merge_tensor_pair(w0, w1)
    M = w0.shape[0]

    def herm_env(n):
        x = orc.crandn((n, M, n), rng) / n
        return 0.5 * (x + x.transpose(2, 1, 0).conj())
    lb, rb = herm_env(D), herm_env(D)
    a0 = orc.crandn((D, 2, D), rng) / np.sqrt(2 * D * D)
    a1 = orc.crandn((D, 2, D), rng) / np.sqrt(2 * D * D)
    qs = np.zeros(2, dtype=int)
    qb = np.zeros(D, dtype=int)
    t0 = time.perf_counter()
    a2 = orc.mps_merge_tensor_pair(a0, a1)
    a2 = orc.expm_krylov(lambda x: orc.apply_local_hamiltonian(x.reshape(a2.shape), h2, lb, rb).reshape(-1),
                         a2.reshape(-1), 0.025j, numiter, hermitian=True).reshape(a2.shape)
    b0, b1, _ = orc.mps_split_tensor_svd(a2, qs, qs, (qb, qb), "right", tol=1e-10, max_bond=D)
    lnew = orc.contraction_operator_step_left(b0, b0, w0, lb)
    b1 = orc.expm_krylov(lambda x: orc.apply_local_hamiltonian(x.reshape(b1.shape), w1, lnew, rb).reshape(-1),
                         b1.reshape(-1), -0.025j, numiter, hermitian=True).reshape(b1.shape)
    sample = time.perf_counter() - t0
    return {"value": sample * (2 * (L - 2) + 1), "unit": "s/sweep", "cores": int(cores),
            "kind": "port",
            "sample": f"one two-site + one one-site local update at D={D}, M={M}, numiter={numiter} "
                      f"({sample:.1f} s), times {2 * (L - 2) + 1} updates per sweep"}


def profile_summary(dirs):
    """Text tables from rocprofv3 csv output (no GPU needed): per kernel the
    calls and durations of *_kernel_stats.csv, and the mean per dispatch of
    every counter in *_counter_collection.csv (e.g. FETCH_SIZE / WRITE_SIZE
    in KiB)."""
    import csv
    import glob
    from collections import defaultdict

    def short(name, width=110):
        name = name.replace("void ", "").replace("HIP_vector_type<double, 2u>", "c128")
        name = name.replace("HIP_vector_type<float, 2u>", "c64")
        return name if len(name) <= width else name[:width] + "..."
    for d in dirs:
        for f in sorted(glob.glob(os.path.join(d, "**", "*_kernel_stats.csv"), recursive=True)):
            rows = list(csv.DictReader(open(f)))
            tot = sum(float(r["TotalDurationNs"]) for r in rows)
            print(f"# {os.path.basename(f)}: total kernel time {tot / 1e6:.2f} ms")
            print(f"{'calls':>8} {'total ms':>10} {'avg us':>10} {'%':>6}  kernel")
            for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:30]:
                print(f"{int(r['Calls']):8d} {float(r['TotalDurationNs']) / 1e6:10.3f} "
                      f"{float(r['AverageNs']) / 1e3:10.2f} {float(r['Percentage']):6.2f}  "
                      f"{short(r['Name'])}")
        for f in sorted(glob.glob(os.path.join(d, "**", "*_counter_collection.csv"), recursive=True)):
            agg = defaultdict(lambda: defaultdict(list))
            for r in csv.DictReader(open(f)):
                agg[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
            print(f"# {os.path.basename(f)}: mean counter value per dispatch")
            print(f"{'calls':>6} {'counter':>12} {'mean':>14}  kernel")
            for k, cs in agg.items():
                for c, v in cs.items():
                    print(f"{len(v):6d} {c:>12} {sum(v) / len(v):14.1f}  {short(k)}")


def _sectored_bonds(L, Dmax, rng):
    """Sorted U(1) bond quantum numbers: bond i draws min(Dmax, 2^min(i, L-i))
    labels from the binomial Sz distribution of min(i, L-i) spins."""
    from math import comb
    qbonds = []
    for i in range(L + 1):
        m = min(i, L - i)
        qs = np.arange(-m, m + 1, 2)
        w = np.array([comb(m, (q + m) // 2) for q in qs], float)
        qbonds.append(np.sort(rng.choice(qs, size=min(Dmax, 2**m), p=w / w.sum())))
    return qbonds


def workload_dmrg1(args):
    """Single-site DMRG, XXZ L=64 (J=1, Delta=1, h=0, zero quantum numbers),
    seeded random MPS with bonds min(D, 2^i, 2^(L-i)), numiter 10, FP64."""
    import torch
    import pytenet_amd as ptn
    L, D, sweeps = args.L or 64, args.D or 512, args.sweeps or 10
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 1.0, 0.0).zero_qnumbers()
    bonds = [min(D, 2**min(i, 40), 2**min(L - i, 40)) for i in range(L + 1)]
    psi = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random",
                  rng=np.random.default_rng(args.seed))
    psi.to_device()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    en = ptn.dmrg_singlesite(H, psi, sweeps, numiter_lanczos=args.numiter or 10)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"workload": f"single-site DMRG, XXZ L={L} Delta=1, D={D} (tapered), "
                        f"{sweeps} sweeps, FP64", "value": round(dt / sweeps, 4),
            "unit": "s/sweep", "total_s": round(dt, 2),
            "site_updates_per_s": round(2 * (L - 1) * sweeps / dt, 2),
            "energies": [float(e) for e in np.asarray(en).real]}


def workload_dmrg2_u1(args):
    """Two-site DMRG with U(1) sectors, XXZ L=128 (J=1, Delta=1, h=0), sorted
    seeded bond labels (about 30 sectors at the centre, bond cap D=2048),
    tol_split 0 (bonds stay at the cap), numiter 25, FP64; one progress line
    per 16 bond splits. `value` is the uninstrumented sweep time; the first
    sweep repeated from the same state with the live launch timing gives the
    per-kernel rates."""
    import torch
    import pytenet_amd as ptn
    from pytenet_amd import dmrg as dm
    L, D, sweeps = args.L or 128, args.D or 2048, args.sweeps or 1
    rng = np.random.default_rng(4)
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 1.0, 0.0)
    qbonds = _sectored_bonds(L, D, rng)
    u, cnt = np.unique(qbonds[L // 2], return_counts=True)
    psi = ptn.MPS(H.qsite, qbonds, fill="random", rng=rng)
    psi.to_device()
    psi_timed = copy.deepcopy(psi)
    split = dm.mps_split_tensor_svd
    t0 = [0.0]
    nsplit = [0]

    def split_progress(*a, **k):
        r = split(*a, **k)
        nsplit[0] += 1
        if nsplit[0] % 16 == 0:
            print(f"[bench] split {nsplit[0]} done, {time.perf_counter() - t0[0]:.1f} s",
                  file=sys.stderr, flush=True)
        return r

    dm.mps_split_tensor_svd = split_progress
    from pytenet_amd import backend
    try:
        torch.cuda.synchronize()
        t0[0] = time.perf_counter()
        en = ptn.dmrg_twosite(H, psi, sweeps, numiter_lanczos=args.numiter or 25, tol_split=0,
                              max_bond=D)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0[0]
        # the first sweep again, from the same state, with the live launch timing
        # on, for the per-kernel rates (the timed sweep above runs
        # uninstrumented; from a converged state the Lanczos solves break down
        # early and their skipped matvecs would inflate the rates)
        backend.timing_begin()
        try:
            ptn.dmrg_twosite(H, psi_timed, 1, numiter_lanczos=args.numiter or 25, tol_split=0,
                             max_bond=D)
            torch.cuda.synchronize()
        finally:
            timings = backend.timing_end()
    finally:
        dm.mps_split_tensor_svd = split
    return {"workload": f"two-site DMRG with U(1) sectors, XXZ L={L} Delta=1, D_max={D}, "
                        f"{sweeps} sweep(s) incl. setup, FP64, sector-grouped products",
            "value": round(dt / sweeps, 2), "unit": "s/sweep", "total_s": round(dt, 2),
            "sectors_centre": int(len(u)), "largest_sector": int(cnt.max()),
            "energies": [float(e) for e in np.asarray(en).real],
            "bond_max": int(max(psi.bond_dims)),
            "gpu_mem_peak_gib": round(torch.cuda.max_memory_allocated() / 2**30, 1),
            "roofline": kernel_roofline(timings),
            "timed_kernels": top_kernels(timings, 10, rates=True)}


def workload_chains(args):
    """256 transverse-field Ising chains (L=32, J=1, h=0, g uniform in
    [0.5, 1.5], seeded), each starting from a random product state padded to
    its own bond dimension D_i drawn uniformly from {32, 64, 128, 256}
    (seeded), imaginary-time single-site TDVP dt=0.05, 20 steps, numiter 10,
    evolved together by tdvp_singlesite_batched (buckets of equal D on their
    own streams, advancing one site step at a time together; tile autotuner
    per bucket) with complex64 contractions and FP64 Lanczos, and in all-FP64;
    the oracle is timed on one chain per D for one step."""
    import copy as cp
    import torch
    import pytenet_amd as ptn
    import oracle as orc
    nch, L, steps, nit = 256, args.L or 32, args.steps_chains, args.numiter or 10
    Ds = [32, 64, 128, 256]
    rng = np.random.default_rng(args.seed)
    Hs, psis = [], []
    for k in range(nch):
        g = rng.uniform(0.5, 1.5)
        H = ptn.ising_1d_mpo(L, 1.0, 0.0, g)
        psi = ptn.MPS.padded_product_state(H.qsite, L, int(rng.choice(Ds)), rng)
        psi.to_device()
        Hs.append(H)
        psis.append(psi)
    counts = {D: sum(1 for p in psis if max(p.bond_dims) == D) for D in Ds}
    out = {"workload": f"{nch} Ising chains L={L}, padded product states, D_i uniform in {Ds} "
                       f"(counts {counts}), single-site TDVP dt=0.05, {steps} steps, numiter {nit}, "
                       "batched", "unit": "site updates/s"}
    upd = nch * steps * 2 * (L - 1)
    ens = {}
    for prec in ("fp32", "fp64"):
        ps = [cp.deepcopy(p) for p in psis]
        torch.cuda.synchronize()
        g0 = ptn.backend.graph_launches()
        t0 = time.perf_counter()
        ptn.tdvp_singlesite_batched(Hs, ps, 0.05, steps, numiter_lanczos=nit, precision=prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ens[prec] = np.array([ptn.mpo_average(p, H).real / ptn.mps_norm(p)**2
                              for p, H in zip(ps, Hs)])
        out[f"{prec}_s"] = round(dt, 2)
        out[f"{prec}_site_updates_per_s"] = round(upd / dt, 1)
        out[f"{prec}_graph_launches"] = ptn.backend.graph_launches() - g0
        print(f"[bench] chains {prec} done ({dt:.1f} s)", file=sys.stderr, flush=True)
    err = np.abs(ens["fp32"] - ens["fp64"]) / np.abs(ens["fp64"])
    out["value"] = out["fp32_site_updates_per_s"]
    out["energy_rel_err_fp32_vs_fp64_max"] = float(err.max())
    out["energy_rel_err_fp32_vs_fp64_median"] = float(np.median(err))
    out["energy_rel_err_fp32_vs_fp64_quantiles"] = [float(np.quantile(err, q))
                                                     for q in (0.5, 0.9, 0.99)]
    worst = int(np.argmax(err))
    out["worst_chain"] = {"D": int(max(psis[worst].bond_dims)),
                          "energy_fp64": float(ens["fp64"][worst]),
                          "energy_fp32": float(ens["fp32"][worst])}
    t_cpu = 0.0
    for D in Ds:
        k = next(k for k in range(nch) if max(psis[k].bond_dims) == D)
        p = orc.MPS(psis[k].qsite, psis[k].qbonds)
        p.a = [a.cpu().numpy() for a in psis[k].a]
        t0 = time.perf_counter()
        orc.tdvp_singlesite(orc.MPO(Hs[k].qsite, Hs[k].qbonds, Hs[k].a), p, 0.05, 1,
                            numiter_lanczos=nit)
        t_cpu += time.perf_counter() - t0
    out["cpu_baseline"] = {"value": round(len(Ds) * 2 * (L - 1) / t_cpu, 2),
                           "unit": "site updates/s", "kind": "port",
                           "sample": "the oracle, one chain per D, one step"}
    return out


def _graded_matrix(n, decades, seed):
    """n x n complex matrix with singular values 10^(-decades i / (n-1)), built
    on the GPU from two seeded random unitaries (the library's QR and product)."""
    import torch
    from pytenet_amd import backend
    dev = backend.device()
    g = torch.Generator(device=dev).manual_seed(seed)
    u, _ = backend.qr(torch.randn(n, n, dtype=torch.complex128, device=dev, generator=g))
    v, _ = backend.qr(torch.randn(n, n, dtype=torch.complex128, device=dev, generator=g))
    s = 10.0 ** (-decades * torch.arange(n, device=dev, dtype=torch.float64) / (n - 1))
    # u diag(s) v^H: the adjoint through the product's conjugation flag on v's transpose
    return backend.matmul((u * s.to(torch.complex128)).contiguous(), v.transpose(0, 1),
                          conjB=True)


def workload_svd(args):
    """The bond split of the two-site TDVP at D: the truncating split
    (ptn_svd_split, cap D, tol 1e-10) of a 2D x 2D complex matrix with a
    spectrum graded over nine decades and of a seeded random one, in FP64 and
    with FP32 Jacobi preconditioning, plus the full SVD (two timed calls each
    after one untimed). With --copy-calibration, a plain 512 MiB device copy
    first (the reference point of HBM byte counters in a profile run)."""
    import torch
    from pytenet_amd import backend
    D = args.D or 1024
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    out = {"workload": f"bond split of a {2 * D} x {2 * D} complex matrix", "unit": "s"}
    g = torch.Generator(device=backend.device()).manual_seed(1)
    mats = {"graded": _graded_matrix(2 * D, 9.0, 7),
            "random": torch.randn(2 * D, 2 * D, dtype=torch.complex128, device="cuda", generator=g)}
    timed = {}
    for name, a in mats.items():
        for pre in (False, True):
            key = f"{name}_{'precond32' if pre else 'fp64'}"
            backend.svd_split(a, "right", 1e-10, D, pre)
            torch.cuda.synchronize()
            backend.timing_begin()
            t0 = time.perf_counter()
            for _ in range(2):
                backend.svd_split(a, "right", 1e-10, D, pre)
            torch.cuda.synchronize()
            out[f"split_{key}_s"] = round((time.perf_counter() - t0) / 2, 4)
            for t in backend.timing_end():
                agg = timed.setdefault(t["name"], {"launches": 0, "total_ms": 0.0, "bytes": 0.0,
                                                   "flop": 0.0})
                for f in ("launches", "total_ms", "bytes", "flop"):
                    agg[f] += t[f]
    out["timed_kernels"] = [{"name": k, "launches": v["launches"],
                             "avg_us": round(1e3 * v["total_ms"] / v["launches"], 2),
                             "bytes_per_launch": v["bytes"] / v["launches"],
                             "flop_per_launch": v["flop"] / v["launches"]}
                            for k, v in sorted(timed.items(), key=lambda kv: -kv[1]["total_ms"])[:8]]
    out["value"] = out["split_graded_precond32_s"]
    t64, sw64, _ = svd_time(D, False)
    t32, sw64p, sw32p = svd_time(D, True)
    out.update({"full_svd_random_fp64_s": round(t64, 4), "full_svd_random_fp64_sweeps": sw64,
                "full_svd_random_precond32_s": round(t32, 4),
                "full_svd_random_precond32_sweeps_fp32_fp64": [sw32p, sw64p]})
    return out


def workload_growth(args):
    """The precision study of the two-site TDVP workload as specified: XXZ
    L=100 (J=1, Delta=0.8, h=-0.1), seeded random MPS with bond 16,
    left-orthonormalised, dt=-0.05j, 40 steps, numiter 8, tol_split 1e-10, the
    bond cap D (default 256; the bonds grow from 16 towards it), run in FP64,
    FP32 and split-bf16 contractions. Reports, per mode, the energy drift
    |E(t) - E(0)| / |E(0)| after every step (real-time evolution conserves the
    energy), and against the FP64 run the final |1 - fidelity| and relative
    energy difference, plus the bond profile and wall time."""
    import torch
    import pytenet_amd as ptn
    L, D, steps = args.L or 100, args.D or 256, args.sweeps or 40
    nit = args.numiter or 8
    H = ptn.heisenberg_xxz_1d_mpo(L, 1.0, 0.8, -0.1).zero_qnumbers()
    bonds = [min(16, 2**min(i, 40), 2**min(L - i, 40)) for i in range(L + 1)]
    psi0 = ptn.MPS(H.qsite, [np.zeros(b, dtype=int) for b in bonds], fill="random",
                   rng=np.random.default_rng(args.seed))
    psi0.to_device()
    psi0.orthonormalize(mode="left")
    e0 = ptn.mpo_average(psi0, H).real
    out = {"workload": f"two-site TDVP, XXZ L={L}, bond 16 -> cap {D}, dt=-0.05j, {steps} steps, "
                       f"numiter {nit}, tol 1e-10", "unit": "s", "E0": e0}
    finals = {}
    for prec in ("fp64", "fp32", "bf16x3", "bf16x6"):
        psi = copy.deepcopy(psi0)
        run = ptn.TwoSiteTDVP(H, psi, -0.05j, nit, 1e-10, precision=prec, max_bond=D)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.setup()
        drift = []
        for k in range(steps):
            run.sweep()
            n = ptn.mps_norm(psi)
            drift.append(abs(ptn.mpo_average(psi, H).real / n**2 - e0) / abs(e0))
            if k % 5 == 4:
                print(f"[bench] growth {prec}: step {k + 1}/{steps}, bond max {max(psi.bond_dims)}, "
                      f"drift {drift[-1]:.2e}", file=sys.stderr, flush=True)
        run.finish()
        torch.cuda.synchronize()
        out[f"{prec}_s"] = round(time.perf_counter() - t0, 2)
        out[f"{prec}_energy_drift"] = [float(f"{x:.3e}") for x in drift[::4]] + [float(f"{drift[-1]:.3e}")]
        out[f"{prec}_bond_max"] = int(max(psi.bond_dims))
        finals[prec] = psi
        print(f"[bench] growth {prec} done ({out[f'{prec}_s']} s)", file=sys.stderr, flush=True)
    ref = finals["fp64"]
    nr = ptn.mps_norm(ref)
    er = ptn.mpo_average(ref, H).real / nr**2
    for prec in ("fp32", "bf16x3", "bf16x6"):
        p = finals[prec]
        npr = ptn.mps_norm(p)
        out[f"{prec}_infidelity_vs_fp64"] = abs(1 - abs(ptn.mps_vdot(ref, p))**2 / (nr**2 * npr**2))
        out[f"{prec}_energy_rel_err_vs_fp64"] = abs(ptn.mpo_average(p, H).real / npr**2 - er) / abs(er)
    out["value"] = out["fp32_s"]
    return out


WORKLOADS = {"growth": workload_growth, "dmrg1": workload_dmrg1, "dmrg2-u1": workload_dmrg2_u1,
             "chains": workload_chains, "svd": workload_svd}


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--profile-summary":
        profile_summary(sys.argv[2:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--L", type=int, default=None)
    ap.add_argument("--D", type=int, default=None)
    ap.add_argument("--numiter", type=int, default=None)
    ap.add_argument("--workload", default="tdvp2", choices=["tdvp2"] + sorted(WORKLOADS),
                    help="tdvp2 (default): the headline line; the others time the further "
                         "workloads (single-site DMRG, two-site DMRG with U(1) sectors, "
                         "many independent chains, the bond SVD) and print one JSON line each")
    ap.add_argument("--sweeps", type=int, default=None)
    ap.add_argument("--autotune", action="store_true",
                    help="switch on the tile autotuner of the complex products")
    ap.add_argument("--steps-chains", type=int, default=20)
    ap.add_argument("--mixed", default="bf16x6",
                    help="contraction precision of the mixed leg: bf16x6 (three bf16 terms per "
                         "FP32 operand, FP32 accuracy), fp32 or bf16x3")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--strict-svd", action="store_true",
                    help="with --extras: also time the mixed sweep with the plain FP64 Jacobi SVD")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="with --workload svd: a plain 512 MiB device copy first (counter reference)")
    ap.add_argument("--extras", action="store_true",
                    help="after the JSON line, time the bond SVD alone and print it to stderr")
    ap.add_argument("--time-budget", type=float, default=420.0,
                    help="wall-clock seconds (from start) the headline run aims to stay within; "
                         "the number of timed sweeps is cut to fit")
    ap.add_argument("--tail-s", type=float, default=80.0,
                    help="seconds reserved after the FP64 leg's sweeps for the fixed tail")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only time the two-site effective-Hamiltonian kernels (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the sweep is sequential over sites and this build "
              "has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import pytenet_amd as ptn

    torch.cuda.set_device(0)
    if args.autotune:
        from pytenet_amd import backend
        backend.set_autotune(True)
    if args.workload != "tdvp2":
        print(json.dumps(WORKLOADS[args.workload](args)), flush=True)
        return
    args.L = args.L or 100
    args.D = args.D or 1024
    args.numiter = args.numiter or 8
    if args.roofline_only:
        # a plain device copy of known size: calibrates the HBM byte counters of a profile run
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        res = {"copy_calibration_bytes": 2 * src.numel() * 8}
        for prec in ("fp64", "fp32", "bf16x3", "bf16x6"):
            ms, tf = heff_roofline(args.D, prec, reps=20)
            res[prec] = {"avg_ms": round(ms, 3), "achieved_tflops": round(tf, 2),
                         "peak": PEAK_TFLOPS[prec], "frac": round(tf / PEAK_TFLOPS[prec], 4)}
        print(json.dumps({"roofline_only": res, "D": args.D}), flush=True)
        return
    prewarm()
    H, psi0 = build_state(args.L, args.D, args.seed)
    print(f"[bench] state ready: L={args.L}, bonds max {max(psi0.bond_dims)} "
          f"({time.perf_counter() - T_START:.0f} s)", file=sys.stderr, flush=True)
    warm = min(args.warmup, 1)

    def decide(t_sweep):
        # timed sweeps per leg that fit the wall-clock budget: both legs run
        # warm + k sweeps (the mixed ones are no slower than the FP64 ones), plus
        # the fixed tail (mixed setup, CPU baseline, roofline bookkeeping)
        left = args.time_budget - (time.perf_counter() - T_START) - args.tail_s
        left -= warm * t_sweep + (0 if warm else t_sweep)
        return int(left // (2 * t_sweep)) + (0 if warm else 1)

    psi64, t64, steps, tim64 = run_leg(H, psi0, "fp64", warm, args.steps, args.D, args.numiter,
                                       decide=decide)
    psimx, tmx, _, timmx = run_leg(H, psi0, args.mixed, warm, steps, args.D, args.numiter)
    ov = ptn.mps_vdot(psi64, psimx)
    n64, nmx = ptn.mps_norm(psi64), ptn.mps_norm(psimx)
    infid = abs(1.0 - abs(ov)**2 / (n64**2 * nmx**2))
    e64 = ptn.mpo_average(psi64, H).real / n64**2
    emx = ptn.mpo_average(psimx, H).real / nmx**2
    rel_e = abs(emx - e64) / abs(e64)
    del psi64, psimx

    nupd = 2 * (args.L - 2) + 1
    roof = kernel_roofline(timmx)
    if roof is not None:
        roof["traffic"] = profiled_traffic(roof["kernel"])
        roof["traffic_source"] = ("rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch of this kernel, "
                                  "copy-calibrated (profiles/" + os.path.basename(TRAFFIC_FILE) + ")")
    out = {
        "metric": METRIC,
        "value": round(tmx, 4),
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(tmx * 1e3, 1),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{MIXED_DTYPE.get(args.mixed, args.mixed)} contractions, f64 Krylov/QR, "
                 "SVD: f32 Jacobi preconditioning sweeps finished by f64 sweeps (f64 accuracy)",
        "data": "synthetic (seeded random MPS, XXZ MPO)",
        "config": {"workload": "two-site TDVP sweep, XXZ L=%d J=1 Delta=0.8 h=-0.1, fixed D=%d "
                               "(tapered), dt=-0.05j, numiter=%d, tol_split=1e-10"
                               % (args.L, args.D, args.numiter),
                   "L": args.L, "D": args.D, "numiter": args.numiter, "seed": args.seed},
        "fp64_sweep_s": round(t64, 4),
        "speedup_vs_fp64": round(t64 / tmx, 3),
        "infidelity_vs_fp64": infid,
        "energy_rel_err_vs_fp64": rel_e,
        "site_updates_per_s": round(nupd / tmx, 3),
        "roofline": roof,
        "fp64_roofline": kernel_roofline(tim64),
        "timed_kernels_mixed": top_kernels(timmx, rates=True),
        "timed_kernels_fp64": top_kernels(tim64, rates=True),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.D, args.L, args.numiter)
    out["wall_s"] = round(time.perf_counter() - T_START, 1)
    print(json.dumps(out), flush=True)
    if args.extras:
        # optional legs after the line: SVD timings, live Jacobi-apply roofline,
        # the mixed sweep with the plain FP64 SVD
        extra = {"svd": workload_svd(args)}
        if args.strict_svd:
            _, tstrict, _, _ = run_leg(H, psi0, args.mixed, warm, steps, args.D, args.numiter,
                                       svd_precond32=False)
            extra["mixed_sweep_fp64_svd_s"] = round(tstrict, 4)
            extra["speedup_vs_fp64_with_fp64_svd"] = round(t64 / tstrict, 3)
        print("[bench] extras " + json.dumps(extra), file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
