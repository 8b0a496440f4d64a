#!/usr/bin/env python3
"""Benchmark of the TPB path on one MI355X.

Workload (default c36): norb=36, n_alpha=n_beta=8, 8000 x 8000 strings
(Hartree-Fock + singles + seeded doubles), 64M determinants, integrals from
150 seeded Cholesky vectors; plain FP64.  A step is one Davidson solve from
the start vector to |E0 W - H W| < eps, with the vector resident in HBM when
the timed region starts.  value = time-to-E0 in seconds (lower is better);
sigma throughput, per-kernel roofline figures and a CPU baseline measured on
a bounded sample with the oracle are reported beside it.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "sbd-pub_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "time-to-E0 (|ΔE|<1e-8 Ha) & sigma GB/s: mixed+learned vs plain FP64, 1 GPU"
HBM_PEAK_GBPS = 8000.0
MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}  # MI355X dense matrix peaks
LDS_PEAK_TBPS = 256 * 256 * 2.4e9 / 1e12  # 256 CUs x 256 B/clk x 2.4 GHz


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(wc, na, nb, e=8.0):
    """Algorithmic bytes per launch of each sigma kernel group, element size
    e (8 FP64, 4 FP32).  same-spin (sparse form): every (entry x other-spin
    column) reads x[J,m], plus diagonal, hii and output streams, plus the
    table entries once; mixed (direct form): one x gather per (alpha single
    x beta single) term, per (alpha single x beta string) and (alpha string
    x beta single) dressing term, plus reading and writing sigma, plus the
    tables once; transposes: one read + one write; the dense same-spin form
    streams x, hii and y for the diagonal and reads x and y once per GEMM."""
    N = na * nb
    ab = wc["alpha_singles"] * wc["beta_singles"]
    dressing = wc["alpha_singles"] * nb + wc["beta_singles"] * na
    return {
        "same_alpha": e * (nb * (wc["alpha_singles"] + wc["alpha_doubles"]) + 3 * N)
        + (8.0 + e) * wc["alpha_singles"] + (4.0 + e) * wc["alpha_doubles"],
        "same_beta": e * (na * (wc["beta_singles"] + wc["beta_doubles"]) + 2 * N)
        + (8.0 + e) * wc["beta_singles"] + (4.0 + e) * wc["beta_doubles"],
        "mixed": e * (ab + dressing + 2 * N) + 8.0 * wc["alpha_singles"] + 4.0 * wc["beta_singles"],
        "same_gemm": e * (3 * N + 2 * 2 * N + na * na + nb * nb),
        # spin-flip symmetric form: the mixed term for the upper triangle
        # (half the gathers); one GEMM plus y = hii*x + G + G^T
        "sym_mixed": e * ((ab + dressing) / 2 + 2 * N) + 8.0 * wc["alpha_singles"]
        + 4.0 * wc["beta_singles"],
        "sym_same": e * (2 * N + na * na + 4 * N),
        "transpose": 2 * e * N,
        "transpose_add": 3 * e * N,
    }


def profiled_traffic(name, width=None, slots=None):
    """HBM bytes per launch of kernel group `name` (at the given mixed-kernel
    configuration, when it has one), from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary in profiles/ (None when not profiled)."""
    path = os.path.join(ROOT, "profiles", "c36_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    want = f"eb={width} nt={slots}" if width is not None else "-"
    for line in open(path):
        f = line.split()
        if f and f[0] == "traffic_bytes" and f[1] == name:
            cfg = f"{f[2]} {f[3]}" if len(f) == 5 else "-"
            if cfg == want:
                return float(f[-1])
    return None


def lds_bytes_mixed(wc, na, nb, n_elec, eb, e=4.0, symmetric=True):
    """LDS bytes one launch of the staged mixed-term kernel reads: per round of
    eb staged alpha singles (plus one spectator round per bra), every slot
    reads two eb-wide values (staged ket row, staged ERI row) per beta single
    and n_elec + 1 eb-wide values for its dressing; the symmetric form visits
    half the slots.  Width 16 is the signed-triangle kernel: 16-byte pairs,
    16 / e values each."""
    if eb == 16:
        eb = int(16 // e)
    rounds = (wc["alpha_singles"] + na) / eb
    slots = nb / 2 if symmetric else nb
    singles = wc["beta_singles"] / 2 if symmetric else wc["beta_singles"]
    return rounds * (2 * e * eb * singles + e * eb * slots * (n_elec + 1))


def profiled_limiter(name, width, slots):
    """Measured on-chip utilisation of kernel group `name` at the given
    mixed-kernel configuration (LDS busy fraction, share of it stalled on
    bank conflicts, VALU busy fraction), from the rocprofv3 SQ-counter
    summary in profiles/ (None when not profiled)."""
    path = os.path.join(ROOT, "profiles", "c36_sq_counters.md")
    if not os.path.exists(path):
        return None
    want = f"eb={width} nt={slots}"
    for line in open(path):
        f = line.split()
        if len(f) == 7 and f[0] == "limiter" and f[1] == name and f"{f[2]} {f[3]}" == want:
            out = {kv.split("=")[0]: float(kv.split("=")[1]) for kv in f[4:]}
            out["source"] = "profiles/c36_sq_counters.md (rocprofv3 SQ counters)"
            return out
    return None


def cpu_calibration():
    """Reference-over-oracle sigma time ratio measured by oracle/calibrate.py
    (profiles/cpu_calibration.md), or None."""
    path = os.path.join(ROOT, "profiles", "cpu_calibration.md")
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) == 2 and f[0] == "ratio_reference_over_oracle":
                return float(f[1])
    return None


def cpu_baseline(inst, n_steps, seconds, nthreads):
    """Oracle sigma on a random sample of alpha rows, scaled to one full
    sigma and multiplied by the GPU solve's sigma count."""
    from oracle import oracle
    orc = oracle.Oracle.from_instance(inst, nthreads)
    n = orc.ndet
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n)
    hii = np.zeros(n)
    y = np.zeros(n)
    rows = rng.permutation(orc.na)
    # random rows in batches of one row per thread until the time budget is used
    done = 0
    t0 = time.perf_counter()
    while done < orc.na and (time.perf_counter() - t0) < seconds:
        orc.sigma_rowlist(x, hii, rows[done:done + nthreads], y)
        done += nthreads
    t = time.perf_counter() - t0
    sample = rows[:min(done, orc.na)]
    orc.close()
    t_sigma = t * orc.na / sample.size
    return dict(value=t_sigma * n_steps, unit="s", cores=nthreads, kind="port",
                sample=f"oracle sigma on {sample.size} of {orc.na} alpha rows (random, seed 0) "
                       f"in {t:.1f} s with {nthreads} threads, scaled x{orc.na / sample.size:.1f} "
                       f"to one sigma ({t_sigma:.1f} s) and x{n_steps} sigma steps of the GPU solve",
                sigma_s=t_sigma)


def factored_leg(op, v, norb):
    """One product per factored form of the mixed term (dense (pq|rs) GEMM
    per batch of beta rows) next to the direct form: device times of the
    mixed term and of its GEMMs, the GEMMs' rate against the MI355X dense
    matrix peak of their input type, and the difference to the direct form."""
    K = norb * norb
    flops = 2.0 * K * K * op.na * op.nb  # per product, one GEMM pass (logical)
    peaks = {"fp64_dgemm": 78.6, "fp32_sgemm": 157.3, "fp32_split_f16x3": 2500.0}
    out = {"flops_per_product": flops}
    for name, form, fp32 in (("direct_fp64", 0, False), ("direct_fp32", 0, True), ("fp64_dgemm", 1, False),
                             ("fp32_sgemm", 1, True), ("fp32_split_f16x3", 2, True)):
        op.set_mixed_form(form)
        f = op.sigma_fp32 if fp32 else op.sigma
        f(v)  # untimed: workspace, FP16 halves of the ERI matrix
        op.profile(True)
        s = f(v)
        st = op.profile_stats()
        op.profile(False)
        if form == 0:
            out[name] = {"mixed_ms": round(st["mixed_f32" if fp32 else "mixed"][1], 3)}
            out["_ref"] = s if not fp32 else out.get("_ref")
            continue
        g = st["fac_gemm_f32" if fp32 else "fac_gemm"]
        mflops = flops * (3 if form == 2 else 1)  # MFMA flops actually issued
        ref = out["_ref"]
        out[name] = {
            "mixed_ms": round(st["mixed_f32" if fp32 else "mixed"][1], 3),
            "gemm_ms": round(g[1], 3), "gemm_calls": g[0],
            "gemm_tflops": round(mflops / (g[1] * 1e-3) / 1e12, 1),
            "mfma_peak_frac": round(mflops / (g[1] * 1e-3) / 1e12 / peaks[name], 3),
            "max_rel_diff_vs_direct_fp64": float(np.max(np.abs(s - ref)) / np.max(np.abs(ref))),
        }
        log(f"factored {name}: {out[name]}")
    op.set_mixed_form(0)
    del out["_ref"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed Davidson solves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed sigma applications")
    ap.add_argument("--workload", default="c36")
    ap.add_argument("--nb", type=int, default=16)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--max-it", type=int, default=200)
    ap.add_argument("--max-time", type=float, default=900.0)
    ap.add_argument("--init", choices=["hf", "random"], default="hf")
    ap.add_argument("--solver", choices=["plain", "thick", "mixed", "block"], default="mixed",
                    help="timed solver: plain = the reference algorithm in FP64; thick = FP64 "
                         "thick restart; mixed = thick restart with FP32 correction products; "
                         "block = block Davidson (--block corrections per iteration, --roots roots) "
                         "with FP32 correction products")
    ap.add_argument("--block", type=int, default=4)
    ap.add_argument("--roots", type=int, default=1, help="roots of the block solver")
    ap.add_argument("--block-basis", type=int, default=0,
                    help="basis size of the block solver (0: as large as memory allows, up to 64)")
    ap.add_argument("--sectors", choices=["lowest", "even", "odd", "off"], default="lowest",
                    help="spin-flip sectors of the optimized solver when the string sets are equal: "
                         "lowest (default) solves the even sector from the start vector, then the "
                         "odd one, and keeps the lower state; even/odd one sector; off the general "
                         "sigma")
    ap.add_argument("--no-symmetry", dest="sectors", action="store_const", const="off",
                    help="same as --sectors off")
    ap.add_argument("--coarse", type=int, default=1000,
                    help="coarse start on the first M strings of each set (optimized solver; 0: off)")
    ap.add_argument("--coarse-eps", type=float, default=1e-2)
    ap.add_argument("--no-graph-run", dest="graph_run", action="store_false",
                    help="skip the extra timed solve with the FP32 sigma core replayed as a hipGraph")
    ap.add_argument("--keep", type=int, default=3,
                    help="Ritz vectors kept at a thick restart (thick / mixed solvers)")
    ap.add_argument("--no-compare", dest="compare", action="store_false",
                    help="skip the plain FP64 reference solve")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    # the host cores one GPU's job gets (16 on the benchmark machine)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--factored", action="store_true",
                    help="also time one product per factored (pq|rs)-GEMM form of the mixed term")
    ap.add_argument("--quiet", action="store_true", help="no per-step solver lines on stderr")
    ap.add_argument("--rdm", action="store_true",
                    help="also compute the 1-/2-RDM of the solution after the timed solves (not timed in value)")
    ap.add_argument("--precond", choices=["diag", "mlp"], default="diag",
                    help="preconditioner of the timed solver (mlp: learned, weights below)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "sbd-pub_amd", "sbdx", "data",
                                                      "precond_mlp.npy"))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import sbdx
    from sbdx import instances

    spec = instances.WORKLOADS[args.workload]
    log(f"building instance {args.workload}: {spec}")
    inst = instances.workload(args.workload)
    t0 = time.perf_counter()
    op = sbdx.TPBOperator.from_instance(inst)
    t_build = time.perf_counter() - t0
    na, nb = op.na, op.nb
    wc = op.work_counts()
    log(f"operator built in {t_build:.2f} s: ndet={op.ndet} counts={wc}")

    same_sets = bool(np.array_equal(inst["astr"], inst["bstr"]))
    symmetric = args.sectors != "off" and same_sets
    if args.init == "hf":
        W0 = sbdx.basis_init_vector(op.ndet)
    else:
        W0 = np.random.default_rng(1).standard_normal(op.ndet)
        W0 /= np.linalg.norm(W0)
    # warmup: also settles the same-spin form (dense GEMM or sparse rows)
    # for each precision before any timed solve
    # (and the symmetric form's mixed-kernel configuration, on the HF vector,
    # which is spin-flip symmetric)
    for _ in range(args.warmup):
        op.sigma(W0)
        if args.solver in ("mixed", "block"):
            op.sigma_fp32(W0)
        if symmetric and args.solver in ("thick", "mixed", "block"):
            e0 = sbdx.basis_init_vector(op.ndet)
            op.sigma_sym(e0)
            if args.solver in ("mixed", "block"):
                op.sigma_sym(e0, fp32=True)
    log(f"warmup done; same-spin form {op.same_spin_forms()}")
    # one FP64 sigma through the C ABI with host buffers in and out (PCIe
    # transfers included; every solve keeps its vectors in HBM instead)
    t0 = time.perf_counter()
    op.sigma(W0)
    sigma_host_ms = (time.perf_counter() - t0) * 1e3
    log(f"sigma with host buffers (PCIe included): {sigma_host_ms:.1f} ms")

    if args.precond == "mlp":
        op.mlp_set_weights(np.load(args.weights))
        log(f"learned preconditioner weights from {args.weights}")

    def run(solver, nsolve, precond="diag"):
        op.configure(solver)
        if solver in ("thick", "mixed"):
            op.set_restart(args.keep)
        if solver in ("thick", "mixed", "block"):
            op.set_symmetry(args.sectors if symmetric else "off")
        if solver in ("thick", "mixed", "block"):
            op.set_coarse_start(args.coarse if args.coarse < min(na, nb) else 0, args.coarse_eps)
        if solver == "block":
            op.set_block_basis(args.block_basis)
        op.set_preconditioner(1 if precond == "mlp" else 0)
        op.set_verbose(0 if args.quiet else 1)
        op.profile(True)
        times, energies, nsteps, conv = [], [], [], []
        roots = sectors = block_info = None
        for k in range(nsolve):
            op.set_vector(W0)
            t0 = time.perf_counter()
            if solver == "block":
                Eb, Wb, n, ok = op.davidson_block(args.roots, args.block,
                                                  W0=W0[None, :] if args.roots == 1 else None,
                                                  max_it=args.max_it, eps=args.eps, max_time=args.max_time)
            else:
                E, n, ok = op.solve(max_it=args.max_it, nb=args.nb, eps=args.eps,
                                    max_time=args.max_time)
            dt = time.perf_counter() - t0
            if solver == "block":
                E = float(Eb[0])
                roots = [float(e) for e in Eb]
                # each root against FP64 refinement (outside the timed
                # region): the FP64 Rayleigh quotient and residual norm of
                # the returned vector, from exact FP64 products
                rq, rr = [], []
                for j in range(args.roots):
                    sj = op.sigma(Wb[j])
                    nw = float(Wb[j] @ Wb[j])
                    rq.append(float(Wb[j] @ sj) / nw)
                    rr.append(float(np.linalg.norm(sj - Eb[j] * Wb[j])))
                block_info = dict(op.block_stats(args.roots), block=args.block, roots=roots,
                                  fp64_rayleigh=rq, abs_dE_vs_fp64=[abs(a - b) for a, b in zip(roots, rq)],
                                  fp64_resid=rr)
                log(f"block roots: {block_info}")
            times.append(dt)
            energies.append(E)
            nsteps.append(n)
            conv.append(ok)
            if solver in ("thick", "mixed") and symmetric:
                sectors = op.sector_stats()
                log(f"spin-flip sectors: {sectors}")
            tr = op.trace()
            extra = (f" coarse={op.coarse_stats()}" if solver in ("thick", "mixed", "block") else "") + \
                (f" fp64-from-step={op.fp64_switch_step()}" if solver in ("thick", "mixed") else "")
            log(f"{solver}/{precond} solve {k}: E0={E:.12f} |r|={tr[-1, 1]:.2e} sigma-products={n} "
                f"converged={ok} time={dt:.3f} s precond-steps={op.precond_stats()}{extra}")
        stats = op.profile_stats()
        op.profile(False)
        return dict(time=float(np.median(times)), E=energies[-1], steps=int(np.median(nsteps)),
                    converged=all(conv), stats=stats, precond_steps=op.precond_stats(), roots=roots,
                    block_info=block_info,
                    sectors=sectors,
                    fp64_switch_step=op.fp64_switch_step() if solver in ("thick", "mixed") else None,
                    coarse=op.coarse_stats() if solver in ("thick", "mixed", "block") else None)

    best = args.solver
    plain = run("plain", 1 if best != "plain" else args.steps) if args.compare or best == "plain" else None
    res = run(best, args.steps, args.precond) if best != "plain" else plain
    graph_res = None
    if args.graph_run and best == "mixed" and symmetric:
        # the same solve with the FP32 symmetric sigma core replayed as a
        # captured hipGraph (profiling off: one timed solve)
        op.set_graphs(True)
        op.configure(best)
        op.set_restart(args.keep)
        op.set_symmetry(args.sectors)
        op.set_coarse_start(args.coarse if args.coarse < min(na, nb) else 0, args.coarse_eps)
        op.set_preconditioner(1 if args.precond == "mlp" else 0)
        op.set_verbose(0)
        op.set_vector(W0)
        t0 = time.perf_counter()
        Eg, ng, okg = op.solve(max_it=args.max_it, nb=args.nb, eps=args.eps, max_time=args.max_time)
        graph_res = dict(time_s=round(time.perf_counter() - t0, 4), sigma_products=ng, converged=okg,
                         E0=Eg)
        op.set_graphs(False)
        log(f"{best}/graph solve: E0={Eg:.12f} sigma-products={ng} converged={okg} "
            f"time={graph_res['time_s']:.3f} s")
    stats = res["stats"]
    times = [res["time"]]
    energies = [res["E"]]
    nsteps = [res["steps"]]
    conv = [res["converged"]]

    gemm_flops = {"same_gemm": 2.0 * na * na * nb + 2.0 * nb * nb * na,  # dense same-spin
                  "sym_same": 2.0 * na * na * nb}                        # symmetric form

    def kernel_table(stats):
        kb64 = kernel_bytes(wc, na, nb, 8.0)
        kb32 = kernel_bytes(wc, na, nb, 4.0)
        out = {}
        for name, (cnt, ms) in stats.items():
            avg = ms / max(cnt, 1)
            ent = dict(launches=cnt, total_ms=round(ms, 3), avg_ms=round(avg, 4))
            base = name[:-4] if name.endswith("_f32") else name
            kb = kb32 if name.endswith("_f32") else kb64
            if base in kb and avg > 0:
                ent["alg_bytes"] = kb[base]
                ent["GBps"] = round(kb[base] / (avg * 1e-3) / 1e9, 1)
            if base in gemm_flops and avg > 0:
                ent["flops"] = gemm_flops[base]
                ent["TFLOPs"] = round(gemm_flops[base] / (avg * 1e-3) / 1e12, 2)
            out[name] = ent
        return out

    SIGMA_PARTS = ("same_gemm", "same_alpha", "transpose", "same_beta", "transpose_add", "mixed",
                   "sym_same", "sym_mixed")

    def sigma_hbm_rate(kernels, suffix, forms):
        """(HBM GB, GB/s) of one product from the profiled per-launch traffic
        of its kernel groups; None when a group was not profiled."""
        names = [n + suffix for n in SIGMA_PARTS if n + suffix in kernels]
        if not any("mixed" in n for n in names) or args.workload != "c36":
            return None
        by = 0.0
        for n in names:
            base = n[:-4] if n.endswith("_f32") else n
            pre = "sym_" if base.startswith("sym_") else ""
            t = (profiled_traffic(n, forms[pre + "mixed_width"], forms[pre + "mixed_slots_per_thread"])
                 if "mixed" in base else profiled_traffic(n))
            if t is None:
                return None
            by += t
        ms = sum(kernels[n]["avg_ms"] for n in names)
        return dict(hbm_bytes=by, ms=round(ms, 3), GBps=round(by / (ms * 1e-3) / 1e9, 1))

    def sigma_rate(kernels, suffix):
        names = [n + suffix for n in SIGMA_PARTS if n + suffix in kernels]
        if not any("mixed" in n for n in names):
            return None, None
        ms = sum(kernels[n]["avg_ms"] for n in names)
        by = sum(kernels[n]["alg_bytes"] for n in names)
        return round(ms, 3), round(by / (ms * 1e-3) / 1e9, 1)

    kernels = kernel_table(stats)
    sig_kernels = [n for n in kernels if "alg_bytes" in kernels[n]]
    dom = max(sig_kernels, key=lambda n: kernels[n]["total_ms"]) if sig_kernels else None
    dk = kernels[dom] if dom else None
    if dom is None:
        # no kernel group with a per-unit byte count was profiled (block
        # products through the multi-vector kernel)
        roofline = None
    elif dom.startswith("same_gemm") or dom.startswith("sym_same"):
        peak = MFMA_PEAK_TFLOPS["f32" if dom.endswith("_f32") else "f64"]
        roofline = dict(kernel=dom, bound="mfma", achieved=dk["TFLOPs"], peak=peak, unit="TFLOP/s",
                        frac=round(dk["TFLOPs"] / peak, 4), traffic=None)
    else:
        # The mixed-term kernel gathers from rows it stages in LDS, so the
        # bytes it moves from HBM are far fewer than the gathers it performs:
        # `achieved` is the HBM traffic the rocprofv3 counters measured for
        # one launch over the launch's live average duration, and the gather
        # count and the LDS bytes it reads are reported beside it.
        f32 = dom.endswith("_f32")
        forms = op.same_spin_forms()["fp32" if f32 else "fp64"]
        pre = "sym_" if dom.startswith("sym_") else ""
        width, slots = forms[pre + "mixed_width"], forms[pre + "mixed_slots_per_thread"]
        traffic = limiter = None
        if args.workload == "c36":
            traffic = profiled_traffic(dom, width, slots)
            limiter = profiled_limiter(dom, width, slots)
        avg_s = dk["avg_ms"] * 1e-3
        achieved = round(traffic / avg_s / 1e9, 1) if traffic else None
        n_elec = spec["nelec_half"]
        lds = lds_bytes_mixed(wc, na, nb, n_elec, max(width, 1), 4.0 if f32 else 8.0, bool(pre))
        roofline = dict(kernel=dom, bound="hbm", achieved=achieved, peak=HBM_PEAK_GBPS,
                        unit="GB/s", frac=round(achieved / HBM_PEAK_GBPS, 4) if achieved else None,
                        traffic=traffic,
                        traffic_source="profiles/c36_hbm_traffic.md (rocprofv3 FETCH_SIZE x2.0 + "
                                       "WRITE_SIZE x1.0, calibrated on streaming kernels)"
                        if traffic else None,
                        avg_ms=dk["avg_ms"],
                        gather_equivalent=dict(bytes=dk["alg_bytes"], GBps=dk["GBps"],
                                               note="every gathered element counted as one read"),
                        lds=dict(bytes=lds, achieved_TBps=round(lds / avg_s / 1e12, 2),
                                 peak_TBps=LDS_PEAK_TBPS, frac=round(lds / avg_s / 1e12 / LDS_PEAK_TBPS, 4),
                                 note="LDS read bytes of the staged gathers; peak = 256 CUs x 256 B/clk "
                                      "x 2.4 GHz"),
                        limiter=limiter)
    s64_ms, _ = sigma_rate(kernels, "")
    s32_ms, _ = sigma_rate(kernels, "_f32")
    if plain is not None:
        pk = kernel_table(plain["stats"])
        s64_ms, _ = sigma_rate(pk, "")
    # sigma GB/s from HBM bytes the counters measured for the product's
    # kernels (profiles/), over the product's live kernel time
    s64_gbps = sigma_hbm_rate(pk if plain is not None else kernels, "", op.same_spin_forms()["fp64"])
    s32_gbps = sigma_hbm_rate(kernels, "_f32", op.same_spin_forms()["fp32"])

    T = res["time"]
    dtypes = {"plain": "f64", "thick": "f64", "mixed": "f32+f64", "block": "f32+f64"}
    out = {
        "metric": METRIC,
        "value": round(T, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * T, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtypes[best],
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: norb={spec['norb']}, n_alpha=n_beta={spec['nelec_half']}, "
                        f"{spec['nstr']}x{spec['nstr']} strings, {op.ndet} determinants, "
                        f"{spec['nchol']} Cholesky vectors, seed {spec['seed']}; Davidson to "
                        f"|E0 W - H W| < {args.eps:g} (nb={args.nb}, start={args.init})",
            "solver": best,
            "subspace": args.nb,
            "restart_keep": args.keep if best in ("thick", "mixed") else None,
            "spin_flip_sectors": args.sectors if symmetric and best in ("thick", "mixed") else "off",
            "sector_results": res.get("sectors"),
            "coarse_start": None if res["coarse"] is None or not args.coarse else
            dict(strings=args.coarse, eps=args.coarse_eps, **res["coarse"]),
            "preconditioner": args.precond,
            "precond_steps": res["precond_steps"],
            "fp64_products_from_step": res.get("fp64_switch_step"),
            "sigma_products": res["steps"],
            "converged": res["converged"],
            "E0": res["E"],
        },
        "plain_fp64": None if plain is None else {
            "time_s": round(plain["time"], 4), "sigma_products": plain["steps"], "E0": plain["E"],
            "converged": plain["converged"]},
        "speedup_vs_plain": None if plain is None else round(plain["time"] / T, 3),
        "hipgraph_solve": graph_res,
        "dE_vs_plain": None if plain is None else abs(res["E"] - plain["E"]),
        "block_roots": res.get("roots"),
        "block_solver": res.get("block_info"),
        "sigma_fp64_ms": s64_ms,
        "sigma_fp64_host_buffers_ms": round(sigma_host_ms, 1),
        "sigma_GBps": s64_gbps["GBps"] if s64_gbps else None,
        "sigma_fp64_hbm": s64_gbps,
        "sigma_fp32_ms": s32_ms,
        "sigma_fp32_GBps": s32_gbps["GBps"] if s32_gbps else None,
        "sigma_fp32_hbm": s32_gbps,
        "sigma_GBps_note": "HBM bytes measured by rocprofv3 counters for the product's kernels "
                           "(profiles/c36_hbm_traffic.md) over their live time",
        "roofline": roofline,
        "kernels": kernels,
        "same_spin_form": op.same_spin_forms(),
        "build_s": round(t_build, 3),
    }
    if args.rdm:
        # one- and two-particle RDMs of the solution, after the timed solves
        # (the vector goes in and the RDMs come out through host buffers)
        Wsol = op.get_vector()
        t0 = time.perf_counter()
        one, two = op.correlation(Wsol)
        L = inst["norb"]
        out["rdm"] = dict(seconds=round(time.perf_counter() - t0, 4),
                          trace_one=float(sum(one[s * L * L + i * (L + 1)] for s in (0, 1) for i in range(L))),
                          note="sbdx_rdm of the solution vector, host buffers included")
        log(f"rdm: {out['rdm']['seconds']} s, trace of the 1-RDM {out['rdm']['trace_one']:.10f}")
    if args.factored:
        out["factored_mixed"] = factored_leg(op, W0 if args.init == "random" else
                                             np.random.default_rng(2).standard_normal(op.ndet), inst["norb"])
    if not args.no_cpu_baseline:
        ref_steps = plain["steps"] if plain is not None else res["steps"]
        log(f"cpu baseline: ~{args.cpu_seconds:.0f} s oracle sample on {args.cpu_threads} threads")
        cb = cpu_baseline(inst, ref_steps, args.cpu_seconds, args.cpu_threads)
        out["cpu_baseline"] = {k: (round(v, 3) if isinstance(v, float) else v)
                               for k, v in cb.items() if k != "sigma_s"}
        out["cpu_baseline"]["sigma_s"] = round(cb["sigma_s"], 3)
        ratio = cpu_calibration()
        if ratio:
            out["cpu_baseline"]["calibration"] = dict(
                reference_over_oracle=ratio, reference_equivalent_value=round(cb["value"] * ratio, 1),
                source="profiles/cpu_calibration.md (reference binary vs oracle, same instances, 8 threads)")
    op.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
