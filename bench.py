"""Benchmark: the LightNE embedding pipeline on one MI355X.

Workload (the headline GPU configuration): Chung-Lu graph, n = 1.14M vertices,
m = 2.99M undirected edges, power-law expected degrees (exponent 2.0, mean
~5.25); NetSMF sparsifier with T = 10 and M = 20 m samples (10 per directed
edge), randomized SVD with k = d + p = 144, d = 128, q = 1, Chebyshev
propagation K = 10, mu = 0.2, theta = 0.5, FP32.  Synthetic, seeded input.

A step is one full embedding: sparsify -> randomized SVD -> U sqrt(S) ->
propagation -> dense tall SVD (the reference's "### Running Time", minus the
file save).  Inputs (graph CSR) are resident in HBM before the timed region.

Prints ONE JSON line.  `value` is the end-to-end embed seconds per step.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "end-to-end embed seconds, sampled edges/s, SpMM HBM GB/s; BF16 vs FP32 at equal err"
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (n, beta, mean_deg, max_deg, graph_seed, T, samples_per_directed_edge, k, d, q, K)
    "c2": dict(n=1_140_000, beta=2.0, mean=5.25, maxdeg=2500.0, gseed=2, T=10, per_edge=10.0, k=144, d=128, q=1,
               K=10, adjoint=True),
    "c1": dict(n=10_312, beta=2.1, mean=65.0, maxdeg=3900.0, gseed=1, T=10, per_edge=50.0, k=144, d=128, q=1,
               K=10),
    # R-MAT scale 24, edge factor 16, (a, b, c) = (.57, .19, .19); T = 5, M = 10 m, q = 2
    "c3": dict(rmat=(24, 16.0, 0.57, 0.19, 0.19), gseed=3, T=5, per_edge=5.0, k=144, d=128, q=2, K=10,
               cpu_frac=1 / 64, cpu_prop_one=True),
    # R-MAT scale 26, edge factor 14, T = 5, M = 10 m; the table takes 8-byte
    # slots and the CSR is built by row ranges to fit one GPU; the propagation
    # writes its result over its input (one n x d buffer less)
    "c5": dict(rmat=(26, 14.0, 0.57, 0.19, 0.19), gseed=5, T=5, per_edge=5.0, k=144, d=128, q=1, K=10,
               cpu_frac=1 / 256, cpu_prop_one=True, inplace=True),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sample_ratio(n, m, T, per_edge):
    return float(np.float32(per_edge * m / (T * n * np.log(n))))


def run_pipeline(L, g, cfg, ratio, precision, ctx, seed=1):
    import torch

    emb, S, st = L.NetSMF(g, walk_len=cfg["T"], rank=cfg["k"], dim=cfg["d"], sample_ratio=ratio, seed=seed,
                          svd_seed=777, power_iters=cfg["q"], precision=precision, ctx=ctx)
    pro, pst = L.chebyshev_expansion(emb, g, cfg["d"], cfg["d"], order=cfg["K"], mu=0.2, s=0.5, precision=precision,
                                     out=emb if cfg.get("inplace") else None, ctx=ctx)
    torch.cuda.synchronize()
    return emb, S, pro, st, pst


def principal_angle(A, B):
    """Largest principal angle between the column spaces of two device n x d
    panels, from FP64 Gram blocks (Cholesky-orthonormalised), row chunks of 1M."""
    import torch

    d = A.shape[1]
    Gaa = torch.zeros(d, d, dtype=torch.float64, device=A.device)
    Gbb, Gab = torch.zeros_like(Gaa), torch.zeros_like(Gaa)
    for r in range(0, A.shape[0], 1 << 20):
        a = A[r:r + (1 << 20)].double()
        b = B[r:r + (1 << 20)].double()
        Gaa += a.T @ a
        Gbb += b.T @ b
        Gab += a.T @ b
    Ra = torch.linalg.cholesky(Gaa.cpu(), upper=True)
    Rb = torch.linalg.cholesky(Gbb.cpu(), upper=True)
    M = torch.linalg.solve_triangular(Ra.T, Gab.cpu(), upper=False)
    M = torch.linalg.solve_triangular(Rb, M, upper=True, left=False)
    s = torch.linalg.svdvals(M).numpy()
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


def cpu_baseline(hg, cfg, ratio, csr_box, frac=0.25):
    """The oracle's CPU restatement of the same pipeline (FP32, all host cores:
    C/OpenMP sampler and SpMM, LAPACK QR/SVD), timed on a bounded sample of the
    same workload and scaled to one full embedding:
      * sampler + trunc-log/CSR on every (1/frac)-th source vertex, scaled by
        (all directed edges / sampled directed edges);
      * randomized SVD: both SpMMs on the rows holding the first `frac` of the
        sparsifier's nonzeros (scaled by nnz), the tall QRs / GEMMs on the
        first `frac` of the rows (scaled by rows), the k x k SVD once; the
        sparsifier is the device's (identical to the oracle's, which the tests
        check);
      * propagation: all 19 sparse products on the full graph, the dense tall
        SVD on the first `frac` of the rows (scaled by rows)."""
    from oracle import oracle as O

    lib = O.load()
    threads = int(lib.orc_max_threads())
    n, m, T, k, d = hg.n, hg.m, cfg["T"], cfg["k"], cfg["d"]
    stride = max(1, int(round(1.0 / frac)))
    e_done = int(np.diff(hg.offsets)[::stride].sum())
    t0 = time.perf_counter()
    per_edge = int(np.ceil(cfg["per_edge"])) + 1
    keys, cnt, wgh, st = O.sample(hg.offsets, hg.edges, 1, T, ratio, 0, threads=threads,
                                  cap_log2=int(np.ceil(np.log2(2 * e_done * per_edge))),
                                  vertex_range=(0, n, stride))
    t_s1 = time.perf_counter() - t0
    log(f"cpu: sampler {t_s1:.2f}s")
    t0 = time.perf_counter()
    O.trunc_log_csr(hg.offsets, keys, wgh)
    t_csr = time.perf_counter() - t0
    log(f"cpu: trunc-log csr {t_csr:.2f}s")
    s1 = m / e_done
    del keys, cnt, wgh
    rp, col, val = csr_box.pop()  # the only reference: freed after the products
    nnz = int(rp[-1])
    r2 = int(np.searchsorted(rp, frac * nnz))
    nnz2 = int(rp[r2])
    sub = (rp[: r2 + 1], col[:nnz2], val[:nnz2])
    Om = O.gaussian(777, 0, n * k).reshape(n, k)
    t0 = time.perf_counter()
    Y = O.spmm_f32(*sub, Om, threads)
    B = O.spmm_f32(*sub, Om, threads)
    # 1 + q products: time two, scale to all
    t_spmm = (time.perf_counter() - t0) * (1 + cfg["q"]) / 2
    log(f"cpu: spmm {t_spmm:.2f}s")
    del rp, col, val, sub
    r3 = int(frac * n)
    Yr, Br = np.ascontiguousarray(Om[:r3]), np.ascontiguousarray(Om[r3:2 * r3])
    t0 = time.perf_counter()
    Q = np.linalg.qr(Yr)[0]
    P = O.gaussian(777, 1, k * k).reshape(k, k)
    Z = np.linalg.qr(Br @ P)[0]
    Cm = Z.T @ Br
    t_rows = time.perf_counter() - t0
    t0 = time.perf_counter()
    Uc, S, _ = np.linalg.svd(Cm)
    t_small = time.perf_counter() - t0
    t0 = time.perf_counter()
    U = Z @ Uc
    t_rows += time.perf_counter() - t0
    del Y, B, Q, Z, U, Yr, Br
    t_svd = t_spmm * nnz / max(nnz2, 1) + t_rows * n / max(r3, 1) + t_small
    log(f"cpu: rsvd scaled {t_svd:.2f}s")
    # stage 3
    off64 = np.ascontiguousarray(hg.offsets, np.uint64)
    edg = np.ascontiguousarray(hg.edges, np.uint32)
    c = O.cheb_coeffs(cfg["K"], 0.5).astype(np.float32)
    a = np.ascontiguousarray(Om[:, :d])
    del Om
    t0 = time.perf_counter()

    def P_(X, alpha, beta, Yb):
        lib.orc_prop_apply_f32(n, off64.ctypes.data, edg.ctypes.data, 0.2, X.ctypes.data, d, alpha, beta,
                               Yb.ctypes.data, threads)

    tmp = np.empty_like(a)
    if cfg.get("cpu_prop_one"):
        # large graphs: one sparse product and one recurrence update on the
        # full graph, scaled to the 2K - 1 products and K - 1 updates
        P_(a, 1.0, 0.0, tmp)
        t_p = time.perf_counter() - t0
        # the recurrence's vector update, timed on the first rows and scaled
        ru = max(1, int(frac * n))
        t1 = time.perf_counter()
        Lx2 = -a[:ru]
        Lx2 -= 2.0 * tmp[:ru]
        conv = c[0] * a[:ru] + c[1] * Lx2
        t_u = (time.perf_counter() - t1) * n / ru
        t_prop = t_p * (2 * cfg["K"]) + t_u * (cfg["K"] - 1)
        del a, Lx2, conv
        z = tmp
        prop_note = f"propagation: 1 of {2 * cfg['K'] - 1} products timed on the full graph ({t_p:.2f} s)"
    else:
        P_(a, 1.0, 0.0, tmp)
        Lx1 = -a
        P_(tmp, 0.5, 1.0, Lx1)
        conv = c[0] * a + c[1] * Lx1
        Lx0 = a
        for i in range(2, cfg["K"]):
            P_(Lx1, 1.0, 0.0, tmp)
            Lx2 = -Lx0
            P_(tmp, 1.0, 1.0, Lx2)
            Lx2 -= 2.0 * Lx1
            conv += c[i] * Lx2
            Lx0, Lx1 = Lx1, Lx2
        y = np.ascontiguousarray((a - conv).astype(np.float64))
        z = np.empty_like(y)
        lib.orc_aggregate_f64(n, off64.ctypes.data, edg.ctypes.data, y.ctypes.data, d, z.ctypes.data, threads)
        t_prop = time.perf_counter() - t0
        prop_note = f"propagation on the full graph ({t_prop:.2f} s)"
    log(f"cpu: propagation {t_prop:.2f}s")
    t0 = time.perf_counter()
    np.linalg.svd(z[:r3].astype(np.float32), full_matrices=False)
    t_tall = (time.perf_counter() - t0) * n / r3
    t_sample = t_s1 * s1
    t_csr_full = t_csr * s1
    total = t_sample + t_csr_full + t_svd + t_prop + t_tall
    return {
        "value": round(total, 3), "unit": "s", "cores": threads, "kind": "port",
        "sample": (f"oracle restatement, FP32, {threads} threads; sampler+CSR on {e_done}/{m} directed edges "
                   f"({t_s1:.2f}+{t_csr:.2f} s, x{s1:.2f}); rSVD SpMMs on {nnz2}/{nnz} nnz ({t_spmm:.2f} s, "
                   f"x{nnz / max(nnz2, 1):.2f}), QR/GEMM on {r3}/{n} rows ({t_rows:.2f} s, x{n / max(r3, 1):.2f}); "
                   f"{prop_note}; tall SVD on {r3}/{n} rows"),
        "sampled_edges_per_s": st["samples"] / t_s1,
        "stages_s": {"sample": round(t_sample, 3), "csr": round(t_csr_full, 3), "svd": round(t_svd, 3),
                     "propagation": round(t_prop + t_tall, 3)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-bf16", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy 1 GiB device-to-device 3 times (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU (the whole pipeline on one MI355X); --gpus {args.gpus} "
              "is not supported", file=sys.stderr)
        sys.exit(2)
    import torch

    import lightne_amd as L

    cfg = CONFIGS[args.config]
    t0 = time.perf_counter()
    if "rmat" in cfg:
        hg = L.rmat(*cfg["rmat"], cfg["gseed"])
    else:
        hg = L.chung_lu(cfg["n"], cfg["beta"], cfg["mean"], cfg["maxdeg"], cfg["gseed"])
    ratio = sample_ratio(hg.n, hg.m, cfg["T"], cfg["per_edge"])
    log(f"graph n={hg.n} m={hg.m} (undirected {hg.m // 2}) generated in {time.perf_counter() - t0:.1f}s; "
        f"sample_ratio={ratio}")
    ctx = L.context(0)
    g = hg.to_device("cuda:0")
    torch.cuda.synchronize()
    free_b, total_b = torch.cuda.mem_get_info(0)
    log(f"device memory: {total_b / 1e9:.1f} GB total, {free_b / 1e9:.1f} GB free with the graph resident")
    for w in range(args.warmup):
        run_pipeline(L, g, cfg, ratio, L.FP32, ctx)
        log(f"warmup {w} done")
    torch.cuda.synchronize()
    times = []
    t_start = time.perf_counter()
    emb = pro = None
    for s in range(args.steps):
        emb = pro = None  # the previous step's outputs go back to the allocator first
        t1 = time.perf_counter()
        emb, S, pro, st, pst = run_pipeline(L, g, cfg, ratio, L.FP32, ctx)
        times.append(time.perf_counter() - t1)
        log(f"step {s}: {times[-1]:.3f}s")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    free_b, total_b = torch.cuda.mem_get_info(0)
    log(f"device memory after the timed steps: {(total_b - free_b) / 1e9:.1f} of {total_b / 1e9:.1f} GB in use")
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    ms_per_step = 1000.0 * wall / args.steps
    sp, rs = st["sparsify"], st["rsvd"]
    n, k, d = hg.n, cfg["k"], cfg["d"]
    nnz = sp["nnz"]
    # SpMM gather model (FP32 panels): row ptrs + (col, val) + gathered rows + written rows
    spmm_bytes = 8 * (n + 1) + nnz * 8 + nnz * k * 4 + n * k * 4
    t_spmm_call = rs["t_spmm"] / max(rs["spmm_calls"], 1)
    spmm_gbs = spmm_bytes / t_spmm_call / 1e9
    sampled_eps = sp["samples"] / sp["t_sample"] if sp["t_sample"] > 0 else None
    # sampler byte model: 20 B / walk step + one slot read and written per
    # sample (8-byte or 16-byte slots) + 12 B / directed edge
    slot_b = 8 if sp["table_layout"] == 2 else 16
    sampler_bytes = 20 * sp["walk_steps"] + 2 * slot_b * sp["samples"] + 12 * hg.m
    sampler_gbs = sampler_bytes / sp["t_sample"] / 1e9
    stages = {"sample": sp["t_sample"], "table": sp["t_table"], "csr": sp["t_csr"], "sparsify_wall": st["t_sparsify"],
              "rsvd_spmm": rs["t_spmm"], "rsvd_qr": rs["t_qr"], "rsvd_gemm": rs["t_gemm"],
              "rsvd_small_svd": rs["t_small_svd"], "rsvd_total": rs["t_total"], "netsmf_total": st["t_total"],
              "prop_cheb": pst["t_cheb"], "prop_aggregate": pst["t_aggregate"], "prop_svd": pst["t_svd"],
              "prop_total": pst["t_total"]}
    # dominant kernel by measured device time; stage-2 SpMM = k_spmm<float,1,false>
    # (+ k_spmm_fixup for rows split over several chunks)
    cands = {"spmm_stage2": (rs["t_spmm"], spmm_bytes * rs["spmm_calls"], rs["spmm_calls"]),
             "sampler": (sp["t_sample"], sampler_bytes, 1)}
    dom = max(cands, key=lambda c: cands[c][0])
    t_dom, bytes_dom, launches = cands[dom]
    achieved = bytes_dom / t_dom / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{args.config}_{dom}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": int(bytes_dom / launches), "t_per_launch_ms": round(1e3 * t_dom / launches, 3)}
    bf16 = None
    if not args.skip_bf16:
        # stage 2 against the FP64 form (same sparsifier, same sketch): top-d
        # principal angle and singular values
        # the FP64 form holds four n x k FP64 panels; the stage-3 check two
        # extra n x d pairs: both only where they fit beside the pipeline
        big = n * k * 8 * 4 > 120e9
        emb64 = S64 = None
        bf16 = {"fp64_reference": None if big else "FP64 form of the randomized SVD, same sparsifier and sketch"}
        if not big:
            log("fp64 reference ...")
            emb64, S64, st64 = L.NetSMF(g, walk_len=cfg["T"], rank=cfg["k"], dim=cfg["d"], sample_ratio=ratio,
                                        seed=1, svd_seed=777, power_iters=cfg["q"], precision=L.FP64, ctx=ctx)
            torch.cuda.synchronize()
            S64 = S64.astype(np.float64)
            # stage 3 on one input (the FP32 stage-2 embedding): z of the BF16
            # iterates against z of the FP32 ones, relative Frobenius error
            _, z32, _ = L.propagate(emb, g, order=cfg["K"], mu=0.2, theta=0.5, precision=L.FP32, ctx=ctx)
            _, z16, _ = L.propagate(emb, g, order=cfg["K"], mu=0.2, theta=0.5, precision=L.BF16, ctx=ctx)
            prop_err = float((torch.linalg.vector_norm(z16 - z32) / torch.linalg.vector_norm(z32)).item())
            del z32, z16
            bf16.update({"fp32_stage2_angle_vs_fp64_rad": principal_angle(emb, emb64),
                         "fp32_sv_rel_err_vs_fp64": float(np.max(np.abs(S[:d] - S64[:d]) / S64[:d])),
                         "spectral_gap_d": float(S64[d - 1] / S64[d]), "stage3_z_rel_err_vs_fp32": prop_err,
                         "fp64_rsvd_s": round(st64["rsvd"]["t_total"], 4)})

        def sv_err(Sx):
            return None if S64 is None else float(np.max(np.abs(Sx[:d].astype(np.float64) - S64[:d]) / S64[:d]))

        def angle64(E):
            return None if emb64 is None else principal_angle(E, emb64)
        forms = [("bf16", L.BF16)] + ([("bf16_power", L.BF16_POWER)] if cfg["q"] > 1 else [])
        if big:  # room for the BF16 runs' own outputs
            emb = pro = None
        for name, prec in forms:
            run_pipeline(L, g, cfg, ratio, prec, ctx)  # warm
            t1 = time.perf_counter()
            embb, Sb, prob, stb, pstb = run_pipeline(L, g, cfg, ratio, prec, ctx)
            tb = time.perf_counter() - t1
            log(f"{name} step: {tb:.3f}s")
            bf16[name] = {"seconds": round(tb, 4), "speedup_vs_fp32": round((ms_per_step / 1e3) / tb, 3),
                          "stage2_angle_vs_fp64_rad": angle64(embb), "sv_rel_err_vs_fp64": sv_err(Sb),
                          "stage2_angle_vs_fp32_rad": None if emb is None else principal_angle(embb, emb),
                          "sv_identical_to_fp32": bool(np.array_equal(Sb, S)),
                          "spmm_ms_per_call": round(1e3 * stb["rsvd"]["t_spmm"] / max(stb["rsvd"]["spmm_calls"], 1),
                                                    3),
                          "prop_cheb_s": pstb["t_cheb"]}
            del embb, prob
        del emb64
        torch.cuda.empty_cache()
    adjoint = None
    if cfg.get("adjoint") and not args.skip_adjoint:
        # differentiation form: forward z = (A + I)(a - sum c_i T_i a) and the
        # hand-written adjoint (dL/da, dL/dmu, dL/dtheta, checkpoint every 3
        # steps) for a ~ N(0, 1), dL/dz ~ N(0, 1)
        gen = torch.Generator(device="cuda:0").manual_seed(4)
        a = torch.randn(hg.n, d, generator=gen, device="cuda:0")
        gz = torch.randn(hg.n, d, generator=gen, device="cuda:0")
        tf, tbk = [], []
        for it in range(3):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            L.propagate(a, g, order=cfg["K"], mu=0.2, theta=0.5, ctx=ctx)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            L.propagate_backward(a, gz, g, order=cfg["K"], mu=0.2, theta=0.5, aggregate=True, checkpoint_every=3,
                                 ctx=ctx)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            if it:  # the first pair warms up
                tf.append(t2 - t1)
                tbk.append(t3 - t2)
        adjoint = {"forward_s": round(float(np.median(tf)), 5), "backward_s": round(float(np.median(tbk)), 5),
                   "checkpoint_every": 3, "d": d, "K": cfg["K"]}
        log(f"adjoint: forward {adjoint['forward_s']}s backward {adjoint['backward_s']}s")
        del a, gz
    cpu = None
    if not args.skip_cpu:
        log("cpu baseline ...")
        emb = pro = None
        torch.cuda.empty_cache()  # torch's cached blocks back to the device for the sparsifier's CSR
        if cfg.get("inplace"):
            # the pipeline context's workspace goes back first; only the rows
            # the CPU products use come to the host
            ctx.close()
            L.api._CTX.pop(0, None)
            ctx = L.context(0)
        csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=cfg["T"], sample_ratio=ratio, ctx=ctx)
        csr_box = [csr.to_numpy(nnz_fraction=cfg.get("cpu_frac", 0.25) if cfg.get("inplace") else None)]
        csr.free()
        log("cpu: sparsifier copied to the host")
        cpu = cpu_baseline(hg, cfg, ratio, csr_box, frac=cfg.get("cpu_frac", 0.25))
    out = {
        "metric": METRIC, "value": round(ms_per_step / 1e3, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (seeded {'R-MAT' if 'rmat' in cfg else 'Chung-Lu'} graph)",
        "config": {"workload": f"{args.config}: {'R-MAT' if 'rmat' in cfg else 'Chung-Lu'} n={hg.n} "
                               f"m_undirected={hg.m // 2} T={cfg['T']} M={sp['samples']} k={k} d={d} q={cfg['q']} "
                               f"K={cfg['K']} mu=0.2 theta=0.5 FP32",
                   "parallelism": "1 GPU"},
        "sampled_edges_per_s": sampled_eps, "sampler_gbs": round(sampler_gbs, 1), "spmm_gbs": round(spmm_gbs, 1),
        "sparsifier_nnz": nnz, "samples": sp["samples"], "distinct_keys": sp["distinct_keys"],
        "table_capacity": sp["table_capacity"], "table_slot_bytes": 8 if sp["table_layout"] == 2 else 16,
        "csr_row_ranges": sp["csr_parts"], "csr_value_cache": bool(sp["value_cache"]), "stages_s": {k_: round(v, 5) for k_, v in stages.items()},
        "bf16": bf16, "adjoint": adjoint, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
