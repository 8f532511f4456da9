"""Benchmark: the LightNE embedding pipeline on one MI355X.

Workload (the headline configuration): R-MAT scale 26 (n = 67.1M vertices),
edge factor 14, (a, b, c) = (.57, .19, .19), symmetrised and deduplicated
(0.92B undirected edges); NetSMF sparsifier with T = 5 and M = 10 m samples (5
per directed edge), built in row ranges to fit one GPU; randomized SVD with
k = d + p = 144, d = 128, q = 1; Chebyshev propagation K = 10, mu = 0.2,
theta = 0.5; dense tall SVD.  FP32 form.  Synthetic, seeded input.
``--config c1|c2|c3`` selects the smaller configurations.

A step is one full embedding: sparsify -> randomized SVD -> U sqrt(S) ->
propagation -> dense tall SVD, with the graph resident in HBM (the
reference's "### Running Time" minus the two file saves; ``--with-save`` times
one more step with the host copies and file writes of both embeddings).

Beside the timed steps the JSON carries: the BF16 form's time and its
accuracy (stage 2 bit-identity with FP32, final-embedding angle against
FP32), the FP32 stage-2 result against the FP64 form of the randomized SVD
(the lean two-panel FP64 form where the explicit one does not fit), the
differentiation form on the 1.14M-vertex Chung-Lu graph (d = 128, K = 10,
checkpoint every 3), the roofline of the dominant kernel, and the CPU
baseline (the oracle's stage-1 restatement plus the reference's MKL calls for
stages 2 and 3) on the host cores.

Prints ONE JSON line.  `value` is the end-to-end embed seconds per step.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "end-to-end embed seconds, sampled edges/s, SpMM HBM GB/s; BF16 vs FP32 at equal err"
HBM_PEAK_GBS = 8000.0
CHUNK_ROWS = 1 << 20

CONFIGS = {
    "c1": dict(n=10_312, beta=2.1, mean=65.0, maxdeg=3900.0, gseed=1, T=10, per_edge=50.0, k=144, d=128, q=1,
               K=10, cpu_frac=1.0),
    "c2": dict(n=1_140_000, beta=2.0, mean=5.25, maxdeg=2500.0, gseed=2, T=10, per_edge=10.0, k=144, d=128, q=1,
               K=10, cpu_frac=1.0),
    # R-MAT scale 24, edge factor 16, (a, b, c) = (.57, .19, .19); T = 5, M = 10 m, q = 2
    "c3": dict(rmat=(24, 16.0, 0.57, 0.19, 0.19), gseed=3, T=5, per_edge=5.0, k=144, d=128, q=2, K=10,
               cpu_frac=1 / 64),
    # R-MAT scale 26, edge factor 14, T = 5, M = 10 m; the table takes 8-byte
    # slots and the CSR is built by row ranges to fit one GPU; the propagation
    # writes its result over its input (one n x d buffer less)
    "c5": dict(rmat=(26, 14.0, 0.57, 0.19, 0.19), gseed=5, T=5, per_edge=5.0, k=144, d=128, q=1, K=10,
               cpu_frac=1 / 256, inplace=True),
}
# the differentiation form's workload: the c2 graph, d = 128, K = 10
ADJOINT = dict(n=1_140_000, beta=2.0, mean=5.25, maxdeg=2500.0, gseed=2, d=128, K=10, checkpoint=3)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sample_ratio(n, m, T, per_edge):
    return float(np.float32(per_edge * m / (T * n * np.log(n))))


def make_graph(L, cfg):
    if "rmat" in cfg:
        return L.rmat(*cfg["rmat"], cfg["gseed"])
    return L.chung_lu(cfg["n"], cfg["beta"], cfg["mean"], cfg["maxdeg"], cfg["gseed"])


def run_pipeline(L, g, cfg, ratio, precision, ctx, after_stage2=None, save_dir=None):
    """One embedding.  ``after_stage2(emb, S)`` runs between the stages and is
    not timed; returns (emb, S, pro, stats, prop stats, seconds)."""
    import torch

    t0 = time.perf_counter()
    emb, S, st = L.NetSMF(g, walk_len=cfg["T"], rank=cfg["k"], dim=cfg["d"], sample_ratio=ratio, seed=1,
                          svd_seed=777, power_iters=cfg["q"], precision=precision, ctx=ctx)
    torch.cuda.synchronize()
    if save_dir is not None:  # LightNE.cc:108-111: -ne_out is written before propagation
        emb.cpu().numpy().tofile(os.path.join(save_dir, "emb.ne"))
    t1 = time.perf_counter()
    if after_stage2 is not None:
        after_stage2(emb, S)
    t2 = time.perf_counter()
    pro, pst = L.chebyshev_expansion(emb, g, cfg["d"], cfg["d"], order=cfg["K"], mu=0.2, s=0.5, precision=precision,
                                     out=emb if cfg.get("inplace") else None, ctx=ctx)
    torch.cuda.synchronize()
    if save_dir is not None:
        pro.cpu().numpy().tofile(os.path.join(save_dir, "emb.pro"))
    t3 = time.perf_counter()
    return emb, S, pro, st, pst, (t1 - t0) + (t3 - t2)


def _dev_rows(X, r0, r1):
    import torch

    if isinstance(X, np.ndarray):
        return torch.from_numpy(X[r0:r1]).to("cuda:0").double()
    return X[r0:r1].double()


def principal_angle(A, B):
    """Largest principal angle between the column spaces of two n x d panels
    (device tensors or host arrays), from FP64 Gram blocks accumulated over
    row chunks on the device (Cholesky-orthonormalised)."""
    import torch

    d = A.shape[1]
    Gaa = torch.zeros(d, d, dtype=torch.float64, device="cuda:0")
    Gbb, Gab = torch.zeros_like(Gaa), torch.zeros_like(Gaa)
    for r in range(0, A.shape[0], CHUNK_ROWS):
        a = _dev_rows(A, r, r + CHUNK_ROWS)
        b = _dev_rows(B, r, r + CHUNK_ROWS)
        Gaa += a.T @ a
        Gbb += b.T @ b
        Gab += a.T @ b
    Ra = torch.linalg.cholesky(Gaa.cpu(), upper=True)
    Rb = torch.linalg.cholesky(Gbb.cpu(), upper=True)
    M = torch.linalg.solve_triangular(Ra.T, Gab.cpu(), upper=False)
    M = torch.linalg.solve_triangular(Rb, M, upper=True, left=False)
    s = torch.linalg.svdvals(M).numpy()
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


def same_bits(A, B) -> bool:
    """A and B (device tensor / host array) hold the same FP32 bits."""
    import torch

    for r in range(0, A.shape[0], CHUNK_ROWS):
        a = A[r:r + CHUNK_ROWS] if not isinstance(A, np.ndarray) else torch.from_numpy(A[r:r + CHUNK_ROWS]).cuda()
        b = B[r:r + CHUNK_ROWS] if not isinstance(B, np.ndarray) else torch.from_numpy(B[r:r + CHUNK_ROWS]).cuda()
        if not torch.equal(a.view(torch.int32), b.view(torch.int32)):
            return False
    return True


def to_host(t) -> np.ndarray:
    return t.cpu().numpy()


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------
def cpu_baseline(hg, cfg, ratio, csr_parts):
    """The reference method on the host cores: stage 1 is the oracle's C/OpenMP
    restatement (the reference's pbbs/MKL build cannot be compiled here);
    stages 2 and 3 make the reference's MKL calls (oracle/mkl_ref.py) when
    Intel MKL is installed, else the oracle's C kernels and numpy LAPACK.

    ``cpu_frac`` = 1 runs the whole workload, unscaled.  Below 1 it is a
    bounded sample, reported as an estimate with every factor named:
      * sampler + trunc-log/CSR on every (1/frac)-th source vertex, scaled by
        (all directed edges / sampled directed edges);
      * randomized SVD: the sparse products on the rows holding the first
        `frac` of the sparsifier's nonzeros (scaled by nnz), the tall QRs and
        GEMMs on the first `frac` of the rows (scaled by rows; dense work
        does not depend on the values), the k x k SVD once;
      * propagation: one of the 19 sparse products on the rows holding `frac`
        of the operator's nonzeros (scaled by nnz and by 19), the vector
        updates and the tall SVD on the first `frac` of the rows (scaled by
        rows)."""
    from oracle import mkl_ref as K
    from oracle import oracle as O

    lib = O.load()
    threads = int(lib.orc_max_threads())
    use_mkl = K.load(threads) is not None
    n, m, T, k, d, q = hg.n, hg.m, cfg["T"], cfg["k"], cfg["d"], cfg["q"]
    frac = cfg["cpu_frac"]
    full = frac >= 1.0
    stride = max(1, int(round(1.0 / frac)))
    e_done = int(np.diff(hg.offsets)[::stride].sum())
    note = []
    # stage 1
    t0 = time.perf_counter()
    per_edge = int(np.ceil(cfg["per_edge"])) + 1
    keys, cnt, wgh, st = O.sample(hg.offsets, hg.edges, 1, T, ratio, 0, threads=threads,
                                  cap_log2=int(np.ceil(np.log2(2 * e_done * per_edge))),
                                  vertex_range=(0, n, stride))
    t_s1 = time.perf_counter() - t0
    log(f"cpu: sampler {t_s1:.2f}s")
    t0 = time.perf_counter()
    rp1, col1, val1, _ = O.trunc_log_csr(hg.offsets, keys, wgh)
    t_csr = time.perf_counter() - t0
    log(f"cpu: trunc-log csr {t_csr:.2f}s")
    del keys, cnt, wgh
    s1 = m / e_done
    if full:
        rp, col, val = rp1, col1, val1
    else:
        del rp1, col1, val1
        rp, col, val = csr_parts  # the device's sparsifier (bit-identical to the oracle's), leading rows
        note.append(f"sampler+CSR on {e_done}/{m} directed edges ({t_s1:.2f}+{t_csr:.2f} s, x{s1:.2f})")
    nnz = int(rp[-1])
    # stage 2
    r2 = n if full else int(np.searchsorted(rp, frac * nnz))
    nnz2 = int(rp[r2])
    rr = n if full else max(1, min(n, int(frac * n)))  # rows of the dense sample
    Om = O.gaussian(777, 0, n * k).reshape(n, k)
    Pm = O.gaussian(777, 1, k * k).reshape(k, k)
    if use_mkl:
        A = K.SparseCSR(rp[:r2 + 1], col[:nnz2], val[:nnz2], ncols=n, hint_k=k, hint_calls=1 + q)
        t0 = time.perf_counter()
        Y = A.mm(Om)
        t_sp = time.perf_counter() - t0
        if not full:  # the dense steps on the first rr rows of a panel of the same shape
            Y = np.ascontiguousarray(Om[:rr])
        t0 = time.perf_counter()
        K.gram_schmidt(Y)
        t_rows = time.perf_counter() - t0
        for _ in range(q - 1):
            t0 = time.perf_counter()
            Y2 = A.mm(Y if full else Om)  # a row sample multiplies an n x k panel of the same shape
            t_sp += time.perf_counter() - t0
            Y = Y2 if full else np.ascontiguousarray(Om[:rr])
            t0 = time.perf_counter()
            K.gram_schmidt(Y)
            t_rows += time.perf_counter() - t0
        t0 = time.perf_counter()
        B = A.mm(Y if full else Om)
        t_sp += time.perf_counter() - t0
        if not full:
            B = np.ascontiguousarray(Om[:rr])
        t0 = time.perf_counter()
        Z = K.sgemm(B, Pm)
        K.gram_schmidt(Z)
        Cm = K.sgemm(Z, B, trans_a=True)
        t_rows += time.perf_counter() - t0
        t0 = time.perf_counter()
        Uc, S = K.gesvd_left(Cm)
        t_small = time.perf_counter() - t0
        t0 = time.perf_counter()
        U = K.sgemm(Z, Uc)
        t_rows += time.perf_counter() - t0
        A.close()
        del Y, B, Z, Cm, A
    else:
        sub = (rp[:r2 + 1], col[:nnz2], val[:nnz2])
        t0 = time.perf_counter()
        Y = O.spmm_f32(*sub, Om, threads)
        for _ in range(q):
            Y = O.spmm_f32(*sub, Om, threads)
        t_sp = time.perf_counter() - t0
        t0 = time.perf_counter()
        Q = np.linalg.qr(Y)[0]
        Z = np.linalg.qr(Y @ Pm)[0]
        Cm = Z.T @ Y
        t_rows = time.perf_counter() - t0
        t0 = time.perf_counter()
        Uc, S, _ = np.linalg.svd(Cm)
        t_small = time.perf_counter() - t0
        t0 = time.perf_counter()
        U = Z @ Uc
        t_rows += time.perf_counter() - t0
        del Y, Q, Z, Cm
    s2n, s2r = nnz / max(nnz2, 1), n / max(rr, 1)
    t_svd = t_sp * s2n + t_rows * s2r + t_small
    # what ran, unscaled: seconds of each timed piece and the work it covered
    measured = {"sampler": {"s": round(t_s1, 3), "directed_edges": e_done, "samples": int(st["samples"])},
                "trunc_log_csr": {"s": round(t_csr, 3), "directed_edges": e_done},
                "rsvd_sparse_products": {"s": round(t_sp, 3), "nnz": nnz2, "products": 1 + q},
                "rsvd_dense": {"s": round(t_rows, 3), "rows": rr},
                "rsvd_small_svd": {"s": round(t_small, 3)}}
    if not full:
        note.append(f"rSVD sparse products on {nnz2}/{nnz} nnz ({t_sp:.2f} s, x{s2n:.2f}), QR/GEMM on {rr}/{n} rows "
                    f"({t_rows:.2f} s, x{s2r:.2f}), k x k SVD once")
    log(f"cpu: rsvd {t_svd:.2f}s")
    del rp, col, val
    # U sqrt(S) (PathEmbed.h:371-408)
    a = np.ascontiguousarray(U[:, :d] * np.sqrt(S[:d])[None, :]) if full else None
    del U
    # stage 3
    off, edg = hg.offsets, hg.edges
    deg_nnz = np.asarray(off, np.int64) + np.arange(n + 1)  # rows of M hold deg + 1 entries
    r3 = n if full else int(np.searchsorted(deg_nnz, frac * deg_nnz[-1]))
    nnz3, nnz3_all = int(deg_nnz[r3]), int(deg_nnz[-1])
    c = O.cheb_coeffs(cfg["K"], float(np.float32(0.5)))
    if use_mkl:
        M = K.prop_operator(off, edg, 0.2, rows_end=None if full else r3)
        if full:
            t0 = time.perf_counter()
            y = K.chebyshev(M, a, c)
            Agg = K.aggregate_operator(off, edg)
            z = Agg.mm(y)
            t_prop = time.perf_counter() - t0
            t0 = time.perf_counter()
            K.gesvd_left(z)
            t_tall = time.perf_counter() - t0
        else:
            # the products: M_rows x (n x d); the updates and the tall SVD on r3 rows
            Xd = Om  # n x k, its first d columns as the n x d operand (row stride k)
            Yr = np.empty((r3, d), np.float32)
            t0 = time.perf_counter()
            M.mm(Xd, 1.0, 0.0, Yr, cols=d)
            t_p = time.perf_counter() - t0
            ar = np.ascontiguousarray(Om[:rr, :d])
            t0 = time.perf_counter()
            conv = np.zeros_like(ar)
            Lx2 = ar.copy()
            for i in range(2, cfg["K"]):  # the recurrence's per-step vector work (SpectralPropagation.h:164-197)
                np.copyto(Lx2, ar)
                K.saxpy(-2.0, ar, Lx2)
                K.saxpy(float(np.float32(c[i])), Lx2, conv)
            t_u = time.perf_counter() - t0
            t0 = time.perf_counter()
            K.gesvd_left(ar.copy())
            t_tall = time.perf_counter() - t0
            s3n, s3r = nnz3_all / max(nnz3, 1), n / max(rr, 1)
            measured["propagation_one_product"] = {"s": round(t_p, 3), "operator_nnz": nnz3}
            measured["propagation_updates"] = {"s": round(t_u, 3), "rows": rr}
            measured["tall_svd"] = {"s": round(t_tall, 3), "rows": rr}
            t_prop = t_p * s3n * (2 * cfg["K"] - 1) + t_u * s3r
            t_tall *= s3r
            note.append(f"propagation: 1 of {2 * cfg['K'] - 1} products on {nnz3}/{nnz3_all} operator nnz "
                        f"({t_p:.2f} s, x{s3n:.2f} x{2 * cfg['K'] - 1}), updates and tall SVD on {rr}/{n} rows "
                        f"(x{s3r:.2f})")
            del Yr, ar, conv, Lx2
        M.close()
    else:
        off64 = np.ascontiguousarray(off, np.uint64)
        edg32 = np.ascontiguousarray(edg, np.uint32)
        a = a if a is not None else np.ascontiguousarray(Om[:, :d])
        t0 = time.perf_counter()
        tmp = np.empty_like(a)

        def P_(X, alpha, beta, Yb):
            lib.orc_prop_apply_f32(n, off64.ctypes.data, edg32.ctypes.data, 0.2, X.ctypes.data, d, alpha, beta,
                                   Yb.ctypes.data, threads)

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
        lib.orc_aggregate_f64(n, off64.ctypes.data, edg32.ctypes.data, y.ctypes.data, d, z.ctypes.data, threads)
        t_prop = time.perf_counter() - t0
        t0 = time.perf_counter()
        np.linalg.svd(z.astype(np.float32), full_matrices=False)
        t_tall = time.perf_counter() - t0
        note.append("propagation with the oracle's C kernels (MKL absent), full graph")
        measured["propagation"] = {"s": round(t_prop, 3)}
        measured["tall_svd"] = {"s": round(t_tall, 3), "rows": n}
    del Om
    log(f"cpu: propagation {t_prop:.2f}s tall svd {t_tall:.2f}s")
    t_sample, t_csr_full = t_s1 * s1, t_csr * s1
    total = t_sample + t_csr_full + t_svd + t_prop + t_tall
    libs = (f"stage 1: oracle C/OpenMP restatement; stages 2-3: reference's MKL calls ({K.version()}, ILP64, "
            f"GNU threads)" if use_mkl else "stage 1: oracle C/OpenMP; stages 2-3: oracle C kernels + numpy LAPACK")
    sample = (f"whole workload, unscaled, {threads} threads; {libs}" if full else
              f"ESTIMATE from bounded samples of the same workload, {threads} threads; {libs}; " + "; ".join(note))
    if full and use_mkl:
        measured["propagation"] = {"s": round(t_prop, 3)}
        measured["tall_svd"] = {"s": round(t_tall, 3), "rows": n}
    return {
        "value": round(total, 3), "unit": "s", "cores": threads, "kind": "port", "estimate": not full,
        "sample": sample, "mkl": use_mkl, "measured_unscaled": measured,
        "sampled_edges_per_s": st["samples"] / t_s1,
        "stages_s": {"sample": round(t_sample, 3), "csr": round(t_csr_full, 3), "svd": round(t_svd, 3),
                     "propagation": round(t_prop + t_tall, 3)},
    }


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c5", choices=sorted(CONFIGS))
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-bf16", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--with-save", action="store_true",
                    help="time one more step that copies both embeddings to the host and writes them to files "
                         "(LightNE.cc:18-27,108-120)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy 1 GiB device-to-device 3 times (a kernel of known traffic "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--gather-calibration", action="store_true",
                    help="after the timed steps, gather 2^25 random 64-byte rows of a 4 GiB table 3 times "
                         "(index_select: 2 GiB of rows and 256 MiB of indices read, 2 GiB written; random access "
                         "of known traffic for scaling FETCH_SIZE on gather kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU (the whole pipeline on one MI355X); --gpus {args.gpus} "
              "is not supported", file=sys.stderr)
        sys.exit(2)
    import torch

    import lightne_amd as L

    cfg = CONFIGS[args.config]
    big = bool(cfg.get("inplace"))  # outputs compared through host copies, FP64 in the lean form
    t0 = time.perf_counter()
    hg = make_graph(L, cfg)
    ratio = sample_ratio(hg.n, hg.m, cfg["T"], cfg["per_edge"])
    log(f"graph n={hg.n} m={hg.m} (undirected {hg.m // 2}) generated in {time.perf_counter() - t0:.1f}s; "
        f"sample_ratio={ratio}")
    ctx = L.context(0)
    g = hg.to_device("cuda:0")
    torch.cuda.synchronize()
    free_b, total_b = torch.cuda.mem_get_info(0)
    log(f"device memory: {total_b / 1e9:.1f} GB total, {free_b / 1e9:.1f} GB free with the graph resident")
    for w in range(args.warmup):
        t1 = time.perf_counter()
        run_pipeline(L, g, cfg, ratio, L.FP32, ctx)
        log(f"warmup {w}: {time.perf_counter() - t1:.3f}s")
    torch.cuda.synchronize()
    times = []
    t_start = time.perf_counter()
    emb = pro = None
    for s in range(args.steps):
        emb = pro = None  # the previous step's outputs go back to the allocator first
        t1 = time.perf_counter()
        emb, S, pro, st, pst, _ = run_pipeline(L, g, cfg, ratio, L.FP32, ctx)
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
    if args.gather_calibration:
        gen = torch.Generator(device="cuda:0").manual_seed(9)
        src = torch.empty((1 << 26, 16), dtype=torch.float32, device="cuda:0").fill_(1.0)
        idx = torch.randint(0, 1 << 26, (1 << 25,), generator=gen, device="cuda:0")
        for _ in range(3):
            out = src.index_select(0, idx)
        torch.cuda.synchronize()
        del src, idx, out
    ms_per_step = 1000.0 * wall / args.steps
    sp, rs = st["sparsify"], st["rsvd"]
    n, k, d = hg.n, cfg["k"], cfg["d"]
    nnz = sp["nnz"]
    # SpMM gather model: row ptrs + (col, val) + gathered rows + written rows.
    # The sketch product Y = A Omega gathers the BF16-stored sketch (2 bytes
    # per element); the later products B = A Q gather FP32 panels.
    spmm_bytes = 8 * (n + 1) + nnz * 8 + nnz * k * 4 + n * k * 4
    sketch_bytes = 8 * (n + 1) + nnz * 8 + nnz * k * 2 + n * k * 4
    fp32_calls = max(rs["spmm_calls"] - 1, 0)
    t_fp32 = rs["t_spmm"] - rs["t_spmm_sketch"]
    t_spmm_call = t_fp32 / fp32_calls if fp32_calls else float("nan")
    spmm_gbs = spmm_bytes / t_spmm_call / 1e9 if fp32_calls else None
    sketch_gbs = sketch_bytes / rs["t_spmm_sketch"] / 1e9 if rs["t_spmm_sketch"] > 0 else None
    sampled_eps = sp["samples"] / sp["t_sample"] if sp["t_sample"] > 0 else None
    # sampler byte model: 20 B / walk step + one slot read and written per
    # sample (8-byte or 16-byte slots) + 12 B / directed edge
    slot_b = 8 if sp["table_layout"] == 2 else 16
    sampler_bytes = 20 * sp["walk_steps"] + 2 * slot_b * sp["samples"] + 12 * hg.m
    sampler_gbs = sampler_bytes / sp["t_sample"] / 1e9
    stages = {"sample": sp["t_sample"], "table": sp["t_table"], "canon": sp["t_canon"], "csr": sp["t_csr"],
              "sparsify_wall": st["t_sparsify"],
              "rsvd_spmm": rs["t_spmm"], "rsvd_qr": rs["t_qr"], "rsvd_gemm": rs["t_gemm"],
              "rsvd_small_svd": rs["t_small_svd"], "rsvd_total": rs["t_total"], "netsmf_total": st["t_total"],
              "prop_cheb": pst["t_cheb"], "prop_aggregate": pst["t_aggregate"], "prop_svd": pst["t_svd"],
              "prop_total": pst["t_total"]}
    # dominant kernel by measured device time (HIP events on the context
    # stream): the stage-2 FP32 SpMM B = A Q, k_spmm<float,1,false> (+
    # k_spmm_fixup for rows split over several chunks), the sketch product
    # k_spmm<bf16,1,false>, or the sampler k_sample
    cands = {"sampler": (sp["t_sample"], sampler_bytes, 1),
             "spmm_sketch": (rs["t_spmm_sketch"], sketch_bytes, 1)}
    if fp32_calls:
        cands["spmm_stage2"] = (t_fp32, spmm_bytes * fp32_calls, fp32_calls)
    dom = max(cands, key=lambda c: cands[c][0])
    t_dom, bytes_dom, launches = cands[dom]
    achieved = bytes_dom / t_dom / 1e9
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", f"{args.config}_{dom}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
        traffic_src = (f"profiles/{os.path.basename(tpath)}: rocprofv3 FETCH_SIZE + WRITE_SIZE of the same kernel "
                       "and workload (separate --pmc runs), not measured in this run")
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "bytes_per_launch": int(bytes_dom / launches), "t_per_launch_ms": round(1e3 * t_dom / launches, 3)}
    emb = pro = None
    torch.cuda.empty_cache()

    # -- optional: one step with both file saves (the reference's Running Time)
    save = None
    if args.with_save:
        tmpd = tempfile.mkdtemp(prefix="lightne_save_")
        need = 2 * n * d * 4
        if shutil.disk_usage(tmpd).free > need + (8 << 30):
            _, _, _, _, _, ts = run_pipeline(L, g, cfg, ratio, L.FP32, ctx, save_dir=tmpd)
            save = {"seconds_with_save": round(ts, 3), "bytes_written": need, "dir_fs": tmpd}
            log(f"step with save: {ts:.3f}s")
        else:
            save = {"skipped": f"less than {need / 1e9:.1f} GB free under {tmpd}"}
        shutil.rmtree(tmpd, ignore_errors=True)

    # -- accuracy: the FP32 outputs of one more step, kept for the comparisons
    keep = {}

    def keep32(e, S_):
        keep["emb"] = to_host(e) if big else e.clone()
        keep["S"] = S_.copy()

    bf16 = None
    if not args.skip_bf16 or not args.skip_fp64:
        r32 = run_pipeline(L, g, cfg, ratio, L.FP32, ctx, after_stage2=keep32)
        keep["pro"] = to_host(r32[2]) if big else r32[2]
        del r32  # with the in-place propagation emb and pro are one tensor
        torch.cuda.empty_cache()
    if not args.skip_bf16:
        bf16 = {}
        chk = {}

        def cmp16(e, S_):
            chk["stage2_bit_identical_to_fp32"] = bool(np.array_equal(S_, keep["S"])) and same_bits(e, keep["emb"])

        run_pipeline(L, g, cfg, ratio, L.BF16, ctx)  # warm
        r16 = run_pipeline(L, g, cfg, ratio, L.BF16, ctx, after_stage2=cmp16)
        pro16, stb, pstb, tb = r16[2], r16[3], r16[4], r16[5]
        del r16
        log(f"bf16 step: {tb:.3f}s")
        bf16.update({"seconds": round(tb, 4), "speedup_vs_fp32": round((ms_per_step / 1e3) / tb, 3),
                     "stage2_bit_identical_to_fp32": chk["stage2_bit_identical_to_fp32"],
                     "final_angle_vs_fp32_rad": principal_angle(pro16, keep["pro"]),
                     "spmm_ms_per_call": round(1e3 * stb["rsvd"]["t_spmm"] / max(stb["rsvd"]["spmm_calls"], 1), 3),
                     "prop_cheb_s": pstb["t_cheb"]})
        del pro16
        if cfg["q"] > 1:
            run_pipeline(L, g, cfg, ratio, L.BF16_POWER, ctx)
            rp_ = run_pipeline(L, g, cfg, ratio, L.BF16_POWER, ctx)
            prp, tp = rp_[2], rp_[5]
            del rp_
            bf16["bf16_power"] = {"seconds": round(tp, 4), "speedup_vs_fp32": round((ms_per_step / 1e3) / tp, 3),
                                  "final_angle_vs_fp32_rad": principal_angle(prp, keep["pro"])}
            del prp
        torch.cuda.empty_cache()
    fp64 = None
    if not args.skip_fp64:
        # stage 2 against the FP64 form of the randomized SVD (same sparsifier,
        # same sketch): the explicit four-panel form where it fits, else the
        # lean two-panel form writing its embedding to host memory
        prec = L.FP64_LEAN if big else L.FP64
        if big:
            ctx.release()  # the FP32 pipeline's workspace
            torch.cuda.empty_cache()
        log("fp64 reference ...")
        try:
            out64 = np.empty((n, d), np.float32) if big else None
            emb64, S64, st64 = L.NetSMF(g, walk_len=cfg["T"], rank=k, dim=d, sample_ratio=ratio, seed=1,
                                        svd_seed=777, power_iters=cfg["q"], precision=prec, out=out64, ctx=ctx)
            torch.cuda.synchronize()
            S64 = S64.astype(np.float64)
            if big:
                ctx.release()
            ang = principal_angle(keep["emb"], emb64)
            sv = float(np.max(np.abs(keep["S"][:d] - S64[:d]) / S64[:d]))
            fp64 = {"form": "lean FP64 (two panels)" if big else "explicit FP64 (the reference's Z/C/U schedule)",
                    "fp32_stage2_angle_vs_fp64_rad": ang, "fp32_sv_rel_err_vs_fp64": sv,
                    "spectral_gap_d": float(S64[d - 1] / S64[d]), "fp64_rsvd_s": round(st64["rsvd"]["t_total"], 3)}
            if bf16 is not None and bf16.get("stage2_bit_identical_to_fp32"):
                bf16["stage2_angle_vs_fp64_rad"] = ang
                bf16["sv_rel_err_vs_fp64"] = sv
            del emb64
        except L.LightNEError as e:
            fp64 = {"error": str(e)}
            log(f"fp64 reference failed: {e}")
        torch.cuda.empty_cache()
    keep.clear()
    adjoint = None
    if not args.skip_adjoint:
        # differentiation form: forward z = (A + I)(a - sum c_i T_i a) and the
        # hand-written adjoint (dL/da, dL/dmu, dL/dtheta, checkpoint every 3
        # steps) for a ~ N(0, 1), dL/dz ~ N(0, 1), on the c2 graph at d = 128
        ha = make_graph(L, ADJOINT) if args.config != "c2" else hg
        ga = ha.to_device("cuda:0") if args.config != "c2" else g
        gen = torch.Generator(device="cuda:0").manual_seed(4)
        a = torch.randn(ha.n, ADJOINT["d"], generator=gen, device="cuda:0")
        gz = torch.randn(ha.n, ADJOINT["d"], generator=gen, device="cuda:0")
        tf, tbk = [], []
        for it in range(3):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            L.propagate(a, ga, order=ADJOINT["K"], mu=0.2, theta=0.5, ctx=ctx)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            L.propagate_backward(a, gz, ga, order=ADJOINT["K"], mu=0.2, theta=0.5, aggregate=True,
                                 checkpoint_every=ADJOINT["checkpoint"], ctx=ctx)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            if it:  # the first pair warms up
                tf.append(t2 - t1)
                tbk.append(t3 - t2)
        adjoint = {"workload": f"Chung-Lu n={ha.n} m_undirected={ha.m // 2} d={ADJOINT['d']} K={ADJOINT['K']}",
                   "forward_s": round(float(np.median(tf)), 5), "backward_s": round(float(np.median(tbk)), 5),
                   "checkpoint_every": ADJOINT["checkpoint"]}
        log(f"adjoint: forward {adjoint['forward_s']}s backward {adjoint['backward_s']}s")
        del a, gz, ga
        torch.cuda.empty_cache()
    cpu = None
    if not args.skip_cpu:
        log("cpu baseline ...")
        parts = None
        if cfg["cpu_frac"] < 1.0:
            # the rows of the sparsifier the sampled CPU products use
            ctx.release()
            torch.cuda.empty_cache()
            csr, _ = L.generate_trunc_log_matrix(g, 1, walk_len=cfg["T"], sample_ratio=ratio, ctx=ctx)
            parts = csr.to_numpy(nnz_fraction=cfg["cpu_frac"])
            csr.free()
            log("cpu: sparsifier rows copied to the host")
        cpu = cpu_baseline(hg, cfg, ratio, parts)
    out = {
        "metric": METRIC, "value": round(ms_per_step / 1e3, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic (seeded {'R-MAT' if 'rmat' in cfg else 'Chung-Lu'} graph)",
        "config": {"workload": f"{args.config}: {'R-MAT scale %d ef %g' % cfg['rmat'][:2] if 'rmat' in cfg else 'Chung-Lu'}"
                               f" n={hg.n} m_undirected={hg.m // 2} T={cfg['T']} M={sp['samples']} k={k} d={d} "
                               f"q={cfg['q']} K={cfg['K']} mu=0.2 theta=0.5 FP32",
                   "parallelism": "1 GPU"},
        "sampled_edges_per_s": sampled_eps, "sampler_gbs": round(sampler_gbs, 1),
        "spmm_gbs": round(spmm_gbs, 1) if spmm_gbs else None,
        "spmm_sketch": {"gbs": round(sketch_gbs, 1) if sketch_gbs else None,
                        "ms": round(1e3 * rs["t_spmm_sketch"], 3), "bytes": int(sketch_bytes),
                        "frac": round(sketch_gbs / HBM_PEAK_GBS, 4) if sketch_gbs else None},
        "sparsifier_nnz": nnz, "samples": sp["samples"], "distinct_keys": sp["distinct_keys"],
        "table_capacity": sp["table_capacity"], "table_slot_bytes": slot_b, "csr_row_ranges": sp["csr_parts"], "csr_row_order": int(sp["row_order"]),
        "csr_value_cache": int(sp["value_cache"]), "sampler_runs": sp["sampler_runs"], "stages_s": {k_: round(v, 5) for k_, v in stages.items()},
        "bf16": bf16, "fp64_reference": fp64, "adjoint": adjoint, "save": save, "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
