"""The CPU oracle pinned against the reference's own code and against what the
mathematics fixes (CPU only)."""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

MASK = (1 << 64) - 1


def py_hash64_2(x):
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK
    return x ^ (x >> 31)


def test_hash64_2_known_answers():
    # 0 is a fixed point of the splitmix finaliser; others against big-int arithmetic
    assert O.hash64_2(0) == 0
    for x in [1, 2, 0xDEADBEEF, 1 << 63, MASK, 0x123456789ABCDEF0]:
        assert O.hash64_2(x) == py_hash64_2(x)


@pytest.mark.parametrize("rec", H.manifest(), ids=lambda r: r["case"])
def test_sampler_matches_reference_harness(rec):
    """Key -> value tables of the reference-header harness (pbbs::random, hash64_2,
    par_table, ligra reader) reproduced bit for bit, both -sample modes; for the
    compressed-format cases the harness read the -c file, so its walks used the
    reference's compressed-graph neighbour lookup."""
    off, edg = H.golden_graph(rec["graph"])
    gd = H.golden_case(rec["case"])
    keys, cnt, wgh, st = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"],
                                  rec["sample"], threads=1, ith=int(rec.get("compressed", False)))
    sc = gd["scalars"]
    assert st["logn"] == sc[0] and st["M"] == sc[1] and st["one_over_lambda"] == sc[2]
    assert st["fractional_bits"] == sc[3] and st["exp_rv_rounded"] == int(gd["scalars_u"][0])
    assert st["overflow"] == 0
    assert len(keys) == rec["nkeys"]
    if rec["full"]:
        np.testing.assert_array_equal(keys, gd["keys"])
        np.testing.assert_array_equal(wgh, gd["vals"])
    else:
        h = hashlib.sha256()
        h.update(keys.tobytes())
        h.update(wgh.tobytes())
        assert h.hexdigest() == rec["sha256_keys_vals"]
        np.testing.assert_array_equal(keys[gd["sub_idx"]], gd["sub_keys"])
    # invariants: counts sum to the emitted samples; -sample 0 weights are the counts
    assert int(cnt.sum()) == st["samples"]
    if rec["sample"] == 0:
        np.testing.assert_array_equal(wgh, cnt.astype(np.float64))
        assert st["samples"] == st["attempts"]


@pytest.mark.parametrize("rec", [r for r in H.manifest()], ids=lambda r: r["case"])
def test_trunc_log_matches_reference_harness(rec):
    off, edg = H.golden_graph(rec["graph"])
    gd = H.golden_case(rec["case"])
    if rec["full"]:
        keys, vals, tl = gd["keys"], gd["vals"], gd["trunclog"]
    else:
        keys, _, vals, _ = O.sample(off, edg, int(rec["seed"]), rec["walk_len"], rec["sample_ratio"],
                                    rec["sample"], threads=1, ith=int(rec.get("compressed", False)))
        tl = None
    rp, col, val, info = O.trunc_log_csr(off, keys, vals, M_exact=float(gd["scalars"][4]))
    assert info["factor"] == gd["scalars"][5]
    assert info["kept_upper"] == rec["kept"]
    # CSR invariants: symmetric, rows sorted, nnz = 2 kept - diagonal
    n = len(off) - 1
    assert rp[0] == 0 and rp[-1] == len(col)
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(np.diff(rows.astype(np.int64) * n + col) > 0)
    diag = int(np.sum(rows == col))
    assert len(col) == 2 * info["kept_upper"] - diag
    import scipy.sparse as sp

    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    assert (A != A.T).nnz == 0
    if tl is not None:
        keep = np.isfinite(tl)
        k = keys[keep]
        u = (k >> np.uint64(32)).astype(np.int64)
        v = (k & np.uint64(0xFFFFFFFF)).astype(np.int64)
        np.testing.assert_array_equal(np.asarray(A[u, v]).ravel(), tl[keep].astype(np.float32))


def test_walk_len_one_closed_form():
    """T = 1: every sample is the source edge itself, so count(u<v) is the sum of
    the two directed edges' exp_rv draws, in {2E, 2E+1, 2E+2}."""
    off, edg = H.golden_graph("cl200")
    n = len(off) - 1
    m = int(off[-1])
    ratio = H.c1_sample_ratio(n, m, T=1, per_edge=3.4)
    keys, cnt, _, st = O.sample(off, edg, 5, 1, ratio, 0, threads=1)
    E = st["exp_rv_rounded"]
    rows = np.repeat(np.arange(n), np.diff(off))
    und = np.unique(np.minimum(rows, edg).astype(np.uint64) << np.uint64(32) | np.maximum(rows, edg).astype(np.uint64))
    np.testing.assert_array_equal(keys, und)
    assert np.all((cnt >= 2 * E) & (cnt <= 2 * E + 2))
    assert st["walk_steps"] == 0


def test_expected_counts_match_netmf_matrix():
    """E[count(x,z)] = (M/2m)(1/T) sum_r [(A (D^-1 A)^(r-1))_xz + sym], i.e. the
    sampled matrix concentrates on the DeepWalk/NetMF matrix (toy graph, many seeds)."""
    off, edg = H.golden_graph("toy")
    n = len(off) - 1
    m = int(off[-1])
    T = 3
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(off))
    A[rows, edg] = 1.0
    P = A / A.sum(1, keepdims=True)
    ratio = float(np.float32(400.0 * m / (T * n * np.log(n))))
    tot = np.zeros((n, n))
    nrep = 12
    for s in range(nrep):
        keys, cnt, _, st = O.sample(off, edg, 1000 + s, T, ratio, 0, threads=1)
        u = (keys >> np.uint64(32)).astype(int)
        v = (keys & np.uint64(0xFFFFFFFF)).astype(int)
        tot[u, v] += cnt
    per_edge = st["one_over_lambda"]
    # directed edge (u,v) of length r: x ~ (P^(k-1))[u], z ~ (P^(r-k))[v], k uniform
    exp = np.zeros((n, n))
    for r in range(1, T + 1):
        for k in range(1, r + 1):
            exp += (np.linalg.matrix_power(P.T, k - 1) @ A @ np.linalg.matrix_power(P, r - k)) / (r * T)
    exp = per_edge * nrep * (exp + exp.T - np.diag(np.diag(exp)))
    iu = np.triu_indices(n)
    e, o = exp[iu], tot[iu]
    assert np.all(o[e == 0] == 0)  # unreachable pairs are never sampled
    big = e > 20
    z = (o[big] - e[big]) / np.sqrt(e[big])  # counts are (close to) Poisson
    assert np.abs(z).max() < 5.0, np.abs(z).max()
    assert 0.3 < np.mean(z ** 2) < 2.0, np.mean(z ** 2)
    assert np.abs(o.sum() - e.sum()) / e.sum() < 0.01


def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors (Random123 kat_vectors)."""
    out = O.philox4x32([0, 0, 0, 0], [0, 0])
    np.testing.assert_array_equal(out, np.array([0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8], np.uint32))
    out = O.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)
    np.testing.assert_array_equal(out, np.array([0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD], np.uint32))
    out = O.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])
    np.testing.assert_array_equal(out, np.array([0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1], np.uint32))


def test_gaussian_moments():
    x = O.gaussian(777, 0, 2_000_000).astype(np.float64)
    assert abs(x.mean()) < 3e-3 and abs(x.var() - 1) < 5e-3
    assert abs(np.mean(x ** 4) - 3.0) < 3e-2
    y = O.gaussian(777, 1, 1000)
    assert not np.array_equal(y, x[:1000].astype(np.float32))


def test_bessel_series_vs_scipy():
    from scipy.special import iv

    for k in range(12):
        for x in [0.0, 0.1, 0.5, 1.0, 2.5]:
            assert abs(O.bessel_iv(k, x) - iv(k, x)) <= 4e-15 * max(1.0, iv(k, x))


def test_spmm_oracle_vs_scipy():
    import scipy.sparse as sp

    rng = np.random.default_rng(0)
    A = sp.random(300, 300, density=0.05, random_state=1, format="csr", dtype=np.float32)
    X = rng.standard_normal((300, 32))
    Y = O.spmm_f64(A.indptr, A.indices, A.data, X)
    np.testing.assert_allclose(Y, A.astype(np.float64) @ X, rtol=1e-12, atol=1e-12)


def test_propagation_oracle_is_truncated_exponential():
    """y = a - sum_{k<K} c_k T_k(X) a with X = M^2/2 - I and c_k from I_k(theta):
    exp(-theta X) = sum c_k T_k(X), so for large K, y -> a - expm(-theta X) a."""
    from scipy.linalg import expm

    off, edg = H.golden_graph("toy")
    n = len(off) - 1
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, 4))
    M = O.prop_matrix(off, edg, float(np.float32(0.2))).toarray()
    X = 0.5 * M @ M - np.eye(n)
    ev = np.linalg.eigvals(X).real
    assert ev.min() >= -1 - 1e-12 and ev.max() < 1
    theta = float(np.float32(0.5))
    y, z = O.propagate(off, edg, a, order=30, mu=0.2, theta=0.5)
    np.testing.assert_allclose(y, a - expm(-theta * X) @ a, atol=1e-12)
    y10, _ = O.propagate(off, edg, a, order=10, mu=0.2, theta=0.5)
    assert np.abs(y10 - y).max() < 1e-8  # K = 10 truncation error at theta = 0.5
    A = np.zeros((n, n))
    A[np.repeat(np.arange(n), np.diff(off)), edg] = 1
    np.testing.assert_allclose(z, (A + np.eye(n)) @ y, atol=1e-12)


def test_rsvd_oracle_exact_rank():
    """A symmetric matrix of rank <= k is recovered to rounding (singular values)."""
    rng = np.random.default_rng(7)
    n, r = 400, 20
    Q = np.linalg.qr(rng.standard_normal((n, r)))[0]
    lam = np.linspace(5, 1, r) * rng.choice([-1, 1], r)
    A = (Q * lam) @ Q.T
    import scipy.sparse as sp

    S_ = sp.csr_matrix(A.astype(np.float32))
    U, S = O.rsvd(S_.indptr, S_.indices, S_.data, 32, seed=777)
    np.testing.assert_allclose(S[:r], np.sort(np.abs(lam))[::-1], rtol=1e-5)
    assert S[r:].max() < 1e-5
    assert H.principal_angles(U[:, :r], Q) < 1e-5


def _sym_graph(n, pairs):
    adj = [[] for _ in range(n)]
    for u, v in pairs:
        adj[u].append(v)
        adj[v].append(u)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(a) for a in adj])
    return off, np.array([v for a in adj for v in a], np.uint32)


def test_ne_zhang_matrix_regular_graph_closed_form():
    """d-regular graph: neg_j = (d * 1/d)^0.75 = 1, sum = n, every entry
    log(1/d) - log(1/n) = log(n/d) (PathEmbed.h:418-441)."""
    n = 12
    off, edg = _sym_graph(n, [(i, (i + 1) % n) for i in range(n)] + [(i, (i + 6) % n) for i in range(6)])
    assert np.all(np.diff(off) == 3)
    rp, col, val = O.ne_zhang_matrix(off, edg)
    assert np.array_equal(rp, off) and np.array_equal(col, edg.astype(np.int32))
    np.testing.assert_allclose(val, np.float32(np.log(n / 3.0)), rtol=1e-7)


def test_ne_zhang_matrix_star_closed_form():
    """Star K_{1,m}: neg_0 = m^0.75, neg_leaf = m^-0.75, S = m^0.75 + m^0.25;
    centre->leaf = log S - 0.25 log m, leaf->centre = log S - 0.75 log m (a
    swapped d_u / d_v or a missing exponent changes both)."""
    m = 7
    off, edg = _sym_graph(m + 1, [(0, i) for i in range(1, m + 1)])
    rp, col, val = O.ne_zhang_matrix(off, edg)
    S = m ** 0.75 + m ** 0.25
    np.testing.assert_allclose(val[:m], np.log(S) - 0.25 * np.log(m), rtol=1e-6)
    np.testing.assert_allclose(val[m:], np.log(S) - 0.75 * np.log(m), rtol=1e-6)


def test_sparse_projection_distribution():
    """Very sparse random projection (mklredsvd.cc:148-189): entries +-sqrt(s)
    with probability 1/s (s = sqrt(n) by default, -sparse_project_s when >= 1),
    signs balanced, E[Omega_ic^2] = 1 (the Johnson-Lindenstrauss scaling)."""
    for n, k, s_param in [(20000, 144, 0.0), (4000, 64, 16.0)]:
        M = O.sparse_projection(n, k, 777, s_param)
        s = float(np.float32(s_param if s_param >= 1 else np.sqrt(n)))
        p = 1.0 / s
        frac = (M != 0).mean()
        sd = np.sqrt(p * (1 - p) / (n * k))
        assert abs(frac - p) < 5 * sd, (frac, p)
        nz = M[M != 0]
        np.testing.assert_allclose(np.abs(nz), np.sqrt(np.float32(s)), rtol=1e-6)
        assert abs((nz > 0).mean() - 0.5) < 5 * 0.5 / np.sqrt(nz.size)
        assert abs((M ** 2).mean() - 1.0) < 5 * np.sqrt((s - 1) / (n * k))
    # different seeds draw different matrices; the same seed the same one
    assert not np.array_equal(O.sparse_projection(500, 32, 1, 4.0), O.sparse_projection(500, 32, 2, 4.0))
    assert np.array_equal(O.sparse_projection(500, 32, 1, 4.0), O.sparse_projection(500, 32, 1, 4.0))


# ---------------------------------------------------------------------------
# propagation operator M = (1 - mu) I - D~^-1 A~ (SpectralPropagation.h:51-75)
# pinned by closed forms that a D-for-D~, sign or diagonal mistake fails
# ---------------------------------------------------------------------------
def _random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    A = np.triu(rng.random((n, n)) < p, 1)
    pairs = list(zip(*np.nonzero(A)))
    off, edg = _sym_graph(n, pairs)
    # shuffle each adjacency list: the operator must not depend on neighbour order
    for i in range(n):
        seg = edg[off[i]:off[i + 1]]
        rng.shuffle(seg)
    return off, edg


@pytest.mark.parametrize("graph", ["cl200", "blogcatalog"])
@pytest.mark.parametrize("mu", [0.2, 0.0, 0.7])
def test_prop_matrix_rows_sum_to_minus_mu(graph, mu):
    """M 1 = (1 - mu) 1 - D~^-1 (A + I) 1 = (1 - mu) - (d + 1)/(d + 1) = -mu for
    every row (D in place of D~, or a dropped self loop, breaks it)."""
    off, edg = H.golden_graph(graph)
    n = len(off) - 1
    M = O.prop_matrix(off, edg, mu)
    np.testing.assert_allclose(M @ np.ones(n), -mu * np.ones(n), atol=1e-12)
    # the row-stochastic part: every off-diagonal entry of row i is -1/(d_i + 1)
    deg = np.diff(off)
    Mc = M.tocsr()
    i = int(np.argmax(deg))
    row = Mc.getrow(i).toarray().ravel()
    assert row[i] == 1.0 - mu - 1.0 / (deg[i] + 1)
    nb = np.asarray(edg[off[i]:off[i + 1]], np.int64)
    np.testing.assert_array_equal(row[nb], -1.0 / (deg[i] + 1))
    assert np.count_nonzero(row) == deg[i] + 1


@pytest.mark.parametrize("n", [7, 12, 31])
def test_prop_matrix_cycle_spectrum(n):
    """Cycle C_n: D~ = 3 I, eig(A) = 2 cos(2 pi j / n), so
    eig(M) = 1 - mu - (1 + 2 cos(2 pi j / n)) / 3 exactly."""
    mu = 0.2
    off, edg = _sym_graph(n, [(i, (i + 1) % n) for i in range(n)])
    M = O.prop_matrix(off, edg, mu).toarray()
    got = np.sort(np.linalg.eigvals(M).real)
    j = np.arange(n)
    want = np.sort(1 - mu - (1 + 2 * np.cos(2 * np.pi * j / n)) / 3)
    np.testing.assert_allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("m", [3, 8, 40])
def test_prop_matrix_star_spectrum(m):
    """Star K_{1,m}: D~^-1 A~ has eigenvalues 1, 1/(m+1) - 1/2 and 1/2 (x m-1),
    so eig(M) = {-mu, 3/2 - mu - 1/(m+1), 1/2 - mu (m-1 times)}."""
    mu = 0.2
    off, edg = _sym_graph(m + 1, [(0, i) for i in range(1, m + 1)])
    M = O.prop_matrix(off, edg, mu).toarray()
    got = np.sort(np.linalg.eigvals(M).real)
    want = np.sort(np.array([-mu, 1.5 - mu - 1.0 / (m + 1)] + [0.5 - mu] * (m - 1)))
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_prop_matrix_spectrum_range():
    """M is similar to the symmetric (1 - mu) I - D~^-1/2 A~ D~^-1/2, whose
    normalised part has spectrum in (-1, 1]: eig(M) in [-mu, 2 - mu), with -mu
    attained (the constant vector)."""
    off, edg = H.golden_graph("toy")
    mu = 0.2
    M = O.prop_matrix(off, edg, mu).toarray()
    deg = np.diff(off) + 1.0
    S = np.diag(np.sqrt(deg)) @ M @ np.diag(1 / np.sqrt(deg))
    np.testing.assert_allclose(S, S.T, atol=1e-14)
    ev = np.linalg.eigvalsh(0.5 * (S + S.T))
    assert abs(ev.min() + mu) < 1e-12 and ev.max() < 2 - mu


@pytest.mark.parametrize("k", [1, 16, 300])
def test_prop_apply_vs_explicit_matrix(k):
    """The implicit C operator (FP64 and FP32 forms, alpha/beta) equals the
    pinned explicit M applied by scipy."""
    off, edg = _random_graph(120, 0.08, 5)
    rng = np.random.default_rng(k)
    X = rng.standard_normal((120, k))
    Y0 = rng.standard_normal((120, k))
    M = O.prop_matrix(off, edg, float(np.float32(0.2)))
    for alpha, beta in [(1.0, 0.0), (0.5, -2.0), (-1.0, 1.0)]:
        got = O.prop_apply(off, edg, X, float(np.float32(0.2)), alpha, beta, Y0, threads=2)
        np.testing.assert_allclose(got, alpha * (M @ X) + beta * Y0, rtol=1e-13, atol=1e-13)
        g32 = O.prop_apply(off, edg, X, 0.2, alpha, beta, Y0, fp32=True, threads=2)
        np.testing.assert_allclose(g32, alpha * (M @ X) + beta * Y0, rtol=2e-5, atol=2e-5)
    # FP32 coefficients are the reference's double expressions rounded to float
    E = np.eye(120)[:, :1]
    col = O.prop_apply(off, edg, E, 0.2, fp32=True, threads=1).ravel()
    d0 = float(off[1] - off[0]) + 1.0
    assert col[0] == np.float32(1.0 - float(np.float32(0.2)) - 1.0 / d0)
    for j in edg[off[0]:off[1]]:
        dj = float(off[j + 1] - off[j]) + 1.0
        assert col[j] == np.float32(-1.0 / dj)


def test_aggregate_vs_scipy():
    import scipy.sparse as sp

    off, edg = _random_graph(150, 0.05, 9)
    n = 150
    rows = np.repeat(np.arange(n), np.diff(off))
    A = sp.csr_matrix((np.ones(len(edg)), (rows, edg.astype(np.int64))), shape=(n, n))
    Y = np.random.default_rng(1).standard_normal((n, 20))
    np.testing.assert_allclose(O.aggregate(off, edg, Y, threads=2), (A + sp.eye(n)) @ Y, rtol=1e-13, atol=1e-13)


def test_spmm_f32_oracle_vs_scipy():
    """FP32 oracle SpMM (the CPU baseline kernel) against scipy in FP64, and the
    FP64 one at k above the oracle's column block."""
    import scipy.sparse as sp

    rng = np.random.default_rng(2)
    A = sp.random(400, 400, density=0.03, random_state=3, format="csr", dtype=np.float32)
    for k in [8, 144, 600]:
        X = rng.standard_normal((400, k)).astype(np.float32)
        ref = A.astype(np.float64) @ X.astype(np.float64)
        np.testing.assert_allclose(O.spmm_f32(A.indptr, A.indices, A.data, X, threads=2), ref, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(O.spmm_f64(A.indptr, A.indices, A.data, X.astype(np.float64), threads=2), ref,
                                   rtol=1e-12, atol=1e-12)
