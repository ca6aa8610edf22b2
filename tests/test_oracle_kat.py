"""The FP64 oracle against the reference's own known answers (tests.cpp) and
the published Philox4x32-10 vectors.  CPU only."""
import numpy as np
import pytest

import oracle as O


def test_philox_kat(kat):
    for v in kat["philox4x32_10"]:
        assert list(O.philox(v["ctr"], v["key"])) == v["out"]


def _feat(kat, binary):
    f = kat["feat"]
    vals = None if binary else np.array(f["vals"])
    return O.csr_pair(f["rows"], f["cols"], vals, f["nrow"], f["ncol"]) if not binary else \
        O.csr_pair(f["rows"], f["cols"], np.ones(len(f["rows"])), f["nrow"], f["ncol"])


def test_A_mul_Bx(kat):
    F, _ = _feat(kat, False)
    B = np.array(kat["A_mul_Bx"]["B"])  # 2 x 4 (nrhs x ncol)
    X = O.csr_A_mul_Bn(F, B.T)           # (nrow, nrhs)
    np.testing.assert_allclose(X.T, np.array(kat["A_mul_Bx"]["X"]), atol=1e-12)


def test_AtA_mul_Bx(kat):
    f = kat["feat"]
    F, Ft = _feat(kat, False)
    B = np.array(kat["AtA_mul_Bx"]["B"])
    reg = kat["AtA_mul_Bx"]["reg"]
    Xd = np.zeros((6, 4))
    for r, c, v in zip(f["rows"], f["cols"], f["vals"]):
        Xd[r, c] = v
    expected = (Xd.T @ Xd @ B.T + reg * B.T).T
    out = O.AtA_mul_B(F, Ft, reg, B.T)
    np.testing.assert_allclose(out.T, expected, atol=1e-12)


def test_compute_uhat_binary(kat):
    F, _ = _feat(kat, True)
    beta = np.array(kat["compute_uhat_bin"]["beta"])   # 3 x 4
    uhat = O.csr_A_mul_Bn(F, beta.T)                   # uhat[:, n] = sum_f F[n,f] beta[:, f]
    np.testing.assert_allclose(uhat.T, np.array(kat["compute_uhat_bin"]["uhat"]), atol=1e-12)


def test_blockcg_one_rhs(kat):
    F, Ft = _feat(kat, True)
    c = kat["blockcg_1"]
    X, it = O.blockcg(F, Ft, c["reg"], np.array(c["B"]).T, c["tol"])
    np.testing.assert_allclose(X.T, np.array(c["X"]), rtol=1e-6)
    assert it <= c["max_iter"]


def test_blockcg_three_rhs_separately(kat):
    F, Ft = _feat(kat, True)
    c = kat["blockcg_3"]
    X, _ = O.blockcg(F, Ft, c["reg"], np.array(c["B"]).T, c["tol"], blocksize=1, excess=0)
    np.testing.assert_allclose(X.T, np.array(c["X"]), rtol=1e-6)


def test_blockcg_three_rhs_jointly(kat):
    # the dense-F reference test (tests.cpp:183-207) solves the 3 rhs in one block
    F, Ft = _feat(kat, True)
    c = kat["blockcg_3"]
    X, _ = O.blockcg(F, Ft, c["reg"], np.array(c["B"]).T, c["tol"])
    np.testing.assert_allclose(X.T, np.array(c["X"]), rtol=1e-6)


def test_posterior_lambda_beta(kat):
    c = kat["posterior_lambda_beta"]
    beta = np.array(c["beta"])  # K x nfeat
    shape, scale = O.posterior_lambda_beta(beta.T, np.array(c["Lambda"]), c["nu"], c["mu"])
    assert shape == pytest.approx(c["shape"], rel=1e-9)
    assert scale == pytest.approx(c["scale"], rel=1e-9)


def test_eval_rmse_welford(kat):
    c = kat["eval_rmse"]
    pred = np.zeros(1)
    var = np.zeros(1)
    for n, st in enumerate(c["steps"]):
        U = np.array([st["U"]])
        V = np.array([st["V"]])
        r1, ra = O.eval_rmse(2, np.array([[0, 0]]), np.array([c["y"]]), c["mean"], [U, V], n, pred, var)
        assert pred[0] == pytest.approx(st["pred"])
        assert var[0] == pytest.approx(st["var"])
        assert r1 == pytest.approx(st["rmse1"])
        assert ra == pytest.approx(st["rmse"])


def test_sparse_mode_layout(kat):
    c = kat["sparse_mode"]
    for m in (0, 1):
        e = c["mode%d" % m]
        rows = O.sparse_mode(np.array(c["C"]), np.array(c["v"]), m, e["size"])
        assert list(rows.row_ptr) == e["row_ptr"]
        assert rows.idx.tolist() == e["indices"]
        assert list(rows.vals) == e["values"]


def test_tensor_test_order(kat):
    c = kat["tensor_test_order"]
    idx, vals = O.tensor_test(np.array(c["Cte"]), np.array(c["vte"]), c["dims"])
    got = np.column_stack([idx, vals])
    np.testing.assert_allclose(got, np.array(c["expected"]))


def test_sparse_from_ijv(kat):
    c = kat["sparse_from_ijv"]
    r, cc, v = O.sum_duplicates(c["rows"], c["cols"], c["vals"], 3, 3)
    dense = np.zeros((3, 3))
    dense[r, cc] = v
    np.testing.assert_allclose(dense, np.array(c["dense"]))
    assert len(v) == 3


def test_eval_rmse_tensor_bruteforce():
    # tests.cpp:667-705: predictions equal the brute-force Hadamard contraction
    rng = np.random.default_rng(3)
    C = np.array([[0, 1, 0], [0, 0, 0], [1, 3, 1], [1, 0, 1], [2, 3, 0]])
    v = np.array([0.1, 0.2, 0.3, 0.4, 0.5])
    dims, K, gmean = [4, 5, 2], 5, 0.9
    fs = [rng.standard_normal((d, K)) for d in dims]
    idx, vals = O.tensor_test(C, v, dims)
    pred, var = np.zeros(5), np.zeros(5)
    O.eval_rmse(K, idx, vals, gmean, fs, 0, pred, var)
    brute = gmean + np.sum(fs[0][idx[:, 0]] * fs[1][idx[:, 1]] * fs[2][idx[:, 2]], axis=1)
    np.testing.assert_allclose(pred, brute, rtol=1e-13)


def test_empty_test_set_rmse_is_nan():
    pred, var = np.zeros(0), np.zeros(0)
    r1, ra = O.eval_rmse(2, np.zeros((0, 2), dtype=np.int32), np.zeros(0), 0.0,
                         [np.ones((1, 2)), np.ones((1, 2))], 0, pred, var)
    assert np.isnan(r1) and np.isnan(ra)


def _check_lower(AA, entries):
    for key, v in entries.items():
        i, j = (int(x) for x in key.split(","))
        assert AA[i, j] == pytest.approx(v, abs=1e-4), key


def test_at_mul_a_binary(kat):  # tests.cpp:16-29
    F, Ft = _feat(kat, True)
    _check_lower(O.at_mul_a(F, Ft), kat["At_mul_A_bin"]["lower"])


def test_at_mul_a_csr(kat):  # tests.cpp:31-58
    F, Ft = _feat(kat, False)
    AA = O.at_mul_a(F, Ft)
    _check_lower(AA, kat["At_mul_A_csr"]["lower"])
    assert np.all(np.triu(AA, 1) == 0.0)


def test_chol_solve_t(kat):  # tests.cpp:209-233
    c = kat["chol_solve_t"]
    X = O.ftf_solve(np.array(c["m"], dtype=float), 0.0, np.array(c["rhs"]).T)
    np.testing.assert_allclose(X.T, np.array(c["xopt"]), rtol=1e-4, atol=1e-5)


def test_dense_prior_ftf(kat):  # tests.cpp:859-885
    F = np.array(kat["dense_prior_ftf"]["F"])
    r, c = np.nonzero(np.ones_like(F))
    A, At = O.csr_pair(r, c, F[r, c], F.shape[0], F.shape[1])
    np.testing.assert_allclose(O.at_mul_a(A, At), np.tril(F.T @ F), atol=1e-15)
