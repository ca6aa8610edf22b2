"""Writes the golden fixtures used by the parity tests.

reference_kat.json -- known answers transcribed from the reference's own
C++ unit tests (lib/macau-cpp/tests.cpp; line numbers beside each entry),
which inline their inputs and expected outputs, and the Philox4x32-10
known-answer vectors published with Random123 (kat_vectors).

Run:  python tests/golden/make_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

# 6 x 4 feature matrix shared by many reference tests (tests.cpp:17-18, 41-43)
FEAT_ROWS = [0, 3, 3, 2, 5, 4, 1, 2, 4]
FEAT_COLS = [1, 0, 2, 1, 3, 0, 1, 3, 2]
FEAT_VALS = [0.6, -0.76, 1.48, 1.19, 2.44, 1.95, -0.82, 0.06, 2.54]


def reference_kat():
    k = {}
    k["feat"] = {"rows": FEAT_ROWS, "cols": FEAT_COLS, "vals": FEAT_VALS, "nrow": 6, "ncol": 4}
    # tests.cpp:66-78  A_mul_Bx<2>(csr)
    k["A_mul_Bx"] = {
        "B": [[-1.38, 1.04, -0.28, -0.18], [0.03, 0.88, 1.32, -0.31]],
        "X": [[0.624, -0.8528, 1.2268, 0.6344, -3.4022, -0.4392],
              [0.528, -0.7216, 1.0286, 1.9308, 3.4113, -0.7564]],
    }
    # tests.cpp:80-100  AtA_mul_Bx<2>(csr), reg = 0.6, dense form of F given
    k["AtA_mul_Bx"] = {"B": [[-1.38, 1.04, -0.28, -0.18], [0.03, 0.88, 1.32, -0.31]], "reg": 0.6}
    # tests.cpp:102-121  compute_uhat on binary CSR
    k["compute_uhat_bin"] = {
        "beta": [[0.56, 0.55, 0.3, -1.78], [1.63, -0.71, 0.8, -0.28], [0.47, 0.37, -1.36, 0.86]],
        "uhat": [[0.55, 0.55, -1.23, 0.86, 0.86, -1.78],
                 [-0.71, -0.71, -0.99, 2.43, 2.43, -0.28],
                 [0.37, 0.37, 1.23, -0.89, -0.89, 0.86]],
    }
    # tests.cpp:123-138  block CG, 1 rhs, reg 0.5, binary F, niter <= 4
    k["blockcg_1"] = {"B": [[0.56, 0.55, 0.3, -1.78]], "reg": 0.5, "tol": 1e-6,
                      "X": [[0.35555556, 0.40709677, -0.16444444, -0.87483871]], "max_iter": 4}
    # tests.cpp:141-161  block CG, 3 rhs solved separately (blocksize 1, excess 0)
    k["blockcg_3"] = {
        "B": [[0.56, 0.55, 0.3, -1.78], [0.34, 0.05, -1.48, 1.11], [0.09, 0.51, -0.63, 1.59]],
        "reg": 0.5, "tol": 1e-6,
        "X": [[0.35555556, 0.40709677, -0.16444444, -0.87483871],
              [1.69333333, -0.12709677, -1.94666667, 0.49483871],
              [0.66, -0.04064516, -0.78, 0.65225806]],
    }
    # tests.cpp:16-29  At_mul_A on the binary CSR (lower triangle entries)
    k["At_mul_A_bin"] = {"lower": {"0,0": 2, "1,1": 3, "2,2": 2, "3,3": 2, "1,0": 0, "2,0": 2, "3,0": 0,
                                   "2,1": 0, "3,1": 1, "3,2": 0}}
    # tests.cpp:31-58  At_mul_A on the valued CSR (Approx entries)
    k["At_mul_A_csr"] = {"lower": {"0,0": 4.3801, "1,1": 2.4485, "2,2": 8.6420, "3,3": 5.9572, "1,0": 0,
                                   "2,0": 3.8282, "3,0": 0, "2,1": 0, "3,1": 0.0714, "3,2": 0}}
    # tests.cpp:209-233  chol_decomp + chol_solve_t on a 3 x 3 lower triangle, 5 rhs
    k["chol_solve_t"] = {
        "m": [[7, 0, 0], [2, 5, 0], [6, 1, 6]],
        "rhs": [[-1.227, -0.890, 0.293], [0.356, -0.733, -1.201], [-0.003, -0.091, -1.467],
                [0.819, 0.725, -0.719], [-0.485, 0.955, 1.707]],
        "xopt": [[-1.67161, 0.151609, 1.69517], [2.10217, -0.545174, -2.21148], [1.80587, -0.34187, -1.99339],
                 [1.71883, -0.180826, -1.80852], [-2.93874, 0.746739, 3.09878]],
    }
    # tests.cpp:859-885  make_dense_prior: FtF = lower(F^T F) of a 3 x 2 F
    k["dense_prior_ftf"] = {"F": [[0.1, 0.3], [0.4, 0.11], [-0.7, 0.23]]}
    # tests.cpp:338-368  At_mul_Bt: B (2 x 6) times column 1 of the feature matrix
    k["At_mul_Bt"] = {"B": [[-0.23, -2.89, -1.04, -0.52, -1.45, -1.42], [-0.16, -0.62, 1.19, 1.12, 0.11, 0.61]],
                      "col": 1, "Y_bin": [-4.16, 0.41], "Y_csr": [0.9942, 1.8285]}
    # tests.cpp:370-403  add_Acol_mul_bt: Z (2 x 6) += b * column 1^T
    k["add_Acol_mul_bt"] = {
        "Z": [[-0.23, -2.89, -1.04, -0.52, -1.45, -1.42], [-0.16, -0.62, 1.19, 1.12, 0.11, 0.61]],
        "b": [-4.16, 0.41], "col": 1,
        "Z_bin": [[-4.39, -7.05, -5.2, -0.52, -1.45, -1.42], [0.25, -0.21, 1.6, 1.12, 0.11, 0.61]],
        "Z_csr": [[-2.726, 0.5212, -5.9904, -0.52, -1.45, -1.42], [0.086, -0.9562, 1.6779, 1.12, 0.11, 0.61]]}
    # tests.cpp:606-614  auc of descending predictions
    k["auc"] = {"test": [1, 0, 1, 0, 1, 0, 0, 1, 0, 1] + [0] * 10,
                "pred": [20.0 - i for i in range(20)], "auc": 0.84}
    # tests.cpp:887-892  inv_norm_cdf (p, expected); p = 0 gives -inf
    k["inv_norm_cdf"] = [[0.5, 0.0], [0.9, 1.2815515655446004], [0.01, -2.3263478740408408]]
    # tests.cpp:894-899  norm_cdf (x, expected)
    k["norm_cdf"] = [[0.0, 0.5], [-1.0, 0.15865525393145707], [-3.0, 0.0013498980316300933],
                     [4.0, 0.99996832875816688]]
    # tests.cpp:901-910  rand_truncnorm(mean, std, low) >= low, 10 draws each
    k["rand_truncnorm_bounds"] = [[0.0, 1.0, 2.0], [0.0, 1.0, 3.0], [0.0, 1.0, 5.0], [0.0, 1.0, 50.0],
                                  [30.0, 2.0, 50.0]]
    # tests.cpp:241-254  posterior_lambda_beta(beta, Lambda_u, nu = 0.01, mu = 0.05)
    k["posterior_lambda_beta"] = {"beta": [[3.0, -2.0, 0.5], [1.0, 0.91, -0.2]],
                                  "Lambda": [[0.5, 0.1], [0.1, 0.3]], "nu": 0.01, "mu": 0.05,
                                  "shape": 3.005, "scale": 0.2631083888}
    # tests.cpp:552-591  eval_rmse Welford over three iterations
    k["eval_rmse"] = {
        "y": 4.5, "mean": 2.0,
        "steps": [
            {"U": [1.0, 0.0], "V": [1.0, 0.0], "pred": 3.0, "var": 0.0, "rmse1": 1.5, "rmse": 1.5},
            {"U": [2.0, 0.0], "V": [1.0, 0.0], "pred": 3.5, "var": 0.5, "rmse1": 0.5, "rmse": 1.0},
            {"U": [2.0, 0.0], "V": [3.0, 0.0], "pred": 5.0, "var": 14.0, "rmse1": 3.5, "rmse": 0.5},
        ],
    }
    # tests.cpp:616-665  SparseMode layout
    k["sparse_mode"] = {
        "C": [[0, 1, 0], [0, 0, 0], [1, 3, 1], [2, 3, 0], [1, 0, 1]],
        "v": [0.1, 0.2, 0.3, 0.4, 0.5],
        "mode0": {"size": 4, "row_ptr": [0, 2, 4, 5, 5], "indices": [[1, 0], [0, 0], [3, 1], [0, 1], [3, 0]],
                  "values": [0.1, 0.2, 0.3, 0.5, 0.4]},
        "mode1": {"size": 4, "row_ptr": [0, 2, 3, 3, 5], "indices": [[0, 0], [1, 1], [0, 0], [1, 1], [2, 0]],
                  "values": [0.2, 0.5, 0.1, 0.3, 0.4]},
    }
    # tests.cpp:707-756  TensorData test-cell order (getTestData)
    k["tensor_test_order"] = {
        "dims": [4, 4, 2],
        "Cte": [[1, 1, 0], [0, 0, 0], [1, 3, 0], [0, 3, 0], [2, 3, 1], [2, 0, 0]],
        "vte": [-0.1, -0.2, -0.3, -0.4, -0.5, -0.6],
        "expected": [[0, 0, 0, -0.2], [0, 3, 0, -0.4], [1, 1, 0, -0.1], [1, 3, 0, -0.3], [2, 3, 1, -0.5],
                     [2, 0, 0, -0.6]],
    }
    # tests.cpp:518-550  sparseFromIJV
    k["sparse_from_ijv"] = {"rows": [0, 1, 2], "cols": [2, 1, 0], "vals": [1.0, 0.0, 2.0],
                            "dense": [[0.0, 0.0, 1.0], [0.0, 0.0, 0.0], [2.0, 0.0, 0.0]]}
    # Random123 kat_vectors: philox4x32 10 rounds (ctr[4], key[2] -> out[4])
    k["philox4x32_10"] = [
        {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
        {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2, "out": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
        {"ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], "key": [0xA4093822, 0x299F31D0],
         "out": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
    ]
    return k


def main():
    with open(os.path.join(HERE, "reference_kat.json"), "w") as f:
        json.dump(reference_kat(), f, indent=1)


if __name__ == "__main__":
    main()
