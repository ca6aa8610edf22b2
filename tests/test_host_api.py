"""Host-side logic of the Python API (macau.pyx) that runs without a GPU:
input validation, train/test splitting, data layout.  CPU only."""
import itertools

import numpy as np
import pandas as pd
import pytest
import scipy.sparse

import macau_amd
from macau_amd.macau import Data


def test_make_train_test():  # test_macau.py:90-95
    X = scipy.sparse.rand(15, 10, 0.2, random_state=1)
    Xtr, Xte = macau_amd.make_train_test(X, 0.5)
    assert X.nnz == Xtr.nnz + Xte.nnz
    assert np.linalg.norm((X - Xtr - Xte).todense()) == 0.0


def test_make_train_test_df():  # test_macau.py:97-112
    idx = list(itertools.product(np.arange(10), np.arange(8), np.arange(3)))
    df = pd.DataFrame(np.asarray(idx), columns=["A", "B", "C"])
    df["value"] = np.arange(10.0 * 8.0 * 3.0)
    Ytr, Yte = macau_amd.make_train_test_df(df, 0.4)
    assert Ytr.shape[0] == df.shape[0] * 0.6
    assert Yte.shape[0] == df.shape[0] * 0.4
    A1 = np.zeros((10, 8, 3))
    A2 = np.zeros((10, 8, 3))
    A1[df.A, df.B, df.C] = df.value
    A2[Ytr.A, Ytr.B, Ytr.C] = Ytr.value
    A2[Yte.A, Yte.B, Yte.C] = Yte.value
    assert np.allclose(A1, A2)


def test_make_train_test_errors():
    with pytest.raises(TypeError):
        macau_amd.make_train_test(np.zeros((3, 3)), 0.5)
    with pytest.raises(TypeError):
        macau_amd.make_train_test(scipy.sparse.rand(3, 3, 0.5), -1)
    with pytest.raises(TypeError):
        macau_amd.make_train_test_df(np.zeros((3, 3)), 0.5)


def test_too_many_sides():  # test_macau.py:75-78
    Y = scipy.sparse.rand(10, 20, 0.2)
    with pytest.raises(ValueError):
        macau_amd.macau(Y, verbose=False, side=[None, None, None])


def test_side_must_be_list():
    Y = scipy.sparse.rand(10, 20, 0.2)
    with pytest.raises(ValueError):
        macau_amd.macau(Y, verbose=False, side="abc")


def test_univariate_argument_errors():  # macau.pyx:194-199, 393-395
    Y = scipy.sparse.rand(10, 20, 0.2)
    with pytest.raises(ValueError, match="probit"):
        macau_amd.macau(Y, verbose=False, univariate=True, precision="probit")
    with pytest.raises(ValueError, match="Unsupported side information type"):
        macau_amd.macau(Y, side=[np.ones((10, 2)), None], verbose=False, univariate=True)


def test_data_sparse_layout():
    Y = scipy.sparse.coo_matrix(([1.0, 2.0, np.nan], ([0, 1, 2], [1, 0, 2])), shape=(3, 3))
    d = Data(Y, None)
    assert list(d.shape) == [3, 3]
    assert len(d.valTrain) == 2  # NaN removed (macau.pyx:236-240)
    assert len(d.valTest) == 0
    assert list(d.colnames) == ["row", "col"]


def test_data_sparse_shape_mismatch():
    with pytest.raises(ValueError):
        Data(scipy.sparse.rand(3, 4, 0.5), scipy.sparse.rand(4, 3, 0.5))
    with pytest.raises(ValueError):
        Data(scipy.sparse.rand(3, 4, 0.5), np.zeros((3, 4)))


def test_data_frame_layout():
    df = pd.DataFrame({"A": [0, 2, 1], "B": [1, 0, 3], "C": [0, 0, 1], "v": [0.5, 1.5, 2.5]})
    te = pd.DataFrame({"A": [3], "B": [0], "C": [2], "v": [1.0]})
    d = Data(df, te)
    assert list(d.shape) == [4, 4, 3]  # max over train and test + 1
    assert d.colnames == ["A", "B", "C"]
    np.testing.assert_array_equal(d.idxTest[2], [2])


def test_data_frame_errors():
    df = pd.DataFrame({"A": [0, 1], "v": [0.5, 1.5]})
    with pytest.raises(ValueError):
        Data(df, None)  # fewer than 2 index columns
    df2 = pd.DataFrame({"A": [0, 1], "B": [1, 0], "v": [0.5, 1.5], "w": [1.0, 2.0]})
    with pytest.raises(ValueError):
        Data(df2, None)  # two float columns
    with pytest.raises(TypeError):
        Data(pd.DataFrame({"A": [0], "B": [0], "v": [1.0]}), scipy.sparse.rand(2, 2, 0.5))
    with pytest.raises(TypeError):
        Data([[1, 2]], None)


def test_blockcg_input_checks():
    X = scipy.sparse.rand(6, 4, 0.5)
    with pytest.raises(ValueError):
        macau_amd.blockcg(X, np.zeros((2, 4)), 0.5)  # not Fortran order
    with pytest.raises(ValueError):
        macau_amd.blockcg(X, np.zeros((2, 5), order="F"), 0.5)
