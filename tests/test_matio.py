"""The MAT v5 reader/writer of the library (vls_mat_*) against scipy.io, on
the reference's own run records (tests/golden/CFL0_5.mat, CFL0_1.mat: the
arrays p00_test_singlephase_relax.m saves) and on round trips.  Host-only: no
GPU call is made."""
import importlib
import os

import numpy as np
import pytest
import scipy.io as sio

vls = importlib.import_module("2018_7_vesicle_amd")
HERE = os.path.dirname(os.path.abspath(__file__))
RECORDS = [os.path.join(HERE, "golden", f) for f in ("CFL0_5.mat", "CFL0_1.mat", "oblate48.mat", "oblate64.mat",
                                                     "oblate80.mat")]
PEAR = "/root/reference/src/7_0_paper/2_0_Vesicle/pear.mat"


@pytest.mark.parametrize("path", RECORDS, ids=[os.path.basename(p)[:-4] for p in RECORDS])
def test_reads_the_reference_run_records(path):
    got = vls.matio.loadmat(path)
    want = {k: v for k, v in sio.loadmat(path).items() if not k.startswith("__")}
    assert sorted(got) == sorted(want)
    for k in want:
        assert got[k].shape == want[k].shape and np.array_equal(got[k], want[k]), k


@pytest.mark.skipif(not os.path.exists(PEAR), reason="the reference checkout is not present")
def test_reads_a_compressed_level_set_field():
    got = vls.matio.loadmat(PEAR)
    want = sio.loadmat(PEAR)["F"]
    assert np.array_equal(got["F"], want)
    f = vls.matio.field(got["F"])
    assert f.shape == (64, 64, 64) and f.flags.c_contiguous
    assert np.array_equal(vls.matio.matlab_array(f), want)


@pytest.mark.parametrize("compress", [False, True])
def test_round_trip_and_scipy_reads_ours(tmp_path, compress):
    rng = np.random.default_rng(5)
    arrays = {"F": rng.normal(size=(5, 6, 7)), "vector": np.arange(11.0), "scalar": -2.25,
              "longer_name_here": rng.normal(size=(3, 4)), "empty": np.zeros((0, 3))}
    p = str(tmp_path / "out.mat")
    vls.matio.savemat(p, arrays, compress=compress)
    back = vls.matio.loadmat(p)
    sp = sio.loadmat(p)
    for k, a in arrays.items():
        a = np.asarray(a, dtype=np.float64)
        a = a.reshape(1, -1) if a.ndim == 1 else (a.reshape(1, 1) if a.ndim == 0 else a)
        assert np.array_equal(back[k], a), k
        assert np.array_equal(sp[k], a), k
    assert [n for n, _ in vls.matio.whosmat(p)] == list(arrays)


def test_reads_narrowed_integer_and_single_storage(tmp_path):
    p = str(tmp_path / "s.mat")
    sio.savemat(p, {"i16": np.arange(-3, 3, dtype=np.int16).reshape(2, 3), "u8": np.uint8([[7, 200]]),
                    "f32": np.float32([[1.5, -2.5, 3.25]]), "i64": np.int64([[2 ** 40]]), "flag": np.array([[True]])},
                do_compression=True)
    got = vls.matio.loadmat(p)
    assert np.array_equal(got["i16"], np.arange(-3, 3).reshape(2, 3))
    assert np.array_equal(got["u8"], [[7, 200]]) and np.array_equal(got["f32"], [[1.5, -2.5, 3.25]])
    assert got["i64"][0, 0] == 2 ** 40 and got["flag"][0, 0] == 1


def test_errors_are_raised(tmp_path):
    with pytest.raises(vls.VLSError, match="cannot open"):
        vls.matio.loadmat(str(tmp_path / "missing.mat"))
    raw = open(RECORDS[0], "rb").read()
    bad = tmp_path / "trunc.mat"
    bad.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(vls.VLSError):
        vls.matio.loadmat(str(bad))
    short = tmp_path / "short.mat"
    short.write_bytes(raw[:100])
    with pytest.raises(vls.VLSError, match="short"):
        vls.matio.loadmat(str(short))
