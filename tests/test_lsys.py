"""The LSYS Newton-system snapshot format (dump_linsys.hpp:1-16, 105-123):
header and record layout written byte for byte as the reference's dumper
lays them out, and a written system read back is the same operator
(dense fp64 comparison) with the same right-hand side; repeated records are
summed, malformed files are refused."""
import numpy as np
import pytest

from ppf_amd import lsys


def random_system(n=40, seed=0, per_row=4):
    rng = np.random.default_rng(seed)
    rp, ci = [0], []
    for i in range(n):
        cols = sorted(set([i] + list(rng.integers(i, n, per_row))))
        ci += cols
        rp.append(len(ci))
    nnz = len(ci)
    return {"row_offset": np.array(rp, np.int32), "col_index": np.array(ci, np.int32),
            "value": rng.normal(0, 1, (nnz, 9)).astype(np.float32),
            "cdiag": rng.normal(0, 1, (n, 9)).astype(np.float32),
            "b": rng.normal(0, 1, 3 * n).astype(np.float32)}


def dense(s):
    rp = np.asarray(s["row_offset"], np.int64)
    ci = np.asarray(s["col_index"], np.int64)
    n = len(rp) - 1
    A = np.zeros((3 * n, 3 * n))
    val = np.asarray(s["value"], np.float64).reshape(-1, 9)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            B = val[k].reshape(3, 3, order="F")
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] += B
            if j != i:
                A[3 * j:3 * j + 3, 3 * i:3 * i + 3] += B.T
    cd = np.asarray(s["cdiag"], np.float64).reshape(n, 9)
    for i in range(n):
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] += cd[i].reshape(3, 3, order="F")
    return A


def test_layout_and_round_trip(tmp_path):
    s = random_system()
    path = tmp_path / "linsys_0.bin"
    lsys.write(path, s)
    raw = path.read_bytes()
    magic, n, noff = np.frombuffer(raw[:12], "<u4")
    assert raw[:4] == b"SYSL"            # 0x4C535953 little-endian
    assert magic == 0x4C535953 and n == 40
    rows = np.repeat(np.arange(40), np.diff(s["row_offset"]))
    assert noff == int(np.sum(s["col_index"] != rows))
    assert len(raw) == 12 + noff * (8 + 36) + 40 * 36 + 40 * 12
    first = np.frombuffer(raw[12:12 + 44], np.uint8)
    r0, c0 = np.frombuffer(first[:8].tobytes(), "<u4")
    assert r0 < c0
    back = lsys.read(path)
    assert np.allclose(dense(back), dense(s), rtol=1e-6, atol=1e-6)
    assert np.array_equal(back["b"], s["b"])


def test_repeated_records_are_summed_and_bad_files_refused(tmp_path):
    rec = np.zeros(3, lsys._REC)
    rec["row"], rec["col"] = [0, 0, 1], [1, 1, 2]
    rec["val"][0], rec["val"][1], rec["val"][2] = 1.0, 2.0, 5.0
    diag = np.tile(np.eye(3, dtype=np.float32).reshape(-1), (3, 1))
    b = np.arange(9, dtype=np.float32)
    path = tmp_path / "dup.bin"
    with open(path, "wb") as f:
        f.write(np.array([lsys.MAGIC, 3, 3], "<u4").tobytes())
        f.write(rec.tobytes() + diag.tobytes() + b.tobytes())
    s = lsys.read(path)
    assert list(s["row_offset"]) == [0, 1, 2, 2] and list(s["col_index"]) == [1, 2]
    assert np.all(s["value"][0] == 3.0) and np.all(s["value"][1] == 5.0)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(np.array([0x12345678, 1, 0], "<u4").tobytes() + bytes(48))
    with pytest.raises(ValueError):
        lsys.read(bad)
    short = tmp_path / "short.bin"
    short.write_bytes(path.read_bytes()[:-4])
    with pytest.raises(ValueError):
        lsys.read(short)
