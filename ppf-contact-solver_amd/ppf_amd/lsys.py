"""The Newton-system snapshot format (LSYS) of the reference's offline dumper
(crates/ppf-cts-solver/src/cpp/main/dump_linsys.hpp:1-16, 105-123): the
assembled system M x = b of one solve, little-endian:

    u32 magic 0x4C535953 ("LSYS"), u32 nrow, u32 n_offdiag
    n_offdiag records: u32 row, u32 col (row < col), 9 f32 (column-major block)
    nrow diagonal blocks: 9 f32 (stored diagonal + per-row diagonal)
    3 nrow f32 right-hand side

Records may repeat a (row, col) pair (the dumper emits the dynamic and the
fixed matrix separately); readers sum them. This module moves the solver's
systems (the dict `Stepper.system()` returns: row_offset, col_index, value,
cdiag, b) to and from that format, so snapshots logged here train and
evaluate like the reference's dumps.
"""
import numpy as np

MAGIC = 0x4C535953
_REC = np.dtype([("row", "<u4"), ("col", "<u4"), ("val", "<f4", (9,))])


def _host(a, dtype):
    """A host copy of a (numpy or torch) array as `dtype`; integer arrays of
    the same width are reinterpreted (the solver keeps indices as int32)."""
    a = np.ascontiguousarray(a.detach().cpu().numpy() if hasattr(a, "detach") else a)
    want = np.dtype(dtype)
    if a.dtype.kind in "iu" and want.kind in "iu" and a.dtype.itemsize == want.itemsize:
        return a.view(want)
    return a.astype(want, copy=False)


def to_lsys(sysd):
    """(header, records, diag (n, 9), b (3n)) of a solver system: the stored
    off-diagonal blocks as records in pattern order, each row's stored
    diagonal block added to its per-row diagonal."""
    rp = _host(sysd["row_offset"], np.uint32).reshape(-1)
    ci = _host(sysd["col_index"], np.uint32).reshape(-1)
    n = len(rp) - 1
    val = _host(sysd["value"], np.float32).reshape(-1, 9)[:len(ci)]
    diag = _host(sysd["cdiag"], np.float32).reshape(n, 9).copy()
    b = _host(sysd["b"], np.float32).reshape(-1)[:3 * n]
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rp).astype(np.int64))
    on = ci == rows
    np.add.at(diag, rows[on], val[on])
    off = ~on
    if np.any(ci[off] < rows[off]):
        raise ValueError("lsys: the pattern stores a block below the diagonal")
    rec = np.zeros(int(off.sum()), _REC)
    rec["row"], rec["col"], rec["val"] = rows[off], ci[off], val[off]
    return (MAGIC, n, len(rec)), rec, diag, b


def write(path, sysd):
    (magic, n, noff), rec, diag, b = to_lsys(sysd)
    with open(path, "wb") as f:
        f.write(np.array([magic, n, noff], "<u4").tobytes())
        f.write(rec.tobytes())
        f.write(diag.astype("<f4").tobytes())
        f.write(b.astype("<f4").tobytes())


def read(path):
    """A system dict (numpy): row_offset, col_index (upper pattern of the
    off-diagonal blocks, columns ascending per row, repeated pairs summed),
    value (nnz, 9), cdiag (n, 9) = the file's diagonal blocks, b (3n), n."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < 12:
        raise ValueError("lsys: file too short")
    magic, n, noff = np.frombuffer(raw[:12].tobytes(), "<u4")
    if magic != MAGIC:
        raise ValueError(f"lsys: bad magic {magic:#x}")
    n, noff = int(n), int(noff)
    need = 12 + noff * _REC.itemsize + 36 * n + 12 * n
    if raw.size != need:
        raise ValueError(f"lsys: {raw.size} bytes, expected {need}")
    o = 12
    rec = np.frombuffer(raw[o:o + noff * _REC.itemsize].tobytes(), _REC)
    o += noff * _REC.itemsize
    diag = np.frombuffer(raw[o:o + 36 * n].tobytes(), "<f4").reshape(n, 9).copy()
    o += 36 * n
    b = np.frombuffer(raw[o:o + 12 * n].tobytes(), "<f4").copy()
    if noff and (np.any(rec["row"] >= rec["col"]) or rec["col"].max() >= n):
        raise ValueError("lsys: a record is not strictly upper or is out of range")
    key = rec["row"].astype(np.uint64) * np.uint64(max(n, 1)) + rec["col"]
    order = np.argsort(key, kind="stable")
    key, recs = key[order], rec[order]
    first = np.ones(len(key), bool)
    first[1:] = key[1:] != key[:-1]
    group = np.cumsum(first) - 1
    vals = np.zeros((int(first.sum()), 9), np.float32)
    np.add.at(vals, group, recs["val"])
    rows, cols = recs["row"][first], recs["col"][first]
    rp = np.zeros(n + 1, np.uint32)
    np.add.at(rp, rows.astype(np.int64) + 1, 1)
    rp = np.cumsum(rp, dtype=np.uint64).astype(np.uint32)
    return {"n": n, "row_offset": rp, "col_index": cols.astype(np.uint32), "value": vals,
            "cdiag": diag, "b": b}
