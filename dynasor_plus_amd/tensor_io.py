"""Sparse tensor container and file formats (FROSTT .tns, .dnb cache, DYFM).

Byte-compatible with the reference formats:
  * .tns  -- text, one nonzero per line "i_0 ... i_{N-1} value", 1-based
             indices, '#' comments, optional SPLATT header "N" then the N mode
             sizes; mode size = max(index) (or the header size, whichever is
             larger); values parsed as double then cast to float32
             (src/dynasor_io.cpp:576-745, 960-1176).
  * .dnb  -- 88-byte header {'DNB1', u32 version=1, u32 value_bytes=4,
             u32 index_bytes=4, u32 N, u64 nnz, u32 dims[8], u32 reserved[7]}
             then vals f32[nnz] then idx u32[N][nnz] (src/dynasor_io.cpp:747-893).
  * DYFM  -- 24-byte header {'DYFM', u32 version=1, u32 value_bytes=4,
             u32 rank, u64 rows} then row-major float32 rows x rank
             (src/dynasor_io.cpp:1287-1348).
Duplicates are kept as stored, like the reference loader.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

MAX_MODES = 8
DNB_MAGIC = b"DNB1"
DNB_HEADER = struct.Struct("<4sIIIIQ8I7I")        # 88 bytes
DYFM_HEADER = struct.Struct("<4sIIIQ")             # 24 bytes
assert DNB_HEADER.size == 88 and DYFM_HEADER.size == 24


class TensorFormatError(ValueError):
    pass


@dataclass
class Tensor:
    """SoA COO tensor: vals[nnz] float32, idx[n][nnz] uint32 (0-based)."""
    dims: tuple
    vals: np.ndarray
    idx: list = field(default_factory=list)

    @property
    def num_modes(self) -> int:
        return len(self.dims)

    @property
    def nnz(self) -> int:
        return int(self.vals.shape[0])

    def __post_init__(self):
        self.dims = tuple(int(d) for d in self.dims)
        self.vals = np.ascontiguousarray(self.vals, dtype=np.float32)
        self.idx = [np.ascontiguousarray(i, dtype=np.uint32) for i in self.idx]
        if not 2 <= len(self.dims) <= MAX_MODES:
            raise TensorFormatError(f"num_modes must be in [2, {MAX_MODES}], got {len(self.dims)}")
        if len(self.idx) != len(self.dims):
            raise TensorFormatError("one index column per mode is required")
        for n, col in enumerate(self.idx):
            if col.shape != self.vals.shape:
                raise TensorFormatError(f"index column {n} has {col.shape[0]} entries, "
                                        f"vals has {self.vals.shape[0]}")


# ---------------------------------------------------------------------------
# .tns text
# ---------------------------------------------------------------------------
def load_tns(path: str | os.PathLike) -> Tensor:
    rows = []
    header_modes = 0
    header_sizes = None
    header_done = False
    num_modes = 0
    with open(path, "r") as fh:
        for line in fh:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            toks = s.split()
            try:
                vals = [float(t) for t in toks]
            except ValueError as exc:
                raise TensorFormatError(f"cannot parse line: {s!r}") from exc
            if not header_done:
                if (len(vals) == 1 and num_modes == 0 and vals[0] == int(vals[0])
                        and 2 <= vals[0] <= MAX_MODES):
                    header_modes = int(vals[0])
                    num_modes = header_modes
                    continue
                if header_modes and header_sizes is None and len(vals) == header_modes:
                    header_sizes = [int(v) for v in vals]
                    header_done = True
                    continue
                header_done = True
            if len(vals) < 2:
                continue
            nm = len(vals) - 1
            if num_modes == 0:
                if nm > MAX_MODES:
                    raise TensorFormatError(f"tensor has {nm} modes (max {MAX_MODES})")
                num_modes = nm
            elif nm != num_modes:
                raise TensorFormatError(
                    f"inconsistent mode count (line expects {num_modes}, got {nm})")
            rows.append(vals)
    if num_modes == 0 or not rows:
        raise TensorFormatError(f"{path}: no nonzeros")
    arr = np.array(rows, dtype=np.float64)
    ix = arr[:, :num_modes].astype(np.int64)
    if (ix <= 0).any():
        raise TensorFormatError(f"non-positive index {int(ix.min())}")
    idx = [(ix[:, n] - 1).astype(np.uint32) for n in range(num_modes)]
    dims = [int(col.max()) + 1 for col in idx]
    if header_sizes:
        dims = [max(d, h) for d, h in zip(dims, header_sizes)]
    vals = arr[:, num_modes].astype(np.float32)
    return Tensor(tuple(dims), vals, idx)


def save_tns(path: str | os.PathLike, T: Tensor, values: np.ndarray | None = None,
             fmt: str = "%.9g") -> None:
    """Write FROSTT text.  `values` (e.g. float64) overrides T.vals."""
    v = T.vals if values is None else values
    cols = [T.idx[n].astype(np.int64) + 1 for n in range(T.num_modes)]
    with open(path, "w") as fh:
        for k in range(T.nnz):
            fh.write(" ".join(str(int(c[k])) for c in cols))
            fh.write(" " + (fmt % float(v[k])) + "\n")


# ---------------------------------------------------------------------------
# .dnb binary cache
# ---------------------------------------------------------------------------
def save_dnb(path: str | os.PathLike, T: Tensor) -> None:
    if T.nnz == 0:
        raise TensorFormatError("cannot write an empty tensor")
    dims = list(T.dims) + [0] * (MAX_MODES - T.num_modes)
    hdr = DNB_HEADER.pack(DNB_MAGIC, 1, 4, 4, T.num_modes, T.nnz, *dims, *([0] * 7))
    tmp = str(path) + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(hdr)
        fh.write(T.vals.astype("<f4").tobytes())
        for col in T.idx:
            fh.write(col.astype("<u4").tobytes())
    os.replace(tmp, path)


def load_dnb(path: str | os.PathLike) -> Tensor:
    with open(path, "rb") as fh:
        raw = fh.read(DNB_HEADER.size)
        if len(raw) != DNB_HEADER.size:
            raise TensorFormatError(f"{path}: truncated header")
        fields = DNB_HEADER.unpack(raw)
        magic, version, vbytes, ibytes, N, nnz = fields[:6]
        dims = fields[6:14]
        if magic != DNB_MAGIC:
            raise TensorFormatError(f"{path}: bad magic {magic!r}")
        if version != 1:
            raise TensorFormatError(f"{path}: unsupported version {version}")
        if vbytes != 4 or ibytes != 4:
            raise TensorFormatError(f"{path}: value_bytes={vbytes} index_bytes={ibytes} (need 4/4)")
        if N == 0 or N > MAX_MODES:
            raise TensorFormatError(f"{path}: num_modes {N} out of range")
        if nnz == 0:
            raise TensorFormatError(f"{path}: nnz is 0")
        vals = np.fromfile(fh, dtype="<f4", count=nnz)
        if vals.shape[0] != nnz:
            raise TensorFormatError(f"{path}: truncated values")
        idx = []
        for n in range(N):
            col = np.fromfile(fh, dtype="<u4", count=nnz)
            if col.shape[0] != nnz:
                raise TensorFormatError(f"{path}: truncated index column {n}")
            idx.append(col)
    return Tensor(tuple(int(d) for d in dims[:N]), vals, idx)


def dnb_cache_is_fresh(tns: str | os.PathLike, dnb: str | os.PathLike) -> bool:
    d = Path(dnb)
    if not d.exists() or d.stat().st_size < DNB_HEADER.size:
        return False
    t = Path(tns)
    if not t.exists():
        return True
    return d.stat().st_mtime >= t.stat().st_mtime


def load_coo_tensor(path: str | os.PathLike, cache_on_success: bool = True) -> Tensor:
    """`<path>.dnb` when fresh, else parse the text and (optionally) write the cache
    (src/dynasor_io.cpp:1184-1261)."""
    dnb = str(path) + ".dnb"
    if cache_on_success and dnb_cache_is_fresh(path, dnb):
        try:
            return load_dnb(dnb)
        except TensorFormatError:
            pass    # stale or invalid cache: reparse the text
    T = load_tns(path)
    if cache_on_success:
        try:
            save_dnb(dnb, T)
        except OSError:
            pass
    return T


# ---------------------------------------------------------------------------
# DYFM factor dumps
# ---------------------------------------------------------------------------
def save_dyfm(path: str | os.PathLike, Y: np.ndarray) -> None:
    Y = np.ascontiguousarray(Y, dtype="<f4")
    rows, rank = Y.shape
    with open(path, "wb") as fh:
        fh.write(DYFM_HEADER.pack(b"DYFM", 1, 4, rank, rows))
        fh.write(Y.tobytes())


def load_dyfm(path: str | os.PathLike) -> np.ndarray:
    with open(path, "rb") as fh:
        raw = fh.read(DYFM_HEADER.size)
        if len(raw) != DYFM_HEADER.size:
            raise TensorFormatError(f"{path}: truncated DYFM header")
        magic, version, vbytes, rank, rows = DYFM_HEADER.unpack(raw)
        if magic != b"DYFM" or version != 1 or vbytes != 4:
            raise TensorFormatError(f"{path}: not a DYFM v1 float32 file")
        data = np.fromfile(fh, dtype="<f4", count=rows * rank)
        if data.shape[0] != rows * rank:
            raise TensorFormatError(f"{path}: truncated DYFM data")
    return data.reshape(rows, rank)
