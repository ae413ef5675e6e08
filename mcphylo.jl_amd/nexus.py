"""Alignment ingestion: the NEXUS subset and the CSV format the reference
reads, and `datafortree`.

`ParseNexus` follows `src/Parser/ParseNexus.jl:10-130` (MrBayes-style,
non-interleaved DATA block; `symbols` from the FORMAT line, otherwise the
sorted alphabet of the matrix without the gap and missing symbols).
`datafortree` follows `src/Parser/Parser.jl:48-85`: one-hot tip columns, an
all-ones column for gap or missing, unknown symbols raise.

The partial array uses the reference's shape (S, C, N) with Julia's
column-major memory order (`order='F'`), so x[s, c, v-1] is state s of site c
at node num v.  Internal-node slices are zero here; the likelihood never reads
them (the reference's `my_repeat` overwrites them, `VectorizedFunctions.jl:26`).
"""

from __future__ import annotations

import numpy as np

from .tree import Node, post_order

__all__ = ["ParseNexus", "ParseCSV", "create_csvdf", "datafortree", "compress_patterns"]


def ParseNexus(filename: str):
    """Returns (ntax, nchar, gap, missing, symbols, df, langs) where df is a
    (ntax x nchar) array of single characters and langs the taxon names."""
    with open(filename, "r") as fh:
        content = fh.read().splitlines()
    if not content or content[0].strip().lower() != "#nexus":
        raise ValueError(f"{filename} is not a Nexus file!")
    while True:
        if not content:
            raise ValueError("no 'begin data;' block")
        line = content.pop(0)
        if line.strip().lower() == "begin data;":
            break
    ntax, nchar, gap, missing, symbols = 0, 0, "-", "?", "NOSYMBOLS"
    while True:
        line = content.pop(0)
        if "matrix" in line.lower():
            break
        for entry in line.split():
            info = [t.lower() for t in entry.split("=")]
            if len(info) != 1:
                val = info[1]
                if val.endswith(";"):
                    val = val[:-1]
                key = info[0]
                if key == "ntax":
                    ntax = int(val)
                elif key == "nchar":
                    nchar = int(val)
                elif key == "gap":
                    gap = val
                elif key == "missing":
                    missing = val
                elif key == "symbols":
                    symbols = val.strip('"')
    langs: list[str] = []
    rows: list[str] = []
    while True:
        line = content.pop(0)
        if line == "":
            continue
        if line.rstrip()[-1:] == ";":
            break
        lang, raw = line.split(None, 1)
        langs.append(lang)
        rows.append(raw.strip())
    width = len(rows[-1])
    if any(len(r) != width for r in rows):
        raise ValueError("rows of the matrix have different lengths")
    df = np.array([list(r) for r in rows], dtype="<U1")
    if symbols == "NOSYMBOLS":
        alph = sorted({ch for ch in df.ravel() if ch != gap and ch != missing})
    else:
        alph = [ch for ch in symbols]
    return ntax, nchar, gap, missing, alph, df, langs


def create_csvdf(filecontent, separator: str = ","):
    """(langs, df) from lines `name<sep>c1<sep>c2...` (`src/Parser/ParseCSV.jl:44-62`):
    the first field of a line is the taxon name, and df holds the first
    character of every other field, one row per line."""
    langs: list[str] = []
    rows: list[list[str]] = []
    for line in filecontent:
        fields = line.split(separator)
        langs.append(fields[0])
        rows.append([f[0] for f in fields[1:]])
    width = len(rows[-1]) if rows else 0
    if any(len(r) != width for r in rows):
        raise ValueError("rows of the matrix have different lengths")
    return langs, np.array(rows, dtype="<U1").reshape(len(rows), width)


def ParseCSV(filename: str, gap: str, miss: str, header: bool = True):
    """MrBayes-convention CSV alignment (`src/Parser/ParseCSV.jl:17-36`).
    Returns (ntax, nchar, gap, miss, symbols, df, langs) like ParseNexus;
    `header` skips the first line; symbols are the sorted alphabet of the
    matrix without the gap and missing symbols (`ParseNexus.jl:38-49`)."""
    with open(filename, "r") as fh:
        content = fh.read().splitlines()
    if header and content:
        content.pop(0)
    langs, df = create_csvdf(content)
    ntax, nchar = df.shape
    alph = sorted({ch for ch in df.ravel() if ch != gap and ch != miss})
    return ntax, nchar, gap, miss, alph, df, langs


def datafortree(df, leave_names, tree: Node, symbols, gap: str, miss: str,
                log_space: bool = False) -> np.ndarray:
    """One-hot tip partials of shape (S, C, N), Fortran order."""
    nodes = post_order(tree)
    N = len(nodes)
    by_name = {v.name: v for v in nodes}
    S = len(symbols)
    df = np.asarray(df)
    C = df.shape[1]
    x = np.zeros((S, C, N), order="F")
    lut = {s: i for i, s in enumerate(symbols)}
    one = 0.0 if log_space else 1.0
    zero = -np.inf if log_space else 0.0
    for row, name in zip(df, leave_names):
        if name not in by_name:
            raise KeyError(f"taxon {name!r} not in tree")
        v = by_name[name].num - 1
        for c, ch in enumerate(row):
            idx = lut.get(str(ch))
            if idx is None:
                if ch == gap or ch == miss:
                    x[:, c, v] = one
                else:
                    raise ValueError(f"unknown symbol {ch}, {symbols}")
            else:
                x[:, c, v] = zero
                x[idx, c, v] = one
    return x


def compress_patterns(x: np.ndarray, leaf_nums):
    """Merge identical site columns over the leaves.  Returns (x_unique,
    weights, inverse) with x_unique of shape (S, C', N) and integer weights
    (counts).  The reference has no compression (every column is one site);
    with these weights the likelihood sums are unchanged up to summation
    order."""
    leaf_idx = np.asarray(leaf_nums) - 1
    S, C, N = x.shape
    keys = np.ascontiguousarray(np.transpose(x[:, :, leaf_idx], (1, 0, 2)).reshape(C, -1))
    _, first, inverse, counts = np.unique(keys, axis=0, return_index=True,
                                          return_inverse=True, return_counts=True)
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(order.shape[0])
    xu = np.asfortranarray(x[:, first[order], :])
    return xu, counts[order].astype(np.float64), remap[inverse.ravel()]
