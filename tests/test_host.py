"""CPU tests of the host side: substitution models, rates, NEXUS/Newick
ingestion, node numbering, PhyloDist constructors, and the C-ABI library
(loads and exports every declared symbol; no compute call needs a GPU).

Golden values: test/likelihood/substitutionmodels.jl:1-118,
test/likelihood/rates.jl:1-18, test/distributions/phylodist.jl:1-95.
"""

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PRIMATES_NEWICK, ROOT, SIMU_NEWICK, load_fixture, mcp

RTOL = 1.5e-8  # Julia's default isapprox tolerance used by the reference tests


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.allclose(a, b, rtol=rtol, atol=rtol * max(1.0, np.max(np.abs(b))))


def same_up_to_column_sign(U, T):
    for j in range(T.shape[1]):
        if not (close(U[:, j], T[:, j]) or close(-U[:, j], T[:, j])):
            return False
    return True


# ---------------------------------------------------------------- models

def test_restriction_golden():
    U, D, Ui, mu = mcp.Restriction([0.3, 0.7], [0.0])
    assert close(U, [[0.7, 0.7], [-0.3, 0.7]])
    assert close(D, [-1.0, 0.0])
    assert close(Ui, [[1.0, -1.0], [0.428571428, 1.0]])
    assert close(mu, 2.380952380952381)
    with pytest.raises(AssertionError):
        mcp.Restriction(np.ones(3) / 3, [0.1])


def test_jc_golden():
    U, D, Ui, mu = mcp.JC([0.15, 0.45, 0.4], [0.0])
    target_Q = np.array([[-2 / 3, 1 / 3, 1 / 3], [1 / 3, -2 / 3, 1 / 3], [1 / 3, 1 / 3, -2 / 3]])
    assert close(U @ np.diag(D) @ Ui, target_Q)
    assert close(mu, 1.5)


def test_gtr_golden():
    U, D, Ui, mu = mcp.GTR([0.5, 0.2, 0.3],
                           [0.060325906174435326, 0.48940696298364417, 2.4502671308419206])
    TU = np.array([[-0.06919383208280398, 0.5545230380047614, 0.5773502691896261],
                   [-0.7729814547853464, -0.6912032799070211, 0.5773502691896257],
                   [0.6306440233282372, -0.46340287673658875, 0.5773502691896257]])
    TD = np.array([-1.3622649268468585, -0.29662234328312154, -5.551115123125783e-17])
    TUi = np.array([[-0.1434321592442829, -0.6409264859501984, 0.7843586451944811],
                    [0.8837782123919058, -0.44064564120807803, -0.443132571183828],
                    [0.8660254037844387, 0.34641016151377513, 0.5196152422706631]])
    assert np.allclose(D, TD, rtol=RTOL, atol=1e-14)
    assert same_up_to_column_sign(U, TU)
    assert close(U @ np.diag(D) @ Ui, TU @ np.diag(TD) @ TUi)
    assert close(mu, 0.6028137161614641)


def test_freeK_golden():
    U, D, Ui, mu = mcp.freeK([0.0], [0.03338775337571049, 0.2519159175077897,
                                     0.8202684796606095, 2.8944278494558904])
    TU = np.array([[-0.9913312346241385, -0.7071067811865475],
                   [0.131386389167909, -0.7071067811865476]])
    TD = np.array([-0.2853036708835002, -6.938893903907228e-18])
    TUi = np.array([[-0.8906959139221835, 0.8906959139221834],
                    [-0.16549879465230727, -1.2487147677207877]])
    assert np.allclose(D, TD, rtol=RTOL, atol=1e-14)
    assert same_up_to_column_sign(U, TU)
    assert close(U @ np.diag(D) @ Ui, TU @ np.diag(TD) @ TUi)
    assert close(mu, 3.5050372710007514)


def test_setmatrix_golden():
    r = mcp.setmatrix([0.1, 0.2, 0.3, 0.4, 0.5, 0.6])
    assert np.array_equal(r, [[0.0, 0.1, 0.2, 0.4], [0.1, 0.0, 0.3, 0.5],
                              [0.2, 0.3, 0.0, 0.6], [0.4, 0.5, 0.6, 0.0]])
    with pytest.raises(ValueError):
        mcp.setmatrix([0.1, 0.2])


# ---------------------------------------------------------------- rates

def test_discrete_gamma_rates_golden():
    assert close(mcp.discrete_gamma_rates(0.5, 0.5, 4),
                 [0.03338775337571049, 0.2519159175077897, 0.8202684796606095, 2.8944278494558904])
    assert close(mcp.discrete_gamma_rates(0.5, 0.5, 4, method="median"),
                 [0.029077754761925846, 0.2807145371399754, 0.92477306511421, 2.7654346429838887])
    assert close(mcp.median_boundaries(0.5, 0.5, 4),
                 [0.024746651492520606, 0.23890238006213632, 0.7870290171790321, 2.353525844604548])
    assert close(mcp.mean_boundaries(0.5, 0.5, 4)[:3],
                 [0.10153104426762159, 0.45493642311957283, 1.3233036969314669])
    # categories have mean one
    assert abs(np.mean(mcp.discrete_gamma_rates(0.7, 0.7, 6)) - 1.0) < 1e-12


# ---------------------------------------------------------------- ingestion

def test_parse_nexus_fixtures():
    ntax, nchar, gap, miss, sym, df, langs = mcp.ParseNexus(os.path.join(GOLDEN, "primates.nex"))
    assert (ntax, nchar, gap, miss) == (12, 898, "-", "?")
    assert list(sym) == ["A", "C", "G", "T"]
    assert df.shape == (12, 898) and langs[0] == "Tarsius_syrichta"
    ntax, nchar, *_ = mcp.ParseNexus(os.path.join(GOLDEN, "simudata.nex"))
    assert (ntax, nchar) == (10, 100)


# Expected values of the reference's own parser tests (test/parser/csv.jl,
# test/parser/parser.jl) on its fixtures example.csv / example2.nex.
CSV_MATRIX = ["00000-00001100-00?00?00", "0-0000?0000??00?0001??0", "0010000000-000100?00000"]
CSV_LANGS = ["Swedish_0", "Welsh_N_0", "Sardinian_N_0"]


def test_parse_csv_matches_reference_tests():
    ntax, nchar, gap, miss, sym, df, langs = mcp.ParseCSV(os.path.join(GOLDEN, "example.csv"),
                                                          "-", "?", False)
    assert (ntax, nchar, gap, miss) == (3, 23, "-", "?")
    assert list(sym) == ["0", "1"] and langs == CSV_LANGS
    assert ["".join(r) for r in df] == CSV_MATRIX
    content = [name + "," + ",".join(row) for name, row in zip(CSV_LANGS, CSV_MATRIX)]
    l1, d1 = mcp.create_csvdf(content)
    l2, d2 = mcp.create_csvdf([c.replace(",", ":") for c in content], ":")
    assert l1 == l2 == CSV_LANGS and np.array_equal(d1, df) and np.array_equal(d2, df)
    # a header line is skipped
    with open(os.path.join(GOLDEN, "example.csv")) as fh:
        first = fh.readline().strip()
    assert mcp.ParseCSV(os.path.join(GOLDEN, "example.csv"), "-", "?", True)[6] == CSV_LANGS[1:]
    assert first.startswith(CSV_LANGS[0])
    # the same alignment as Nexus parses to the same tuple
    nx = mcp.ParseNexus(os.path.join(GOLDEN, "example2.nex"))
    assert nx[:5] == (3, 23, "-", "?", ["0", "1"]) and nx[6] == CSV_LANGS
    assert np.array_equal(nx[5], df)


def test_datafortree_on_csv_matches_reference_tip_columns():
    # make_tree_with_data's expected one-hot columns of nodes 1 and 2
    # (Sardinian_N_0 and Welsh_N_0), gaps and missing as all-ones
    dim1 = np.array([[1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1],
                     [0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0]], float)
    dim2 = np.array([[1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1],
                     [0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 0, 1, 0, 0, 0, 1, 1, 1, 0]], float)
    ntax, nchar, gap, miss, sym, df, langs = mcp.ParseCSV(os.path.join(GOLDEN, "example.csv"),
                                                          "-", "?", False)
    t = mcp.ParseNewick("((Sardinian_N_0:0.1,Welsh_N_0:0.2):0.1,Swedish_0:0.3);")
    x = mcp.datafortree(df, langs, t, sym, gap, miss)
    by_name = {v.name: v.num for v in mcp.post_order(t)}
    assert by_name["Sardinian_N_0"] == 1 and by_name["Welsh_N_0"] == 2
    assert np.array_equal(x[:, :, 0], dim1) and np.array_equal(x[:, :, 1], dim2)


def test_newick_numbering_and_blv():
    t = mcp.ParseNewick(SIMU_NEWICK)
    po = mcp.post_order(t)
    assert len(po) == 18 and t.num == 18 and po[-1] is t
    by_name = {v.name: v.num for v in po}
    for k in range(10):
        assert by_name[str(k)] == k + 1          # leaves first, Newick order
    for k in range(10, 17):
        assert by_name[str(k)] == k + 1          # internal nodes in post-order
    blv = mcp.get_branchlength_vector(t)
    assert blv.shape == (17,) and blv[0] == 0.110833 and blv[16] == 0.0209771
    mcp.set_branchlength_vector(t, blv * 2)
    assert np.array_equal(mcp.get_branchlength_vector(t), blv * 2)
    t2 = mcp.ParseNewick(mcp.to_newick(t))
    assert np.array_equal(mcp.get_branchlength_vector(t2), blv * 2)
    assert np.array_equal(mcp.tree_arrays(t2).parent, mcp.tree_arrays(t).parent)


def test_datafortree_one_hot_and_gaps():
    t, x = load_fixture("primates")
    assert x.shape == (4, 898, 22)
    leaves = mcp.tree_arrays(t).leaf_nums - 1
    cols = x[:, :, leaves]
    s = cols.sum(axis=0)
    assert set(np.unique(s)) <= {1.0, 4.0}
    internal = np.setdiff1d(np.arange(22), leaves)
    assert np.all(x[:, :, internal] == 0.0)


def test_compress_patterns():
    t, x = load_fixture("primates")
    leaf = mcp.tree_arrays(t).leaf_nums
    xu, w, inv = mcp.compress_patterns(x, leaf)
    assert w.sum() == x.shape[1]
    assert np.array_equal(xu[:, inv, :][:, :, leaf - 1], x[:, :, leaf - 1])


def test_tree_from_parent_roundtrip():
    w = mcp.synthetic.make_workload("custom", seed=3, n=15, S=4, C=5, R=1)
    blv = mcp.get_branchlength_vector(w.tree)
    t = mcp.tree_from_parent(w.ta.parent, blv)
    assert np.array_equal(mcp.get_branchlength_vector(t), blv)


# ---------------------------------------------------------------- PhyloDist API

def test_phylodist_constructors_cpu():
    t, _ = load_fixture("primates")
    pden = np.ones(4) / 4
    pd = mcp.PhyloDist(t, pden, [1.0], [1.0], mcp.JC)
    pd4 = mcp.PhyloDist(t, pden, 1.0, 1.0, mcp.JC)
    assert pd == pd4
    assert mcp.size(pd) == (4, 1, 22)
    assert mcp.minimum(pd) == -np.inf and mcp.maximum(pd) == np.inf
    pd5 = mcp.PhyloDist(t, [1.0], [1.0], mcp.freeK)
    assert pd5.substitution_model is mcp.freeK and pd5.nbase == 1


def test_multiple_phylodist_dimension_errors_cpu():
    t, _ = load_fixture("primates")
    trees = [t, t]
    freqs = np.full((4, 2), 0.25)
    bad = np.array([[1.0, 2.0, 3.0]])
    with pytest.raises(mcp.McpDimensionMismatch):
        mcp.MultiplePhyloDist(trees, np.full((4, 3), 0.25), [1.0], [1.0], mcp.JC)
    with pytest.raises(mcp.McpDimensionMismatch):
        mcp.MultiplePhyloDist(trees, freqs, bad, [1.0], mcp.JC)
    with pytest.raises(mcp.McpDimensionMismatch):
        mcp.MultiplePhyloDist(trees, freqs, [1.0], bad, mcp.JC)
    mpd = mcp.MultiplePhyloDist(trees, freqs, np.ones((1, 2)), [1.0], mcp.JC)
    assert mcp.size(mpd) == (4, 1, 22, 2)


# ---------------------------------------------------------------- C ABI

def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "mcphylo_gpu.h")).read()
    declared = set(re.findall(r"\b(mcp_[a-z_]+)\s*\(", header))
    assert declared == set(mcp._lib.EXPORTED)
    lib = mcp.load_library()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.mcp_last_error() == b""


def test_library_rejects_null_arguments_without_gpu():
    lib = mcp.load_library()
    assert lib.mcp_plan_create(None, None) == mcp._lib.MCP_ERR_INVALID
    assert b"null" in lib.mcp_last_error()
    assert lib.mcp_logl(None, None, None) == mcp._lib.MCP_ERR_INVALID


def test_no_cpu_fallback_when_library_missing(monkeypatch):
    monkeypatch.setattr(mcp._lib, "_lib", None)
    monkeypatch.setattr(mcp._lib, "LIB_PATH", os.path.join(ROOT, "does-not-exist.so"))
    with pytest.raises(RuntimeError, match="not built"):
        mcp.load_library()


def test_header_is_c99_and_a_c_program_links_against_the_library(tmp_path):
    """The boundary is plain C: the header compiles as C99 with warnings as
    errors, and a C caller links against the shared library and gets the
    reference-style argument errors (no GPU call is made)."""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "caller.c"
    src.write_text(
        '#include <stdio.h>\n#include <string.h>\n#include "mcphylo_gpu.h"\n'
        "int main(void) {\n"
        "    mcp_plan* p = 0;\n"
        "    double ll = 0.0;\n"
        "    if (mcp_plan_create(&p, NULL) != MCP_ERR_INVALID) return 1;\n"
        '    if (strstr(mcp_last_error(), "null") == NULL) return 2;\n'
        "    if (mcp_logl(NULL, NULL, &ll) != MCP_ERR_INVALID) return 3;\n"
        '    printf("ok\\n");\n'
        "    return 0;\n}\n")
    exe = tmp_path / "caller"
    lib = os.path.dirname(mcp._lib.LIB_PATH)
    subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", inc, str(src), "-o",
                    str(exe), "-L", lib, "-lmcphylo_hip", "-Wl,-rpath," + lib], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", (out.returncode, out.stderr)
