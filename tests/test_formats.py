"""CPU tests of the file formats (include/rcm_formats.h): MatrixMarket and .hicsr read/write, COO -> CSR and
prepareRCM with values, and the Graph<T>::load pipeline.

Parity is pinned three ways: against the reference's own loadMTX / storeMTX / loadCSR / storeCSR / convert /
prepareRCM / Graph::load compiled from its sources (oracle/_ref, present in the build container only), against
byte layouts derived from the reference's struct definitions (src/CSR.cpp:44-104), and against independent
numpy restatements (a stable lexicographic sort for convert).
"""
from __future__ import annotations

import os
import shutil
import struct

import numpy as np
import pytest

import oracle_lib as O
import parallelbatchrcm_amd as P

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref absent")

DTYPES = [(np.float32, False), (np.float64, True)]


def write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)
    return str(path)


def random_csr(rng, n, m=None, density=0.05, dtype=np.float64, dup=False, sym=False):
    m = n if m is None else m
    mask = rng.random((n, m)) < density
    if sym:
        mask = mask | mask.T
    r, c = np.nonzero(mask)
    if dup and r.size:
        k = rng.integers(0, r.size, size=max(1, r.size // 5))
        r = np.concatenate([r, r[k]])
        c = np.concatenate([c, c[k]])
    v = rng.integers(-50, 50, size=r.size).astype(dtype) / dtype(4)
    return P.convert_coo(n, m, r, c, v)


# ---------------------------------------------------------------------------------------------------------
# MatrixMarket
# ---------------------------------------------------------------------------------------------------------
GOOD_MTX = {
    "general_real": "%%MatrixMarket matrix coordinate real general\n% comment\n3 4 4\n1 1 1.5\n3 4 -2\n2 2 1e3\n1 3 .25\n",
    "symmetric": "%%MatrixMarket matrix coordinate real symmetric\n4 4 5\n1 1 2\n2 1 -1\n3 2 -1\n4 3 -1.5e-2\n4 4 7\n",
    "pattern_sym": "%%MatrixMarket matrix coordinate pattern symmetric\n5 5 4\n2 1\n3 1\n5 4\n3 3\n",
    "integer_general": "%%MatrixMarket matrix coordinate integer general\n2 2 3\n1 2 7\n2 1 -3\n2 2 +4\n",
    "complex_hermitian": "%%MatrixMarket matrix coordinate complex Hermitian\n3 3 3\n1 1 2.0 0.0\n2 1 1.0 -1.0\n3 2 0.5 0.5\n",
    "crlf_tabs_blank": "%%MatrixMarket matrix coordinate real general\r\n%c\r\n3 3 3\r\n\t1\t1\t1.0\r\n\r\n   \r\n2 3 2.0\r\n3 1 3.0 trailing words\r\n",
    "fewer_entries": "%%MatrixMarket matrix coordinate real general\n4 4 9\n1 1 1\n4 4 2\n",
    "unsorted_dups_free": "%%MatrixMarket matrix coordinate real general\n3 3 5\n3 3 1\n1 2 2\n3 1 3\n1 1 4\n2 2 5\n",
    "float_rounding": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 0.1\n2 2 3.14159265358979323846\n",
    "no_trailing_newline": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n2 2 2",
    "empty_matrix": "%%MatrixMarket matrix coordinate real general\n3 3 0\n",
}

BAD_MTX_REFERENCE_REJECTS = {
    "array_form": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "unknown_type": "%%MatrixMarket matrix coordinate quaternion general\n1 1 1\n1 1 1\n",
    "skew": "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n",
    "lowercase_hermitian": "%%MatrixMarket matrix coordinate complex hermitian\n2 2 1\n2 1 1 0\n",
    "row_out_of_bounds": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
    "col_out_of_bounds": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 1\n",
    "missing_value": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n",
    "bad_size_line": "%%MatrixMarket matrix coordinate real general\nthree 3 3\n",
    "blank_before_size": "%%MatrixMarket matrix coordinate real general\n\n2 2 1\n1 1 1\n",
    "not_a_number": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 x 1\n",
}

# Inputs the reference reads past its buffers or with undefined behaviour; rejected here.
BAD_MTX_OURS_ONLY = {
    "zero_index": "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n",
    "too_many_entries": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1\n2 2 1\n",
    "short_banner": "%%MatrixMarket matrix coordinate real\n2 2 1\n1 1 1\n",
    "no_size_line": "%%MatrixMarket matrix coordinate real general\n% only comments\n",
    "negative_index": "%%MatrixMarket matrix coordinate real general\n2 2 1\n-1 1 1\n",
}


@needs_ref
@pytest.mark.parametrize("name", sorted(GOOD_MTX))
@pytest.mark.parametrize("dtype,is_double", DTYPES)
def test_mtx_load_matches_reference(tmp_path, name, dtype, is_double):
    path = write(tmp_path / f"{name}.mtx", GOOD_MTX[name])
    ours, info = P.loadMTX(path, dtype)
    ref = O.ref_mtx_load(path, is_double)
    assert ref is not None, "the reference rejected a well-formed file"
    rows, cols, rp, ci, vals, rinfo = ref
    assert (ours.rows, ours.cols) == (rows, cols)
    assert info == rinfo
    np.testing.assert_array_equal(ours.row_offsets, rp)
    np.testing.assert_array_equal(ours.col_ids, ci)
    assert ours.data.dtype == dtype
    np.testing.assert_array_equal(ours.data.astype(np.float64), vals)


@pytest.mark.parametrize("name", sorted(BAD_MTX_REFERENCE_REJECTS))
def test_mtx_errors_as_reference(tmp_path, name):
    path = write(tmp_path / f"{name}.mtx", BAD_MTX_REFERENCE_REJECTS[name])
    with pytest.raises(P.RcmError) as e:
        P.loadMTX(path)
    assert e.value.status == P.RCM_EINVAL
    if O.ref_available():
        assert O.ref_mtx_load(path) is None, "the reference accepts this file"


@pytest.mark.parametrize("name", sorted(BAD_MTX_OURS_ONLY))
def test_mtx_rejects_undefined_inputs(tmp_path, name):
    path = write(tmp_path / f"{name}.mtx", BAD_MTX_OURS_ONLY[name])
    with pytest.raises(P.RcmError) as e:
        P.loadMTX(path)
    assert e.value.status == P.RCM_EINVAL


def test_mtx_missing_file_is_eio(tmp_path):
    with pytest.raises(P.RcmError) as e:
        P.loadMTX(tmp_path / "absent.mtx")
    assert e.value.status == P.RCM_EIO


def test_mtx_error_names_the_line(tmp_path):
    path = write(tmp_path / "bad.mtx", "%%MatrixMarket matrix coordinate real general\n% c\n2 2 2\n1 1 1\n2 9 1\n")
    with pytest.raises(P.RcmError, match="line 5"):
        P.loadMTX(path)


def test_mtx_symmetric_expansion_closed_form(tmp_path):
    # symmetric lower triangle of a path graph: every off-diagonal entry is mirrored, diagonals are not
    n = 6
    lines = [f"{i + 1} {i + 1} 2" for i in range(n)] + [f"{i + 2} {i + 1} -1" for i in range(n - 1)]
    path = write(tmp_path / "path.mtx",
                 f"%%MatrixMarket matrix coordinate real symmetric\n{n} {n} {len(lines)}\n" + "\n".join(lines) + "\n")
    m, info = P.loadMTX(path, np.float64)
    assert info == P.MTX_SYMMETRIC
    dense = np.zeros((n, n))
    for r in range(n):
        for k in range(m.row_offsets[r], m.row_offsets[r + 1]):
            dense[r, m.col_ids[k]] += m.data[k]
    expect = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    np.testing.assert_array_equal(dense, expect)
    assert m.nnz == 3 * n - 2


def test_mtx_large_parallel_parse_matches_serial_semantics(tmp_path):
    # big enough for many parse chunks; compared with a numpy restatement of loadMTX + convert
    rng = np.random.default_rng(5)
    n, k = 3000, 400_000
    r = rng.integers(1, n + 1, k)
    c = rng.integers(1, n + 1, k)
    v = rng.integers(-999, 999, k) / 8.0
    body = "".join(f"{a} {b} {x}\n" for a, b, x in zip(r, c, v))
    path = write(tmp_path / "big.mtx", f"%%MatrixMarket matrix coordinate real general\n{n} {n} {k}\n" + body)
    m, _ = P.loadMTX(path, np.float64)
    order = np.lexsort((np.arange(k), c, r))  # stable on (row, col), duplicates in file order
    np.testing.assert_array_equal(m.col_ids, c[order] - 1)
    np.testing.assert_array_equal(m.data, v[order])
    np.testing.assert_array_equal(m.row_offsets, np.concatenate([[0], np.cumsum(np.bincount(r - 1, minlength=n))]))


@needs_ref
@pytest.mark.parametrize("dtype,is_double", DTYPES)
def test_mtx_store_bytes_match_reference(tmp_path, dtype, is_double):
    rng = np.random.default_rng(11)
    m = random_csr(rng, 60, 45, density=0.1)
    m.data = m.data * 1.37 - 0.001  # values that need rounding
    ours = tmp_path / "ours.mtx"
    refp = tmp_path / "ref.mtx"
    P.storeMTX(m, ours, dtype)
    assert O.ref_mtx_store(str(refp), m.rows, m.cols, m.row_offsets, m.col_ids, m.data, is_double) == 0
    assert ours.read_bytes() == refp.read_bytes()


def test_mtx_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    m = random_csr(rng, 80, density=0.08)
    m.data = np.round(m.data, 3)
    P.storeMTX(m, tmp_path / "m.mtx", np.float64)
    back, info = P.loadMTX(tmp_path / "m.mtx", np.float64)
    assert info == 0
    np.testing.assert_array_equal(back.row_offsets, m.row_offsets)
    np.testing.assert_array_equal(back.col_ids, m.col_ids)
    np.testing.assert_allclose(back.data, m.data, atol=5e-7)  # "%f" keeps 6 decimals


# ---------------------------------------------------------------------------------------------------------
# .hicsr
# ---------------------------------------------------------------------------------------------------------
def hicsr_layout(m, value_bytes):
    """Expected bytes of storeCSR from the reference's struct definitions (src/CSR.cpp:44-104, :153-168):
    CSRIOHeader = char[9] magic + 7 padding bytes + 8 uint64; State<T> = T scaling, bool transpose, padding."""
    head = b"Hi\x01Compsd" + b"\0" * 7
    head += struct.pack("<8Q", value_bytes, 0, 4, 0, 4, m.rows, m.cols, m.nnz)
    if value_bytes == 4:
        state = struct.pack("<f?3x", 1.0, False)
        data = np.asarray(m.data, np.float32).tobytes()
    else:
        state = struct.pack("<d?7x", 1.0, False)
        data = np.asarray(m.data, np.float64).tobytes()
    return head + state + data + m.col_ids.astype("<u4").tobytes() + m.row_offsets.astype("<u4").tobytes()


@pytest.mark.parametrize("dtype,vb", [(np.float32, 4), (np.float64, 8)])
def test_hicsr_layout(tmp_path, dtype, vb):
    rng = np.random.default_rng(2)
    m = random_csr(rng, 50, 70, density=0.1)
    P.storeCSR(m, tmp_path / "m.hicsr", dtype)
    blob = (tmp_path / "m.hicsr").read_bytes()
    assert len(blob) == 80 + (8 if vb == 4 else 16) + vb * m.nnz + 4 * m.nnz + 4 * (m.rows + 1)
    assert blob == hicsr_layout(m, vb)
    back = P.loadCSR(tmp_path / "m.hicsr", dtype)
    np.testing.assert_array_equal(back.row_offsets, m.row_offsets)
    np.testing.assert_array_equal(back.col_ids, m.col_ids)
    np.testing.assert_array_equal(back.data, m.data.astype(dtype))
    assert back.cols == m.cols


@needs_ref
@pytest.mark.parametrize("dtype,is_double", DTYPES)
def test_hicsr_matches_reference_both_ways(tmp_path, dtype, is_double):
    rng = np.random.default_rng(9)
    m = random_csr(rng, 120, 90, density=0.07)
    vb = 8 if is_double else 4
    # reference writes, we read; mask the bytes the reference leaves uninitialised (struct padding)
    refp = str(tmp_path / "ref.hicsr")
    assert O.ref_hicsr_store(refp, m.rows, m.cols, m.row_offsets, m.col_ids, m.data, is_double) == 0
    ref_blob = bytearray(open(refp, "rb").read())
    ours_blob = bytearray(hicsr_layout(m, vb))
    pad = list(range(9, 16)) + list(range(80 + vb + 1, 80 + 2 * vb))
    for i in pad:
        ref_blob[i] = ours_blob[i] = 0
    assert ref_blob == ours_blob
    back = P.loadCSR(refp, dtype)
    np.testing.assert_array_equal(back.col_ids, m.col_ids)
    np.testing.assert_array_equal(back.data, m.data.astype(dtype))
    # we write, the reference reads
    ours = str(tmp_path / "ours.hicsr")
    P.storeCSR(m, ours, dtype)
    rows, cols, rp, ci, vals = O.ref_hicsr_load(ours, is_double)
    assert (rows, cols) == (m.rows, m.cols)
    np.testing.assert_array_equal(rp, m.row_offsets)
    np.testing.assert_array_equal(ci, m.col_ids)
    np.testing.assert_array_equal(vals, m.data.astype(dtype).astype(np.float64))


def test_hicsr_errors(tmp_path):
    rng = np.random.default_rng(1)
    m = random_csr(rng, 30, density=0.2)
    good = tmp_path / "m.hicsr"
    P.storeCSR(m, good, np.float32)
    blob = good.read_bytes()
    cases = {
        "bad_magic": b"Hx" + blob[2:],
        "truncated_header": blob[:40],
        "truncated_data": blob[:-3],
        "bad_index_size": blob[:32] + struct.pack("<Q", 8) + blob[40:],
    }
    for name, b in cases.items():
        p = tmp_path / f"{name}.hicsr"
        p.write_bytes(b)
        with pytest.raises(P.RcmError) as e:
            P.loadCSR(p, np.float32)
        assert e.value.status == P.RCM_EINVAL, name
        if O.ref_available() and name != "bad_index_size":  # the reference does not check index sizes
            assert O.ref_hicsr_load(str(p)) is None, name
    # a float file read as double: the reference's "matching type" check
    with pytest.raises(P.RcmError) as e:
        P.loadCSR(good, np.float64)
    assert e.value.status == P.RCM_EINVAL
    if O.ref_available():
        assert O.ref_hicsr_load(str(good), True) is None
    with pytest.raises(P.RcmError) as e:
        P.loadCSR(tmp_path / "absent.hicsr")
    assert e.value.status == P.RCM_EIO


# ---------------------------------------------------------------------------------------------------------
# COO -> CSR and prepareRCM with values
# ---------------------------------------------------------------------------------------------------------
def test_convert_is_stable_lexicographic():
    rng = np.random.default_rng(4)
    n, k = 200, 5000
    r = rng.integers(0, n, k)
    c = rng.integers(0, 150, k)
    v = np.arange(k, dtype=np.float64)  # distinct values expose any reordering of duplicates
    m = P.convert_coo(n, 150, r, c, v)
    order = np.lexsort((np.arange(k), c, r))
    np.testing.assert_array_equal(m.col_ids, c[order])
    np.testing.assert_array_equal(m.data, v[order])
    np.testing.assert_array_equal(m.row_offsets, np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))]))
    with pytest.raises(P.RcmError):
        P.convert_coo(3, 3, [0, 3], [0, 0], None)


def dense_sum(m, n):
    d = np.zeros((n, n))
    for r in range(m.rows):
        for k in range(m.row_offsets[r], m.row_offsets[r + 1]):
            d[r, m.col_ids[k]] += m.data[k]
    return d


@needs_ref
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("check_sym", [True, False])
def test_prepare_values_matches_reference(seed, check_sym):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20, 120))
    m = random_csr(rng, n, density=float(rng.uniform(0.005, 0.05)), sym=not check_sym)
    # every row non-empty on the diagonal so the reference's repair loop terminates
    rr = np.concatenate([np.repeat(np.arange(n), np.diff(m.row_offsets)), np.arange(n)])
    cc = np.concatenate([m.col_ids, np.arange(n)])
    vv = np.concatenate([m.data, np.full(n, 9.0)])
    keep = np.ones(rr.size, bool)
    _, first = np.unique(rr.astype(np.int64) * n + cc, return_index=True)
    keep[:] = False
    keep[first] = True  # no duplicate entries: the reference's unstable sort then has a unique result
    m = P.convert_coo(n, n, rr[keep], cc[keep], vv[keep])
    ours, links = P.prepare_rcm(m, check_sym)
    rows, cols, rp, ci, vals = O.ref_prepare_values(m.rows, m.cols, m.row_offsets, m.col_ids, m.data, check_sym,
                                                    True)
    assert (ours.rows, ours.cols) == (rows, cols)
    np.testing.assert_array_equal(ours.row_offsets, rp)
    np.testing.assert_array_equal(ours.col_ids, ci)
    np.testing.assert_array_equal(ours.data, vals)
    # pattern part agrees with the pattern-only host repair as well
    pat = P.Graph(P.CSR(m.row_offsets.copy(), m.col_ids.copy()))
    added = pat.prepareRCM(check_sym)
    assert added == links
    np.testing.assert_array_equal(pat.csr.col_ids, ours.col_ids)


def test_prepare_values_closed_form():
    # two 2-node components {0,1} and {2,3}: one link (1, 2) with value 1 both ways
    m = P.convert_coo(4, 4, [0, 0, 1, 2, 2, 3], [0, 1, 0, 2, 3, 2], [5.0, -1, -1, 5, -2, -2])
    out, links = P.prepare_rcm(m, True)
    assert links == 1
    expect = np.array([[5, -1, 0, 0], [-1, 0, 1, 0], [0, 1, 5, -2], [0, 0, -2, 0]], float)
    np.testing.assert_array_equal(dense_sum(out, 4), expect)
    # a one-sided entry gets its transpose with the same value
    m = P.convert_coo(3, 3, [0, 1, 2, 2], [0, 1, 2, 0], [1.0, 1, 1, 4.5])
    out, links = P.prepare_rcm(m, True)
    d = dense_sum(out, 3)
    assert d[0, 2] == d[2, 0] == 4.5
    assert links == 1  # node 1 is its own component: linked to node 0


# ---------------------------------------------------------------------------------------------------------
# Graph<T>::load
# ---------------------------------------------------------------------------------------------------------
def disconnected_mtx(symmetric: bool):
    # three components: {1,2,3} {4,5} {6}; values distinct; symmetric banner lists the lower triangle
    ent = [(1, 1, 4), (2, 1, -1), (3, 2, -1.5), (2, 2, 4), (3, 3, 4), (4, 4, 3), (5, 4, -0.5), (5, 5, 3), (6, 6, 2)]
    if not symmetric:
        ent = ent + [(1, 2, -1), (2, 3, -1.5)]  # (5, 4) stays one-sided: prepareRCM(true) adds (4, 5)
    kind = "symmetric" if symmetric else "general"
    return f"%%MatrixMarket matrix coordinate real {kind}\n6 6 {len(ent)}\n" + "".join(f"{a} {b} {v}\n" for a, b, v in ent)


@needs_ref
@pytest.mark.parametrize("symmetric", [True, False])
@pytest.mark.parametrize("dtype,is_double", DTYPES)
def test_graph_load_matches_reference(tmp_path, symmetric, dtype, is_double):
    text = disconnected_mtx(symmetric)
    (tmp_path / "ours").mkdir()
    (tmp_path / "ref").mkdir()
    ours_file = write(tmp_path / "ours" / "g.mtx", text)
    ref_file = write(tmp_path / "ref" / "g.mtx", text)
    g = P.Graph.load(ours_file, dtype)
    assert g.source == P.SOURCE_MTX and g.links == 2
    ret, ref = O.ref_graph_load(ref_file, is_double)
    assert ret == 0
    rows, cols, rp, ci, vals = ref
    np.testing.assert_array_equal(g.csr.row_offsets, rp)
    np.testing.assert_array_equal(g.csr.col_ids, ci)
    np.testing.assert_array_equal(g.csr.data.astype(np.float64), vals)
    # the caches have the same content (padding bytes aside) and names
    ext = "d_" if is_double else ""
    vb = 8 if is_double else 4
    pad = list(range(9, 16)) + list(range(80 + vb + 1, 80 + 2 * vb))
    for suffix in (".hicsr", ".rcm.hicsr"):
        a = bytearray((tmp_path / "ours" / f"g.mtx{ext}{suffix}").read_bytes())
        b = bytearray((tmp_path / "ref" / f"g.mtx{ext}{suffix}").read_bytes())
        for i in pad:
            a[i] = b[i] = 0
        assert a == b, suffix
    # second load comes from the .rcm.hicsr cache, unchanged
    g2 = P.Graph.load(ours_file, dtype)
    assert g2.source == P.SOURCE_RCM_HICSR and g2.links == 0
    np.testing.assert_array_equal(g2.csr.col_ids, g.csr.col_ids)
    np.testing.assert_array_equal(g2.csr.data, g.csr.data)
    # with only the plain .hicsr present, it is read and prepared again
    os.remove(tmp_path / "ours" / f"g.mtx{ext}.rcm.hicsr")
    g3 = P.Graph.load(ours_file, dtype)
    assert g3.source == P.SOURCE_HICSR
    if symmetric:  # the .hicsr path prepares with check_sym = true; the symmetric matrix needs nothing new
        np.testing.assert_array_equal(g3.csr.col_ids, g.csr.col_ids)


def test_graph_load_errors(tmp_path):
    with pytest.raises(P.RcmError) as e:
        P.Graph.load(tmp_path / "nothing.mtx")
    assert e.value.status == P.RCM_EIO
    # a one-sided entry makes a 2 x 3 matrix square through symmetrisation (max(rows, cols) = 3) ...
    p = write(tmp_path / "rect.mtx", "%%MatrixMarket matrix coordinate real general\n2 3 2\n1 1 1\n2 3 1\n")
    g = P.Graph.load(p, write_cache=False)
    assert (g.csr.rows, g.csr.cols) == (3, 3)
    assert sorted(os.listdir(tmp_path)) == ["rect.mtx"]  # no caches written without write_cache
    # ... but with nothing to mirror it stays 2 x 3: the reference's "Matrix must be symmetric!" return
    p = write(tmp_path / "diag.mtx", "%%MatrixMarket matrix coordinate real general\n2 3 2\n1 1 1\n2 2 1\n")
    with pytest.raises(P.RcmError) as e:
        P.Graph.load(p, write_cache=False)
    assert e.value.status == P.RCM_EINVAL
    if O.ref_available():
        ref_dir = tmp_path / "ref"
        ref_dir.mkdir()
        for name in ("rect.mtx", "diag.mtx"):
            shutil.copy(tmp_path / name, ref_dir / name)
        ret, ref = O.ref_graph_load(str(ref_dir / "rect.mtx"))
        assert ret == 0 and (ref[0], ref[1]) == (3, 3)
        np.testing.assert_array_equal(ref[3], g.csr.col_ids)
        ret, _ = O.ref_graph_load(str(ref_dir / "diag.mtx"))
        assert ret == 1


def test_graph_load_then_engine_inputs(tmp_path):
    # a loaded graph is a valid engine input: symmetric, connected, sorted columns
    rng = np.random.default_rng(8)
    m = random_csr(rng, 300, density=0.01)
    n = m.rows
    rr = np.concatenate([np.repeat(np.arange(n), np.diff(m.row_offsets)), np.arange(n)])
    m = P.convert_coo(n, n, rr, np.concatenate([m.col_ids, np.arange(n)]), np.ones(rr.size))  # full diagonal
    P.storeMTX(m, tmp_path / "r.mtx", np.float64)
    g = P.Graph.load(tmp_path / "r.mtx", np.float64)
    n = g.csr.rows
    rp, ci = g.csr.row_offsets, g.csr.col_ids
    rows = np.repeat(np.arange(n), np.diff(rp))
    pairs = set(zip(rows.tolist(), ci.tolist()))
    assert all((c, r) in pairs for r, c in pairs)
    order, _ = O.bfs(rp, ci, int(rows[0]))
    assert len(order) == n
    for r in range(n):
        seg = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(seg) >= 0)
