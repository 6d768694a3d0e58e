"""Generator, reference file formats and the C-ABI library on the CPU (no GPU calls)."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def damping_counts(nx, ny, nz):
    """Closed form: identity rows plus one second-difference row per interior
    triple along each axis (damping_binary.py:50 without cross terms)."""
    rows = nx * ny * nz + (nx - 2) * ny * nz + nx * (ny - 2) * nz + nx * ny * (nz - 2)
    nnz = nx * ny * nz + 3 * ((nx - 2) * ny * nz + nx * (ny - 2) * nz + nx * ny * (nz - 2))
    return rows, nnz


def test_library_exports_every_declared_symbol(P):
    lib = P.load()
    header = open(os.path.join(ROOT, "include", "plsqr_hip.h")).read()
    declared = set(re.findall(r"\b(plsqr_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(P.SIGNATURES), declared ^ set(P.SIGNATURES)


@pytest.mark.parametrize("grid", [(16, 16, 16), (67, 67, 67), (7, 5, 3), (3, 3, 3), (20, 31, 48)])
def test_damping_closed_form_counts(P, grid):
    nx, ny, nz = grid
    if nx * ny * nz > 100000:
        rows, nnz = damping_counts(nx, ny, nz)
        assert (rows, nnz) == (1176118, 2926828)  # the 67^3 sizes quoted for the 2M-ray workload
        return
    s = P.generate(nx, ny, nz, 4, 0, 1, 0.0, 0.1)
    assert (s.m_d, s.nnz_d) == damping_counts(nx, ny, nz)


def test_damping_row_order_and_values(P):
    nx, ny, nz = 5, 4, 3
    s = P.generate(nx, ny, nz, 4, 0, 1, 0.0, 0.25, w=2.0)
    rp, ci, v = s.d_row_ptr, s.d_col_idx, s.d_val
    # cell (0,0,0): identity, xx, yy, zz in that order (damping_binary.py:84-289)
    assert list(ci[rp[0]:rp[1]]) == [0] and v[rp[0]] == 0.25
    assert list(ci[rp[1]:rp[2]]) == [0, 1, 2] and list(v[rp[1]:rp[2]]) == [2.0, -4.0, 2.0]
    assert list(ci[rp[2]:rp[3]]) == [0, nx, 2 * nx]
    assert list(ci[rp[3]:rp[4]]) == [0, nx * ny, 2 * nx * ny]
    # cell (nx-2, 0, 0) has no xx row: identity then yy
    k = 0
    r = 0
    while k < nx - 2:  # walk to the cell
        r += 1 + 1 + 1 + 1  # identity, xx, yy, zz for interior cells of the first line
        k += 1
    assert list(ci[rp[r]:rp[r + 1]]) == [nx - 2]
    assert list(ci[rp[r + 1]:rp[r + 2]]) == [nx - 2, nx - 2 + nx, nx - 2 + 2 * nx]
    # column layout holds the same entries, rows ascending inside each column
    cp, ri, cv = s.dc_col_ptr, s.dc_row_idx, s.dc_val
    for c in range(s.n):
        rows = ri[cp[c]:cp[c + 1]]
        assert np.all(np.diff(rows) > 0)
        for j, rr in enumerate(rows):
            e = rp[rr] + list(ci[rp[rr]:rp[rr + 1]]).index(c)
            assert v[e] == cv[cp[c] + j]


def test_kernel_rows_are_ray_segments(P):
    s = P.generate(16, 16, 16, 2000, 0, 5, 0.0, 0.1)
    assert s.k_val.min() > 0.0 and s.k_val.max() <= np.sqrt(3.0) * (1 + 1e-12)
    cnt = np.bincount(s.k_row_idx, minlength=s.m_k)
    assert cnt.min() >= 16 and cnt.max() <= 48          # N .. 3N voxels between opposite faces
    length = np.bincount(s.k_row_idx, weights=s.k_val, minlength=s.m_k)
    assert length.min() >= 16.0 - 1e-9                  # spans the full grid along one axis
    assert length.max() <= 16.0 * np.sqrt(3.0) + 1e-9
    for c in range(0, s.n, 97):                         # rows ascend inside each column
        assert np.all(np.diff(s.k_row_idx[s.k_col_ptr[c]:s.k_col_ptr[c + 1]]) > 0)


def test_generator_independent_of_thread_count(P, monkeypatch):
    monkeypatch.setenv("PLSQR_HOST_THREADS", "1")
    a = P.generate(20, 20, 20, 30000, 1, 9, 0.02, 0.05)
    monkeypatch.setenv("PLSQR_HOST_THREADS", "7")
    b = P.generate(20, 20, 20, 30000, 1, 9, 0.02, 0.05)
    for k in ("k_col_ptr", "k_row_idx", "k_val", "b", "x_true"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_gen_rhs_matches_generate(P):
    s = P.generate(12, 12, 12, 3000, 0, 21, 0.03, 0.1)
    b, xt = P.gen_rhs(s, 12, 12, 12, 21, 0.03)
    assert np.array_equal(b, s.b) and np.array_equal(xt, s.x_true)
    b2, _ = P.gen_rhs(s, 12, 12, 12, 22, 0.03)
    assert not np.array_equal(b2, s.b)


def test_reference_files_round_trip(P, tmp_path):
    s = P.generate(9, 8, 7, 500, 1, 3, 0.01, 0.2)
    names = s.write_reference_files(str(tmp_path))
    t = P.System.read_reference_files(str(tmp_path), names, s.m_k, s.m_d, s.n)
    for k in ("k_col_ptr", "k_row_idx", "k_val", "d_row_ptr", "d_col_idx", "d_val", "dc_col_ptr",
              "dc_row_idx", "dc_val", "b"):
        assert np.array_equal(getattr(s, k), getattr(t, k)), k
    # .info: 1-based columns, start = 0 for empty columns (Ker2PLSQR3.c:505-511)
    lines = open(tmp_path / names["ker_i"]).read().split("\n")
    cnt = np.diff(s.k_col_ptr)
    start = 1
    for c in range(s.n):
        ci, k, st = map(int, lines[c].split())
        assert ci == c + 1 and k == cnt[c] and st == (0 if k == 0 else start)
        start += k
    assert (cnt == 0).any()  # few rays from the top: some columns are empty
    # 16-byte records, 1-based, column-major
    rec = np.fromfile(tmp_path / names["ker_f"], dtype=[("r", "<i4"), ("c", "<i4"), ("v", "<f8")])
    assert rec.size == s.nnz_k
    assert np.array_equal(rec["r"] - 1, s.k_row_idx) and np.array_equal(rec["v"], s.k_val)
    assert np.all(np.diff(rec["c"]) >= 0)


def test_x_file_round_trip(P, tmp_path):
    x = np.random.default_rng(0).standard_normal(1000)
    p = str(tmp_path / "p_resultx_itn=3_NProc=1_float64_mpiio.bin")
    P.write_x(p, x)
    assert os.path.getsize(p) == 8000
    assert np.array_equal(P.read_x(p, 1000), x)


def test_loader_errors_are_codes_not_exits(P, tmp_path):
    s = P.generate(5, 5, 5, 50, 0, 3, 0.01, 0.2)
    names = s.write_reference_files(str(tmp_path))
    # row index out of range (plsqr3.c:1606-1615 exits; here PLSQR_ERR_BOUNDS)
    rec = np.fromfile(tmp_path / names["ker_f"], dtype=[("r", "<i4"), ("c", "<i4"), ("v", "<f8")])
    rec["r"][3] = s.m_k + 5
    rec.tofile(tmp_path / "bad.bin")
    bad = dict(names, ker_f="bad.bin")
    with pytest.raises(P.PlsqrError) as e:
        P.System.read_reference_files(str(tmp_path), bad, s.m_k, s.m_d, s.n)
    assert e.value.code == 2
    # truncated kernel file
    rec[: s.nnz_k // 2].tofile(tmp_path / "short.bin")
    with pytest.raises(P.PlsqrError) as e:
        P.System.read_reference_files(str(tmp_path), dict(names, ker_f="short.bin"), s.m_k, s.m_d, s.n)
    assert e.value.code == 7
    # missing file
    with pytest.raises(P.PlsqrError) as e:
        P.System.read_reference_files(str(tmp_path), dict(names, b_k="nope.list"), s.m_k, s.m_d, s.n)
    assert e.value.code == 3
    # b shorter than row_k
    with pytest.raises(P.PlsqrError) as e:
        P.System.read_reference_files(str(tmp_path), names, s.m_k + 3, s.m_d, s.n)
    assert e.value.code in (2, 7)


def test_operator_rejects_bad_indices_before_touching_the_gpu(P):
    """Host validation runs first, so this error path needs no device."""
    s = P.generate(5, 5, 5, 50, 0, 3, 0.01, 0.2)
    ri = s.k_row_idx.copy()
    ri[7] = s.m_k  # out of range
    bad = P.System.from_arrays(s.m_k, s.m_d, s.n, s.k_col_ptr, ri, s.k_val, s.d_row_ptr,
                               s.d_col_idx, s.d_val, s.dc_col_ptr, s.dc_row_idx, s.dc_val, s.b)
    with pytest.raises(P.PlsqrError) as e:
        P.Operator(bad)
    assert e.value.code == 2 and "out of range" in str(e.value)
    val = s.k_val.copy()
    val[0] = np.nan
    bad = P.System.from_arrays(s.m_k, s.m_d, s.n, s.k_col_ptr, s.k_row_idx, val, s.d_row_ptr,
                               s.d_col_idx, s.d_val, s.dc_col_ptr, s.dc_row_idx, s.dc_val, s.b)
    with pytest.raises(P.PlsqrError) as e:
        P.Operator(bad)
    assert e.value.code == 8
    cp = s.k_col_ptr.copy()
    cp[3], cp[4] = cp[4], cp[3]
    bad = P.System.from_arrays(s.m_k, s.m_d, s.n, cp, s.k_row_idx, s.k_val, s.d_row_ptr,
                               s.d_col_idx, s.d_val, s.dc_col_ptr, s.dc_row_idx, s.dc_val, s.b)
    with pytest.raises(P.PlsqrError) as e:
        P.Operator(bad)
    assert e.value.code in (2, 7)


@pytest.mark.parametrize("segwin", ["1", "64", "512"])
def test_tiled_layout_self_check_on_host(P, monkeypatch, segwin):
    """The bricked mode-2 layout is built on the host before any device call;
    with PLSQR_TILE_CHECK=1 its per-slice column sums are compared with the
    CSC's there, and the mode-1 layout's ray sums (segments taken in ray
    order, or sorted by length inside windows of 64 or 512) with the CSC's
    row sums.  Without a GPU the creation then stops at hipGetDevice, so a
    HIP error (not a self-check error) means the layout passed."""
    monkeypatch.setenv("PLSQR_TILE_CHECK", "1")
    monkeypatch.setenv("PLSQR_SEGWIN", segwin)
    for args in ((16, 16, 16, 20000, 0, 2013, 0.01, 0.1), (9, 7, 20, 3000, 1, 4, 0.02, 0.05)):
        s = P.generate(*args)
        for grid in ("auto", None):
            try:
                op = P.Operator(s, grid=grid)
            except P.PlsqrError as e:
                assert e.code == 4 and "self-check" not in str(e), str(e)
            else:  # a GPU is present: the operator was built
                op.close()


def kernel_rows(s, seed):
    """The kernel of s as rows (Ker2PLSQR3's input): row pointers, 1-based
    columns shuffled inside every row, values."""
    import scipy.sparse as sp
    A = sp.csc_matrix((s.k_val, s.k_row_idx, s.k_col_ptr), shape=(s.m_k, s.n)).tocsr()
    rng = np.random.default_rng(seed)
    cols, vals = A.indices.copy(), A.data.copy()
    for r in range(s.m_k):
        b, e = A.indptr[r], A.indptr[r + 1]
        p = rng.permutation(e - b)
        cols[b:e], vals[b:e] = cols[b:e][p], vals[b:e][p]
    return A.indptr.astype(np.int32), (cols + 1).astype(np.int32), vals


def test_kernel_relayout_rejects_bad_rows_on_the_host(P):
    """Ker2PLSQR3's input checks run before any device call."""
    rp = np.array([0, 2, 3], np.int32)
    with pytest.raises(P.PlsqrError) as e:
        P.kernel_rows_to_csc(2, 4, rp, np.array([1, 5, 2], np.int32), np.ones(3))
    assert e.value.code == 2  # column 5 > n
    with pytest.raises(P.PlsqrError) as e:
        P.kernel_rows_to_csc(2, 4, np.array([0, 2, 1], np.int32), np.array([1, 2, 3], np.int32), np.ones(3))
    assert e.value.code == 7  # row_ptr decreases
    with pytest.raises(P.PlsqrError) as e:
        P.kernel_rows_to_csc(2, 4, rp, np.array([1, 2, 3], np.int32), np.array([1.0, np.nan, 2.0]))
    assert e.value.code == 8


def test_ker2plsqr3_missing_list_is_an_io_error(P, tmp_path):
    with pytest.raises(P.PlsqrError) as e:
        P.ker2plsqr3(str(tmp_path), "no_such.list", 10, "out.bin", "out.info")
    assert e.value.code == 3


@pytest.mark.gpu
def test_kernel_relayout_reproduces_the_csc(P, gpu_available):
    """Rows with their columns shuffled, re-laid on the GPU: the kernel CSC
    comes back bit for bit (rows ascending inside every column, as
    Ker2PLSQR3.c:595-606 writes them)."""
    print("kernel relayout", flush=True)
    for args, seed in (((16, 16, 16, 20000, 0, 2013, 0.01, 0.1), 1), ((9, 7, 20, 3000, 1, 4, 0.02, 0.05), 2)):
        s = P.generate(*args)
        rp, c1, v = kernel_rows(s, seed)
        cp, ri, vo = P.kernel_rows_to_csc(s.m_k, s.n, rp, c1, v)
        assert np.array_equal(cp, s.k_col_ptr) and np.array_equal(ri, s.k_row_idx) and np.array_equal(vo, s.k_val)


@pytest.mark.gpu
def test_ker2plsqr3_files_load_as_the_reference_system(P, gpu_available, tmp_path):
    """One row file per ray plus the list, through the tool, then the
    reference loader: the same system as the generator's."""
    print("ker2plsqr3 files", flush=True)
    s = P.generate(9, 7, 5, 600, 0, 21, 0.01, 0.1)
    rp, c1, v = kernel_rows(s, 3)
    lines = []
    for r in range(s.m_k):
        rec = np.zeros(rp[r + 1] - rp[r], dtype=[("aux", "<i4"), ("col", "<i4"), ("val", "<f8")])
        rec["aux"], rec["col"], rec["val"] = r + 1, c1[rp[r]:rp[r + 1]], v[rp[r]:rp[r + 1]]
        rec.tofile(str(tmp_path / f"row{r}.bin"))
        lines.append(f"row{r}.bin {rec.size}")
    (tmp_path / "rows.list").write_text("\n".join(lines) + "\n")
    names = s.write_reference_files(str(tmp_path))
    P.ker2plsqr3(str(tmp_path), "rows.list", s.n, "matrix_bycol.mm2.bin", "matrix_bycol.mm2.info")
    names2 = dict(names, ker_f="matrix_bycol.mm2.bin", ker_i="matrix_bycol.mm2.info")
    t = P.System.read_reference_files(str(tmp_path), names2, s.m_k, s.m_d, s.n)
    assert np.array_equal(t.k_col_ptr, s.k_col_ptr) and np.array_equal(t.k_row_idx, s.k_row_idx)
    assert np.array_equal(t.k_val, s.k_val)
    assert (tmp_path / "matrix_bycol.mm2.bin").read_bytes() == (tmp_path / names["ker_f"]).read_bytes()
    assert (tmp_path / "matrix_bycol.mm2.info").read_text() == (tmp_path / names["ker_i"]).read_text()


def test_damping_cross_rows_and_parameter_blocks(P):
    """damping_binary.py's full stencil: per cell identity, xx, yy, zz, then
    xy, xz, yz (0.5, -0.5, -0.5, 0.5 at the corners of a 3 x 3 square,
    :291-479), for every parameter block (block p on columns p*V.., :93-95,
    rows after block p-1's); the row counts of :50 and nonzeros of :61; the
    column layout the transpose of the row layout."""
    import scipy.sparse as sp
    nx, ny, nz, npar = 5, 4, 3, 2
    w = (0.3, 1.1, 1.2, 1.3, 2.1, 2.2, 2.3)
    d = P.gen_damping(nx, ny, nz, n_param=npar, weights=w, cross=True)
    V = nx * ny * nz
    per = (V + (nx - 2) * ny * nz + nx * (ny - 2) * nz + nx * ny * (nz - 2) + (nx - 2) * (ny - 2) * nz
           + (nx - 2) * ny * (nz - 2) + nx * (ny - 2) * (nz - 2))
    nnz = (V + 3 * ((nx - 2) * ny * nz + nx * (ny - 2) * nz + nx * ny * (nz - 2))
           + 4 * ((nx - 2) * (ny - 2) * nz + (nx - 2) * ny * (nz - 2) + nx * (ny - 2) * (nz - 2)))
    assert (d.m_d, d.nnz_d, d.n) == (npar * per, npar * nnz, npar * V)
    rp, ci, v = d.d_row_ptr, d.d_col_idx, d.d_val
    for p in range(npar):
        r = p * per  # cell (0,0,0) of block p: identity, xx, yy, zz, xy, xz, yz
        o = p * V
        rows = [(list(ci[rp[r + k]:rp[r + k + 1]] - o), list(v[rp[r + k]:rp[r + k + 1]])) for k in range(7)]
        assert rows[0] == ([0], [0.3])
        assert rows[1] == ([0, 1, 2], [1.1, -2.2, 1.1])
        assert rows[2] == ([0, nx, 2 * nx], [1.2, -2.4, 1.2])
        assert rows[3] == ([0, nx * ny, 2 * nx * ny], [1.3, -2.6, 1.3])
        assert rows[4] == ([0, 2, 2 * nx, 2 * nx + 2], [0.5 * 2.1, -0.5 * 2.1, -0.5 * 2.1, 0.5 * 2.1])
        assert rows[5] == ([0, 2, 2 * nx * ny, 2 * nx * ny + 2], [0.5 * 2.2, -0.5 * 2.2, -0.5 * 2.2, 0.5 * 2.2])
        assert rows[6] == ([0, 2 * nx, 2 * nx * ny, 2 * nx * ny + 2 * nx],
                           [0.5 * 2.3, -0.5 * 2.3, -0.5 * 2.3, 0.5 * 2.3])
        # the last cell of the block has only its identity row
        assert list(ci[rp[r + per - 1]:rp[r + per]]) == [o + V - 1]
    R = sp.csr_matrix((v, ci, rp), shape=(d.m_d, d.n))
    Cc = sp.csc_matrix((d.dc_val, d.dc_row_idx, d.dc_col_ptr), shape=(d.m_d, d.n))
    assert (R != Cc).nnz == 0
    for c in range(d.n):
        rows_c = d.dc_row_idx[d.dc_col_ptr[c]:d.dc_col_ptr[c + 1]]
        assert np.all(np.diff(rows_c) > 0)
    # without cross rows and with one block: the generator's own damping
    g = P.generate(nx, ny, nz, 4, 0, 1, 0.0, 0.3, w=1.1)
    e = P.gen_damping(nx, ny, nz, n_param=1, weights=(0.3, 1.1, 1.1, 1.1, 0, 0, 0), cross=False)
    for k in ("d_row_ptr", "d_col_idx", "d_val", "dc_col_ptr", "dc_row_idx", "dc_val"):
        assert np.array_equal(getattr(e, k), getattr(g, k)), k


def test_damping_files_round_trip(P, tmp_path):
    """The four files plsqr_write_damping_files writes load back through the
    reference loader (with a kernel) as the same matrix."""
    s = P.generate(5, 4, 3, 30, 0, 3, 0.01, 0.1)
    dn = P.write_damping_files(str(tmp_path), 5, 4, 3, n_param=1, weights=(0.1, 1, 1, 1, 0.5, 0.5, 0.5))
    names = s.write_reference_files(str(tmp_path), prefix="k_")
    d = P.gen_damping(5, 4, 3, n_param=1, weights=(0.1, 1, 1, 1, 0.5, 0.5, 0.5))
    r = P.System.read_reference_files(str(tmp_path), dict(names, **dn), s.m_k, d.m_d, s.n)
    for k in ("d_row_ptr", "d_col_idx", "d_val", "dc_col_ptr", "dc_row_idx", "dc_val"):
        assert np.array_equal(getattr(r, k), getattr(d, k)), k
    assert np.array_equal(r.k_val, s.k_val) and np.array_equal(r.b, s.b)
