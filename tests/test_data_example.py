"""The reference's own example dataset end to end (data/ of PLSQR3.2013.09).

100 real sensitivity-kernel rows (signed values around 1e-6, about 12,700
nonzeros per row) over 302,940 columns = 99 x 153 x 10 cells x 2
parameters; the damping of data/README step (5): identity, xx, yy, zz and
the xy, xz, yz cross rows for both parameters, all weights 1.0; b from
measure_fake.list; 100 iterations (data/LSQR_run/LSQR3.pbs:10).

The fixtures (tests/golden/make_data_example.py) hold the dataset and the
sha256 digests of what the reference's own programs made of it: the row
files of ker_ascii2bin, the column-sorted matrix of Ker2PLSQR3, and PLSQR3's
x after 100 iterations.

CPU: the fixture decodes to the dataset byte for byte; plsqr_ker_ascii2bin
writes the reference tool's row files byte for byte; the damping has the
dataset's own row and nonzero counts; the oracle reproduces the reference's
x bit for bit.  GPU: plsqr_ker2plsqr3 writes the reference tool's matrix
files byte for byte, the exact-order solve gives the reference's x bit for
bit, and the default operator stays within the oracle's own sensitivity.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, rel
import data_example as DX

M_D, NNZ_D = 1910672, 5895716  # data/load_balance.sh:1


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


@pytest.fixture(scope="module")
def inputs():
    with np.load(os.path.join(GOLDEN, "data_example_inputs.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def golden():
    return load_golden("data_example_itn100")


@pytest.fixture(scope="module")
def files(P, inputs, tmp_path_factory):
    """The dataset decoded, and the row files plsqr_ker_ascii2bin makes of it."""
    d = str(tmp_path_factory.mktemp("data_example"))
    adir = DX.decode(inputs, d)
    kb = os.path.join(d, "ker_bin")
    os.makedirs(kb)
    P.ker_ascii2bin(os.path.join(d, "ker.list"), adir, kb, "ker_bin.list")
    return d, adir, kb


def system(P, inputs):
    """Kernel CSC from the rows (rows ascending in each column) + the damping."""
    import scipy.sparse as sp
    ptr, cols, vals = DX.rows(inputs)
    A = sp.csr_matrix((vals, cols - 1, ptr), shape=(len(ptr) - 1, DX.N_COL)).tocsc()
    A.sort_indices()
    d = P.gen_damping(*DX.GRID, n_param=DX.N_PARAM, weights=(1.0,) * 7, cross=True)
    return P.System.from_arrays(A.shape[0], d.m_d, DX.N_COL, A.indptr, A.indices, A.data, d.d_row_ptr,
                                d.d_col_idx, d.d_val, d.dc_col_ptr, d.dc_row_idx, d.dc_val, DX.b_values(inputs),
                                grid=(DX.GRID[0], DX.GRID[1], DX.GRID[2] * DX.N_PARAM))


def test_fixture_is_the_dataset(inputs, files):
    d, adir, _ = files
    got = [sha(os.path.join(adir, str(n))) for n in inputs["names"]]
    assert got == [str(h) for h in inputs["text_sha256"]]
    assert len(got) == 100 and int(inputs["ptr"][-1]) == 1266788  # data/load_balance.sh:1


def test_ascii_rows_become_the_reference_tools_row_files(inputs, golden, files):
    _, _, kb = files
    got = [sha(os.path.join(kb, str(n) + ".bin")) for n in inputs["names"]]
    assert got == [str(h) for h in golden["a2b_sha256"]]
    lines = open(os.path.join(kb, "ker_bin.list")).read().split("\n")
    ptr = inputs["ptr"]
    assert lines[0] == f"{inputs['names'][0]}.bin {ptr[1] - ptr[0]}"  # as data/ker_bin.list


def test_ascii_reader_rejects_a_broken_entry(P, tmp_path):
    (tmp_path / "a").mkdir()
    (tmp_path / "a" / "r0").write_text("12 1 1 1 1.5e-06\n13 2 1 1 oops\n")
    (tmp_path / "list").write_text("1\nr0\n")
    with pytest.raises(P.PlsqrError) as e:
        P.ker_ascii2bin(str(tmp_path / "list"), str(tmp_path / "a"), str(tmp_path))
    assert e.value.code == 7 and "entry 2" in str(e.value)
    (tmp_path / "list").write_text("2\nr0\n")
    with pytest.raises(P.PlsqrError) as e:
        P.ker_ascii2bin(str(tmp_path / "list"), str(tmp_path / "a"), str(tmp_path))
    assert e.value.code == 7


def test_damping_has_the_datasets_counts(P):
    d = P.gen_damping(*DX.GRID, n_param=DX.N_PARAM, weights=(1.0,) * 7, cross=True)
    assert (d.m_d, d.nnz_d, d.n) == (M_D, NNZ_D, DX.N_COL)
    rp = d.d_row_ptr
    assert rp[-1] == NNZ_D and set(np.diff(rp).tolist()) == {1, 3, 4}


def test_oracle_reproduces_the_reference_x(P, O, inputs, golden):
    s = system(P, inputs)
    assert (s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d) == tuple(golden["dims"])
    x, _, _ = O.Oracle(s).lsqr(s.b, 100, trace=False)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(golden["x_sha256"]), \
        rel(x[::997], golden["x_sample"])


@pytest.mark.gpu
def test_gpu_relayout_writes_the_reference_tools_matrix(P, golden, files, gpu_available):
    _, _, kb = files
    P.ker2plsqr3(kb, "ker_bin.list", DX.N_COL, "matrix_bycol.mm2.bin", "matrix_bycol.mm2.info")
    assert sha(os.path.join(kb, "matrix_bycol.mm2.bin")) == str(golden["ker_bin_sha256"])
    assert sha(os.path.join(kb, "matrix_bycol.mm2.info")) == str(golden["ker_info_sha256"])


@pytest.mark.gpu
def test_gpu_solves_the_dataset_as_the_reference(P, O, inputs, golden, files, gpu_available):
    """From the files: the matrix written by plsqr_ker2plsqr3, the damping
    files of plsqr_write_damping_files and measure_fake.list, loaded by
    plsqr_read_reference_files as the reference's CLI names them."""
    d, _, kb = files
    if not os.path.exists(os.path.join(kb, "matrix_bycol.mm2.bin")):
        P.ker2plsqr3(kb, "ker_bin.list", DX.N_COL, "matrix_bycol.mm2.bin", "matrix_bycol.mm2.info")
    dn = P.write_damping_files(kb, *DX.GRID, n_param=DX.N_PARAM, weights=(1.0,) * 7, cross=True)
    with open(os.path.join(kb, "measure_fake.list"), "w") as f:
        f.write(str(inputs["b_text"]))
    s = P.System.read_reference_files(kb, dict(ker_f="matrix_bycol.mm2.bin", ker_i="matrix_bycol.mm2.info",
                                               b_k="measure_fake.list", **dn), 100, M_D, DX.N_COL)
    assert (s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d) == tuple(golden["dims"])
    with P.Operator(s, P.OP_EXACT_ORDER) as op:
        x, info = op.lsqr(s.b, 100)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(golden["x_sha256"]), \
        rel(x[::997], golden["x_sample"])
    # the default operator: the columns bricked as a 99 x 153 x 20 grid (the
    # two parameter blocks stacked in z)
    orc = O.Oracle(s)
    xo, _, _ = orc.lsqr(s.b, 100, trace=False)
    assert np.array_equal(xo, x)
    b2 = s.b * (1 + 1e-15 * np.random.default_rng(1).standard_normal(s.b.size))
    xp, _, _ = orc.lsqr(b2, 100, trace=False)
    sens = rel(xp, xo)
    for grid in ((99, 153, 20), None):
        with P.Operator(s, grid=grid) as op:
            xg, ig = op.lsqr(s.b, 100)
        print(f"default operator (grid {grid}): x vs reference {rel(xg, xo):.3e}, oracle sensitivity {sens:.3e}",
              flush=True)
        assert rel(xg, xo) <= max(1e-10, 100 * sens)
