"""Fixtures of the reference's example dataset (data/ of PLSQR3.2013.09).

Runs in the build container only: it reads /root/reference/data and runs the
reference's own programs compiled by oracle/build_ref.sh.  It writes

  data_example_inputs.npz   the dataset itself, compactly (tests/data_example.py)
  data_example_itn100.npz   what the reference's chain makes of it:
    a2b_sha256    sha256 of each row file ker_ascii2bin writes
                  (PLSQR3_tools/kernel_format/ker_ascii2bin.c, mpiexec -np 1)
    col_info_sha256, ker_bin_sha256, ker_info_sha256
                  Ker2PLSQR3_preprocess's col_info.txt and Ker2PLSQR3's
                  matrix_bycol.mm2.bin / .info (mpiexec -np 2, as
                  data/Ker2PLSQR3.pbs runs it) from those row files, with the
                  row-file list the dataset ships (data/ker_bin.list)
    x_sha256, x_norm, x_sample
                  PLSQR3 itself (FP64 build, mpiexec -np 1, -itn 100 as
                  data/LSQR_run/LSQR3.pbs:10) on that matrix, the damping
                  files of step (5) of data/README (damping_binary.py 1 1 damp
                  99 153 10 2 1.0 x 7) and measure_fake.list; x read from
                  p_resultx_itn=100_NProc=1_float64_mpiio.bin
    alpha, beta   the values the reference prints per iteration
    dims          m_k, m_d, n, nnz_k, nnz_d

damping_binary.py is Python 2 (no python2 here); its files are written by
this library's plsqr_write_damping_files, whose counts the test checks
against the dataset's own numbers (data/load_balance.sh:1: 1,910,672 rows,
5,895,716 nonzeros) and whose values (weights 1.0, 0.5, -2, -0.5) FP32
holds exactly, so the tool's FP32 column copy would be the same.

    python tests/golden/make_data_example.py
"""
import hashlib
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import data_example as DX  # noqa: E402
from make_golden import load_pkg, parse  # noqa: E402

DATA = "/root/reference/data"
REF = os.path.join(ROOT, "oracle", "_ref")
MPIEXEC = os.environ.get("MPIEXEC", "/opt/conda/bin/mpiexec")
ITN = 100


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r.stdout


def main():
    if not os.path.exists(os.path.join(REF, "Ker2PLSQR3")):
        subprocess.run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    P = load_pkg()
    inputs = DX.encode(DATA)
    np.savez_compressed(os.path.join(HERE, "data_example_inputs.npz"), **inputs)
    names = [str(n) for n in inputs["names"]]
    with tempfile.TemporaryDirectory() as d:
        kb = os.path.join(d, "ker_bin")
        os.makedirs(kb)
        run([MPIEXEC, "-np", "1", os.path.join(REF, "ker_ascii2bin"), os.path.join(DATA, "ker.list"),
             os.path.join(DATA, "ker_ASCII"), kb], d)
        a2b = [sha(os.path.join(kb, n + ".bin")) for n in names]
        shutil.copy(os.path.join(DATA, "ker_bin.list"), kb)
        run([os.path.join(REF, "Ker2PLSQR3_preprocess"), kb, "ker_bin.list", str(DX.N_COL)], d)
        run([MPIEXEC, "-np", "2", os.path.join(REF, "Ker2PLSQR3"), kb, "ker_bin.list", str(DX.N_COL),
             os.path.join(kb, "col_info.txt")], d)
        # LSQR3.pbs runs in data/ with the matrix moved there (data/README steps 4-8)
        for f in ("matrix_bycol.mm2.bin", "matrix_bycol.mm2.info"):
            shutil.move(os.path.join(kb, f), os.path.join(d, f))
        shutil.copy(os.path.join(DATA, "measure_fake.list"), d)
        dn = P.write_damping_files(d, *DX.GRID, n_param=DX.N_PARAM, weights=(1.0,) * 7, cross=True)
        s = P.System.read_reference_files(d, dict(ker_f="matrix_bycol.mm2.bin", ker_i="matrix_bycol.mm2.info",
                                                  b_k="measure_fake.list", **dn),
                                          100, 1910672, DX.N_COL)
        out = run([MPIEXEC, "-np", "1", os.path.join(REF, "plsqr3_ref_fp64"), "-dir", d,
                   "-ker_f", "matrix_bycol.mm2.bin", "-ker_i", "matrix_bycol.mm2.info",
                   "-damp_f", dn["damp_f"], "-damp_f_bycol", dn["damp_f_bycol"], "-damp_i", dn["damp_i"],
                   "-damp_i_bycol", dn["damp_i_bycol"], "-b_k", "measure_fake.list", "-row_k", "100",
                   "-row_d", str(s.m_d), "-col", str(DX.N_COL), "-itn", str(ITN)], d)
        res = parse(out, d, ITN, DX.N_COL)
        x = res["x"]
        np.savez(os.path.join(HERE, f"data_example_itn{ITN}.npz"),
                 a2b_sha256=np.array(a2b), col_info_sha256=np.array(sha(os.path.join(kb, "col_info.txt"))),
                 ker_bin_sha256=np.array(sha(os.path.join(d, "matrix_bycol.mm2.bin"))),
                 ker_info_sha256=np.array(sha(os.path.join(d, "matrix_bycol.mm2.info"))),
                 x_sha256=np.array(hashlib.sha256(x.tobytes()).hexdigest()),
                 x_norm=np.array(np.linalg.norm(x)), x_sample=x[::997].copy(),
                 alpha=res["alpha"], beta=res["beta"], log=res["log"], itn=np.array(ITN),
                 dims=np.array([s.m_k, s.m_d, s.n, s.nnz_k, s.nnz_d], dtype=np.int64))
        print(f"m_k={s.m_k} m_d={s.m_d} n={s.n} nnz_k={s.nnz_k} nnz_d={s.nnz_d}; |x|={np.linalg.norm(x):.6e}; "
              f"loop {res['loop_time_s']} s")


if __name__ == "__main__":
    main()
