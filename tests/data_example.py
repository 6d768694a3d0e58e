"""The reference's example dataset (data/ of PLSQR3.2013.09) as a compact fixture.

The dataset is 100 ASCII kernel rows (data/ker_ASCII/*.APBPnz, 1,266,788
entries "col ix iy iz value" over 302,940 columns = 99 x 153 x 10 cells x 2
parameters, data/README:1-6), their list (data/ker.list) and a right-hand
side (data/measure_fake.list).  The fixture keeps what the text holds and
nothing more:

  * col: per row, the first column then differences (columns ascend within
    a row); ix, iy, iz follow from col (col - 1 = p*V + (iz-1)*ny*nx +
    (iy-1)*nx + (ix-1), V = nx*ny*nz, p the parameter);
  * value: the "%e" text as a signed 7-digit mantissa and a decimal exponent.

decode() writes the files back; the fixture also stores each file's sha256,
so a test can confirm the reconstruction is byte-identical to the dataset.
"""
import hashlib
import os

import numpy as np

GRID = (99, 153, 10)   # data/README:2-5
N_PARAM = 2
N_COL = 99 * 153 * 10 * 2


def _row_text(cols, mant, expo):
    nx, ny, nz = GRID
    V = nx * ny * nz
    c0 = (cols - 1) % V
    ix = c0 % nx + 1
    iy = (c0 // nx) % ny + 1
    iz = c0 // (nx * ny) + 1
    out = []
    for c, i, j, k, m, e in zip(cols.tolist(), ix.tolist(), iy.tolist(), iz.tolist(), mant.tolist(), expo.tolist()):
        sign = "-" if m < 0 else ""
        a = abs(m)
        out.append(f"{c} {i} {j} {k} {sign}{a // 1000000}.{a % 1000000:06d}e{e:+03d}\n")
    return "".join(out)


def encode(data_dir):
    """Arrays of the fixture from the dataset directory (data/)."""
    list_text = open(os.path.join(data_dir, "ker.list")).read()
    tok = list_text.split()
    names = tok[1:1 + int(tok[0])]
    ptr, dcol, mant, expo, sha = [0], [], [], [], []
    for name in names:
        text = open(os.path.join(data_dir, "ker_ASCII", name)).read()
        rows = [ln.split() for ln in text.splitlines()]
        cols = np.array([int(r[0]) for r in rows], np.int64)
        m = np.array([int(r[4].split("e")[0].replace(".", "")) for r in rows], np.int64)
        e = np.array([int(r[4].split("e")[1]) for r in rows], np.int64)
        assert (np.diff(cols) > 0).all()
        assert _row_text(cols, m, e) == text, name  # the encoding is lossless
        dcol.append(np.diff(np.concatenate([[0], cols])))
        mant.append(m)
        expo.append(e)
        ptr.append(ptr[-1] + len(rows))
        sha.append(hashlib.sha256(text.encode()).hexdigest())
    return dict(names=np.array(names), ptr=np.array(ptr, np.int64),
                dcol=np.concatenate(dcol).astype(np.int32), mant=np.concatenate(mant).astype(np.int32),
                expo=np.concatenate(expo).astype(np.int8), text_sha256=np.array(sha),
                list_text=np.array(list_text), b_text=np.array(open(os.path.join(data_dir, "measure_fake.list")).read()))


def decode(z, out_dir):
    """Writes ker.list, ker_ASCII/<names> and measure_fake.list under out_dir;
    returns the ASCII directory."""
    adir = os.path.join(out_dir, "ker_ASCII")
    os.makedirs(adir, exist_ok=True)
    with open(os.path.join(out_dir, "ker.list"), "w") as f:
        f.write(str(z["list_text"]))
    with open(os.path.join(out_dir, "measure_fake.list"), "w") as f:
        f.write(str(z["b_text"]))
    ptr = z["ptr"]
    for r, name in enumerate(z["names"]):
        lo, hi = int(ptr[r]), int(ptr[r + 1])
        cols = np.cumsum(z["dcol"][lo:hi].astype(np.int64))
        text = _row_text(cols, z["mant"][lo:hi].astype(np.int64), z["expo"][lo:hi].astype(np.int64))
        with open(os.path.join(adir, str(name)), "w") as f:
            f.write(text)
    return adir


def rows(z):
    """(row_ptr, 1-based columns, values) of the kernel rows, values parsed
    from their decimal text as strtod does (correctly rounded)."""
    ptr = z["ptr"].astype(np.int64)
    cols = np.empty(int(ptr[-1]), np.int64)
    for r in range(len(ptr) - 1):
        lo, hi = int(ptr[r]), int(ptr[r + 1])
        cols[lo:hi] = np.cumsum(z["dcol"][lo:hi].astype(np.int64))
    vals = np.array([float(f"{m}e{int(e) - 6}") for m, e in zip(z["mant"].tolist(), z["expo"].tolist())])
    return ptr, cols, vals


def b_values(z):
    return np.array([float(v) for v in str(z["b_text"]).split()])
