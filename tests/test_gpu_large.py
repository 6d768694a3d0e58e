"""Parity at the production shape: 67^3 grid, 2,000,000 rays from the top face.

The oracle runs single-threaded, so full 100-iteration solves are not
compared here; instead:
  * one aprod of each mode against the oracle (1e-12 relative), and mode 1
    on the mode-1 tile shapes other than the default,
  * mode 1 with length-sorted segment windows bitwise the default's,
  * the adjoint identity <A x, y> = <x, A^T y> on the device operator,
  * bitwise run-to-run reproducibility,
  * the first 3 LSQR iterations against the oracle (1e-10 relative in x),
  * a 100-iteration solve: rnorm from the recurrence against the true
    residual norm computed in FP64 on the device.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def big(P, O, gpu_available):
    print("generating grid67_rays2m", flush=True)
    s = P.generate_config("grid67_rays2m")
    print(f"m_k={s.m_k} m_d={s.m_d} n={s.n} nnz_k={s.nnz_k}", flush=True)
    op = P.Operator(s)
    yield s, op, O.Oracle(s)
    op.close()


def test_big_aprod_vs_oracle(big):
    s, op, orc = big
    rng = np.random.default_rng(8)
    x = rng.standard_normal(s.n)
    y = rng.standard_normal(s.m_k + s.m_d)
    for mode in (1, 2):
        print(f"mode {mode}", flush=True)
        xg, yg = x.copy(), y.copy()
        op.aprod(mode, xg, yg)
        xo, yo = x.copy(), y.copy()
        orc.aprod(mode, xo, yo)
        got, want = (yg, yo) if mode == 1 else (xg, xo)
        assert rel(got, want) < 1e-12, (mode, rel(got, want))
        xg2, yg2 = x.copy(), y.copy()
        op.aprod(mode, xg2, yg2)
        assert np.array_equal(xg2, xg) and np.array_equal(yg2, yg)


@pytest.mark.parametrize("stack", ["1", "2"])
def test_big_mode1_tile_shapes_vs_oracle(big, P, monkeypatch, stack):
    """Mode 1 at full size on the tile shapes other than the default (two
    stacked bricks, 4 x 8 x 16 voxels with one count per segment; four
    bricks, 8 x 8 x 16 with three): FP64 against the oracle as the default
    operator is held, FP32 within 1e-6."""
    s, _, orc = big
    rng = np.random.default_rng(12)
    x = rng.standard_normal(s.n)
    y = rng.standard_normal(s.m_k + s.m_d)
    yo = y.copy()
    orc.aprod(1, x.copy(), yo)
    monkeypatch.setenv("PLSQR_M1_STACK", stack)
    print(f"mode 1 on tile shape {stack}", flush=True)
    with P.Operator(s) as op:
        yg = y.copy()
        op.aprod(1, x.copy(), yg)
        y32 = y.astype(np.float32)
        op.aprod_f32(1, x.astype(np.float32), y32)
    assert rel(yg, yo) < 1e-12, (stack, rel(yg, yo))
    assert rel(y32.astype(np.float64), yo) < 1e-6, (stack, rel(y32.astype(np.float64), yo))


def test_big_sorted_windows_bitwise(big, P, monkeypatch):
    """Mode 1 with each item's segments dealt to the lanes sorted by length
    inside windows of 1024 (PLSQR_SEGWIN; the partials keep ray order
    through 16-bit destinations, items here hold up to ~15k segments) is
    bitwise the default operator's, FP64 and FP32."""
    s, op, _ = big
    rng = np.random.default_rng(13)
    x = rng.standard_normal(s.n)
    y = rng.standard_normal(s.m_k + s.m_d)
    ref, ref32 = y.copy(), y.astype(np.float32)
    op.aprod(1, x.copy(), ref)
    op.aprod_f32(1, x.astype(np.float32), ref32)
    monkeypatch.setenv("PLSQR_SEGWIN", "1024")
    print("mode 1 with sorted windows", flush=True)
    with P.Operator(s) as opw:
        yg, y32 = y.copy(), y.astype(np.float32)
        opw.aprod(1, x.copy(), yg)
        opw.aprod_f32(1, x.astype(np.float32), y32)
    assert np.array_equal(yg, ref) and np.array_equal(y32, ref32)


def test_big_adjoint_identity(big):
    s, op, _ = big
    rng = np.random.default_rng(9)
    x = rng.standard_normal(s.n)
    y = rng.standard_normal(s.m_k + s.m_d)
    ax = np.zeros(s.m_k + s.m_d)
    op.aprod(1, x.copy(), ax)
    aty = np.zeros(s.n)
    op.aprod(2, aty, y.copy())
    lhs, rhs = ax @ y, x @ aty
    assert abs(lhs - rhs) <= 1e-11 * (np.linalg.norm(ax) * np.linalg.norm(y))


def test_big_first_iterations_vs_oracle(big):
    s, op, orc = big
    print("lsqr 3 iterations", flush=True)
    x, info, tr = op.lsqr(s.b, 3, trace=True)
    xo, res, tro = orc.lsqr(s.b, 3)
    assert rel(x, xo) < 1e-10, rel(x, xo)
    np.testing.assert_allclose(tr, tro, rtol=1e-10)


def test_big_hundred_iterations_residual(big):
    s, op, _ = big
    print("lsqr 100 iterations", flush=True)
    x, info = op.lsqr(s.b, 100)
    x2, info2 = op.lsqr(s.b, 100)
    assert np.array_equal(x, x2)
    rn, an, af = op.residual_norms(s.b, x)
    np.testing.assert_allclose(info["rnorm"], rn, rtol=1e-6)
    assert an / (af * rn) <= 10 * info["arnorm"] + 1e-12
    print(f"100 its: {info['elapsed_ms']:.1f} ms, test2 {info['arnorm']:.3e}, "
          f"true ||A^T r||/(||A||_F ||r||) {an / (af * rn):.3e}", flush=True)
