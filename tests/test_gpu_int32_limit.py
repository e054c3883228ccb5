"""The largest matrix the reference's int32 CSR interface can describe:
nnz = 2^31 - 1 nonzeros (m = 2^24 rows of 127-128 entries and a last row
of 8,388,735, n = 2^25 columns), about 26 GB per copy of the matrix, held
twice (A and A') plus tiles and the solve's scaled copies in one GPU's HBM.

Column of position k is (a k + b) mod 2^25 with a odd, so every column is
hit 64 times (one of them 63) and no row repeats a column; a value depends
only on its column and is a multiple of 1/2, so every row and column sum
is exact in FP64 whatever the order and A 1, A' 1 are checked bit for bit
against closed forms / an in-order host sum.  Both layouts (A on 4 XCD
column panels with A' untiled, and A untiled with A' on 4 panels: the long
last row then runs the whole-workgroup path up to position 2^31 - 1) are
uploaded in turn, checked for adjointness with random vectors, and solved
for 20 iterations; the two solves must agree to 1e-9."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M, N = 1 << 24, 1 << 25
NNZ = (1 << 31) - 1
MUL, ADD = 0x9E3779B1 & (N - 1) | 1, 12345


def column_value(j):
    """the value every entry of column j holds: a nonzero multiple of 1/2"""
    t = np.bitwise_and(j, 15).astype(np.float64)
    t -= 7.5
    t *= 0.5
    return t


def build(m=M, n=N, nnz=NNZ):
    """row pointers, columns, values (rows of 127/128 entries, the rest in
    the last row) and the closed-form column counts"""
    lens = np.empty(m, dtype=np.int64)
    lens[:-1] = 127 + (np.arange(m - 1) & 1)
    lens[-1] = nnz - int(lens[:-1].sum())
    assert 0 < lens[-1] < n
    rp = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    assert rp[-1] == nnz
    ci = np.empty(nnz, dtype=np.int32)
    step = 1 << 26
    for s in range(0, nnz, step):
        k = np.arange(s, min(s + step, nnz), dtype=np.int64)
        k *= MUL
        k += ADD
        np.bitwise_and(k, n - 1, out=k)
        ci[s:s + len(k)] = k
    # k -> (MUL k + ADD) mod n is a bijection on every n consecutive
    # positions: each column is hit once per full period, plus once more if
    # it is the image of a position in the last, partial period
    counts = np.full(n, nnz // n, dtype=np.int64)
    tail = np.arange(nnz // n * n, nnz, dtype=np.int64)
    counts[(tail * MUL + ADD) & (n - 1)] += 1
    return rp.astype(np.int32), ci, column_value(ci), lens, counts


@pytest.mark.timeout(1200)
def test_int32_limit_matrix_products_and_solve(dp):
    rp, ci, v, lens, counts = build()
    print(f"built nnz={NNZ} last row {lens[-1]}", flush=True)
    row_sums = np.add.reduceat(v, rp[:-1].astype(np.int64))
    col_sums = counts * column_value(np.arange(N, dtype=np.int64))
    c = (np.arange(N) % 7) - 3.0
    prob = dp.Problem.from_csr(M, N, rp, ci, v, c=c, con_lb=np.full(M, -1.0),
                               con_ub=np.full(M, 1.0), var_lb=np.zeros(N), var_ub=np.ones(N))
    del ci, v
    assert prob.nnz == NNZ
    rng = np.random.default_rng(5)
    xr, yr = rng.standard_normal(N), rng.standard_normal(M)
    params = dp.default_params(presolve=False, sv_max_iter=20,
                               termination_criteria__iteration_limit=20)
    sols = []
    for panels in ((4, 1), (1, 4)):
        res = dp.Resident(prob, spmv_cap=4096, panels=panels)
        assert res.panels == panels and res.nnz == NNZ
        print(f"uploaded {panels}", flush=True)
        assert np.array_equal(res.spmv(np.ones(N)), row_sums)
        assert np.array_equal(res.spmv(np.ones(M), transpose=True), col_sums)
        ax, aty = res.spmv(xr), res.spmv(yr, transpose=True)
        lhs, rhs = float(yr @ ax), float(aty @ xr)
        scale = np.linalg.norm(yr) * np.linalg.norm(ax) + np.linalg.norm(aty) * np.linalg.norm(xr)
        assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)
        # the long last row against a host dot product
        seg = slice(int(rp[-2]), int(rp[-1]))
        cols = (np.arange(seg.start, seg.stop, dtype=np.int64) * MUL + ADD) & (N - 1)
        want = float(column_value(cols) @ xr[cols])
        assert abs(ax[-1] - want) <= 1e-12 * float(np.abs(column_value(cols)) @ np.abs(xr[cols]))
        print(f"products {panels} ok", flush=True)
        out, _ = res.solve(params)
        assert out.status == "ITERATION_LIMIT" and out.iterations == 20
        assert np.all(np.isfinite(out.x)) and np.all(np.isfinite(out.y))
        assert np.all(out.x >= 0.0) and np.all(out.x <= 1.0)
        assert np.isfinite(out.primal_obj)
        sols.append((out.x.copy(), out.y.copy(), out.primal_obj))
        print(f"solve {panels}: obj {out.primal_obj!r}", flush=True)
        res.free()
    (x0, y0, o0), (x1, y1, o1) = sols
    assert np.linalg.norm(x0 - x1) <= 1e-9 * max(1.0, np.linalg.norm(x0))
    assert np.linalg.norm(y0 - y1) <= 1e-9 * max(1.0, np.linalg.norm(y0))
    assert abs(o0 - o1) <= 1e-9 * max(1.0, abs(o0))
