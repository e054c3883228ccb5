"""Host-side boundary tests (CPU only): exported C ABI, struct layouts,
create_lp_problem converters, MPS reader rules, CLI argument handling, and the
SpMV block / column-panel plans.  No GPU compute is called."""
import ctypes as C
import gzip
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from helpers import random_csr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _decls(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//.*", "", txt)
    txt = re.sub(r"^\s*#.*$", "", txt, flags=re.M)   # preprocessor lines
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt)
    return [n for n in names if n not in ("if", "for", "while", "sizeof", "return")]


def test_library_exports_every_declared_symbol(dp):
    lib = C.CDLL(dp.LIB_PATH)
    for h in ("cupdlpx.h", "mps_parser.h", "dpdlp_ext.h", "dpdlp_dev.h"):
        for name in _decls(os.path.join(ROOT, "include", h)):
            assert hasattr(lib, name), f"{name} ({h}) not exported"
    for name in dp.EXPORTED_SYMBOLS:
        assert hasattr(lib, name)


def test_struct_sizes_match_reference_abi(dp):
    assert C.sizeof(dp.LpProblem) == 104
    assert C.sizeof(dp.PdhgParameters) == 192
    assert C.sizeof(dp.CupdlpxResult) == 208
    assert C.sizeof(dp.MatrixDesc) == 56
    assert dp.PdhgParameters.termination_criteria.offset == 40
    assert dp.PdhgParameters.restart_params.offset == 88
    assert dp.PdhgParameters.permute_block_size.offset == 184
    assert dp.CupdlpxResult.termination_reason.offset == 184
    assert dp.CupdlpxResult.feasibility_iteration.offset == 200


def test_default_parameters(dp):
    p = dp.default_params()
    assert p.l_inf_ruiz_iterations == 10 and p.has_pock_chambolle_alpha
    assert p.pock_chambolle_alpha == 1.0 and p.bound_objective_rescaling
    assert p.termination_evaluation_frequency == 200 and p.sv_max_iter == 5000
    assert p.sv_tol == 1e-4 and p.reflection_coefficient == 1.0
    tc = p.termination_criteria
    assert (tc.eps_optimal_relative, tc.eps_feasible_relative, tc.eps_infeasible) == (1e-4, 1e-4, 1e-10)
    assert tc.time_sec_limit == 3600.0 and tc.iteration_limit == 2**31 - 1
    assert tc.eps_feas_polish_relative == 1e-6
    rp = p.restart_params
    assert (rp.artificial_restart_threshold, rp.sufficient_reduction_for_restart,
            rp.necessary_reduction_for_restart) == (0.36, 0.2, 0.5)
    assert (rp.k_p, rp.k_i, rp.k_d, rp.i_smooth) == (0.99, 0.01, 0.0, 0.3)
    assert p.optimality_norm == dp.NORM_L2 and p.presolve and not p.feasibility_polishing


def _csr_of(dense, tol=0.0):
    m, n = dense.shape
    rp, ci, v = [0], [], []
    for i in range(m):
        for j in range(n):
            if abs(dense[i, j]) > tol:
                ci.append(j)
                v.append(dense[i, j])
        rp.append(len(ci))
    return np.array(rp), np.array(ci), np.array(v)


def test_converters_agree(dp):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((7, 9)) * (rng.random((7, 9)) < 0.4)
    A[2, 3] = 1e-13  # below the dense default tolerance
    m, n = A.shape
    rp, ci, v = _csr_of(A, 1e-12)
    ref = dict(row_ptr=rp, col_ind=ci, vals=v)
    pd = dp.Problem.from_dense(A)
    pr = dp.Problem.from_csr(m, n, rp, ci, v)
    rows, cols = np.nonzero(np.abs(A) > 1e-12)
    order = rng.permutation(len(rows))  # COO in arbitrary order
    pc = dp.Problem.from_coo(m, n, rows[order], cols[order], A[rows, cols][order])
    ccp, cri, cv = _csr_of(A.T.copy(), 1e-12)
    pcs = dp.Problem.from_csc(m, n, ccp, cri, cv)
    for p in (pd, pr, pcs):
        a = p.arrays()
        for k in ref:
            assert np.array_equal(a[k], ref[k]), k
    a = pc.arrays()  # COO keeps input order within rows: same rows, same sets
    assert np.array_equal(a["row_ptr"], rp)
    for i in range(m):
        s = slice(rp[i], rp[i + 1])
        assert sorted(zip(a["col_ind"][s], a["vals"][s])) == sorted(zip(ci[s], v[s]))


def test_create_defaults_and_errors(dp):
    A = np.array([[1.0, 0.0], [2.0, 3.0]])
    p = dp.Problem.from_dense(A)
    a = p.arrays()
    assert np.all(a["c"] == 0) and np.all(a["var_lb"] == -np.inf) and np.all(a["var_ub"] == np.inf)
    assert np.all(a["con_lb"] == -np.inf) and np.all(a["con_ub"] == np.inf)
    assert a["obj_const"] == 0.0
    with pytest.raises(ValueError):
        dp.Problem.from_coo(2, 2, [0, 5], [0, 1], [1.0, 2.0])   # row out of range
    with pytest.raises(ValueError):
        dp.Problem.from_coo(2, 2, [0, 1], [0, 7], [1.0, 2.0])   # column out of range
    with pytest.raises(ValueError):
        dp.Problem.from_csc(2, 2, [0, 1, 2], [0, 9], [1.0, 2.0])
    d = dp.MatrixDesc(m=1, n=1, fmt=7, zero_tolerance=0.0)
    assert not dp.lib.create_lp_problem(None, C.byref(d), None, None, None, None, None)
    # tolerance filters CSC / COO only when positive
    p = dp.Problem.from_coo(1, 3, [0, 0, 0], [0, 1, 2], [1e-9, 0.5, 0.0], zero_tolerance=1e-6)
    assert p.nnz == 1
    p = dp.Problem.from_coo(1, 3, [0, 0, 0], [0, 1, 2], [1e-9, 0.5, 0.0])
    assert p.nnz == 3


def test_set_start_values_copies(dp):
    p = dp.Problem.from_dense(np.eye(2))
    x = np.array([1.0, 2.0])
    p.set_start(primal=x)
    x[0] = 99.0
    got = np.ctypeslib.as_array(p.ptr.contents.primal_start, shape=(2,))
    assert list(got) == [1.0, 2.0]
    assert not p.ptr.contents.dual_start
    p.set_start(dual=[3.0, 4.0])
    assert not p.ptr.contents.primal_start


def _expect(path):
    e = json.load(open(path))
    f = lambda xs: np.array([float(v) for v in xs])
    return e, f


@pytest.mark.parametrize("gz", [False, True])
def test_mps_rules_golden(dp, tmp_path, gz):
    src = os.path.join(GOLD, "rules.mps")
    path = src
    if gz:
        path = str(tmp_path / "rules.mps.gz")
        with open(src, "rb") as fi, gzip.open(path, "wb") as fo:
            shutil.copyfileobj(fi, fo)
    p = dp.Problem.from_mps(path)
    a = p.arrays()
    e, f = _expect(os.path.join(GOLD, "rules_expected.json"))
    assert (a["m"], a["n"]) == (e["m"], e["n"])
    assert list(a["row_ptr"]) == e["row_ptr"] and list(a["col_ind"]) == e["col_ind"]
    assert np.array_equal(a["vals"], f(e["vals"])) and np.array_equal(a["c"], f(e["c"]))
    assert a["obj_const"] == e["obj_const"]
    for k in ("con_lb", "con_ub", "var_lb", "var_ub"):
        assert np.array_equal(a[k], f(e[k])), k


def test_mps_objsense_quirk_and_missing_file(dp):
    p = dp.Problem.from_mps(os.path.join(GOLD, "objsense_quirk.mps"))
    # a lone "MAX" line is taken as a section header (src/mps_parser.c:397-434)
    assert list(p.arrays()["c"]) == [1.0]
    with pytest.raises(ValueError):
        dp.Problem.from_mps("/nonexistent/file.mps")


def test_mps_capi_example(dp):
    a = dp.Problem.from_mps(os.path.join(GOLD, "capi_example.mps")).arrays()
    dense = np.zeros((3, 2))
    for i in range(3):
        dense[i, a["col_ind"][a["row_ptr"][i]:a["row_ptr"][i + 1]]] = a["vals"][a["row_ptr"][i]:a["row_ptr"][i + 1]]
    assert np.array_equal(dense, [[1, 2], [0, 1], [3, 2]])
    assert list(a["con_lb"]) == [5.0, -np.inf, -np.inf] and list(a["con_ub"]) == [5.0, 2.0, 8.0]
    assert np.all(np.isinf(a["var_lb"])) and np.all(np.isinf(a["var_ub"]))


def test_cli_arguments(dp, tmp_path):
    cli = dp.CLI_PATH
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    assert subprocess.run([cli, "-h"], capture_output=True).returncode == 0
    r = subprocess.run([cli, "--opt_norm", "l3", "a.mps", str(tmp_path)], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "opt_norm" in r.stderr
    r = subprocess.run([cli, str(tmp_path / "missing.mps"), str(tmp_path)], capture_output=True,
                       text=True)
    assert r.returncode == 1


def test_cli_process_grid_options(dp, tmp_path):
    """The reference's process-grid options (cli.c:470-547) are accepted with
    its validation; one GPU holds the whole matrix, so only the 1,1 grid
    passes (the run then stops at the missing file, after the options)."""
    cli, missing = dp.CLI_PATH, str(tmp_path / "missing.mps")

    def run(*args):
        return subprocess.run([cli, *args, missing, str(tmp_path)], capture_output=True,
                              text=True)
    r = run("--grid_size", "1,1", "--partition_method", "nnz", "--permute_method", "block",
            "--permute_block_size", "64")
    assert r.returncode == 1 and "Failed to read" in r.stderr
    for args, msg in ((("--grid_size", "2,2"), "one (1,1)"), (("--grid_size", "0,1"), "positive"),
                      (("--grid_size", "3"), "Invalid grid_size"),
                      (("--partition_method", "rows"), "partition_method"),
                      (("--permute_method", "shuffle"), "permute_method"),
                      (("--permute_block_size", "0"), "permute_block_size")):
        r = run(*args)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)
        assert "Failed to read" not in r.stderr


@pytest.mark.parametrize("cap", [1024, 2048, 4096])
def test_block_plan_invariants(dp, cap):
    rng = np.random.default_rng(1)
    lens = np.minimum((rng.pareto(1.1, 4000) * 5).astype(int), 9000)
    lens[::97] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    blk = dp.plan_blocks(rp, cap)
    assert blk[0] == 0 and blk[-1] == 4000 and np.all(np.diff(blk) >= 1)
    for b in range(len(blk) - 1):
        r0, r1 = blk[b], blk[b + 1]
        nz = rp[r1] - rp[r0]
        assert r1 - r0 <= 256
        assert nz <= cap or r1 - r0 == 1


@pytest.mark.parametrize("P", [2, 4, 8])
def test_panel_tiles_reconstruct_spmv(dp, P):
    L = dp.lib
    rng = np.random.default_rng(P)
    m, n = 1500, 2200
    lens = np.minimum((rng.pareto(1.2, m) * 4).astype(int) + 1, 2100)
    rp, ci, v = random_csr(m, n, lens, 5)
    Pi = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    blk = np.zeros(m + 2, np.int32)
    nb = L.dpdlp_plan_blocks_panels(m, n, Pi(rp), Pi(ci), 1024, P, 1024, Pi(blk))
    blk = blk[:nb + 1]
    toff = np.zeros(nb * P + 1, np.int32)
    rofs = np.zeros(P * (m + nb), np.uint16)
    perm = np.zeros(len(v), np.int32)
    pci = np.zeros(len(v), np.int32)
    assert L.dpdlp_build_panels(m, n, Pi(rp), Pi(ci), Pi(blk), nb, P, Pi(toff),
                                rofs.ctypes.data_as(C.POINTER(C.c_ushort)), Pi(perm), Pi(pci)) == 0
    assert sorted(perm) == list(range(len(v))) and np.array_equal(pci, ci[perm])
    x = rng.standard_normal(n)
    y = np.zeros(m)
    W = -(-n // P)
    for b in range(nb):
        r0, r1 = blk[b], blk[b + 1]
        R = r1 - r0
        assert R <= 1024
        for k in range(P):
            t0, t1 = toff[b * P + k], toff[b * P + k + 1]
            assert np.all(pci[t0:t1] // W == k)
            if R == 1:
                y[r0] += np.dot(v[perm[t0:t1]], x[pci[t0:t1]])
                continue
            assert t1 - t0 <= 1024
            ro = rofs[P * (r0 + b) + k * (R + 1): P * (r0 + b) + k * (R + 1) + R + 1].astype(int)
            for i in range(R):
                s = slice(t0 + ro[i], t0 + ro[i + 1])
                y[r0 + i] += np.dot(v[perm[s]], x[pci[s]])
    ref = np.zeros(m)
    np.add.at(ref, np.repeat(np.arange(m), np.diff(rp)), v * x[ci])
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("threads", ["1", "6"])
def test_column_counts_and_panel_plan_any_thread_count(dp, monkeypatch, threads):
    """A' row pointers (column counts, split over worker threads by column
    range) equal numpy's; the panel block plan (per-row panel counts on
    worker threads) is the same at every thread count."""
    monkeypatch.setenv("DPDLP_HOST_THREADS", threads)
    L = dp.lib
    rng = np.random.default_rng(7)
    m, n = 3000, 5000
    lens = np.minimum((rng.pareto(1.1, m) * 6).astype(int), 4000)
    rp, ci, _ = random_csr(m, n, lens, 3)
    Pi = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    # column counts over 2^21 entries (enough for the threaded count)
    big_ci = rng.integers(0, n, 1 << 21).astype(np.int32)
    big_rp = np.linspace(0, len(big_ci), m + 1).astype(np.int32)
    trp = np.zeros(n + 1, np.int32)
    L.dpdlp_host_col_ptr(m, n, Pi(big_rp), Pi(big_ci), Pi(trp))
    assert np.array_equal(trp, np.concatenate([[0], np.cumsum(np.bincount(big_ci, minlength=n))]))
    blk = np.zeros(m + 2, np.int32)
    nb = L.dpdlp_plan_blocks_panels(m, n, Pi(rp), Pi(ci), 1024, 8, 1024, Pi(blk))
    monkeypatch.setenv("DPDLP_HOST_THREADS", "1")
    blk1 = np.zeros(m + 2, np.int32)
    nb1 = L.dpdlp_plan_blocks_panels(m, n, Pi(rp), Pi(ci), 1024, 8, 1024, Pi(blk1))
    assert nb == nb1 and np.array_equal(blk[:nb + 1], blk1[:nb1 + 1])


@pytest.mark.parametrize("threads", ["1", "5"])
def test_host_transpose_is_a_stable_sort(dp, monkeypatch, threads):
    """The host transpose (worker threads own column ranges) equals a stable
    sort of the nonzeros by column: each A' row in increasing row order, the
    layout csr2csc produces (pdlp_core_op.cu:111-128)."""
    monkeypatch.setenv("DPDLP_HOST_THREADS", threads)
    rng = np.random.default_rng(11)
    m, n, nnz = 50_000, 70_000, 1 << 21
    rp = np.sort(rng.integers(0, nnz + 1, m - 1)).astype(np.int32)
    rp = np.concatenate([[0], rp, [nnz]]).astype(np.int32)
    ci = rng.integers(0, n, nnz).astype(np.int32)
    v = rng.standard_normal(nnz)
    t_rp = np.zeros(n + 1, np.int32)
    t_ci = np.zeros(nnz, np.int32)
    t_v = np.zeros(nnz)
    Pi = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    assert dp.lib.dpdlp_host_transpose(m, n, Pi(rp), Pi(ci), dp._ptr(v), Pi(t_rp), Pi(t_ci),
                                       dp._ptr(t_v)) == 0
    order = np.argsort(ci, kind="stable")
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(rp))
    assert np.array_equal(t_rp, np.concatenate([[0], np.cumsum(np.bincount(ci, minlength=n))]))
    assert np.array_equal(t_ci, rows[order]) and np.array_equal(t_v, v[order])


def test_generator_rules(dp):
    p, opt = dp.generate(3000, 7000, 45000, row_rule=1, seed=5)
    assert p.nnz == 45000
    a = p.arrays()
    lens = np.diff(a["row_ptr"])
    assert lens.min() >= 1 and lens.max() > 10 * lens.mean()  # heavy tail
    for i in range(0, 3000, 97):  # distinct sorted columns
        s = a["col_ind"][a["row_ptr"][i]:a["row_ptr"][i + 1]]
        assert np.all(np.diff(s) > 0)
    q, _ = dp.generate(2000, 4000, 0, row_rule=2, seed=6, poisson_lambda=8.0)
    ql = np.diff(q.arrays()["row_ptr"])
    assert ql.min() >= 1 and abs(ql.mean() - 9.0) < 0.3
    # feasible by construction: bounds consistent, c'x* finite
    assert np.all(a["con_lb"] <= a["con_ub"]) and np.isfinite(opt)
    p2, opt2 = dp.generate(3000, 7000, 45000, row_rule=1, seed=5)
    assert opt2 == opt and np.array_equal(p2.arrays()["vals"], a["vals"])


def test_extension_struct_layouts_match_c(dp, tmp_path):
    """ctypes mirrors of the extension structs agree with the C compiler's
    sizes and offsets (include/dpdlp_ext.h)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    mirrors = {"dpdlp_solve_opts_t": dp.SolveOpts, "dpdlp_gen_spec_t": dp.GenSpec,
               "dpdlp_batch_opts_t": dp.BatchOpts, "dpdlp_serve_spec_t": dp.ServeSpec,
               "dpdlp_serve_report_t": dp.ServeReport}
    lines = ['#include "dpdlp_ext.h"', "#include <stddef.h>", "#include <stdio.h>",
             "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = mirrors[cname]
        want = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert int(val) == want, (cname, field, int(val), want)


def test_mps_number_formats_parse_exactly(dp, tmp_path):
    """COLUMNS / RHS values in many textual forms read back as the correctly
    rounded doubles (what strtod gives)."""
    rng = np.random.default_rng(7)
    forms = []
    for k in range(400):
        x = rng.standard_normal() * 10.0 ** rng.integers(-30, 30)
        f = ["%.3g", "%.6g", "%.12g", "%.15g", "%.17g", "%e", "%.10E", "%.1f", "%+.8g"][k % 9]
        forms.append(f % x)
    forms += ["1.", ".5", "-0.0", "+.25e-2", "1e22", "1e23", "123456789012345",
              "1234567890123456789", "0.000001", "5E-324", "1e-400", "1.5e+300"]
    lines = ["NAME fmt", "ROWS", " N obj", " L r0", "COLUMNS"]
    for j, t in enumerate(forms):
        lines.append(f" x{j} r0 {t}")
    lines += ["RHS", " rhs r0 1.0", "ENDATA"]
    path = tmp_path / "fmt.mps"
    path.write_text("\n".join(lines) + "\n")
    p = dp.Problem.from_mps(str(path))
    vals = p.arrays()["vals"]
    want = np.array([float(t) for t in forms])
    assert vals.shape == want.shape
    assert np.array_equal(vals.view(np.int64), want.view(np.int64))


STD_NORMALS = r"""
#include <random>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
  std::mt19937 gen(std::atoi(argv[1]));
  std::normal_distribution<double> dist(0.0, 1.0);
  const int n = std::atoi(argv[2]);
  for (int i = 0; i < n; ++i) std::printf("%a\n", dist(gen));
  return 0;
}
"""


def test_power_iteration_start_vector_is_the_std_library_draw(dp, tmp_path):
    """utils.cu:29-30,80-82: the library's start vector (engine run once,
    logarithms on worker threads) equals std::mt19937 + std::
    normal_distribution<double> bit for bit, for odd and even lengths and on
    1, 3 and 8 host threads."""
    src, exe = tmp_path / "g.cc", tmp_path / "g"
    src.write_text(STD_NORMALS)
    subprocess.run(["g++", "-O2", "-o", str(exe), str(src)], check=True)
    lib = C.CDLL(dp.LIB_PATH)
    fill = lib.dpdlp_normal_fill
    fill.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_double)]
    longest = 200_001
    for seed in (1, 7):
        out = subprocess.run([str(exe), str(seed), str(longest)], check=True,
                             capture_output=True, text=True).stdout.split()
        ref = np.array([float.fromhex(s) for s in out])
        for threads in ("1", "3", "8"):
            os.environ["DPDLP_HOST_THREADS"] = threads
            try:
                for count in (0, 1, 2, 3, 16384 * 2 + 1, longest):
                    got = np.full(count + 1, 7.0)
                    fill(seed, count, got.ctypes.data_as(C.POINTER(C.c_double)))
                    assert np.array_equal(got[:count].view(np.int64),
                                          ref[:count].view(np.int64)), (seed, threads, count)
                    assert got[count] == 7.0  # nothing written past the end
            finally:
                del os.environ["DPDLP_HOST_THREADS"]


def test_host_worker_cap_is_per_thread(dp):
    """dpdlp_host_threads_cap limits the calling thread's workers only (the
    serving loader and large-LP lane use it); 0 lifts the cap."""
    import threading
    lib = C.CDLL(dp.LIB_PATH)
    lib.dpdlp_host_threads.restype = C.c_int
    lib.dpdlp_host_threads_cap.argtypes = [C.c_int]
    os.environ["DPDLP_HOST_THREADS"] = "8"
    try:
        assert lib.dpdlp_host_threads() == 8
        lib.dpdlp_host_threads_cap(3)
        assert lib.dpdlp_host_threads() == 3
        other = []
        t = threading.Thread(target=lambda: other.append(lib.dpdlp_host_threads()))
        t.start()
        t.join()
        assert other == [8]
        lib.dpdlp_host_threads_cap(1)
        assert lib.dpdlp_host_threads() == 1
        lib.dpdlp_host_threads_cap(0)
        assert lib.dpdlp_host_threads() == 8
    finally:
        lib.dpdlp_host_threads_cap(0)
        del os.environ["DPDLP_HOST_THREADS"]


def test_start_vector_restarts_from_the_seed_every_call(dp):
    """Every solve draws its power-iteration start vector from a fresh
    mt19937(1) (solver.c and batch.c call dpdlp_normal_fill(1, m, ...)): two
    draws in one process are the same bits, and a shorter draw is a prefix of a
    longer one (no state carried between calls).  The reference keeps one
    process-wide generator (utils.cu:29-30), so only its first solve in a
    process starts from this vector."""
    lib = C.CDLL(dp.LIB_PATH)
    fill = lib.dpdlp_normal_fill
    fill.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_double)]
    P = C.POINTER(C.c_double)
    a, b, c = np.zeros(5000), np.zeros(5000), np.zeros(1234)
    fill(1, 5000, a.ctypes.data_as(P))
    fill(1, 1234, c.ctypes.data_as(P))
    fill(1, 5000, b.ctypes.data_as(P))
    assert np.array_equal(a.view(np.int64), b.view(np.int64))
    assert np.array_equal(c.view(np.int64), a[:1234].view(np.int64))
    src = open(os.path.join(ROOT, "d-pdlp_amd", "csrc", "host", "solver.c")).read()
    assert "dpdlp_normal_fill(1u, m, h_q)" in src


def test_batch_size_classes_and_layout_cache(dp, tmp_path):
    """The autotuner's size classes (nonzeros in factors of 4, mean row
    lengths of A and A' in factors of 2) and its cache: when every class of
    the LPs is in the cache for this device and precision, the layouts come
    from the cache and nothing is timed (so this runs without a GPU, whose
    device signature is then "unknown")."""
    A = np.array([[1.0, 2.0, 0.0], [0.0, 1.0, 1.0]])
    p_small = dp.Problem.from_dense(A, c=np.ones(3))            # nnz 4, 2 per row, 4/3 per col
    cls_small = dp.lib.dpdlp_batch_size_class(p_small.ptr)
    assert cls_small == 1 * 256 + 1 * 16 + 0
    p_big, _ = dp.batch_member(3)                                # m 3607, n 10760, nnz 32342
    cls_big = dp.lib.dpdlp_batch_size_class(p_big.ptr)
    assert cls_big == 7 * 256 + 3 * 16 + 1
    cache = tmp_path / "layouts.txt"
    cache.write_text(
        f"dpdlp-layout-v1 unknown:fp64 {cls_small} 1024 64 2048 128 1.0 2.0\n"
        f"dpdlp-layout-v1 unknown:mixed {cls_big} 1024 64 1024 64 1.0 2.0\n"
        f"dpdlp-layout-v1 other:fp64 {cls_big} 1024 64 1024 64 1.0 2.0\n"
        f"dpdlp-layout-v1 unknown:fp64 {cls_big} 2048 256 4096 128 1.0 2.0\n")
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present: its device signature is not 'unknown'")
    lays, rep = dp.batch_autotune([p_small, p_big, p_small], cache_path=str(cache))
    assert rep["classes"] == 2 and rep["cached"] == 2 and rep["tuned"] == 0
    assert lays == [(1024, 64, 2048, 128), (2048, 256, 4096, 128), (1024, 64, 2048, 128)]
    # no LPs: nothing to tune; bad arguments: an error, no crash
    lays, rep = dp.batch_autotune([], cache_path=str(cache))
    assert lays == [] and rep["classes"] == 0
    assert dp.lib.dpdlp_batch_autotune(None, 0, None, 0, None, None, None) == -1
    assert "bad arguments" in dp.last_error()
    assert dp.lib.dpdlp_batch_size_class(None) == -1


@pytest.mark.parametrize("P", [1, 2, 3, 4, 8])
def test_transpose_panel_plan_from_a_equals_plan_on_the_transpose(dp, P):
    """The row blocks of a tiled A' are planned from A's CSR without a host
    transpose (panel counts of A''s rows added up from A's rows): the same
    blocks as the planner over the stable transpose gives, for every cap
    (long A' rows alone in their block) and row limit, with empty rows and
    columns and a column dense enough (70000 entries) to saturate the
    16-bit panel counts."""
    rng = np.random.default_rng(P)
    Pi = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    cases = []
    m, n = 4000, 3000
    lens = np.minimum((rng.pareto(1.2, m) * 3).astype(int), n) * (rng.random(m) < 0.8)
    rp = np.zeros(m + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int32)
    cases.append((m, n, rp, ci))
    m2, n2 = 70000, 50
    ci2 = np.zeros(m2, np.int32)                       # every row: column 0 only
    cases.append((m2, n2, np.arange(m2 + 1, dtype=np.int32), ci2))
    for m, n, rp, ci in cases:
        rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(rp))
        order = np.argsort(ci, kind="stable")
        t_rp = np.zeros(n + 1, np.int32)
        np.cumsum(np.bincount(ci, minlength=n), out=t_rp[1:])
        t_ci = np.ascontiguousarray(rows[order])
        for cap in (8, 64, 1024, 4096):
            for rmax in (1024, 16):
                a = np.zeros(n + 2, np.int32)
                b = np.zeros(n + 2, np.int32)
                na = dp.lib.dpdlp_plan_blocks_panels(n, m, Pi(t_rp), Pi(t_ci), cap, P, rmax, Pi(a))
                nb = dp.lib.dpdlp_plan_blocks_panels_T(m, n, Pi(rp), Pi(ci), cap, P, rmax, Pi(b))
                assert na == nb and np.array_equal(a[:na + 1], b[:nb + 1]), (m, cap, rmax)
    assert dp.lib.dpdlp_plan_blocks_panels_T(2, 2, Pi(np.zeros(3, np.int32)),
                                             Pi(np.zeros(1, np.int32)), 64, 9, 16,
                                             Pi(np.zeros(4, np.int32))) == -1
