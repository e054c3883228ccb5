"""The MPS reader parses COLUMNS, RHS, RANGES and BOUNDS bodies in chunks on
several threads.  Random files that use every rule of the reference reader
(src/mps_parser.c:365-777: comments and blank lines, odd/even token counts,
continuation lines across chunk edges, the six-token cap, 'MARKER' lines,
unknown rows, columns and bound types, repeated and non-contiguous column
names, objective entries and the objective RHS, ranges on every row type, a
second COLUMNS section, CR-LF endings) must give exactly what a line-by-line
reading gives: the restatement below, written from those rules, is the
checker, and every thread count and chunk size must match it bit for bit.
The command line's solution writer, which formats on the same threads, must
give printf's text."""
import math
import os
import random
import re

import numpy as np
import pytest

DELIMS = " \t\n\r"


def _tokens(line):
    """strtok_r(" \\t\\n\\r") with at most six tokens."""
    out, cur = [], []
    for ch in line:
        if ch in DELIMS:
            if cur:
                out.append("".join(cur))
                cur = []
                if len(out) == 6:
                    return out
        else:
            cur.append(ch)
    if cur and len(out) < 6:
        out.append("".join(cur))
    return out


_NUM = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")


def _atof(tok):
    """atof on the names and decimals these files hold: the longest decimal
    prefix, 0.0 when there is none (a row name read as a value)."""
    mt = _NUM.match(tok)
    return float(mt.group(0)) if mt else 0.0


def reference_read(text):
    """Line-by-line reading by the reference's rules (mps_parser.c:365-777)."""
    sec = None
    pend = []            # ROWS: (type, name)
    rows, types, obj_row = {}, [], None
    rows_done = False
    cols, obj, lb, ub = {}, [], [], []
    clb, cub = [], []
    coo = []             # (row, col, value) in file order
    cur_col, cur_idx = None, -1
    obj_const, maximize = 0.0, False
    for line in text.split("\n"):
        if not line or line[0] in "*\r":
            continue
        tk = _tokens(line)
        if not tk:
            continue
        if len(tk) == 1 and tk[0][0].isascii() and tk[0][0].isalpha():
            nxt = {"ROWS": "R", "COLUMNS": "C", "RHS": "H", "RANGES": "G", "BOUNDS": "B",
                   "OBJSENSE": "O", "ENDATA": "E"}.get(tk[0])
            if sec == "R" and nxt != "R" and not rows_done:
                ix = next((i for i, (t, _) in enumerate(pend) if t == "N"), 0 if pend else -1)
                if ix >= 0:
                    obj_row = pend[ix][1]
                for i, (t, name) in enumerate(pend):
                    if i == ix or t not in "ELG":
                        continue
                    if name not in rows:
                        rows[name] = len(types)
                        types.append(t)
                clb = [-math.inf if t == "L" else 0.0 for t in types]
                cub = [math.inf if t == "G" else 0.0 for t in types]
                rows_done = True
            sec = nxt
            if sec == "E":
                break
            continue
        nt = len(tk)
        if sec == "O":
            if tk[0] in ("MAX", "MAXIMIZE"):
                maximize = True
        elif sec == "R":
            if nt >= 2:
                pend.append((tk[0][0], tk[1]))
        elif sec == "C":
            if nt < 2 or tk[1] == "'MARKER'":
                continue
            if nt % 2:
                if cur_col != tk[0]:
                    cur_col, cur_idx = tk[0], -1
                first = 1
            else:
                if cur_col is None:
                    return None
                first = 0
            if cur_idx < 0:
                if cur_col not in cols:
                    cols[cur_col] = len(obj)
                    obj.append(0.0)
                    lb.append(0.0)
                    ub.append(math.inf)
                cur_idx = cols[cur_col]
            for i in range(first, nt - 1, 2):
                v = _atof(tk[i + 1])
                if obj_row is not None and tk[i] == obj_row:
                    obj[cur_idx] += v
                elif tk[i] in rows:
                    coo.append((rows[tk[i]], cur_idx, v))
        elif sec == "H":
            for i in range(1, nt - 1, 2):
                v = _atof(tk[i + 1])
                if obj_row is not None and tk[i] == obj_row:
                    obj_const = -v
                    continue
                r = rows.get(tk[i])
                if r is None:
                    continue
                if types[r] == "L":
                    cub[r] = v
                elif types[r] == "G":
                    clb[r] = v
                else:
                    clb[r] = cub[r] = v
        elif sec == "G":
            for i in range(1, nt - 1, 2):
                R = _atof(tk[i + 1])
                r = rows.get(tk[i])
                if r is None:
                    continue
                t = types[r]
                rhs = cub[r] if t == "L" else clb[r]
                if t == "G":
                    cub[r] = rhs + abs(R)
                elif t == "L":
                    clb[r] = rhs - abs(R)
                elif R >= 0:
                    cub[r] = rhs + R
                else:
                    clb[r] = rhs + R
        elif sec == "B":
            if nt < 3 or tk[2] not in cols:
                continue
            j = cols[tk[2]]
            v = _atof(tk[3]) if nt > 3 else 0.0
            t = tk[0]
            if t == "LO":
                lb[j] = v
            elif t == "UP":
                ub[j] = v
            elif t == "FX":
                lb[j] = ub[j] = v
            elif t == "FR":
                lb[j], ub[j] = -math.inf, math.inf
            elif t == "MI":
                lb[j] = -math.inf
            elif t == "PL":
                ub[j] = math.inf
            elif t == "BV":
                lb[j], ub[j] = 0.0, 1.0
    m, n = len(types), len(obj)
    rp = np.zeros(m + 1, dtype=np.int64)
    for r, _, _ in coo:
        rp[r + 1] += 1
    rp = np.cumsum(rp)
    pos = rp[:-1].copy()
    ci = np.zeros(len(coo), dtype=np.int64)
    val = np.zeros(len(coo))
    for r, c, v in coo:          # stable: file order within each row
        ci[pos[r]] = c
        val[pos[r]] = v
        pos[r] += 1
    sign = -1.0 if maximize else 1.0
    return dict(m=m, n=n, row_ptr=rp, col_ind=ci, vals=val,
                c=np.array(obj) * sign, var_lb=np.array(lb), var_ub=np.array(ub),
                con_lb=np.array(clb), con_ub=np.array(cub),
                obj_const=(-obj_const if maximize else obj_const))


def _num(rng):
    k = rng.random()
    if k < 0.1:
        return "0"
    if k < 0.2:
        return str(rng.randint(-9, 9))
    if k < 0.3:
        return f"{rng.uniform(-1e3, 1e3):.3e}"
    return f"{rng.gauss(0, 3):.6g}"


def random_mps(seed, ncol=300, nrow=60):
    rng = random.Random(seed)
    eol = "\r\n" if seed % 3 == 0 else "\n"
    L = ["NAME rnd", "ROWS", " N obj"]
    if seed % 2:
        L.append(" N obj2")                      # a second N row is dropped
    rnames = [f"r{i}" for i in range(nrow)]
    for r in rnames:
        L.append(f" {rng.choice('ELGGL')} {r}")
    L.append(" X oddtype")                      # not E/L/G: never a row
    names = rnames + ["obj", "nosuch", "obj2"]
    colnames = [f"x{j}" for j in range(ncol)]

    def columns(cset):
        L.append("COLUMNS")
        for j in cset:
            c = colnames[j]
            L.append(f"    {c} {rng.choice(names)} {_num(rng)}")
            for _ in range(rng.randint(0, 6)):
                k = rng.random()
                if k < 0.35:
                    L.append(f"    {c} {rng.choice(names)} {_num(rng)}")
                elif k < 0.55:
                    L.append(f"    {c}\t{rng.choice(names)} {_num(rng)}   {rng.choice(names)} {_num(rng)}")
                elif k < 0.7:                   # continuation lines (even token count)
                    L.append(f"  {rng.choice(names)} {_num(rng)}")
                elif k < 0.75:
                    L.append(f"  {rng.choice(names)} {_num(rng)} {rng.choice(names)} {_num(rng)}")
                elif k < 0.78:                  # seven tokens: capped at six, so a continuation
                    L.append(f" {c} {rng.choice(names)} {_num(rng)} {rng.choice(names)} "
                             f"{_num(rng)} {rng.choice(names)} {_num(rng)}")
                elif k < 0.82:
                    L.append(f"    MARKER                 'MARKER'                 'INTORG'")
                elif k < 0.86:
                    L.append("* a comment")
                elif k < 0.89:
                    L.append("")
                elif k < 0.92:
                    L.append("   5")                 # one token, not a letter: skipped
                else:
                    L.append(f"    {c} {rng.choice(names)} {_num(rng)}")
    order = list(range(ncol))
    columns(order[: ncol * 2 // 3])
    L.append("RHS")
    for r in rng.sample(names, len(names) // 2):
        L.append(f"    rhs {r} {_num(rng)}" + (f" {rng.choice(names)} {_num(rng)}" if rng.random() < 0.3 else ""))
    L.append("    rhs obj 2.5")
    if seed % 4 == 1:                           # a second COLUMNS section, repeats included
        columns(order[ncol * 2 // 3:] + rng.sample(order, 5))
    L.append("RANGES")
    for r in rng.sample(rnames, nrow // 3):
        L.append(f"    rng {r} {_num(rng)}")
    L.append("BOUNDS")
    for _ in range(ncol):
        t = rng.choice(["LO", "UP", "FX", "FR", "MI", "PL", "BV", "UI", "SC"])
        c = rng.choice(colnames + ["nocol"])
        L.append(f" {t} bnd {c}" + ("" if rng.random() < 0.1 else f" {_num(rng)}"))
    if seed % 5 == 2:
        L += ["OBJSENSE", "    MAX  extra"]       # a data line of two tokens sets MAX
    L.append("ENDATA")
    return eol.join(L) + eol


def _same(a, ref):
    assert a["row_ptr"].tolist() == ref["row_ptr"].tolist()
    assert a["col_ind"].tolist() == ref["col_ind"].tolist()
    for k in ("vals", "c", "var_lb", "var_ub", "con_lb", "con_ub"):
        assert np.array_equal(a[k], ref[k]), k


@pytest.mark.parametrize("seed", range(8))
def test_chunked_reader_equals_line_by_line(dp, tmp_path, monkeypatch, seed):
    text = random_mps(seed)
    path = tmp_path / f"r{seed}.mps"
    path.write_bytes(text.encode())
    ref = reference_read(text)
    assert ref is not None and ref["m"] > 0 and len(ref["vals"]) > 0
    for threads, chunk in ((1, 1 << 30), (4, 1), (3, 17), (8, 64), (2, 400)):
        monkeypatch.setenv("DPDLP_HOST_THREADS", str(threads))
        monkeypatch.setenv("DPDLP_MPS_CHUNK", str(chunk))
        p = dp.Problem.from_mps(str(path))
        a = p.arrays()
        assert (p.m, p.n) == (ref["m"], ref["n"]), (threads, chunk)
        _same(a, ref)
        assert a.get("obj_const", 0.0) == ref["obj_const"]


@pytest.mark.parametrize("seed", range(8))
def test_streamed_gzip_reader_equals_line_by_line(dp, tmp_path, monkeypatch, seed):
    """A gzipped file is inflated by a producer thread while the sections are
    parsed in windows of whole lines behind it: every window size (one line
    per window up to the whole file), thread count and chunk size, and the
    whole-file path (DPDLP_MPS_STREAM=0), must give the line-by-line result."""
    import gzip
    text = random_mps(seed + 100, ncol=400, nrow=80)
    path = tmp_path / f"r{seed}.mps.gz"
    with gzip.open(path, "wb") as f:
        f.write(text.encode())
    ref = reference_read(text)
    assert ref is not None and len(ref["vals"]) > 0
    cases = ((1, 1 << 30, 1, "1"), (4, 1, 37, "1"), (3, 17, 1000, "1"), (8, 64, 1 << 24, "1"),
             (2, 400, 5000, "0"))
    for threads, chunk, window, stream in cases:
        monkeypatch.setenv("DPDLP_HOST_THREADS", str(threads))
        monkeypatch.setenv("DPDLP_MPS_CHUNK", str(chunk))
        monkeypatch.setenv("DPDLP_MPS_WINDOW", str(window))
        monkeypatch.setenv("DPDLP_MPS_STREAM", stream)
        p = dp.Problem.from_mps(str(path))
        a = p.arrays()
        assert (p.m, p.n) == (ref["m"], ref["n"]), (threads, chunk, window)
        _same(a, ref)
        assert a.get("obj_const", 0.0) == ref["obj_const"]


def test_streamed_gzip_reader_errors(dp, tmp_path, monkeypatch):
    """Errors of a streamed input: a value pair before any column name fails
    in a late window (the producer is stopped and joined), and a missing
    .gz file is refused."""
    import gzip
    body = "".join(f" r1 {i}.0\n" for i in range(5000))
    path = tmp_path / "bad.mps.gz"
    with gzip.open(path, "wt") as f:
        f.write("NAME X\nROWS\n N obj\n L r1\nCOLUMNS\n" + body + "RHS\nENDATA\n")
    monkeypatch.setenv("DPDLP_MPS_WINDOW", "512")
    with pytest.raises(ValueError):
        dp.Problem.from_mps(str(path))
    with pytest.raises(ValueError):
        dp.Problem.from_mps(str(tmp_path / "missing.mps.gz"))


def bgzf_bytes(data, block=65280, level=6, with_size=True):
    """Block gzip (BGZF): members of at most `block` input bytes, each with a
    "BC" extra subfield holding its total size - 1 (RFC 1952 FEXTRA), then
    the empty end member.  with_size=False writes the subfield only in the
    first member (the rest plain gzip members)."""
    import struct
    import zlib
    out = bytearray()
    pieces = [data[i:i + block] for i in range(0, len(data), block)] + [b""]
    for k, piece in enumerate(pieces):
        co = zlib.compressobj(level, zlib.DEFLATED, -15)
        payload = co.compress(piece) + co.flush()
        trailer = struct.pack("<II", zlib.crc32(piece) & 0xffffffff, len(piece) & 0xffffffff)
        if with_size or k == 0:
            size = 12 + 6 + len(payload) + 8
            head = (b"\x1f\x8b\x08\x04" + b"\x00" * 4 + b"\x00\xff" + struct.pack("<H", 6) +
                    b"BC" + struct.pack("<HH", 2, size - 1))
        else:
            head = b"\x1f\x8b\x08\x00" + b"\x00" * 4 + b"\x00\xff"
        out += head + payload + trailer
    return bytes(out)


@pytest.mark.parametrize("seed", range(4))
def test_block_gzip_inflated_in_parallel_equals_line_by_line(dp, tmp_path, monkeypatch, seed):
    """A block-gzipped file's members are located from their size fields and
    inflated on the worker threads, each into its place: every thread count
    and block size gives the line-by-line result, as do a file whose later
    members lack the size field (the serial reader takes it) and the plain
    gzip of the same text."""
    text = random_mps(seed + 200, ncol=400, nrow=80).encode()
    ref = reference_read(text.decode())
    assert ref is not None and len(ref["vals"]) > 0
    cases = [(1, 65280, True), (3, 4000, True), (8, 997, True), (5, 1, True), (4, 3000, False)]
    for threads, block, sized in cases:
        path = tmp_path / f"b{seed}_{block}_{int(sized)}.mps.gz"
        path.write_bytes(bgzf_bytes(text, block=block, with_size=sized))
        monkeypatch.setenv("DPDLP_HOST_THREADS", str(threads))
        p = dp.Problem.from_mps(str(path))
        a = p.arrays()
        assert (p.m, p.n) == (ref["m"], ref["n"]), (threads, block, sized)
        _same(a, ref)
        assert a.get("obj_const", 0.0) == ref["obj_const"]


def test_block_gzip_corrupt_member_is_refused(dp, tmp_path, monkeypatch):
    """A member whose CRC-32 or size does not match its inflated bytes makes
    the read fail (NULL from read_mps_file), on any thread count."""
    text = random_mps(7, ncol=200, nrow=40).encode()
    good = bgzf_bytes(text, block=2000)
    # flip one bit of the second member's CRC-32: the member's total size is
    # in its header, the CRC the 8 bytes before its end
    size0 = int.from_bytes(good[16:18], "little") + 1
    size1 = int.from_bytes(good[size0 + 16:size0 + 18], "little") + 1
    bad = bytearray(good)
    bad[size0 + size1 - 8] ^= 1
    path = tmp_path / "bad.mps.gz"
    path.write_bytes(bytes(bad))
    for threads in (1, 4):
        monkeypatch.setenv("DPDLP_HOST_THREADS", str(threads))
        with pytest.raises(ValueError):
            dp.Problem.from_mps(str(path))
    # a wrong ISIZE in the last data member
    bad = bytearray(good)
    last_end = len(good) - 28  # the empty end member is 28 bytes
    bad[last_end - 4] ^= 1
    path.write_bytes(bytes(bad))
    with pytest.raises(ValueError):
        dp.Problem.from_mps(str(path))
    # plain gzip (one stream, inflated serially): a wrong CRC-32 is a read
    # failure on both the streamed and the whole-file path, not a short file
    import gzip
    one = bytearray(gzip.compress(text))
    one[-8] ^= 1
    path.write_bytes(bytes(one))
    for stream in ("1", "0"):
        monkeypatch.setenv("DPDLP_MPS_STREAM", stream)
        with pytest.raises(ValueError):
            dp.Problem.from_mps(str(path))


def test_chunked_reader_rejects_pair_before_any_column(dp, tmp_path, monkeypatch):
    """A continuation line before any column name fails (mps_parser.c:616-621),
    also when it is the first line of a later chunk."""
    body = "".join(f" r1 {i}.0\n" for i in range(50))
    path = tmp_path / "bad.mps"
    path.write_text("NAME X\nROWS\n N obj\n L r1\nCOLUMNS\n" + body + "RHS\nENDATA\n")
    for threads, chunk in ((1, 1 << 30), (4, 8)):
        monkeypatch.setenv("DPDLP_HOST_THREADS", str(threads))
        monkeypatch.setenv("DPDLP_MPS_CHUNK", str(chunk))
        with pytest.raises(ValueError):
            dp.Problem.from_mps(str(path))


def test_restatement_on_the_c_api_example():
    """The checker itself on the docs/C_API.md LP (x1 + x2, x1 + 2 x2 = 5,
    x2 <= 2, 3 x1 + 2 x2 <= 8, x free)."""
    here = os.path.join(os.path.dirname(__file__), "golden", "capi_example.mps")
    ref = reference_read(open(here).read())
    assert (ref["m"], ref["n"]) == (3, 2)
    dense = np.zeros((3, 2))
    for i in range(3):
        for k in range(ref["row_ptr"][i], ref["row_ptr"][i + 1]):
            dense[i, ref["col_ind"][k]] = ref["vals"][k]
    assert dense.tolist() == [[1, 2], [0, 1], [3, 2]]
    assert ref["c"].tolist() == [1, 1]
    assert ref["con_lb"].tolist() == [5, -math.inf, -math.inf]
    assert ref["con_ub"].tolist() == [5, 2, 8]
    assert ref["var_lb"].tolist() == [-math.inf] * 2


def test_solution_writer_matches_printf(dp, tmp_path, monkeypatch):
    """dpdlp_write_vector (the command line's solution files, cli.c:59-79):
    one "%.10g" per line, in order, whatever the thread count."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 1, 150_000) * 10.0 ** rng.integers(-30, 30, 150_000),
                        [0.0, -0.0, 1.0, -1.0, 1e308, -1e-308, 5e-324, 123456789012.0,
                         math.inf, -math.inf]])
    want = "".join("%.10g\n" % v for v in x)
    for threads in ("1", "5"):
        monkeypatch.setenv("DPDLP_HOST_THREADS", threads)
        path = tmp_path / f"v{threads}.txt"
        assert dp.lib.dpdlp_write_vector(str(path).encode(), dp._ptr(x), len(x)) == 0
        assert path.read_text() == want
    empty = tmp_path / "empty.txt"
    assert dp.lib.dpdlp_write_vector(str(empty).encode(), None, 0) == 0
    assert empty.read_text() == ""
    assert dp.lib.dpdlp_write_vector(str(tmp_path / "no" / "dir.txt").encode(), dp._ptr(x), 3) == -1
