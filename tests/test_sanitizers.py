"""Host C under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds tests/sanitize_harness.c together with the library's host sources
that need no device (MPS reader, create_lp_problem's converters, generator,
transpose, row-block and panel planners, logging) and the FP64 oracle with
gcc -fsanitize=address,undefined, then runs it over the MPS fixtures, a
gzip-compressed copy, files the reference's reader accepts leniently, a
missing file and a malformed file.  The MPS reader's worker threads also run
under ThreadSanitizer on a file cut into many small chunks."""
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "d-pdlp_amd", "csrc", "host")
SOURCES = ["mps_parser.c", "api.c", "util.c", "lpgen.c", "panels.c", "logging.c"]
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _asan_works(tmp):
    src = os.path.join(tmp, "probe.c")
    with open(src, "w") as f:
        f.write("int main(void){return 0;}\n")
    exe = os.path.join(tmp, "probe")
    r = subprocess.run(["gcc", "-fsanitize=address,undefined", src, "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return False
    return subprocess.run([exe], capture_output=True).returncode == 0


def test_host_code_under_sanitizers(tmp_path):
    if shutil.which("gcc") is None or not _asan_works(str(tmp_path)):
        pytest.skip("gcc with AddressSanitizer is not usable here")
    exe = str(tmp_path / "harness")
    cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I", os.path.join(ROOT, "include"), "-I", HOST, "-I", os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "sanitize_harness.c"),
           os.path.join(ROOT, "oracle", "pdlp_oracle.c")]
    cmd += [os.path.join(HOST, s) for s in SOURCES]
    cmd += ["-o", exe, "-lz", "-lm", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]

    good = [os.path.join(GOLDEN, f) for f in sorted(os.listdir(GOLDEN)) if f.endswith(".mps")]
    gz = str(tmp_path / "capi_example.mps.gz")
    with open(os.path.join(GOLDEN, "capi_example.mps"), "rb") as fi, gzip.open(gz, "wb") as fo:
        fo.write(fi.read())
    # accepted as the reference accepts them (mps_parser.c:397-425, 633-644):
    # lines of an unknown section are skipped, entries of unknown rows dropped
    odd1 = tmp_path / "unknown_section.mps"
    odd1.write_text("NAME X\nROWS\n N obj\nNOSUCHSECTION\n junk 1 2\nENDATA\n")
    odd2 = tmp_path / "unknown_row.mps"
    odd2.write_text("NAME X\nROWS\n N obj\n L r1\nCOLUMNS\n x r9 1.0\nENDATA\n")
    good += [str(odd1), str(odd2)]
    # rejected: a value pair before any column name (mps_parser.c:616-621)
    bad = tmp_path / "pair_without_column.mps"
    bad.write_text("NAME X\nROWS\n N obj\n L r1\nCOLUMNS\n x r1\n")
    missing = str(tmp_path / "does_not_exist.mps")
    # block-gzipped copies: one whole, one with a flipped CRC-32 (refused)
    import test_mps_parallel as tmp_mod
    with open(os.path.join(GOLDEN, "capi_example.mps"), "rb") as fi:
        blocks = tmp_mod.bgzf_bytes(fi.read(), block=40)
    bg = tmp_path / "capi_example_bgzf.mps.gz"
    bg.write_bytes(blocks)
    size0 = int.from_bytes(blocks[16:18], "little") + 1
    corrupt = bytearray(blocks)
    corrupt[size0 - 8] ^= 1
    bg_bad = tmp_path / "corrupt_bgzf.mps.gz"
    bg_bad.write_bytes(bytes(corrupt))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe] + good + [gz, str(bg), "--", missing, str(bad), str(bg_bad)],
                       capture_output=True, text=True, env=env, timeout=300)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-6000:]
    assert "harness ok" in r.stdout
    assert "runtime error" not in report and "AddressSanitizer" not in report


def _tsan_works(tmp):
    src = os.path.join(tmp, "tprobe.c")
    with open(src, "w") as f:
        f.write("#include <pthread.h>\nstatic void *f(void *a) { return a; }\n"
                "int main(void) { pthread_t t; pthread_create(&t, 0, f, 0);"
                " return pthread_join(t, 0); }\n")
    exe = os.path.join(tmp, "tprobe")
    r = subprocess.run(["gcc", "-fsanitize=thread", src, "-o", exe, "-lpthread"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return False
    return subprocess.run([exe], capture_output=True).returncode == 0


def test_mps_reader_threads_under_thread_sanitizer(tmp_path):
    if shutil.which("gcc") is None or not _tsan_works(str(tmp_path)):
        pytest.skip("gcc with ThreadSanitizer is not usable here")
    import test_mps_parallel as tmp_mod
    exe = str(tmp_path / "harness_tsan")
    cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fsanitize=thread",
           "-I", os.path.join(ROOT, "include"), "-I", HOST, "-I", os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests", "sanitize_harness.c"),
           os.path.join(ROOT, "oracle", "pdlp_oracle.c")]
    cmd += [os.path.join(HOST, s) for s in SOURCES]
    cmd += ["-o", exe, "-lz", "-lm", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    files = []
    for seed in range(4):
        f = tmp_path / f"rnd{seed}.mps"
        text = tmp_mod.random_mps(seed, ncol=2000, nrow=300)
        f.write_text(text)
        files.append(str(f))
        # the gzipped copy is inflated by a producer thread while the parser
        # consumes 2 KiB windows behind it
        g = tmp_path / f"rnd{seed}.mps.gz"
        with gzip.open(g, "wt") as fh:
            fh.write(text)
        files.append(str(g))
        # a block-gzipped copy: members inflated in parallel on the workers
        b = tmp_path / f"rnd{seed}_bgzf.mps.gz"
        b.write_bytes(tmp_mod.bgzf_bytes(text.encode(), block=1500))
        files.append(str(b))
    env = dict(os.environ, DPDLP_HOST_THREADS="6", DPDLP_MPS_CHUNK="256",
               DPDLP_MPS_WINDOW="2048", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, "--mps-only"] + files, capture_output=True, text=True, env=env,
                       timeout=300)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-6000:]
    assert "harness ok" in r.stdout and "ThreadSanitizer" not in report
