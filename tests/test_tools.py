"""Command-line tools that need no GPU: print_*, purge_journal, retry, and
the argument validation of bench_* (which ends before any device call)."""
import json
import os
import subprocess
import time

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gatlas_amd", "bin")


def _tool(name):
    p = os.path.join(BIN, name)
    if not os.path.exists(p):
        pytest.skip("%s not built (run __graft_entry__.build())" % name)
    return p


def _run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=60, **kw)


def test_print_matmul_compat_and_mfma():
    r = _run([_tool("print_matmul"), "-T", "float4", "-n", "1024", "-g", "8", "-y", "8", "-x", "5"])
    assert r.returncode == 0, r.stderr
    assert "matmulbufferfloat4" in r.stdout and "BH=8" in r.stdout
    # global work items (N/VL, M/bh) = (256, 128), local (gs, gs) (KernelBaseMatmul.cpp:364-376)
    assert "global work-items (256, 128), local (8, 8)" in r.stdout
    r = _run([_tool("print_matmul"), "-T", "floatmfma", "-n", "8192", "-g", "2", "-y", "2", "-x", "2", "-G"])
    assert r.returncode == 0, r.stderr
    assert "macro tile 128 x 128" in r.stdout and "instantiated=yes" in r.stdout


def test_print_matmul_rejects_invalid():
    # M % (gs*bh) != 0
    r = _run([_tool("print_matmul"), "-T", "float4", "-n", "1000", "-g", "16", "-y", "8", "-x", "0"])
    assert r.returncode == 1
    assert "error: invalid parameters" in r.stderr
    r = _run([_tool("print_matmul"), "-T", "float4", "-n", "1024", "-g", "17", "-y", "8", "-x", "0"])
    assert r.returncode == 1 and "from 1 to 16" in r.stderr
    r = _run([_tool("print_matmul"), "-T", "nosuch", "-n", "1024", "-g", "8", "-y", "8", "-x", "0"])
    assert r.returncode == 1 and "invalid kernel type" in r.stderr


def test_print_matvec_and_saxpy():
    r = _run([_tool("print_matvec"), "-T", "float4", "-n", "16384", "-g", "64", "-y", "8", "-x", "1"])
    assert r.returncode == 0, r.stderr
    assert "matvecbufferfloat4" in r.stdout
    r = _run([_tool("print_matvec"), "-T", "floatnative", "-n", "16384", "-g", "256", "-y", "4", "-x", "1"])
    assert r.returncode in (0, 1)
    r = _run([_tool("print_saxpy"), "-T", "float4", "-m", str(1 << 26), "-Y", "256", "-G", "0", "-y", "1", "-g", "0",
              "-x", "0"])
    assert r.returncode == 0, r.stderr
    assert "saxpybufferfloat4" in r.stdout
    r = _run([_tool("print_saxpy"), "-T", "float4", "-m", "102", "-Y", "256", "-G", "0", "-y", "1", "-g", "0",
              "-x", "0"])
    assert r.returncode == 1 and "multiple of 4" in r.stderr


@pytest.mark.parametrize("mode", ["purge0", "purge1"])
def test_purge_journal_tool_matches_reference_fixture(tmp_path, mode):
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_journal.json")))[mode]
    f = tmp_path / "journal.txt"
    f.write_text(fx["file_before"])
    args = [_tool("purge_journal"), "-j", str(f)] + (["-d"] if mode == "purge1" else [])
    r = _run(args)
    assert r.returncode == 0, r.stderr
    assert "has %d bad keys" % fx["purge_result"]["bad"] in r.stdout
    assert f.read_text() == fx["file_after"]


def test_bench_tools_validate_before_device():
    r = _run([_tool("bench_matmul"), "-d", "gpu", "-T", "float4", "-n", "1024"])  # no journal
    assert r.returncode == 1
    assert "journal file must be specified" in r.stderr and "***DONE***" in r.stderr
    r = _run([_tool("bench_matmul"), "-d", "gpu", "-j", "/tmp/x", "-T", "float4", "-n", "1024", "-e", "-g", "8"])
    assert r.returncode == 1 and "expectation maximization" in r.stderr


def test_retry_respawns_until_done(tmp_path):
    counter = tmp_path / "n"
    counter.write_text("0")
    script = tmp_path / "flaky.sh"
    script.write_text("#!/bin/bash\nn=$(cat %s); n=$((n+1)); echo $n > %s\n"
                      "if [ $n -ge 3 ]; then echo '***DONE***'; else echo crash; exit 139; fi\n" % (counter, counter))
    script.chmod(0o755)
    r = _run([_tool("retry"), str(script)], env=dict(os.environ, GAT_RETRY_MAX="5"))
    assert r.returncode == 0
    assert counter.read_text().strip() == "3"
    counter.write_text("-10")
    r = _run([_tool("retry"), str(script)], env=dict(os.environ, GAT_RETRY_MAX="2"))
    assert r.returncode == 1


def test_retry_watchdog_ends_a_hung_attempt(tmp_path):
    script = tmp_path / "hang.sh"
    script.write_text("#!/bin/bash\nsleep 30\n")
    script.chmod(0o755)
    t0 = time.time()
    r = _run([_tool("retry"), str(script)], env=dict(os.environ, GAT_RETRY_MAX="2", GAT_RETRY_TIMEOUT="1"))
    assert r.returncode == 1
    assert time.time() - t0 < 25
