#!/usr/bin/env python3
"""Writes the golden fixtures of the parameter interface, journal and ranking
helpers by running the reference's own code (oracle/_ref/ref_driver, built
from the reference sources in place by oracle/Makefile.ref).

Run where the reference checkout is present (the tests read only the fixtures):
    python tests/golden/make_ref_golden.py
Outputs (committed): ref_params.jsonl, ref_journal.json, ref_rank.jsonl.
"""
import json
import os
import random
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def build():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-f", "Makefile.ref"])


def run(lines):
    out = subprocess.run([DRIVER], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return [json.loads(x) for x in out.stdout.splitlines()]


def param_cases(rng):
    cases = []
    dims = [8, 16, 24, 32, 40, 64, 96, 100, 128, 192, 256, 1024, 2048, 3520, 8192]
    for _ in range(1400):
        variant = rng.choice(["buffer", "buffer", "image"])
        vl = rng.choice([1, 2, 4]) if variant == "buffer" else rng.choice([4, 2])
        g = rng.choice(range(1, 17))
        gw = g if rng.random() < 0.85 else rng.choice(range(1, 17))
        p = [rng.choice([0, 1, 1, 1, 3]), rng.choice([0, 1]), rng.choice(dims), rng.choice(dims), rng.choice(dims),
             rng.choice([0, 1]), rng.choice([0, 1]), g, gw, rng.choice(range(1, 13)), vl,
             rng.choice([0, 1, 5, 6, 11, 12, 13, 23, 24, 30])]
        cases.append("matmul %s %s" % (variant, " ".join(map(str, p))))
    mvdims = [64, 128, 256, 512, 1000, 1024, 4096, 16384]
    for _ in range(700):
        variant = rng.choice(["buffer", "buffer", "image"])
        vl = rng.choice([1, 2, 4])
        p = [rng.choice([0, 1, 1, 2]), rng.choice([0, 1]), rng.choice(mvdims), rng.choice(mvdims), rng.choice([0, 1]),
             rng.choice([1, 8, 16, 32, 64, 65, 100, 128, 256]), rng.choice(range(1, 65)), vl,
             rng.choice([0, 1, 2, 3, 4])]
        cases.append("matvec %s %s" % (variant, " ".join(map(str, p))))
    for _ in range(700):
        variant = rng.choice(["buffer", "buffer", "image"])
        vl = rng.choice([1, 2, 4])
        oned = rng.random() < 0.5
        M = rng.choice([64, 256, 1000, 1024, 4096, 1 << 20, 1 << 26])
        N = vl if oned else rng.choice([vl, 8, 16, 64, 100, 128])
        gH = rng.choice([1, 2, 4, 8, 16, 64, 65, 128, 256])
        gW = 0 if oned else rng.choice([0, 1, 2, 4, 8, 16])
        bh = rng.choice(range(1, 17))
        bw = 0 if oned else rng.choice(range(0, 9))
        p = [rng.choice([0, 1, 1, 2]), M, N, gH, gW, bh, bw, vl, rng.choice([0, 1, 2, 3, 4])]
        cases.append("saxpy %s %s" % (variant, " ".join(map(str, p))))
    return cases


def journal_script(path):
    # A trial sequence exercising every run state, repeated keys, failed
    # checks (time 0), replay queries and both purge modes.
    s = ["jfile %s" % path]
    k1 = "matmulbufferfloat4 12 1 0 1024 1024 1024 0 0 16 16 8 4 5"
    k2 = "matmulbufferfloat4 12 1 0 1024 1024 1024 0 0 16 16 4 4 0"
    k3 = "matvecbufferfloat4 9 1 1 16384 16384 0 64 8 4 1"
    k4 = "saxpybufferfloat1 9 1 268435456 1 256 0 4 0 1 0"
    for key, states in ((k1, [-1, -2, -3, -4, 1234]), (k2, [-1, -2, -3]), (k3, [-1]),
                        (k4, [-1, -2, -3, -4, 0])):
        for v in states:
            s.append("jtake %s %d" % (key, v))
    s += ["jtake %s -1" % k1, "jtake %s -2" % k1, "jtake %s -3" % k1, "jtake %s -4" % k1, "jtake %s 1200" % k1]
    s += ["jload", "jgood"]
    for key in (k1, k2, k3, k4, "matmulbufferfloat2 12 1 0 8 8 8 0 0 2 2 2 2 0"):
        s.append("jstate %s" % key)
        for t in (0, 1, 2):
            s.append("jtime %s %d" % (key, t))
    return s


def main():
    if not os.path.isdir("/root/reference"):
        sys.exit("the reference checkout is needed to regenerate the fixtures")
    build()
    rng = random.Random(20101203)
    cases = param_cases(rng)
    res = run(cases)
    with open(os.path.join(HERE, "ref_params.jsonl"), "w") as f:
        for c, r in zip(cases, res):
            fam, var = c.split()[:2]
            r["family"], r["variant"] = fam, var
            f.write(json.dumps(r, separators=(",", ":")) + "\n")

    fixtures = {}
    for mode in (0, 1):
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "journal.txt")
            script = journal_script(path)
            out = run(script)
            before = open(path).read()
            purge = run(["jfile %s" % path, "jload", "jpurge %d" % mode])
            after = open(path).read()
        fixtures["purge%d" % mode] = {
            "script": [line.replace(path, "JOURNAL") for line in script],
            "outputs": out, "file_before": before, "purge_result": purge[-1], "file_after": after}
    with open(os.path.join(HERE, "ref_journal.json"), "w") as f:
        json.dump(fixtures, f, indent=1)

    rank_cases = []
    for _ in range(300):
        n = rng.randint(1, 12)
        ok = [rng.choice([0, 1, 1]) for _ in range(n)]
        avg = [rng.choice([1.0, 2.0, 2.5, 3.0, 7.25, 9.0, 0.5]) for _ in range(n)]
        tim = [rng.choice([10, 20, 30, 40, 50, 60]) for _ in range(n)]
        which = rng.choice(["rank", "mark", "markavg"])
        a = rng.randint(0, 4) if which == "rank" else rng.randint(0, 5)
        vals = tim if which == "mark" else avg
        rank_cases.append("%s %d %d %s %s" % (which, a, n, " ".join(map(str, ok)), " ".join(map(repr, vals))))
    res = run(rank_cases)
    with open(os.path.join(HERE, "ref_rank.jsonl"), "w") as f:
        for c, r in zip(rank_cases, res):
            f.write(json.dumps({"cmd": c, "out": r}, separators=(",", ":")) + "\n")
    print("wrote %d param, %d rank fixtures" % (len(cases), len(rank_cases)))


if __name__ == "__main__":
    main()
