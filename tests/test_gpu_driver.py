"""The driver surface on the GPU: the `dakc` CLI (file in, k, table out; the
reference's main, dakc/src/main/main.cpp:21-55, with its compile-time KMERLEN
and READLEN as flags) and the smallest include/dakc.hpp driver
(fqreader -> kmercounter -> serialize, main.cpp:40-49 with the classes
swapped).  Both write the serialized table of a golden read set, compared
byte for byte with the brute-force fixtures (tests/golden/expected.json)."""
import hashlib
import json
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

PKG = Path(__file__).resolve().parents[1] / "dakc_amd"
CLI = PKG / "dakc"
HPP = PKG / "hpp_example"


def _golden(golden_dir):
    return json.loads((golden_dir / "expected.json").read_text())


def _run(cmd, timeout=120):
    return subprocess.run([str(c) for c in cmd], capture_output=True, text=True, timeout=timeout)


def test_cli_tables_match_the_golden_fixtures(golden_dir, tmp_path):
    assert CLI.exists(), "dakc_amd/dakc is not built (python -m dakc_amd.build)"
    for name, e in _golden(golden_dir).items():
        out = tmp_path / f"{name}.bin"
        cmd = [CLI, "-f", golden_dir / e["reads"], "-k", e["k"], "-r", e["read_len"], "-p", e["num_pes"],
               "-b", e["owner_bin"], "-o", out] + (["-c"] if e["canonical"] else [])
        r = _run(cmd)
        assert r.returncode == 0, (name, r.stdout, r.stderr)
        assert hashlib.sha256(out.read_bytes()).hexdigest() == e["sha256"], name
        # the reference's closing prints (kcounter.cpp:416-430 reports the same two totals)
        assert f"global_kmers: {e['total']}\n" in r.stdout, (name, r.stdout)
        assert f"global_distinct_kmers: {e['distinct']}\n" in r.stdout, (name, r.stdout)


def test_hpp_driver_tables_match_the_golden_fixtures(golden_dir, tmp_path):
    assert HPP.exists(), "dakc_amd/hpp_example is not built (python -m dakc_amd.build)"
    for name, e in _golden(golden_dir).items():
        out = tmp_path / f"{name}.bin"
        r = _run([HPP, golden_dir / e["reads"], e["k"], e["read_len"], int(e["canonical"]), e["num_pes"],
                  e["owner_bin"], out])
        assert r.returncode == 0, (name, r.stdout, r.stderr)
        assert hashlib.sha256(out.read_bytes()).hexdigest() == e["sha256"], name
        assert r.stdout.strip() == f"total {e['total']} distinct {e['distinct']}", name


def test_cli_errors(golden_dir, tmp_path):
    # a missing input file exits with 2, as fqreader.cpp:31-36 does
    r = _run([CLI, "-f", tmp_path / "absent.txt", "-k", 31, "-r", 150])
    assert r.returncode == 2 and "could not open" in r.stderr
    # k outside 1..64 and a record width that is not the file's: status errors, exit 1
    reads = golden_dir / "reads_a.txt"
    for args in (["-k", 0], ["-k", 65], ["-r", 149]):
        r = _run([CLI, "-f", reads, "-r", 150, *args] if args[0] == "-k" else [CLI, "-f", reads, *args])
        assert r.returncode == 1 and "dakc:" in r.stderr, (args, r.stdout, r.stderr)
    r = _run([HPP, tmp_path / "absent.txt", 31, 150, 0, 1, -1, tmp_path / "o.bin"])
    assert r.returncode == 1 and "could not open" in r.stderr
