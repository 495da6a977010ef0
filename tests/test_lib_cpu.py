"""CPU tests of the product library: it loads, exports every symbol that
include/dakc.h declares, and its host-side logic (status strings, read
generator) behaves; no GPU compute is called here."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from dakc_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "dakc.h").read_text()
    return sorted(set(re.findall(r"\b(dakc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_lib.LIB_PATH))
    decl = declared_symbols()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(decl) == sorted(_lib.EXPORTS)


def test_abi_version_and_status_strings():
    lib = _lib.load()
    assert lib.dakc_abi_version() == 2
    for s in range(11):
        assert _lib.status_str(s)


def test_generator_format_and_determinism():
    p = _lib.gen_params(5, 300, 150, min_len=100, genome_len=5000, sub_rate=0.01, n_rate=0.002)
    a = _lib.generate_reads(p, threads=1)
    b = _lib.generate_reads(p, threads=4)
    assert np.array_equal(a, b)
    recs = a.reshape(300, 151)
    assert (recs[:, 150] == ord("\n")).all()
    body = recs[:, :150]
    assert set(np.unique(body).tolist()) <= set(b"ACGTNM")
    # M only as right padding; lengths in [100, 150]
    for r in body:
        s = bytes(r).decode()
        ln = len(s.rstrip("M"))
        assert 100 <= ln <= 150 and "M" not in s[:ln]
    # slices are independent: reads [100, 200) alone equal the middle third
    q = _lib.gen_params(5, 100, 150, min_len=100, genome_len=5000, sub_rate=0.01, n_rate=0.002,
                        first_read=100)
    assert np.array_equal(_lib.generate_reads(q), a[100 * 151:200 * 151])


def test_generator_genome_reads_are_genome_substrings():
    # no errors, one strand: every read is a substring of the genome the
    # generator samples from, so reads overlap heavily at this coverage
    p = _lib.gen_params(9, 400, 100, genome_len=1000, both_strands=False)
    recs = _lib.generate_reads(p).reshape(400, 101)[:, :100]
    reads = [bytes(r).decode() for r in recs]
    # reconstruct genome positions greedily: any two reads sharing a 30-mer agree
    from collections import defaultdict
    seen = defaultdict(set)
    for s in reads:
        for i in range(0, 71):
            seen[s[i:i + 30]].add(s[i + 30:i + 31] if i < 70 else "")
    conflicts = sum(1 for v in seen.values() if len(v - {""}) > 1)
    assert conflicts == 0
    assert len(seen) < 0.2 * 400 * 71   # heavy overlap: a 1 kbp genome


def test_generator_iid_base_frequencies():
    p = _lib.gen_params(1, 2000, 150)
    body = _lib.generate_reads(p).reshape(2000, 151)[:, :150]
    freq = np.bincount(body.reshape(-1), minlength=256)[[65, 67, 71, 84]] / body.size
    assert np.allclose(freq, 0.25, atol=0.005)


def test_generator_rejects_bad_params():
    lib = _lib.load()
    p = _lib.gen_params(1, 10, 150, genome_len=50)  # genome shorter than a read
    buf = np.empty(10 * 151, np.uint8)
    assert lib.dakc_generate_reads(C.byref(p), buf.ctypes.data_as(C.c_void_p), buf.size, 1) == 1


def test_kcounter_reports_missing_gpu_cleanly():
    # on a host without a GPU, context creation fails with a status, not a crash
    from dakc_amd.kcounter import gpu_available
    assert gpu_available() in (True, False)


def test_ctypes_structs_match_the_header(tmp_path):
    # the ctypes layouts of the binding are the C layouts of include/dakc.h
    structs = {"dakc_job": _lib.Job, "dakc_table": _lib.Table, "dakc_opts": _lib.Opts,
               "dakc_stats": _lib.Stats, "dakc_serve_stats": _lib.ServeStats,
               "dakc_gen_params": _lib.GenParams, "dakc_plan_info": _lib.PlanInfo,
               "dakc_hitter_stats": _lib.HitterStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dakc.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_error_bits_map_to_statuses(tmp_path):
    # the host's mapping of a job's error word (dk_status.hpp, compiled here on
    # its own): a look-back wait that gave up is a device error whatever else
    # is set; then malformed input; then an overflow (the LSD recount)
    src = tmp_path / "st.cpp"
    src.write_text('#include <cstdio>\n#include "dk_status.hpp"\n'
                   'int main() { for (unsigned e = 0; e < 16; ++e) std::printf("%u %d\\n", e, dk::status_of_err(e)); }\n')
    exe = tmp_path / "st"
    import subprocess
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'dakc_amd' / 'csrc'}", f"-I{ROOT / 'include'}", str(src),
                    "-o", str(exe)], check=True)
    got = dict(tuple(map(int, l.split())) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                   text=True).stdout.split("\n") if l.strip())
    DEVICE, FORMAT, OVERFLOW = 5, 2, -100
    for e in range(16):
        want = DEVICE if e & 8 else FORMAT if e & 3 else OVERFLOW if e & 4 else 0
        assert got[e] == want, e


def test_drivers_are_built_and_print_usage_without_a_gpu():
    """The CLI and the dakc.hpp example link against the in-tree library and
    answer usage requests before any device call."""
    import subprocess
    from pathlib import Path

    pkg = Path(__file__).resolve().parents[1] / "dakc_amd"
    r = subprocess.run([str(pkg / "dakc"), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage: dakc -f" in r.stdout
    r = subprocess.run([str(pkg / "dakc")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stdout
    r = subprocess.run([str(pkg / "hpp_example")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
