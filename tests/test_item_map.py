"""The final level's item numbering (fin_item_of, shared by the kernels and
the host): compiled for the host with hipcc and run over 2016 plan
geometries (first-level bits, second-level digits, owners, per-class packs
of the canonical sizing); every (bucket, digit) must be covered exactly once,
in order."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not available")
def test_items_cover_every_digit_once(tmp_path):
    exe = tmp_path / "item_map_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "-I", str(ROOT / "dakc_amd" / "csrc"),
                    str(ROOT / "tests" / "item_map_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout
    assert out.strip().endswith("0 bad"), out
