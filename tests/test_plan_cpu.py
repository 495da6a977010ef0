"""CPU tests of the job planner (dakc_plan_job: host only, no device call):
the MSD geometry never splits more key bits than a k-mer has, owner bins
counted all at once stay on the MSD path at genome scale, and the second
level's groups fit what the final level can gather."""
import pytest

from dakc_amd import _lib


def job(nreads, L, k, P=1, b=-1, canonical=False):
    # a device pointer is only recorded, never read, by the planner
    return _lib.Job(reads=0x1000, nbytes=nreads * (L + 1), format=_lib.FMT_FIXED, read_len=L, k=k,
                    canonical=1 if canonical else 0, num_pes=P, owner_bin=b, reads_on_device=1,
                    table_on_device=0, hash_seed=0)


@pytest.mark.parametrize("nreads", [1_000, 100_000, 3_000_000, 20_000_000])
@pytest.mark.parametrize("P,b", [(1, -1), (64, -1), (64, 0), (256, -1), (7, -1)])
def test_msd_geometry_within_key_bits(nreads, P, b):
    for L in (150, 250):
        for k in range(1, 65):
            if nreads * max(L - k + 1, 0) >= 2**32:
                continue
            for canonical in (False, True):
                pl = _lib.plan_job(job(nreads, L, k, P, b, canonical))
                if pl["path"] != 2:
                    continue
                assert pl["b1"] + pl["b2"] + pl["b3"] <= 2 * k, (L, k, pl)
                assert pl["first_level_buckets"] <= 1024
                assert 1 <= pl["l2_groups"] <= 1024
                assert pl["final_items"] >= 1 and pl["final_threads"] in (256, 512, 1024)


def test_small_k_at_scale_takes_the_lsd_path():
    # k = 1..3: at most 64 distinct keys; no split fits their counts in LDS
    for k in (1, 2, 3):
        assert _lib.plan_job(job(1_000_000, 150, k))["path"] == 1


@pytest.mark.parametrize("P", [4, 8, 64, 256])
def test_all_owner_bins_at_genome_scale_stay_msd(P):
    # 20M reads x 150, k = 31, every owner bin at once: the owner digit splits
    # the k-mers P ways, so the key bits per owner block shrink by log2(P)
    pl = _lib.plan_job(job(20_000_000, 150, 31, P, -1))
    assert pl["path"] == 2, pl
    keys_per_item = 20_000_000 * 120 / pl["final_items"]
    assert keys_per_item < 15360  # fits a 1024-thread final workgroup's LDS


def test_plan_reports_the_lsd_path_on_request():
    pl = _lib.plan_job(job(1_000_000, 150, 31), algo=_lib.ALGO_LSD)
    assert pl["path"] == 1 and pl["radix_passes"] == 8 and pl["device_bytes"] > 2 * 120_000_000 * 8


def test_plan_rejects_bad_jobs():
    with pytest.raises(_lib.DakcError):
        _lib.plan_job(job(10, 150, 0))
    with pytest.raises(_lib.DakcError):
        _lib.plan_job(job(40_000_000, 150, 21))  # 2^32 or more window slots
