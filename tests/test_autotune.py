"""Autotuner bookkeeping (CPU): candidate sets and the JSON cache."""
from sph_fem_2d_cutting_gpu_amd import _lib, autotune


def test_candidate_sets():
    c32 = autotune.candidates(_lib.FP32C)
    assert len(c32) == 27
    assert {"pair_block": 256, "pair_chunk": 32, "update_block": 256, "pair_kernel": _lib.PAIRS_LIST,
            "resort_one_block": 0} in c32
    assert {c["pair_chunk"] for c in c32} == {16, 32, 48}
    staged = [c for c in c32 if c["pair_kernel"] == _lib.PAIRS_STAGED]
    assert len(staged) == 6 and {c["pair_chunk"] for c in staged} == {16, 32, 48}
    # the fused kernels run the update themselves: one candidate per pair-kernel shape
    fused = [c for c in c32 if c["pair_kernel"] == _lib.PAIRS_LIST_FUSED]
    assert len(fused) == 6 and {(c["pair_block"], c["pair_chunk"]) for c in fused} == {
        (b, k) for b in (128, 256) for k in (16, 32, 48)}
    sfused = [c for c in c32 if c["pair_kernel"] == _lib.PAIRS_STAGED_FUSED]
    assert len(sfused) == 3 and all(c["pair_block"] == 256 for c in sfused)
    c64 = autotune.candidates(_lib.FP64)     # the list chunk is an FP32C parameter only
    assert len(c64) == 4 and all(c["pair_chunk"] == 32 and c["pair_kernel"] == _lib.PAIRS_LIST for c in c64)
    assert {c["pair_block"] for c in c64} == {128, 256}


def test_cache_roundtrip(tmp_path):
    path = str(tmp_path / "tuning.json")
    key = autotune.cache_key(4_000_000, _lib.FP32C, "gfx950")
    assert key == "gfx950|fp32c|2^21"
    autotune.save_cache(path, {key: {"pair_block": 128, "pair_chunk": 48, "update_block": 256}})
    assert autotune.load_cache(path)[key]["pair_chunk"] == 48
    assert autotune.load_cache(str(tmp_path / "absent.json")) == {}
