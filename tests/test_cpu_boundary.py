"""CPU-side checks of the boundary: the C-ABI library loads and exports every symbol the
header declares; host-side validation mirrors Knn::new."""
import os
import re
import subprocess

import pytest

import diskann_amd as dk
from diskann_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "diskann_mi355x.h")).read()
    pat = r"^\s*(?:const\s+)?(?:char|void|dkm_status)\s*\*?\s*(dkm_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(pat, txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    out = subprocess.check_output(["nm", "-D", "--defined-only", _native.LIB_PATH]).decode()
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    declared = header_symbols()
    assert declared, "no declarations parsed"
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    assert sorted(_native.EXPORTS) == declared
    for s in declared:
        getattr(lib, s)
    assert b"gfx950" in lib.dkm_version()


def test_knn_validation_matches_reference():
    with pytest.raises(dk.KnnSearchError):
        dk.Knn(0, 10)
    with pytest.raises(dk.KnnSearchError):
        dk.Knn(5, 0)
    with pytest.raises(dk.KnnSearchError):
        dk.Knn(100, 32)
    with pytest.raises(dk.KnnSearchError):
        dk.Knn(10, 10, 0)
    k = dk.Knn(10, 10)
    assert (k.k, k.L, k.beam_width) == (10, 10, 1)


def test_metric_numbering_matches_reference_repr_c():
    assert [int(m) for m in (dk.Metric.COSINE, dk.Metric.INNER_PRODUCT, dk.Metric.L2,
                             dk.Metric.COSINE_NORMALIZED)] == [0, 1, 2, 3]


def test_library_chunk_offsets_match_partition_kat():
    # ChunkOffsets::partition KAT (diskann-quantization/src/views.rs:528-548): host code of
    # the library, no device needed
    import json
    from diskann_amd import quantization as qz
    kats = json.load(open(os.path.join(ROOT, "tests", "golden", "kats.json")))
    for dim, m, exp in kats["partition"]:
        assert list(qz.chunk_offsets(dim, m)) == exp


def test_pq_train_draw_stream_is_documented():
    # the Python restatement of the library's SplitMix64 draw stream; pinned on the GPU
    # against dkm_pq_train's pivots (tests/test_gpu_pq_train.py)
    from diskann_amd import quantization as qz
    u = qz.train_draws(1, 3, 5)
    assert u.shape == (3, 5) and ((u >= 0) & (u < 1)).all()
    assert len(set(u.ravel().tolist())) == 15
