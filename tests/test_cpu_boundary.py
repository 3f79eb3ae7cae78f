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
