"""Spherical-quantizer test helpers: the committed 1-bit golden baselines
(tests/golden/spherical_1bit.json) and the serialized-quantizer reader."""
import json
import os

from diskann_amd.spherical import parse_serialized as parse_quantizer  # noqa: F401

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "spherical_1bit.json")
METRICS = {"squared_l2": 2, "inner_product": 1, "cosine": 0}   # oracle metric numbers


def load():
    return json.load(open(GOLDEN))
