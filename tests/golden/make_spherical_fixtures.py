"""Writes tests/golden/spherical_1bit.json from the reference's own test data: the 16 x 8
dataset of diskann-quantization/src/spherical/iface.rs (test_dataset) and the 1-bit
golden baselines diskann-quantization/test/generated/spherical/1bit_*_double_hadamard.json
(serialized quantizer, compressed vectors, data and query distances). Reads the
reference's files as text / JSON only; run from the repo root with the reference at
/root/reference."""
import json
import os
import re

REF = "/root/reference/diskann-quantization"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spherical_1bit.json")


def dataset():
    src = open(os.path.join(REF, "src/spherical/iface.rs")).read()
    body = src[src.index("fn test_dataset() -> Matrix<f32> {"):]
    body = body[body.index("vec![") + 5: body.index("];")]
    vals = [float(v) for v in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", re.sub(r"//.*", "", body))]
    assert len(vals) == 16 * 8, len(vals)
    return [vals[8 * r: 8 * r + 8] for r in range(16)]


def main():
    out = {"dataset": dataset(), "baselines": {}}
    for metric in ("squared_l2", "inner_product", "cosine"):
        name = "1bit_%s_double_hadamard.json" % metric
        d = json.load(open(os.path.join(REF, "test/generated/spherical", name)))
        keep = {k: d[k] for k in ("nbits", "metric", "transform", "pre_scale", "dim", "full_dim",
                                  "serialized_quantizer", "compressed_vectors",
                                  "data_distances", "query_distances")}
        if "rescaled_query_distances" in d:
            keep["rescaled_query_distances"] = d["rescaled_query_distances"]
        out["baselines"][metric] = keep
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
