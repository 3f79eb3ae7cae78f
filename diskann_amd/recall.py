"""k-recall@n as the reference benchmark scores it (diskann-benchmark-core/src/recall.rs:
146-300, `knn`): per query, the ground-truth set is the first min(len(gt row), k) ids,
extended by the following ground-truth ids whose distance equals the k-th one (ties at the
boundary count as hits); hits = |gt set & first n results|, capped at k; the mean over
queries is formed from integer hit counts grouped by their denominator."""
from collections import defaultdict

import numpy as np


def knn_recall(groundtruth, results, recall_k, recall_n, groundtruth_distances=None):
    if recall_k > recall_n:
        raise ValueError("recall_k (%d) > recall_n (%d)" % (recall_k, recall_n))
    gt = [np.asarray(r) for r in groundtruth]
    res = [np.asarray(r) for r in results]
    if len(gt) != len(res):
        raise ValueError("rows mismatch: %d results, %d ground-truth rows" % (len(res), len(gt)))
    gd = None
    if groundtruth_distances is not None:
        gd = [np.asarray(r) for r in groundtruth_distances]
        if len(gd) != len(gt) or any(len(a) != len(b) for a, b in zip(gd, gt)):
            raise ValueError("ground-truth distances do not match the ground truth")
    hits_by_k = defaultdict(int)
    scored = 0
    for i in range(len(res)):
        k = min(len(gt[i]), recall_k)
        if k == 0:
            continue
        want = set(gt[i][:k].tolist())
        if gd is not None:
            last = gd[i][k - 1]
            for d, g in zip(gd[i][k:], gt[i][k:]):
                if d == last:
                    want.add(int(g))
                else:
                    break
        got = set(res[i][:recall_n].tolist())
        hits_by_k[k] += min(len(want & got), k)
        scored += 1
    if scored == 0:
        return 0.0
    return float(sum(h / float(k * scored) for k, h in sorted(hits_by_k.items())))
