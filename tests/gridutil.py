"""Synthetic grid workload of the reference's graph tests
(diskann/src/graph/test/synthetic.rs:12-410, Provider::grid in
diskann/src/graph/test/provider.rs:323-338).

Rows are generated with the last coordinate varying fastest; nodes are connected to
their neighbours along every axis (lower then upper neighbour per axis, axes in order).
The start point has id u32::MAX in the reference; here it is stored as the extra row
N (internal ids are positions, so results never contain it), holds the vector
[size]*dim and has the single neighbour N-1.
"""
import itertools

import numpy as np


def grid(dim, size):
    pts = np.array(list(itertools.product(range(size), repeat=dim)), dtype=np.float32)
    n = pts.shape[0]
    max_degree = 2 * dim
    adj = np.zeros((n + 1, max_degree), dtype=np.uint32)
    deg = np.zeros(n + 1, dtype=np.uint32)
    strides = [size ** (dim - 1 - a) for a in range(dim)]
    for i in range(n):
        c = pts[i].astype(np.int64)
        lst = []
        for a in range(dim):
            if c[a] > 0:
                lst.append(i - strides[a])
            if c[a] < size - 1:
                lst.append(i + strides[a])
        adj[i, :len(lst)] = lst
        deg[i] = len(lst)
    vectors = np.vstack([pts, np.full((1, dim), float(size), dtype=np.float32)])
    adj[n, 0] = n - 1
    deg[n] = 1
    return vectors, adj, deg, np.array([n], dtype=np.uint32)
