"""Friends-of-friends oracles (test infrastructure only).

* ``fof_bfs`` restates the reference's internal FoF, fof.py:18-79: seeds are visited in
  index order, a seed whose ball (scipy KDTree ``query_ball_point``, fof.py:40) holds only
  itself is skipped (fof.py:43), every friend is expanded breadth-first (fof.py:51-74), and
  each finished group is emitted as a sorted list (fof.py:76).  The result is the list of
  connected components of size >= 2 of the graph with an edge wherever scipy's FP64 test
  ((dx*dx + dy*dy) + dz*dz) <= ll*ll holds, ordered by minimum member index.
* ``fof_labels_scipy``: the same components through ``cKDTree.query_pairs`` and
  ``scipy.sparse.csgraph.connected_components`` (scipy is the reference's own third-party
  dependency for this arithmetic, fof.py:8,20); labels = minimum index of the component.
* ``fof_labels_brute``: O(N^2) numpy with the FP64 predicate written out, for tiny N.
* ``fof_parallel`` restates friends_of_friends_parallel, fof.py:95-176 (subdomain split
  with a 0.1 Mpc buffer, strict ``> minp`` size filter, centre-of-mass ownership test),
  with exact FoF inside each subdomain.
"""
from collections import deque

import numpy as np


def fof_bfs(data, linking_length):
    """fof.py:18-79 restated: list of sorted member lists (components of size >= 2)."""
    from scipy.spatial import KDTree

    data = np.asarray(data)
    tree = KDTree(data)
    n = data.shape[0]
    seen = np.zeros(n, dtype=bool)
    groups = []
    for seed in range(n):
        if seen[seed]:
            continue
        first = tree.query_ball_point(data[seed], linking_length)
        if len(first) <= 1:
            continue
        seen[seed] = True
        members = [seed]
        queue = deque()
        for f in first:
            if not seen[f]:
                seen[f] = True
                members.append(f)
                queue.append(f)
        while queue:
            p = queue.popleft()
            for f in tree.query_ball_point(data[p], linking_length):
                if not seen[f]:
                    seen[f] = True
                    members.append(f)
                    queue.append(f)
        groups.append(sorted(members))
    return groups


def _canonical(labels):
    """Map arbitrary component ids to the minimum member index of each component."""
    labels = np.asarray(labels)
    n = labels.shape[0]
    if n == 0:
        return np.zeros(0, np.int64)
    mins = np.full(labels.max() + 1, n, dtype=np.int64)
    np.minimum.at(mins, labels, np.arange(n, dtype=np.int64))
    return mins[labels]


def fof_labels_scipy(x, y, z, linking_length):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    from scipy.spatial import cKDTree

    data = np.stack([np.asarray(x, np.float64), np.asarray(y, np.float64),
                     np.asarray(z, np.float64)], axis=1)
    n = data.shape[0]
    if n == 0:
        return np.zeros(0, np.int64)
    pairs = cKDTree(data).query_pairs(linking_length, output_type="ndarray")
    g = coo_matrix((np.ones(len(pairs), np.int8), (pairs[:, 0], pairs[:, 1])),
                   shape=(n, n))
    _, comp = connected_components(g, directed=False)
    return _canonical(comp)


def fof_labels_brute(x, y, z, linking_length):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    n = x.shape[0]
    ll2 = float(linking_length) * float(linking_length)
    parent = np.arange(n)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i in range(n):
        dx = x[i] - x
        dy = y[i] - y
        dz = z[i] - z
        d2 = (dx * dx + dy * dy) + dz * dz
        for j in np.nonzero(d2 <= ll2)[0]:
            a, b = find(i), find(int(j))
            if a != b:
                parent[max(a, b)] = min(a, b)
    return np.array([find(i) for i in range(n)], dtype=np.int64)


def labels_to_groups(labels, min_size=2):
    """Groups of >= min_size members as sorted lists, ordered by minimum member index
    (the order fof.py:34 produces by visiting seeds in index order)."""
    labels = np.asarray(labels, np.int64)
    if labels.size == 0:
        return []
    order = np.argsort(labels, kind="stable")
    sl = labels[order]
    cuts = np.nonzero(np.diff(sl))[0] + 1
    out = []
    for seg in np.split(order, cuts):
        if len(seg) >= min_size:
            out.append(seg.tolist())
    return out


def fof_parallel(st_x, st_y, st_z, linking_length, L, minp, st_mass, ncores,
                 fof_fn=None):
    """fof.py:95-176 restated.  ``ncores`` must already be one of 1, 8, 64, 512
    (fof.py:98-99 picks the largest not above the thread count)."""
    from .reductions import center_of_mass

    if fof_fn is None:
        def fof_fn(d, ll):
            return labels_to_groups(fof_labels_scipy(d[:, 0], d[:, 1], d[:, 2], ll))
    st_x = np.asarray(st_x)
    st_y = np.asarray(st_y)
    st_z = np.asarray(st_z)
    nsides = int(np.round(ncores ** (1 / 3)))
    sub_L = L / nsides
    buffer = 0.1
    boxes = []
    for ix in range(nsides):
        for iy in range(nsides):
            for iz in range(nsides):
                lo = (-L / 2 + ix * sub_L, -L / 2 + iy * sub_L, -L / 2 + iz * sub_L)
                hi = (-L / 2 + (ix + 1) * sub_L, -L / 2 + (iy + 1) * sub_L,
                      -L / 2 + (iz + 1) * sub_L)
                boxes.append((lo, hi))
    allidx = np.arange(st_x.shape[0])
    final = []
    for lo, hi in boxes:
        mask = ((st_x >= lo[0] - buffer) & (st_x < hi[0] + buffer)
                & (st_y >= lo[1] - buffer) & (st_y < hi[1] + buffer)
                & (st_z >= lo[2] - buffer) & (st_z < hi[2] + buffer))
        sub = allidx[mask]
        d = np.stack([st_x[mask], st_y[mask], st_z[mask]], axis=1)
        for g in fof_fn(d, linking_length):
            if len(g) > minp:
                glob = sub[np.asarray(g, np.int64)]
                cx, cy, cz, _ = center_of_mass(glob, st_x, st_y, st_z, st_mass)
                if (lo[0] <= cx <= hi[0] and lo[1] <= cy <= hi[1] and lo[2] <= cz <= hi[2]):
                    final.append(glob)
    return final
