"""Oracle for the host label stage -- TEST INFRASTRUCTURE ONLY.

Restates, in plain Python + numpy with the reference's dtypes, the label
pipeline fed by the MST:

    hdbscan.clusters_from_spanning_tree             hdbscan.py:497-574
      np.lexsort by (weight, src, dst)              hdbscan.py:511-519
      cluster_trees.mst_to_linkage_tree             cluster_trees.py:17-95
      cluster_trees.condense_tree                   cluster_trees.py:130-352
      cluster_trees.cluster_tree_from_condensed_tree  673-683
      cluster_trees.extract_eom_clusters            cluster_trees.py:656-771
      cluster_trees.cluster_tree_leaves             cluster_trees.py:369-374
      cluster_trees.simplify_hierarchy              cluster_trees.py:774-831
      cluster_trees.cluster_epsilon_search          cluster_trees.py:834-903
      cluster_trees.get_cluster_label_vector        cluster_trees.py:944-996
      cluster_trees.get_point_membership_strength_vector  999-1031

Only tests/ and the bench's CPU-baseline leg may import it.

Pinned by (tests/test_labels_oracle.py): the reference's own known-answer
tests (test_hdbscan.py:169-186 label/noise counts at two epsilons, the
3-cluster blob counts, test_edge_cases.py) and label agreement with
sklearn.cluster.HDBSCAN on tie-free data (sklearn's min_samples counts the
point itself, so it is given min_samples + 1).  The float32 rounding of
lambdas and of the EOM scores follows the reference's dtypes; whether numba's
fastmath changes a last bit there is not pinned by any reference test.
"""
from __future__ import annotations

import ctypes

import numpy as np


def sort_mst(mst: np.ndarray) -> np.ndarray:
    """hdbscan.py:511-519: rows by weight, then source, then target."""
    return mst[np.lexsort((mst.T[1], mst.T[0], mst.T[2]))]


def mst_to_linkage_tree(sorted_mst: np.ndarray, sample_weights=None) -> np.ndarray:
    """cluster_trees.py:66-95 with create_linkage_merge_data (17-25),
    linkage_merge_find (38-52) and linkage_merge_join (55-62); with
    sample_weights, mst_to_linkage_tree_w_sample_weights (98-127): sizes
    are the weights concatenated with float32 zeros, summed in that dtype.
    Python's negative indexing matches numba's for a -1 endpoint."""
    m = sorted_mst.shape[0]
    n = m + 1
    parent = np.full(2 * n - 1, -1, dtype=np.intp)
    if sample_weights is None:
        size = np.concatenate((np.ones(n, dtype=np.intp), np.zeros(n - 1, dtype=np.intp)))
    else:
        size = np.concatenate((np.asarray(sample_weights), np.zeros(n - 1, dtype=np.float32)))
    nxt = n
    result = np.empty((m, 4))

    def find(node):
        relabel = node
        while parent[node] != -1 and parent[node] != node:
            node = parent[node]
        parent[node] = node
        while parent[relabel] != node:
            nr = parent[relabel]
            parent[relabel] = node
            relabel = nr
        return node

    for index in range(m):
        left = int(np.intp(sorted_mst[index, 0]))
        right = int(np.intp(sorted_mst[index, 1]))
        delta = sorted_mst[index, 2]
        lc = find(left)
        rc = find(right)
        if lc > rc:
            result[index][0] = lc
            result[index][1] = rc
        else:
            result[index][1] = lc
            result[index][0] = rc
        result[index][2] = delta
        result[index][3] = size[lc] + size[rc]
        size[nxt] = size[lc] + size[rc]
        parent[lc] = nxt
        parent[rc] = nxt
        nxt += 1
    return result


def bfs_from_hierarchy(hierarchy, bfs_root, num_points):
    """cluster_trees.py:98-112"""
    to_process = [bfs_root]
    result = []
    while to_process:
        result.extend(to_process)
        nxt = []
        for node in to_process:
            if node >= num_points:
                i = node - num_points
                nxt.append(int(hierarchy[i, 0]))
                nxt.append(int(hierarchy[i, 1]))
        to_process = nxt
    return result


class CondensedTree:
    def __init__(self, parent, child, lambda_val, child_size):
        self.parent = parent
        self.child = child
        self.lambda_val = lambda_val
        self.child_size = child_size


def condense_tree(hierarchy, min_cluster_size=10, max_cluster_size=np.inf, lambda_method="hdbscan",
                  sample_weights=None):
    """cluster_trees.py:205-352 with eliminate_branch (115-145), unit sample
    weights; lambda 1/d for "hdbscan" (192-197), exp(-d) for "plscan"
    (200-202), stored as float32."""
    root = 2 * hierarchy.shape[0]
    num_points = hierarchy.shape[0] + 1
    next_label = num_points + 1
    node_list = bfs_from_hierarchy(hierarchy, root, num_points)
    relabel = np.zeros(root + 1, dtype=np.int64)
    relabel[root] = num_points
    parents = np.ones(root, dtype=np.int64)
    children = np.empty(root, dtype=np.int64)
    lambdas = np.empty(root, dtype=np.float32)
    sizes = np.ones(root, dtype=np.float32)
    ignore = np.zeros(root + 1, dtype=np.bool_)
    if sample_weights is None:
        sample_weights = np.ones(num_points, dtype=np.float32)
    idx = 0

    def eliminate(branch_node, parent_node, lambda_value, idx):
        if branch_node < num_points:
            parents[idx] = parent_node
            children[idx] = branch_node
            lambdas[idx] = lambda_value
            return idx + 1
        for sub in bfs_from_hierarchy(hierarchy, branch_node, num_points):
            if sub < num_points:
                children[idx] = sub
                parents[idx] = parent_node
                lambdas[idx] = lambda_value
                idx += 1
            else:
                ignore[sub] = True
        return idx

    for node in node_list:
        if ignore[node] or node < num_points:
            continue
        parent_node = relabel[node]
        l, r, d, _ = hierarchy[node - num_points]
        left = np.int64(l)
        right = np.int64(r)
        if lambda_method == "hdbscan":
            lambda_value = (1.0 / d) if d > 0.0 else np.inf
        else:
            lambda_value = np.exp(-d)
        left_count = np.float32(hierarchy[left - num_points, 3]) if left >= num_points else sample_weights[left]
        right_count = (np.float32(hierarchy[right - num_points, 3]) if right >= num_points
                       else sample_weights[right])
        if left < num_points and right_count >= min_cluster_size:
            relabel[right] = parent_node
            parents[idx] = parent_node
            children[idx] = left
            lambdas[idx] = lambda_value
            idx += 1
        elif left_count < min_cluster_size and right_count >= min_cluster_size:
            relabel[right] = parent_node
            idx = eliminate(left, parent_node, lambda_value, idx)
        elif left_count >= min_cluster_size and right_count < min_cluster_size:
            relabel[left] = parent_node
            idx = eliminate(right, parent_node, lambda_value, idx)
        elif left_count < min_cluster_size and right_count < min_cluster_size:
            idx = eliminate(left, parent_node, lambda_value, idx)
            idx = eliminate(right, parent_node, lambda_value, idx)
        elif left_count > max_cluster_size and right_count > max_cluster_size:
            relabel[left] = parent_node
            relabel[right] = parent_node
        else:
            relabel[left] = next_label
            parents[idx] = parent_node
            children[idx] = next_label
            lambdas[idx] = lambda_value
            sizes[idx] = left_count
            next_label += 1
            idx += 1
            relabel[right] = next_label
            parents[idx] = parent_node
            children[idx] = next_label
            lambdas[idx] = lambda_value
            sizes[idx] = right_count
            next_label += 1
            idx += 1
    return CondensedTree(parents[:idx], children[:idx], lambdas[:idx], sizes[:idx])


def mask_tree(t, mask):
    return CondensedTree(t.parent[mask], t.child[mask], t.lambda_val[mask], t.child_size[mask])


def cluster_tree_from_condensed_tree(t):
    """cluster_trees.py:673-674"""
    return mask_tree(t, t.child_size > 1)


def score_condensed_tree_nodes(t):
    """cluster_trees.py:656-669: float32 accumulation in insertion order."""
    root = t.parent[0]
    result = {int(root): np.float32(0.0)}
    for i in range(t.parent.shape[0]):
        if t.child_size[i] > 1:
            result[int(t.child[i])] = np.float32(-t.lambda_val[i] * t.child_size[i])
        p = int(t.parent[i])
        result[p] = np.float32(result[p] + np.float32(t.lambda_val[i] * t.child_size[i]))
    return result


def extract_eom_clusters(t, ct, max_cluster_size=np.inf, allow_single_cluster=False):
    """cluster_trees.py:694-771"""
    node_scores = score_condensed_tree_nodes(t)
    if len(ct.parent) > 0:
        node_sizes = {int(c): np.float32(s) for c, s in zip(ct.child, ct.child_size.astype(np.float32))}
        node_sizes[int(ct.parent.min())] = np.float32(ct.parent.min() - 1)
    else:
        node_sizes = {-1: np.float32(0.0)}
    selected = {node: False for node in node_scores}
    if len(ct.parent) == 0:
        return np.zeros(0, dtype=np.int64)
    kids = {}
    for p, c in zip(ct.parent, ct.child):
        kids.setdefault(int(p), []).append(int(c))

    def unselect_below(node):
        for c in kids.get(node, []):
            unselect_below(c)
            selected[c] = False

    def eom(node):
        current_score = max(float(node_scores[node]), 0.0)
        current_size = node_sizes[node]
        child_total = 0.0
        for c in kids.get(node, []):
            child_total += eom(c)
        if child_total > current_score or current_size > max_cluster_size:
            return child_total
        selected[node] = True
        unselect_below(node)
        return current_score

    root = int(ct.parent.min())
    if allow_single_cluster:
        eom(root)
    elif len(node_scores) > 1:
        for c in kids.get(root, []):
            eom(c)
    return np.asarray([node for node, s in selected.items() if s], dtype=np.int64)


def cluster_tree_leaves(ct, n_points):
    """cluster_trees.py:369-374"""
    n_nodes = ct.child.max() + 1
    leaf = np.ones(n_nodes - n_points, dtype=np.bool_)
    leaf[ct.parent - n_points] = False
    return np.nonzero(leaf)[0] + n_points


def segments_in_branch(parents, children, segment):
    """cluster_trees.py:882-903, including its searchsorted lookup for
    segments that are not parents (it then picks up the pair of children of
    the next larger parent id)."""
    result = set()
    to_process = {int(segment)}
    while to_process:
        result |= to_process
        child_set = set()
        for seg in to_process:
            idx = int(np.searchsorted(parents, seg))
            if idx >= len(parents):
                continue
            child_set.add(int(children[idx]))
            child_set.add(int(children[idx + 1]))
        to_process = child_set
    return result


def traverse_upwards(ct, min_epsilon, root, segment):
    """cluster_trees.py:866-879"""
    parent = ct.parent[ct.child == segment][0]
    if parent == root:
        return int(root)
    parent_lam = ct.lambda_val[ct.child == parent][0]
    parent_death_eps = (1.0 / parent_lam) if parent_lam > 0.0 else np.inf
    if parent_death_eps >= min_epsilon:
        return int(parent)
    return traverse_upwards(ct, min_epsilon, root, parent)


def cluster_epsilon_search(clusters, ct, min_epsilon=0.0):
    """cluster_trees.py:834-863"""
    selected = []
    processed = set()
    order = np.argsort(ct.parent, kind="stable")
    parents = ct.parent[order]
    children = ct.child[order]
    root = ct.parent.min()
    for cluster in clusters:
        idx = int(np.searchsorted(ct.child, cluster))
        lam = ct.lambda_val[idx]
        death_eps = (1.0 / lam) if lam > 0.0 else np.inf
        if death_eps < min_epsilon:
            if int(cluster) not in processed:
                parent = traverse_upwards(ct, min_epsilon, root, cluster)
                selected.append(parent)
                processed |= segments_in_branch(parents, children, parent)
        else:
            selected.append(int(cluster))
    return np.asarray(selected, dtype=np.int64)


def get_cluster_label_vector(t, clusters, cluster_selection_epsilon, n_samples):
    """cluster_trees.py:944-996 with a union-by-rank disjoint set
    (disjoint_set.py:11-42)."""
    if len(t.parent) == 0 or len(clusters) == 0:
        return np.full(n_samples, -1, dtype=np.intp)
    root_cluster = t.parent.min()
    result = np.full(n_samples, -1, dtype=np.intp)
    label_map = {int(c): n for n, c in enumerate(np.sort(clusters))}
    size = max(t.parent.max() + 1, t.child.max() + 1)
    par = np.arange(size, dtype=np.int64)
    rank = np.zeros(size, dtype=np.int64)

    def find(x):
        while par[x] != x:
            x = par[x]
        return x

    def union(x, y):
        x, y = find(x), find(y)
        if x == y:
            return
        if rank[x] < rank[y]:
            x, y = y, x
        par[y] = x
        if rank[x] == rank[y]:
            rank[x] += 1

    cset = set(int(c) for c in clusters)
    for n in range(t.parent.shape[0]):
        child = int(t.child[n])
        parent = int(t.parent[n])
        if child not in cset:
            if child < n_samples and t.lambda_val[n] == 0.0:
                continue
            union(parent, child)
    for n in range(root_cluster):
        cluster = find(n)
        if cluster < root_cluster:
            result[n] = -1
        elif cluster == root_cluster:
            if len(cset) == 1:
                max_lambda = t.lambda_val[t.parent == cluster].max()
                cur_lambda = t.lambda_val[t.child == n]
                if cluster_selection_epsilon > 0.0:
                    result[n] = label_map[int(cluster)] if cur_lambda >= 1 / cluster_selection_epsilon else -1
                elif cur_lambda >= max_lambda:
                    result[n] = label_map[int(cluster)]
                else:
                    result[n] = -1
            else:
                result[n] = -1
        else:
            result[n] = label_map[int(cluster)]
    return result


def get_point_membership_strength_vector(t, clusters, labels):
    """cluster_trees.py:999-1031"""
    result = np.zeros(labels.shape[0], dtype=np.float32)
    cset = set(int(c) for c in clusters)
    deaths = {c: 0.0 for c in cset}
    for n in range(t.parent.shape[0]):
        c = int(t.parent[n])
        if c in cset and t.child_size[n] == 1:
            deaths[c] = max(deaths[c], float(t.lambda_val[n]))
    root_cluster = t.parent.min()
    index_map = {n: int(c) for n, c in enumerate(np.sort(clusters))}
    for n in range(t.child.shape[0]):
        point = int(t.child[n])
        if point >= root_cluster or labels[point] < 0:
            continue
        cluster = index_map[int(labels[point])]
        max_lambda = deaths[cluster]
        if max_lambda == 0.0 or not np.isfinite(t.lambda_val[n]):
            result[point] = 1.0
        else:
            lam = min(float(t.lambda_val[n]), max_lambda)
            result[point] = lam / max_lambda
    return result


def clusters_from_spanning_tree(mst, min_cluster_size=10, cluster_selection_method="eom",
                                max_cluster_size=np.inf, allow_single_cluster=False,
                                cluster_selection_epsilon=0.0, sample_weights=None):
    """hdbscan.py:497-574, not semi-supervised, without persistence
    simplification.  Returns (labels, membership_strengths, linkage_tree,
    condensed_tree, sorted_mst)."""
    n_points = mst.shape[0] + 1
    sorted_mst = sort_mst(mst)
    linkage = mst_to_linkage_tree(sorted_mst, sample_weights)
    ct_full = condense_tree(linkage, min_cluster_size=min_cluster_size, sample_weights=sample_weights)
    ct = cluster_tree_from_condensed_tree(ct_full)
    if cluster_selection_method == "eom":
        selected = extract_eom_clusters(ct_full, ct, max_cluster_size=max_cluster_size,
                                        allow_single_cluster=allow_single_cluster)
    elif cluster_selection_method == "leaf":
        selected = np.empty(0, dtype=np.int64) if ct.parent.shape[0] == 0 else cluster_tree_leaves(ct, n_points)
    else:
        raise ValueError(f"Invalid cluster_selection_method {cluster_selection_method}")
    if len(selected) > 1 and cluster_selection_epsilon > 0.0:
        selected = cluster_epsilon_search(selected, ct, min_epsilon=cluster_selection_epsilon)
    labels = get_cluster_label_vector(ct_full, selected, cluster_selection_epsilon, n_samples=n_points)
    strengths = get_point_membership_strength_vector(ct_full, selected, labels)
    return labels, strengths, linkage, ct_full, sorted_mst


def clusters_from_spanning_tree_c(mst, min_cluster_size=10, cluster_selection_method="eom",
                                  max_cluster_size=np.inf, allow_single_cluster=False):
    """The same stage compiled (oracle/labels_c.c, one thread): unit sample
    weights, cluster_selection_epsilon 0.  Returns (labels, strengths,
    linkage, condensed tree)."""
    from . import oracle as O

    mst = np.ascontiguousarray(mst, dtype=np.float64)
    n = mst.shape[0] + 1
    labels = np.empty(n, np.int64)
    strengths = np.empty(n, np.float32)
    linkage = np.empty((max(n - 1, 1), 4), np.float64)
    cap = 2 * n
    ct = (np.empty(cap, np.int64), np.empty(cap, np.int64), np.empty(cap, np.float32), np.empty(cap, np.float32))
    rows = ctypes.c_int64(0)
    method = {"eom": 0, "leaf": 1}[cluster_selection_method]
    r = O.lib().orc_labels(n, mst, int(min_cluster_size), float(max_cluster_size), method,
                           1 if allow_single_cluster else 0, labels, strengths, linkage, *ct, ctypes.byref(rows))
    if r < 0:
        raise ValueError("invalid arguments")
    k = rows.value
    return labels, strengths, linkage[:n - 1], CondensedTree(ct[0][:k], ct[1][:k], ct[2][:k], ct[3][:k])
