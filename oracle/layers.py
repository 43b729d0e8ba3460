"""Oracle for the PLSCAN layer stage -- TEST INFRASTRUCTURE ONLY.

Restates, in plain Python + numpy with the reference's dtypes, what
layer_clusters.build_cluster_layers (layer_clusters.py:353-495) does with
the MST that compute_minimum_spanning_tree returns:

    edge normalisation by the core distances     layer_clusters.py:387-397
    np.argsort of the weights (NumPy's, unstable) layer_clusters.py:398
    mst_to_linkage_tree                          cluster_trees.py:66-95
    _binary_search_for_n_clusters                layer_clusters.py:71-153
    condense_tree(lambda_method="plscan")        cluster_trees.py:200-352
    extract_leaves                               cluster_trees.py:356-366
    min_cluster_size_barcode                     layer_clusters.py:183-212
    compute_total_persistence                    layer_clusters.py:215-248
    find_peaks                                   layer_clusters.py:26-68
    estimate_cluster_similarity / jaccard        layer_clusters.py:265-306
    select_diverse_peaks                         layer_clusters.py:309-350
    extract_clusters_by_id                       layer_clusters.py:251-262
    _build_cluster_tree                          layer_clusters.py:498-525

The functions compiled by numba call numba's own ``np.argsort``
(quicksort; numba/misc/quicksort.py, numba >= 0.56, not vendored by the
reference): ``numba_argsort`` restates that published algorithm -- median
of three, Hoare-style partition with the pivot parked at the end, the
larger side pushed, insertion sort below 15 elements.

Pinned by (tests/test_layers.py): the reference's own PLSCAN tests
(test_plscan.py: 3 clusters on the 300-point blobs, one layer with
max_layers=1, base_n_clusters=4 within one, layer ordering, tree coverage,
strengths in [0, 1]); find_peaks against scipy.signal.find_peaks (whose
algorithm it is derived from); numba_argsort against a stable sort where
no ties exist and against insertion sort below 15 elements, and by a hand
trace of run_quicksort on 16 keys with ties (partition, pushed side,
insertion sorts); compute_total_persistence against its definition summed
in float64.
"""
from __future__ import annotations

import numpy as np

from . import labels as L

SMALL_QUICKSORT = 15


def _lt_float(a, b):
    # numba's argsort of floats orders NaN last
    return a < b or (np.isnan(b) and not np.isnan(a))


def numba_argsort(A, is_float=None):
    """numba's np.argsort(kind="quicksort") (numba/misc/quicksort.py,
    run_quicksort1 + partition + insertion_sort)."""
    A = np.asarray(A)
    if is_float is None:
        is_float = np.issubdtype(A.dtype, np.floating)
    lt = _lt_float if is_float else (lambda a, b: a < b)
    vals = A.tolist()
    R = list(range(len(vals)))
    if len(vals) < 2:
        return np.asarray(R, dtype=np.intp)

    def insertion_sort(low, high):
        if high <= low:
            return
        for i in range(low + 1, high + 1):
            k = R[i]
            v = vals[k]
            j = i
            while j > low and lt(v, vals[R[j - 1]]):
                R[j] = R[j - 1]
                j -= 1
            R[j] = k

    def partition(low, high):
        mid = (low + high) >> 1
        if lt(vals[R[mid]], vals[R[low]]):
            R[low], R[mid] = R[mid], R[low]
        if lt(vals[R[high]], vals[R[mid]]):
            R[high], R[mid] = R[mid], R[high]
        if lt(vals[R[mid]], vals[R[low]]):
            R[low], R[mid] = R[mid], R[low]
        pivot = vals[R[mid]]
        R[high], R[mid] = R[mid], R[high]
        i = low
        j = high - 1
        while True:
            while i < high and lt(vals[R[i]], pivot):
                i += 1
            while j >= low and lt(pivot, vals[R[j]]):
                j -= 1
            if i >= j:
                break
            R[i], R[j] = R[j], R[i]
            i += 1
            j -= 1
        R[i], R[high] = R[high], R[i]
        return i

    stack = [(0, len(vals) - 1)]
    while stack:
        low, high = stack.pop()
        while high - low >= SMALL_QUICKSORT:
            i = partition(low, high)
            if high - i > i - low:
                if high > i:
                    stack.append((i + 1, high))
                high = i - 1
            else:
                if i > low:
                    stack.append((low, i - 1))
                low = i + 1
        insertion_sort(low, high)
    return np.asarray(R, dtype=np.intp)


def find_peaks(x):
    """layer_clusters.py:26-68 (scipy's plateau-aware local maxima)."""
    mids = []
    i = 1
    i_max = x.shape[0] - 1
    while i < i_max:
        if x[i - 1] < x[i]:
            ia = i + 1
            while ia < i_max and x[ia] == x[i]:
                ia += 1
            if x[ia] < x[i]:
                mids.append((i + ia - 1) // 2)
                i = ia
        i += 1
    return np.asarray(mids, dtype=np.intp)


def extract_leaves(t):
    """cluster_trees.py:356-366"""
    n_nodes = int(t.parent.max()) + 1
    n_points = int(t.parent.min())
    leaf = np.ones(n_nodes, dtype=np.bool_)
    leaf[:n_points] = False
    for p, s in zip(t.parent, t.child_size):
        if s > 1:
            leaf[p] = False
    return np.nonzero(leaf)[0]


def min_cluster_size_barcode(ct, n_points, min_size):
    """layer_clusters.py:183-212"""
    n_nodes = int(ct.child[-1]) - n_points + 1
    parents = np.empty(n_nodes, dtype=np.int32)
    lambda_deaths = np.empty(n_nodes, dtype=np.float32)
    size_deaths = np.empty(n_nodes, dtype=np.float32)
    size_births = np.full(n_nodes, min_size, dtype=np.float32)
    lambda_deaths[0] = 0
    size_deaths[0] = n_points
    parents[0] = n_points
    n_rows = ct.child.shape[0]
    for idx in range(n_rows - 1, 0, -2):
        o = int(ct.child[idx]) - n_points
        parents[o - 1:o + 1] = ct.parent[idx]
        lv = ct.lambda_val[idx]
        if lv > 0.0:
            lambda_deaths[o - 1:o + 1] = np.exp(-1.0 / np.float64(lv))
        else:
            lambda_deaths[o - 1:o + 1] = 0.0
        death_size = ct.child_size[idx - 1:idx + 1].min()
        size_deaths[o - 1:o + 1] = death_size
        size_births[int(ct.parent[idx]) - n_points] = max(size_births[o - 1], size_births[o], death_size)
    return size_births, size_deaths, parents, lambda_deaths


def compute_total_persistence(births, deaths, lambda_deaths):
    """layer_clusters.py:215-248, float32 arithmetic throughout."""
    sizes = np.unique(births)
    tp = np.zeros(sizes.shape[0], dtype=np.float32)
    for i in range(1, len(births)):
        b = births[i]
        d = deaths[i]
        ld = lambda_deaths[i]
        if d <= b:
            continue
        bi = 0
        for j in range(len(sizes)):
            if sizes[j] >= b:
                bi = j
                break
        di = len(sizes)
        for j in range(len(sizes)):
            if sizes[j] >= d:
                di = j
                break
        v = np.float32(d - np.float32(b * ld))
        for k in range(bi, di):
            tp[k] = np.float32(tp[k] + v)
    return sizes, tp


def estimate_cluster_similarity(births, deaths, birth_a, birth_b):
    """layer_clusters.py:265-306"""
    a = set(i for i in range(len(births)) if births[i] <= birth_a and deaths[i] > birth_a)
    b = [i for i in range(len(births)) if births[i] <= birth_b and deaths[i] > birth_b]
    inter = sum(1 for i in b if i in a)
    union = len(a | set(b))
    return inter / union if union > 0 else 0.0


def select_diverse_peaks(peaks, tp, sizes, births, deaths, min_similarity_threshold=0.2, max_layers=10):
    """layer_clusters.py:309-350"""
    if len(peaks) == 0:
        return np.empty(0, dtype=np.int64)
    order = numba_argsort(tp[peaks], is_float=True)[::-1]
    sorted_peaks = peaks[order]
    sel, sel_births = [], []
    for peak in sorted_peaks:
        if len(sel) >= max_layers:
            break
        birth_size = sizes[peak]
        diverse = True
        for sb in sel_births:
            if estimate_cluster_similarity(births, deaths, birth_size, sb) > min_similarity_threshold:
                diverse = False
                break
        if diverse:
            sel.append(int(peak))
            sel_births.append(float(birth_size))
    return np.asarray(sel, dtype=np.int64)


def _leaf_labels(tree, leaves, n_samples):
    clusters = L.get_cluster_label_vector(tree, leaves, 0.0, n_samples)
    strengths = L.get_point_membership_strength_vector(tree, leaves, clusters)
    return clusters, strengths


def binary_search_for_n_clusters(linkage, approx_n_clusters, n_samples):
    """layer_clusters.py:71-153 (hdbscan lambdas, Python's round())."""
    lo, hi = 2, n_samples // 2

    def count(mcs):
        return len(extract_leaves(L.condense_tree(linkage, mcs)))

    hi_n = count(hi)
    lo_n = count(lo)
    while hi - lo > 1:
        mid = int(round((lo + hi) / 2.0))
        if mid == lo or mid == hi:
            break
        mid_n = count(mid)
        if mid_n < approx_n_clusters:
            hi, hi_n = mid, mid_n
        elif mid_n >= approx_n_clusters:
            lo, lo_n = mid, mid_n
    if abs(lo_n - approx_n_clusters) < abs(hi_n - approx_n_clusters):
        t = L.condense_tree(linkage, lo)
        leaves = extract_leaves(t)
        return (leaves, *_leaf_labels(t, leaves, n_samples))
    if abs(lo_n - approx_n_clusters) > abs(hi_n - approx_n_clusters):
        t = L.condense_tree(linkage, hi)
        leaves = extract_leaves(t)
        return (leaves, *_leaf_labels(t, leaves, n_samples))
    tl = L.condense_tree(linkage, lo)
    ll = extract_leaves(tl)
    cl = L.get_cluster_label_vector(tl, ll, 0.0, n_samples)
    tu = L.condense_tree(linkage, hi)
    lu = extract_leaves(tu)
    cu = L.get_cluster_label_vector(tu, lu, 0.0, n_samples)
    if np.sum(cl >= 0) > np.sum(cu >= 0):
        return ll, cl, L.get_point_membership_strength_vector(tl, ll, cl)
    return lu, cu, L.get_point_membership_strength_vector(tu, lu, cu)


def normalize_edges(edges, core_distances):
    """layer_clusters.py:387-397 (in place on a copy)."""
    edges = edges.copy()
    finite_core = core_distances[np.isfinite(core_distances)]
    if len(finite_core) > 0:
        mean_core = np.mean(finite_core)
        min_core = np.min(finite_core)
    else:
        mean_core = 0.0
        min_core = 0.0
    if mean_core > min_core and np.isfinite(mean_core):
        edges[:, 2] -= min_core
        edges[:, 2] /= mean_core - min_core
    return edges


def cluster_layers_from_mst(edges, core_distances, n_samples, base_min_cluster_size=10, base_n_clusters=None,
                            layer_similarity_threshold=0.2, max_layers=10):
    """layer_clusters.py:386-495 after the MST.  Returns the five values
    of build_cluster_layers plus, for tests, the normalised sorted MST."""
    edges = normalize_edges(edges, core_distances)
    sorted_mst = edges[np.argsort(edges.T[2])]
    linkage = L.mst_to_linkage_tree(sorted_mst)
    mcs = base_min_cluster_size
    if base_n_clusters is not None:
        leaves, clusters, strengths = binary_search_for_n_clusters(linkage, base_n_clusters, n_samples)
        cs = np.bincount(clusters[clusters >= 0])
        mcs = max(1, np.min(cs)) if len(cs) > 0 else base_min_cluster_size
        tree = L.condense_tree(linkage, mcs, lambda_method="plscan")
    else:
        tree = L.condense_tree(linkage, base_min_cluster_size, lambda_method="plscan")
        leaves = extract_leaves(tree)
        clusters, strengths = _leaf_labels(tree, leaves, n_samples)
    ct = L.mask_tree(tree, tree.child >= n_samples)
    if len(ct.child) > 0 and ct.child[-1] >= n_samples:
        births, deaths, parents, lambda_deaths = min_cluster_size_barcode(ct, n_samples, mcs)
        sizes, tp = compute_total_persistence(births, deaths, lambda_deaths)
        peaks = find_peaks(tp)
    else:
        births = deaths = sizes = tp = np.array([])
        peaks = np.array([], dtype=np.int64)
    layers, strength_layers, scores = [clusters], [strengths], [0.0]
    for peak in select_diverse_peaks(peaks, tp, sizes, births, deaths, layer_similarity_threshold, max_layers - 1):
        best = sizes[peak]
        sel = np.where((births <= best) & (deaths > best))[0] + n_samples
        lab = L.get_cluster_label_vector(tree, sel, 0.0, int(tree.parent[0]))
        layers.append(lab)
        strength_layers.append(L.get_point_membership_strength_vector(tree, sel, lab))
        scores.append(tp[peak])
    order = np.argsort([layer.max() + 1 for layer in layers])[::-1]
    return ([layers[i] for i in order], [strength_layers[i] for i in order], [scores[i] for i in order], sizes, tp,
            sorted_mst)


def build_layer_cluster_tree(layers):
    """layer_clusters.py:498-554"""
    labels = np.vstack(layers)
    mapping = []
    found = [{-1} for _ in range(labels.shape[0])]
    for up in range(1, labels.shape[0]):
        uniq = np.unique(labels[up])
        order = numba_argsort(labels[up], is_float=False)
        for lo in range(up - 1, -1, -1):
            groups = np.split(labels[lo][order], np.cumsum(np.bincount(labels[up] + 1))[:-1])
            if len(groups) > uniq.shape[0]:
                groups = groups[1:]
            for i, label in enumerate(uniq):
                if label >= 0:
                    for child in groups[i]:
                        if child >= 0 and child not in found[lo]:
                            mapping.append((up, int(label), lo, int(child)))
                            found[lo].add(int(child))
    for lo in range(labels.shape[0] - 1, -1, -1):
        for child in range(labels[lo].max() + 1):
            if child not in found[lo]:
                mapping.append((labels.shape[0], 0, lo, child))
    result = {}
    for pl, pc, cl, cc in mapping:
        result.setdefault((pl, pc), []).append((cl, cc))
    return result
