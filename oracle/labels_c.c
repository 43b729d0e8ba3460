/*
 * Compiled CPU restatement of the host label stage -- TEST INFRASTRUCTURE.
 *
 * The same steps as oracle/labels.py, in plain C with one thread, the way
 * the reference runs them compiled by numba: it stands for the reference's
 * label stage in the CPU baseline, where the Python restatement would
 * understate it.  Unit sample weights, cluster_selection_epsilon = 0, no
 * persistence simplification:
 *
 *   np.lexsort of the MST rows by (weight, source, target)  hdbscan.py:511-519
 *   mst_to_linkage_tree (+ linkage_merge_find / _join)     cluster_trees.py:17-95
 *   bfs_from_hierarchy, eliminate_branch, condense_tree    cluster_trees.py:130-352
 *   cluster_tree_from_condensed_tree                       cluster_trees.py:673-674
 *   score_condensed_tree_nodes, eom_recursion,
 *   unselect_below_node, extract_eom_clusters             cluster_trees.py:656-771
 *   cluster_tree_leaves ("leaf" selection)                 cluster_trees.py:369-374
 *   get_cluster_label_vector (disjoint_set.py:11-42)       cluster_trees.py:944-996
 *   max_lambdas, get_point_membership_strength_vector      cluster_trees.py:999-1031
 *
 * Dtypes follow the reference: float64 linkage rows, float32 lambdas and
 * sizes in the condensed tree, float32 node scores summed in row order,
 * float64 sums in the EOM recursion and float64 strength ratios stored as
 * float32.  Pinned against oracle/labels.py row for row (condensed tree,
 * labels, strengths) in tests/test_labels_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const double *w;
    const double *s;
    const double *d;
} sort_keys;

static const sort_keys *g_keys;

/* np.lexsort((dst, src, w)): by weight, then source, then target; stable */
static int cmp_rows(const void *a, const void *b) {
    const int64_t i = *(const int64_t *)a, j = *(const int64_t *)b;
    const sort_keys *k = g_keys;
    if (k->w[i] != k->w[j]) return k->w[i] < k->w[j] ? -1 : 1;
    if (k->s[i] != k->s[j]) return k->s[i] < k->s[j] ? -1 : 1;
    if (k->d[i] != k->d[j]) return k->d[i] < k->d[j] ? -1 : 1;
    return i < j ? -1 : (i > j);
}

/* linkage_merge_find: root, then the path relabelled to it.  A node -1
 * (the MST's endpoint when min_samples >= n) keeps its value and indexes
 * the last entry, as numba's wrap-around indexing does. */
static int64_t merge_find(int64_t *parent, int64_t len, int64_t node) {
#define AT(x) parent[(x) < 0 ? (x) + len : (x)]
    int64_t relabel = node;
    while (AT(node) != -1 && AT(node) != node) node = AT(node);
    AT(node) = node;
    while (AT(relabel) != node) {
        const int64_t next = AT(relabel);
        AT(relabel) = node;
        relabel = next;
    }
    return node;
#undef AT
}

typedef struct {
    int64_t *parent, *child;
    float *lambda, *size;
    int64_t rows;
} ctree;

/* bfs_from_hierarchy as a FIFO walk: the same order as its level lists */
static int64_t bfs(const double *H, int64_t root, int64_t num_points, int64_t *out) {
    int64_t head = 0, tail = 0;
    out[tail++] = root;
    while (head < tail) {
        const int64_t node = out[head++];
        if (node >= num_points) {
            const int64_t i = node - num_points;
            out[tail++] = (int64_t)H[4 * i];
            out[tail++] = (int64_t)H[4 * i + 1];
        }
    }
    return tail;
}

static void emit(ctree *t, int64_t parent, int64_t child, float lam, float size) {
    t->parent[t->rows] = parent;
    t->child[t->rows] = child;
    t->lambda[t->rows] = lam;
    t->size[t->rows] = size;
    t->rows++;
}

static void eliminate_branch(ctree *t, const double *H, int64_t branch, int64_t parent_node, float lam,
                             int64_t num_points, unsigned char *ignore, int64_t *scratch) {
    if (branch < num_points) {
        emit(t, parent_node, branch, lam, 1.0f);
        return;
    }
    const int64_t m = bfs(H, branch, num_points, scratch);
    for (int64_t q = 0; q < m; ++q) {
        const int64_t sub = scratch[q];
        if (sub < num_points) emit(t, parent_node, sub, lam, 1.0f);
        else ignore[sub] = 1;
    }
}

static void condense(const double *H, int64_t n, int64_t mcs, double max_cluster_size, ctree *t) {
    const int64_t m = n - 1, root = 2 * m, num_points = n;
    int64_t next_label = num_points + 1;
    int64_t *node_list = malloc(sizeof(int64_t) * (size_t)(root + 1));
    int64_t *scratch = malloc(sizeof(int64_t) * (size_t)(root + 1));
    int64_t *relabel = calloc((size_t)(root + 1), sizeof(int64_t));
    unsigned char *ignore = calloc((size_t)(root + 1), 1);
    const int64_t count = bfs(H, root, num_points, node_list);
    relabel[root] = num_points;
    t->rows = 0;
    for (int64_t q = 0; q < count; ++q) {
        const int64_t node = node_list[q];
        if (ignore[node] || node < num_points) continue;
        const int64_t parent_node = relabel[node];
        const double *row = H + 4 * (node - num_points);
        const int64_t left = (int64_t)row[0], right = (int64_t)row[1];
        const double d = row[2];
        const float lam = d > 0.0 ? (float)(1.0 / d) : INFINITY;
        const float lc = left >= num_points ? (float)H[4 * (left - num_points) + 3] : 1.0f;
        const float rc = right >= num_points ? (float)H[4 * (right - num_points) + 3] : 1.0f;
        const double M = (double)mcs;
        if (left < num_points && rc >= M) {
            relabel[right] = parent_node;
            emit(t, parent_node, left, lam, 1.0f);
        } else if (lc < M && rc >= M) {
            relabel[right] = parent_node;
            eliminate_branch(t, H, left, parent_node, lam, num_points, ignore, scratch);
        } else if (lc >= M && rc < M) {
            relabel[left] = parent_node;
            eliminate_branch(t, H, right, parent_node, lam, num_points, ignore, scratch);
        } else if (lc < M && rc < M) {
            eliminate_branch(t, H, left, parent_node, lam, num_points, ignore, scratch);
            eliminate_branch(t, H, right, parent_node, lam, num_points, ignore, scratch);
        } else if (lc > max_cluster_size && rc > max_cluster_size) {
            relabel[left] = parent_node;
            relabel[right] = parent_node;
        } else {
            relabel[left] = next_label;
            emit(t, parent_node, next_label, lam, lc);
            next_label++;
            relabel[right] = next_label;
            emit(t, parent_node, next_label, lam, rc);
            next_label++;
        }
    }
    free(node_list);
    free(scratch);
    free(relabel);
    free(ignore);
}

/* the cluster tree (rows with child_size > 1) as child lists in row order */
typedef struct {
    int64_t *first, *next_sib, *kid, *last;  /* per node: first/last kid entry; per entry: next */
    int64_t entries;
} kids_t;

static void add_kid(kids_t *k, int64_t parent, int64_t child) {
    const int64_t e = k->entries++;
    k->kid[e] = child;
    k->next_sib[e] = -1;
    if (k->first[parent] < 0) k->first[parent] = e;
    else k->next_sib[k->last[parent]] = e;
    k->last[parent] = e;
}

static void unselect_below(const kids_t *k, int64_t node, unsigned char *selected) {
    for (int64_t e = k->first[node]; e >= 0; e = k->next_sib[e]) {
        unselect_below(k, k->kid[e], selected);
        selected[k->kid[e]] = 0;
    }
}

static double eom(const kids_t *k, int64_t node, const float *score, const float *nsize, unsigned char *selected,
                  double max_cluster_size) {
    /* max(score, 0.0) as Python evaluates it: the score unless 0.0 is larger */
    const double sc = (double)score[node];
    const double current = 0.0 > sc ? 0.0 : sc;
    const double size = (double)nsize[node];
    double child_total = 0.0;
    for (int64_t e = k->first[node]; e >= 0; e = k->next_sib[e])
        child_total += eom(k, k->kid[e], score, nsize, selected, max_cluster_size);
    if (child_total > current || size > max_cluster_size) return child_total;
    selected[node] = 1;
    unselect_below(k, node, selected);
    return current;
}

static int cmp_i64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : (x > y);
}

/* disjoint_set.py: union by rank; the roots do not depend on compression */
static int64_t ds_find(int64_t *parent, int64_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

static void ds_union(int64_t *parent, int64_t *rank, int64_t x, int64_t y) {
    x = ds_find(parent, x);
    y = ds_find(parent, y);
    if (x == y) return;
    if (rank[x] < rank[y]) {
        const int64_t t = x;
        x = y;
        y = t;
    }
    parent[y] = x;
    if (rank[x] == rank[y]) rank[x]++;
}

/*
 * mst: (n-1) rows (source, target, distance) as float64, in any order.
 * method 0: "eom", 1: "leaf".  labels [n] int64, strengths [n] float32;
 * linkage [(n-1)*4] and the condensed tree (capacity 2n) may be NULL.
 * Returns the number of selected clusters, or -1 on bad arguments.
 */
int64_t orc_labels(int64_t n, const double *mst, int64_t mcs, double max_cluster_size, int method,
                   int allow_single_cluster, int64_t *labels, float *strengths, double *linkage, int64_t *ct_parent,
                   int64_t *ct_child, float *ct_lambda, float *ct_size, int64_t *ct_rows) {
    if (n < 2 || !mst || !labels || !strengths || mcs < 1) return -1;
    const int64_t m = n - 1;
    /* 1. rows by (weight, source, target) */
    double *w = malloc(sizeof(double) * m), *s = malloc(sizeof(double) * m), *d = malloc(sizeof(double) * m);
    int64_t *order = malloc(sizeof(int64_t) * m);
    for (int64_t i = 0; i < m; ++i) {
        s[i] = mst[3 * i];
        d[i] = mst[3 * i + 1];
        w[i] = mst[3 * i + 2];
        order[i] = i;
    }
    sort_keys keys = {w, s, d};
    g_keys = &keys;
    qsort(order, (size_t)m, sizeof(int64_t), cmp_rows);
    /* 2. single linkage */
    double *H = malloc(sizeof(double) * 4 * m);
    int64_t *parent = malloc(sizeof(int64_t) * (2 * n - 1));
    int64_t *size = calloc((size_t)(2 * n - 1), sizeof(int64_t));
    for (int64_t i = 0; i < 2 * n - 1; ++i) parent[i] = -1;
    for (int64_t i = 0; i < n; ++i) size[i] = 1;
    int64_t next = n;
    for (int64_t r = 0; r < m; ++r) {
        const int64_t e = order[r];
        const int64_t len = 2 * n - 1;
        const int64_t lc = merge_find(parent, len, (int64_t)s[e]), rc = merge_find(parent, len, (int64_t)d[e]);
        const int64_t li = lc < 0 ? lc + len : lc, ri = rc < 0 ? rc + len : rc;
        H[4 * r] = (double)(lc > rc ? lc : rc);
        H[4 * r + 1] = (double)(lc > rc ? rc : lc);
        H[4 * r + 2] = w[e];
        H[4 * r + 3] = (double)(size[li] + size[ri]);
        size[next] = size[li] + size[ri];
        parent[li] = next;
        parent[ri] = next;
        next++;
    }
    if (linkage) memcpy(linkage, H, sizeof(double) * 4 * m);
    free(w);
    free(s);
    free(d);
    free(order);
    free(parent);
    free(size);
    /* 3. condensed tree */
    ctree t;
    t.parent = malloc(sizeof(int64_t) * 2 * n);
    t.child = malloc(sizeof(int64_t) * 2 * n);
    t.lambda = malloc(sizeof(float) * 2 * n);
    t.size = malloc(sizeof(float) * 2 * n);
    condense(H, n, mcs, max_cluster_size, &t);
    free(H);
    if (ct_rows) *ct_rows = t.rows;
    if (ct_parent) {
        memcpy(ct_parent, t.parent, sizeof(int64_t) * t.rows);
        memcpy(ct_child, t.child, sizeof(int64_t) * t.rows);
        memcpy(ct_lambda, t.lambda, sizeof(float) * t.rows);
        memcpy(ct_size, t.size, sizeof(float) * t.rows);
    }
    /* 4. selection over the cluster tree (rows with child_size > 1); node
     * ids reach n + 2 * (number of splits), up to ~3n */
    int64_t top = 0;
    for (int64_t r = 0; r < t.rows; ++r) {
        if (t.parent[r] + 1 > top) top = t.parent[r] + 1;
        if (t.child[r] + 1 > top) top = t.child[r] + 1;
    }
    const int64_t nodes = top + 1;
    int64_t *selected_list = malloc(sizeof(int64_t) * nodes);
    int64_t nsel = 0;
    int64_t ct_count = 0, ct_root = INT64_MAX, max_child = -1;
    for (int64_t r = 0; r < t.rows; ++r)
        if (t.size[r] > 1.0f) {
            ct_count++;
            if (t.parent[r] < ct_root) ct_root = t.parent[r];
            if (t.child[r] > max_child) max_child = t.child[r];
        }
    if (ct_count > 0 && method == 0) {
        float *score = calloc((size_t)nodes, sizeof(float));
        float *nsize = calloc((size_t)nodes, sizeof(float));
        unsigned char *has_score = calloc((size_t)nodes, 1);
        unsigned char *selected = calloc((size_t)nodes, 1);
        int64_t *order_nodes = malloc(sizeof(int64_t) * nodes); /* dict insertion order */
        int64_t nscored = 0;
        const int64_t root = t.parent[0];
        score[root] = 0.0f;
        has_score[root] = 1;
        order_nodes[nscored++] = root;
        for (int64_t r = 0; r < t.rows; ++r) {
            const float ls = t.lambda[r] * t.size[r];
            if (t.size[r] > 1.0f) {
                if (!has_score[t.child[r]]) order_nodes[nscored++] = t.child[r];
                has_score[t.child[r]] = 1;
                score[t.child[r]] = -ls;
            }
            if (!has_score[t.parent[r]]) {
                has_score[t.parent[r]] = 1;
                order_nodes[nscored++] = t.parent[r];
                score[t.parent[r]] = 0.0f;
            }
            score[t.parent[r]] = score[t.parent[r]] + ls;
        }
        kids_t k;
        k.first = malloc(sizeof(int64_t) * nodes);
        k.last = malloc(sizeof(int64_t) * nodes);
        k.kid = malloc(sizeof(int64_t) * (ct_count + 1));
        k.next_sib = malloc(sizeof(int64_t) * (ct_count + 1));
        k.entries = 0;
        for (int64_t i = 0; i < nodes; ++i) k.first[i] = -1;
        for (int64_t r = 0; r < t.rows; ++r)
            if (t.size[r] > 1.0f) {
                add_kid(&k, t.parent[r], t.child[r]);
                nsize[t.child[r]] = t.size[r];
            }
        nsize[ct_root] = (float)(ct_root - 1);
        if (allow_single_cluster) {
            eom(&k, ct_root, score, nsize, selected, max_cluster_size);
        } else if (nscored > 1) {
            for (int64_t e = k.first[ct_root]; e >= 0; e = k.next_sib[e])
                eom(&k, k.kid[e], score, nsize, selected, max_cluster_size);
        }
        for (int64_t q = 0; q < nscored; ++q)
            if (selected[order_nodes[q]]) selected_list[nsel++] = order_nodes[q];
        free(score);
        free(nsize);
        free(has_score);
        free(selected);
        free(order_nodes);
        free(k.first);
        free(k.last);
        free(k.kid);
        free(k.next_sib);
    } else if (ct_count > 0 && method == 1) {
        /* leaves of the cluster tree: nodes n..max_child that are no parent */
        unsigned char *leaf = malloc((size_t)nodes);
        memset(leaf, 1, (size_t)nodes);
        for (int64_t r = 0; r < t.rows; ++r)
            if (t.size[r] > 1.0f) leaf[t.parent[r]] = 0;
        for (int64_t c = n; c <= max_child; ++c)
            if (leaf[c]) selected_list[nsel++] = c;
        free(leaf);
    }
    /* 5. labels */
    for (int64_t i = 0; i < n; ++i) {
        labels[i] = -1;
        strengths[i] = 0.0f;
    }
    if (t.rows > 0 && nsel > 0) {
        qsort(selected_list, (size_t)nsel, sizeof(int64_t), cmp_i64); /* label = rank among sorted clusters */
        int64_t root_cluster = INT64_MAX;
        for (int64_t r = 0; r < t.rows; ++r)
            if (t.parent[r] < root_cluster) root_cluster = t.parent[r];
        int64_t *label_of = malloc(sizeof(int64_t) * (top + 1));
        unsigned char *is_sel = calloc((size_t)(top + 1), 1);
        for (int64_t q = 0; q < nsel; ++q) {
            label_of[selected_list[q]] = q;
            is_sel[selected_list[q]] = 1;
        }
        int64_t *dsp = malloc(sizeof(int64_t) * top), *dsr = calloc((size_t)top, sizeof(int64_t));
        for (int64_t i = 0; i < top; ++i) dsp[i] = i;
        for (int64_t r = 0; r < t.rows; ++r) {
            const int64_t c = t.child[r];
            if (is_sel[c]) continue;
            if (c < n && t.lambda[r] == 0.0f) continue;
            ds_union(dsp, dsr, t.parent[r], c);
        }
        float max_root_lambda = -INFINITY;
        float *point_lambda = malloc(sizeof(float) * n);
        for (int64_t r = 0; r < t.rows; ++r) {
            if (t.parent[r] == root_cluster && t.lambda[r] > max_root_lambda) max_root_lambda = t.lambda[r];
            if (t.child[r] < n) point_lambda[t.child[r]] = t.lambda[r];
        }
        for (int64_t p = 0; p < root_cluster; ++p) {
            const int64_t c = ds_find(dsp, p);
            if (c < root_cluster) labels[p] = -1;
            else if (c == root_cluster)
                labels[p] = (nsel == 1 && is_sel[c] && point_lambda[p] >= max_root_lambda) ? label_of[c] : -1;
            else labels[p] = label_of[c];
        }
        /* 6. membership strengths */
        double *death = calloc((size_t)(top + 1), sizeof(double));
        for (int64_t r = 0; r < t.rows; ++r) {
            const int64_t c = t.parent[r];
            if (is_sel[c] && t.size[r] == 1.0f && (double)t.lambda[r] > death[c]) death[c] = t.lambda[r];
        }
        for (int64_t r = 0; r < t.rows; ++r) {
            const int64_t p = t.child[r];
            if (p >= root_cluster || labels[p] < 0) continue;
            const double ml = death[selected_list[labels[p]]];
            if (ml == 0.0 || !isfinite(t.lambda[r])) strengths[p] = 1.0f;
            else {
                const double lam = (double)t.lambda[r] < ml ? (double)t.lambda[r] : ml;
                strengths[p] = (float)(lam / ml);
            }
        }
        free(label_of);
        free(is_sel);
        free(dsp);
        free(dsr);
        free(point_lambda);
        free(death);
    }
    free(selected_list);
    free(t.parent);
    free(t.child);
    free(t.lambda);
    free(t.size);
    return nsel;
}
