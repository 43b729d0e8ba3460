// Boruvka on a CSR graph (core_graph.boruvka_mst, core_graph.py:172-208)
// and the core graph it usually runs on (knn_mst_union, flatten_to_csr,
// sort_by_lens, core_graph.py:13-104).
//
// Rounds: the device scans the rows (each point's first remaining child is
// its candidate; core_graph.py:111-129), reduces them per component with a
// packed atomicMin on (ordered weight, point) -- the first point at the
// minimum wins, as the reference's strict `<` in point order does -- and
// marks each component's first point with a candidate, which fixes the
// order of the reference's typed dict.  The host then merges those few
// edges in that order with the reference's union-by-rank disjoint set
// (core_graph.py:131-148, disjoint_set.py), so cycle skips, row order and
// the final component roots are the reference's; the device unites the kept
// edges in its own labels and drops every child that joined its row's
// component (update_graph_components, core_graph.py:151-169).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "../../include/fh.h"
#include "fh_numba_sort.h"

namespace {

inline unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// float -> unsigned order-preserving key (negative lens values included;
// -0 and +0 compare equal, as the reference's `<` sees them)
__device__ __forceinline__ uint32_t ordered(float f) {
    if (f == 0.0f) f = 0.0f;
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ int dev_find(int *par, int x) {
    for (;;) {
        const int p = par[x];
        if (p == x) return x;
        const int g = par[p];
        if (g != p) atomicCAS(&par[x], p, g);
        x = p;
    }
}

// lock-free union, larger label under smaller
__device__ void dev_unite(int *par, int a, int b) {
    for (;;) {
        a = dev_find(par, a);
        b = dev_find(par, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&par[b], b, a) == b) return;
    }
}

__global__ void csr_init_kernel(int64_t n, int *comp, int *par) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    comp[p] = (int)p;
    par[p] = (int)p;
}

__global__ void csr_reset_kernel(int64_t n, unsigned long long *best, int *first) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    best[p] = ~0ull;
    first[p] = 0x7fffffff;
}

// core_graph.py:115-129: a row's candidate is its entry at indptr[p]
// (indices[start] is read even when the row is empty, like the reference,
// unless start is the end of the arrays)
__global__ void csr_cand_kernel(int64_t n, int64_t nnz, const int *__restrict__ indptr,
                                const int *__restrict__ indices, const float *__restrict__ w,
                                const int *__restrict__ comp, unsigned long long *best, int *first) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t s = indptr[p];
    if (s >= nnz || indices[s] == -1) return;
    const int c = comp[p];
    const unsigned long long key = ((unsigned long long)ordered(w[s]) << 32) | (unsigned int)p;
    atomicMin(best + c, key);
    atomicMin(first + c, (int)p);
}

__global__ void csr_first_flags_kernel(int64_t n, const int *__restrict__ comp, const int *__restrict__ first,
                                       int *flag) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    flag[p] = first[comp[p]] == (int)p ? 1 : 0;
}

// components in dict order (their first point), with the chosen edge
__global__ void csr_gather_kernel(int64_t n, const int *__restrict__ flag, const int *__restrict__ offs,
                                  const int *__restrict__ comp, const unsigned long long *__restrict__ best,
                                  const int *__restrict__ indptr, const int *__restrict__ indices, int *eu, int *ev,
                                  float *ew) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !flag[p]) return;
    const unsigned long long key = best[comp[p]];
    const int u = (int)(unsigned int)(key & 0xffffffffull);
    const int o = offs[p];
    eu[o] = u;
    ev[o] = indices[indptr[u]];
    ew[o] = unordered((uint32_t)(key >> 32));
}

__global__ void csr_union_kernel(int k, const int *__restrict__ ku, const int *__restrict__ kv,
                                 const int *__restrict__ comp, int *par) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    dev_unite(par, comp[ku[i]], comp[kv[i]]);
}

__global__ void csr_relabel_kernel(int64_t n, int *comp, int *par) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    comp[p] = dev_find(par, comp[p]);
}

// update_graph_components (core_graph.py:151-169)
__global__ void csr_filter_kernel(int64_t n, const int *__restrict__ indptr, int *indices, float *w,
                                  const int *__restrict__ comp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = indptr[p], e = indptr[p + 1];
    const int cp = comp[p];
    int cnt = 0;
    for (int i = s; i < e; ++i) {
        const int c = indices[i];
        if (c == -1) break;
        if (comp[c] != cp) {
            indices[s + cnt] = c;
            w[s + cnt] = w[i];
            ++cnt;
        }
    }
    for (int i = s + cnt; i < e; ++i) {
        indices[i] = -1;
        w[i] = __builtin_huge_valf();
    }
}

// exclusive scan of n flags (small helper: block sums then a host-free
// second pass over the block totals)
__global__ void csr_block_scan_kernel(int64_t n, const int *__restrict__ flag, int *offs, int *block_sum) {
    __shared__ int s[256];
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int v = p < n ? flag[p] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (p < n) offs[p] = s[threadIdx.x] - v;
    if (threadIdx.x == 255) block_sum[blockIdx.x] = s[255];
}

__global__ void csr_sums_scan_kernel(int nblocks, int *block_sum, int *total) {
    // one thread: the number of blocks is n / 256
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int acc = 0;
    for (int b = 0; b < nblocks; ++b) {
        const int v = block_sum[b];
        block_sum[b] = acc;
        acc += v;
    }
    *total = acc;
}

__global__ void csr_add_kernel(int64_t n, int *offs, const int *__restrict__ block_sum) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    offs[p] += block_sum[blockIdx.x];
}

template <class T>
struct Dev {
    T *p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) { return hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * (count ? count : 1)); }
};

// disjoint_set.py: union by rank; ds_find's tuple assignment leaves the
// parents unchanged, so roots depend only on the unions.
struct RankDS {
    std::vector<int32_t> parent, rank;
    explicit RankDS(int64_t n) : parent(n), rank(n, 0) {
        for (int64_t i = 0; i < n; ++i) parent[i] = (int32_t)i;
    }
    int32_t find(int32_t x) const {
        while (parent[x] != x) x = parent[x];
        return x;
    }
    void unite(int32_t x, int32_t y) {
        x = find(x);
        y = find(y);
        if (x == y) return;
        if (rank[x] < rank[y]) std::swap(x, y);
        parent[y] = x;
        if (rank[x] == rank[y]) rank[x] += 1;
    }
};

}  // namespace

#define CSR_HIP(call)                          \
    do {                                       \
        if ((call) != hipSuccess) return FH_EDEVICE; \
    } while (0)

namespace fh {

int csr_boruvka(int device, int64_t n, const int32_t *indptr, const int32_t *indices, const float *weights,
                double *edges, int64_t *n_edges, int64_t *point_components, int64_t *n_components) {
    if (n < 1 || !indptr || !edges || !n_edges || !point_components || !n_components) return FH_EINVAL;
    if (n >= (int64_t(1) << 31)) return FH_EUNSUPPORTED;
    const int64_t nnz = indptr[n];
    if (nnz < 0 || (nnz > 0 && (!indices || !weights))) return FH_EINVAL;
    try {
        CSR_HIP(hipSetDevice(device));
        hipStream_t st = nullptr;
        Dev<int> d_indptr, d_indices, comp, par, first, flag, offs, bsum, eu, ev, ku, kv, small;
        Dev<float> d_w, ew;
        Dev<unsigned long long> best;
        const int64_t nblk = (n + 255) / 256;
        CSR_HIP(d_indptr.alloc(n + 1));
        CSR_HIP(d_indices.alloc(nnz));
        CSR_HIP(d_w.alloc(nnz));
        CSR_HIP(comp.alloc(n));
        CSR_HIP(par.alloc(n));
        CSR_HIP(first.alloc(n));
        CSR_HIP(flag.alloc(n));
        CSR_HIP(offs.alloc(n));
        CSR_HIP(bsum.alloc(nblk));
        CSR_HIP(eu.alloc(n));
        CSR_HIP(ev.alloc(n));
        CSR_HIP(ew.alloc(n));
        CSR_HIP(ku.alloc(n));
        CSR_HIP(kv.alloc(n));
        CSR_HIP(small.alloc(1));
        CSR_HIP(best.alloc(n));
        CSR_HIP(hipMemcpy(d_indptr.p, indptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice));
        if (nnz > 0) {
            CSR_HIP(hipMemcpy(d_indices.p, indices, sizeof(int) * nnz, hipMemcpyHostToDevice));
            CSR_HIP(hipMemcpy(d_w.p, weights, sizeof(float) * nnz, hipMemcpyHostToDevice));
        }
        hipLaunchKernelGGL(csr_init_kernel, dim3(nb(n, 256)), dim3(256), 0, st, n, comp.p, par.p);
        RankDS ds(n);
        int64_t ncomp = n, out = 0;
        std::vector<int> hu, hv, hku, hkv;
        std::vector<float> hw;
        while (ncomp > 1) {
            hipLaunchKernelGGL(csr_reset_kernel, dim3(nb(n, 256)), dim3(256), 0, st, n, best.p, first.p);
            hipLaunchKernelGGL(csr_cand_kernel, dim3(nb(n, 256)), dim3(256), 0, st, n, nnz, d_indptr.p, d_indices.p,
                               d_w.p, comp.p, best.p, first.p);
            hipLaunchKernelGGL(csr_first_flags_kernel, dim3(nb(n, 256)), dim3(256), 0, st, n, comp.p, first.p,
                               flag.p);
            hipLaunchKernelGGL(csr_block_scan_kernel, dim3((unsigned)nblk), dim3(256), 0, st, n, flag.p, offs.p,
                               bsum.p);
            hipLaunchKernelGGL(csr_sums_scan_kernel, dim3(1), dim3(1), 0, st, (int)nblk, bsum.p, small.p);
            hipLaunchKernelGGL(csr_add_kernel, dim3((unsigned)nblk), dim3(256), 0, st, n, offs.p, bsum.p);
            hipLaunchKernelGGL(csr_gather_kernel, dim3(nb(n, 256)), dim3(256), 0, st, n, flag.p, offs.p, comp.p,
                               best.p, d_indptr.p, d_indices.p, eu.p, ev.p, ew.p);
            CSR_HIP(hipGetLastError());
            int cnt = 0;
            CSR_HIP(hipMemcpy(&cnt, small.p, sizeof(int), hipMemcpyDeviceToHost));
            hu.resize(cnt);
            hv.resize(cnt);
            hw.resize(cnt);
            if (cnt > 0) {
                CSR_HIP(hipMemcpy(hu.data(), eu.p, sizeof(int) * cnt, hipMemcpyDeviceToHost));
                CSR_HIP(hipMemcpy(hv.data(), ev.p, sizeof(int) * cnt, hipMemcpyDeviceToHost));
                CSR_HIP(hipMemcpy(hw.data(), ew.p, sizeof(float) * cnt, hipMemcpyDeviceToHost));
            }
            // merge_components (core_graph.py:131-148)
            hku.clear();
            hkv.clear();
            for (int i = 0; i < cnt; ++i) {
                const int32_t a = ds.find(hu[i]), b = ds.find(hv[i]);
                if (a == b) continue;
                edges[3 * out + 0] = (double)hu[i];
                edges[3 * out + 1] = (double)hv[i];
                edges[3 * out + 2] = (double)hw[i];
                ++out;
                ds.unite(a, b);
                hku.push_back(hu[i]);
                hkv.push_back(hv[i]);
            }
            const int kept = (int)hku.size();
            if (kept == 0) break;  // core_graph.py:194-195
            CSR_HIP(hipMemcpy(ku.p, hku.data(), sizeof(int) * kept, hipMemcpyHostToDevice));
            CSR_HIP(hipMemcpy(kv.p, hkv.data(), sizeof(int) * kept, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(csr_union_kernel, dim3(nb(kept, 256)), dim3(256), 0, st, kept, ku.p, kv.p, comp.p,
                               par.p);
            hipLaunchKernelGGL(csr_relabel_kernel, dim3(nb(n, 256)), dim3(256), 0, st, n, comp.p, par.p);
            hipLaunchKernelGGL(csr_filter_kernel, dim3(nb(n, 256)), dim3(256), 0, st, n, d_indptr.p, d_indices.p,
                               d_w.p, comp.p);
            CSR_HIP(hipGetLastError());
            ncomp -= kept;
        }
        CSR_HIP(hipDeviceSynchronize());
        for (int64_t i = 0; i < n; ++i) point_components[i] = ds.find((int32_t)i);
        *n_edges = out;
        *n_components = ncomp;
    } catch (const std::bad_alloc &) {
        return FH_ENOMEM;
    }
    return FH_OK;
}

}  // namespace fh

// knn_mst_union + flatten_to_csr + sort_by_lens (core_graph.py:13-104):
// each point's children in insertion order (its kNN row, then MST edges
// whose parent it is and whose child is not there yet), weight = max of the
// two lens values, distance = max of the two core distances (the MST
// distance for MST-only children), stored as float32, then every row
// ordered by numba's argsort of its weights.  capacity >= n*k + m always
// suffices.  Returns the number of entries or a negative FH_E* code.
extern "C" int64_t fh_core_graph(int64_t n, int32_t k, const int32_t *neighbors, const float *core_distances,
                                 int64_t m, const double *mst, const double *lens, int32_t *indptr,
                                 int32_t *indices, float *weights, float *distances, int64_t capacity) {
    if (n < 1 || k < 0 || m < 0 || !core_distances || !lens || !indptr) return FH_EINVAL;
    if ((k > 0 && !neighbors) || (m > 0 && !mst)) return FH_EINVAL;
    try {
        struct Child {
            int32_t c;
            double w, d;
        };
        std::vector<std::vector<Child>> g(n);
        auto put = [](std::vector<Child> &row, int32_t c, double w, double d, bool overwrite) {
            for (Child &x : row)
                if (x.c == c) {
                    if (overwrite) {
                        x.w = w;
                        x.d = d;
                    }
                    return;
                }
            row.push_back({c, w, d});
        };
        // Python's max(a, b): b only when b > a
        auto pmax = [](double a, double b) { return b > a ? b : a; };
        for (int64_t p = 0; p < n; ++p) {
            const double pl = lens[p];
            const double pd = core_distances[p];
            for (int32_t t = 0; t < k; ++t) {
                const int32_t c = neighbors[p * k + t];
                if (c < 0) continue;
                put(g[p], c, pmax(pl, lens[c]), pmax(pd, (double)core_distances[c]), true);
            }
        }
        for (int64_t e = 0; e < m; ++e) {
            const int32_t pa = (int32_t)mst[3 * e], ch = (int32_t)mst[3 * e + 1];
            put(g[pa], ch, pmax(lens[pa], lens[ch]), mst[3 * e + 2], false);
        }
        int64_t nnz = 0;
        for (int64_t p = 0; p < n; ++p) nnz += (int64_t)g[p].size();
        if (nnz > capacity || nnz >= (int64_t(1) << 31)) return nnz > capacity ? nnz : FH_EUNSUPPORTED;
        if (nnz > 0 && (!indices || !weights || !distances)) return FH_EINVAL;
        indptr[0] = 0;
        std::vector<float> rw;
        for (int64_t p = 0; p < n; ++p) {
            const int32_t s = indptr[p];
            const std::vector<Child> &row = g[p];
            rw.resize(row.size());
            for (size_t j = 0; j < row.size(); ++j) rw[j] = (float)row[j].w;
            const std::vector<int64_t> order = fh::numba_argsort(rw.data(), (int64_t)rw.size());
            for (size_t j = 0; j < row.size(); ++j) {
                const Child &x = row[order[j]];
                indices[s + j] = x.c;
                weights[s + j] = (float)x.w;
                distances[s + j] = (float)x.d;
            }
            indptr[p + 1] = s + (int32_t)row.size();
        }
        return nnz;
    } catch (const std::bad_alloc &) {
        return FH_ENOMEM;
    }
}
