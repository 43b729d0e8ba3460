// numba's np.argsort(kind="quicksort") restated for the host: the order it
// leaves equal keys in is what the reference's numba-compiled sorts return.
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <utility>
#include <vector>

namespace fh {

// numba's np.argsort(kind="quicksort") (numba/misc/quicksort.py:
// run_quicksort1, partition, insertion_sort): median-of-three pivot parked
// at the end, Hoare-style scan, larger side pushed, insertion sort below 15
// elements.  Equal keys keep the order this procedure leaves them in, which
// is what the reference's numba-compiled sorts return.  Float keys order NaN
// last, as numba's float comparison for sorting does.
template <class T>
std::vector<int64_t> numba_argsort(const T *a, int64_t n) {
    auto lt = [](T x, T y) {
        if constexpr (std::is_floating_point<T>::value)
            return x < y || (std::isnan(y) && !std::isnan(x));
        else
            return x < y;
    };
    std::vector<int64_t> R(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) R[i] = i;
    if (n < 2) return R;
    auto A = [&](int64_t r) { return a[R[r]]; };
    auto insertion = [&](int64_t low, int64_t high) {
        if (high <= low) return;
        for (int64_t i = low + 1; i <= high; ++i) {
            const int64_t k = R[i];
            const T v = a[k];
            int64_t j = i;
            while (j > low && lt(v, a[R[j - 1]])) {
                R[j] = R[j - 1];
                --j;
            }
            R[j] = k;
        }
    };
    auto partition = [&](int64_t low, int64_t high) {
        const int64_t mid = (low + high) >> 1;
        if (lt(A(mid), A(low))) std::swap(R[low], R[mid]);
        if (lt(A(high), A(mid))) std::swap(R[high], R[mid]);
        if (lt(A(mid), A(low))) std::swap(R[low], R[mid]);
        const T pivot = A(mid);
        std::swap(R[high], R[mid]);
        int64_t i = low, j = high - 1;
        for (;;) {
            while (i < high && lt(A(i), pivot)) ++i;
            while (j >= low && lt(pivot, A(j))) --j;
            if (i >= j) break;
            std::swap(R[i], R[j]);
            ++i;
            --j;
        }
        std::swap(R[i], R[high]);
        return i;
    };
    constexpr int64_t kSmall = 15;
    std::vector<std::pair<int64_t, int64_t>> stack;
    stack.emplace_back(0, n - 1);
    while (!stack.empty()) {
        int64_t low = stack.back().first, high = stack.back().second;
        stack.pop_back();
        while (high - low >= kSmall) {
            const int64_t i = partition(low, high);
            if (high - i > i - low) {
                if (high > i) stack.emplace_back(i + 1, high);
                high = i - 1;
            } else {
                if (i > low) stack.emplace_back(low, i - 1);
                low = i + 1;
            }
        }
        insertion(low, high);
    }
    return R;
}

}  // namespace fh
