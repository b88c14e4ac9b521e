// dvg_sort.h -- the comparison sort the reference's BVH build uses, for host and device.
//
// The reference orders BVH leaves with std::sort (scene.cpp:602-611 segment y-centres,
// scene.cpp:632-645 and 676 Morton codes), built against GNU libstdc++.  std::sort is not
// stable, so leaves with equal keys end up in an order fixed by the library's algorithm,
// and that order decides which of two equidistant segments a closest-point query
// reports.  To build BVHs on the device with the same leaf order, this file implements
// that algorithm (Musser's introsort as the SGI STL / libstdc++ formulate it):
//   * ranges longer than 16 elements are partitioned around the median of (second,
//     middle, last) moved to the front, the right part first, with a depth budget of
//     2 floor(log2 n) partitions, after which a range is heap-sorted;
//   * a final insertion pass sorts the first 16 elements and inserts the rest one by one.
// Only comparison outcomes steer it, so sorting element indices with a comparator on
// their keys -- or (key, index) pairs with a comparator on the key -- yields the
// permutation std::sort yields for the elements themselves.
// tests/csrc/sort_check.cpp compares it with std::sort on inputs full of ties.
#pragma once

#if defined(__HIPCC__)
#define DVG_SORT_HD __host__ __device__ inline
#else
#define DVG_SORT_HD inline
#endif

namespace dvg {
namespace stl {

constexpr int kInsertionLimit = 16;

DVG_SORT_HD int floor_log2(int n) {
    int k = 0;
    while (n > 1) {
        n >>= 1;
        k++;
    }
    return k;
}

template <class T, class Less>
DVG_SORT_HD void sift_up(T *a, int hole, int top, T v, Less &less) {
    int parent = (hole - 1) / 2;
    while (hole > top && less(a[parent], v)) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = v;
}

// move the larger child up until the hole reaches the bottom, then sift v up from there
template <class T, class Less>
DVG_SORT_HD void adjust_heap(T *a, int hole, int len, T v, Less &less) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (less(a[child], a[child - 1])) child--;
        a[hole] = a[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[hole] = a[child - 1];
        hole = child - 1;
    }
    sift_up(a, hole, top, v, less);
}

template <class T, class Less>
DVG_SORT_HD void heap_sort(T *a, int n, Less &less) {
    if (n < 2) return;
    for (int parent = (n - 2) / 2;; parent--) {
        adjust_heap(a, parent, n, a[parent], less);
        if (parent == 0) break;
    }
    for (int last = n - 1; last > 0; last--) {
        T v = a[last];
        a[last] = a[0];
        adjust_heap(a, 0, last, v, less);
    }
}

template <class T>
DVG_SORT_HD void swap_at(T *a, int i, int j) {
    T t = a[i];
    a[i] = a[j];
    a[j] = t;
}

// median of a[i], a[j], a[k] swapped into a[r]
template <class T, class Less>
DVG_SORT_HD void median_to(T *a, int r, int i, int j, int k, Less &less) {
    int m;
    if (less(a[i], a[j])) {
        if (less(a[j], a[k]))
            m = j;
        else if (less(a[i], a[k]))
            m = k;
        else
            m = i;
    } else if (less(a[i], a[k])) {
        m = i;
    } else if (less(a[j], a[k])) {
        m = k;
    } else {
        m = j;
    }
    swap_at(a, r, m);
}

// Hoare partition of [lo, hi) around the pivot at a[p] (p < lo); returns the cut
template <class T, class Less>
DVG_SORT_HD int partition(T *a, int lo, int hi, int p, Less &less) {
    for (;;) {
        while (less(a[lo], a[p])) lo++;
        hi--;
        while (less(a[p], a[hi])) hi--;
        if (!(lo < hi)) return lo;
        swap_at(a, lo, hi);
        lo++;
    }
}

template <class T, class Less>
DVG_SORT_HD void linear_insert(T *a, int i, Less &less) {
    T v = a[i];
    int j = i;
    while (j > 0 && less(v, a[j - 1])) {
        a[j] = a[j - 1];
        j--;
    }
    a[j] = v;
}

template <class T, class Less>
DVG_SORT_HD void insertion_sort(T *a, int lo, int hi, Less &less) {
    for (int i = lo + 1; i < hi; i++) {
        if (less(a[i], a[lo])) {
            T v = a[i];
            for (int j = i; j > lo; j--) a[j] = a[j - 1];
            a[lo] = v;
        } else {
            linear_insert(a, i, less);
        }
    }
}

// The partition phase of sort() on [lo, hi) with `depth` partitions left: ranges
// longer than kInsertionLimit are partitioned (heap-sorted once the depth budget is
// spent) until every range left has at most kInsertionLimit elements.  Ranges are
// disjoint and each is processed from its own contents and depth, so they may be
// processed in any order (or at once) with the same result.  `stack` holds
// sort_stack_ints(hi - lo) ints: the pending right parts, at most one per depth level.
template <class T, class Less>
DVG_SORT_HD void partition_phase(T *a, int lo, int hi, int depth, Less &less, int *stack) {
    int top = 0;
    for (;;) {
        while (hi - lo > kInsertionLimit) {
            if (depth == 0) {
                heap_sort(a + lo, hi - lo, less);
                break;
            }
            depth--;
            median_to(a, lo, lo + 1, lo + (hi - lo) / 2, hi - 1, less);
            int cut = partition(a, lo + 1, hi, lo, less);
            // [cut, hi) is sorted with the same remaining depth; continue with [lo, cut)
            stack[top++] = cut;
            stack[top++] = hi;
            stack[top++] = depth;
            hi = cut;
        }
        if (top == 0) break;
        depth = stack[--top];
        hi = stack[--top];
        lo = stack[--top];
    }
}

// The final pass of sort(): insertion sort of the whole array.  Insertion sort is
// stable, so for a strict weak order `less` it leaves element i of the array at
// #{j : less(a[j], a[i])} + #{j < i : a[j] equivalent to a[i]}.
template <class T, class Less>
DVG_SORT_HD void final_insertion(T *a, int n, Less &less) {
    if (n > kInsertionLimit) {
        insertion_sort(a, 0, kInsertionLimit, less);
        for (int i = kInsertionLimit; i < n; i++) linear_insert(a, i, less);
    } else {
        insertion_sort(a, 0, n, less);
    }
}

// Sorts a[0..n) with `less(x, y)` on element values; `stack` holds sort_stack_ints(n)
// ints.
template <class T, class Less>
DVG_SORT_HD void sort(T *a, int n, Less less, int *stack) {
    if (n < 2) return;
    partition_phase(a, 0, n, 2 * floor_log2(n), less, stack);
    final_insertion(a, n, less);
}

// partition(a, lo, hi, p) restated through its stoppers, the form a workgroup computes
// in parallel.  Left stoppers are the positions i of [lo, hi) with !less(a[i], a[p]) in
// increasing order L1 < L2 < ...; right stoppers the positions with !less(a[p], a[i]) in
// decreasing order R1 > R2 > ..., followed by p itself (which stops the right scan).
// The k-th round of the Hoare loop stops at L_k and R_k -- the positions between the
// previous stops are unmodified -- and swaps them while L_k < R_k; as L rises and R
// falls, the rounds that swap are k = 1..M with M the number of k for which L_k < R_k.
// The scan that ends the loop stops at L_{M+1}, or earlier at R_M, which holds a left
// stopper after its swap: the cut is min(L_{M+1}, R_M).  Lpos and Rpos hold hi - lo + 2
// ints.  (A left stopper always exists after median_to; with none the cut is hi.)
template <class T, class Less>
DVG_SORT_HD int partition_by_stoppers(T *a, int lo, int hi, int p, Less &less, int *Lpos, int *Rpos) {
    int nl = 0, nr = 0;
    for (int i = lo; i < hi; i++)
        if (!less(a[i], a[p])) Lpos[++nl] = i;
    for (int i = hi - 1; i >= lo; i--)
        if (!less(a[p], a[i])) Rpos[++nr] = i;
    Rpos[++nr] = p;
    int m = 0;
    for (int k = 1; k <= nl && k <= nr; k++) m += Lpos[k] < Rpos[k] ? 1 : 0;
    for (int k = 1; k <= m; k++) swap_at(a, Lpos[k], Rpos[k]);
    int cut = m + 1 <= nl ? Lpos[m + 1] : hi;
    if (m >= 1 && Rpos[m] < cut) cut = Rpos[m];
    return cut;
}

// ints the `stack` argument of sort() needs for n elements
DVG_SORT_HD int sort_stack_ints(int n) { return 3 * (2 * floor_log2(n > 1 ? n : 2) + 2); }

}  // namespace stl
}  // namespace dvg
