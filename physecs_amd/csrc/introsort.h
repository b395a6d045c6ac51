// The introsort of the GNU C++ library's std::sort (bits/stl_algo.h, bits/stl_heap.h:
// __introsort_loop with a 16-element threshold and a 2*floor(log2 n) depth limit, median
// of three moved to the front, unguarded partition, heap sort when the depth limit is
// reached, final insertion sort), restated over an index array. std::sort is not stable:
// the order it leaves among equal keys is a property of this exact sequence of compares
// and moves, and the reference orders its triangle contacts with it
// (src/CollisionTriangleMesh.cpp:925), so the device repeats it step for step.
//
// Arr provides int get(int i), void set(int i, int v) and bool less(int va, int vb)
// comparing two stored values.
#pragma once

#ifndef PGX_HD
#define PGX_HD inline   // host-only builds of this header
#endif

namespace pgx {

template <class Arr>
PGX_HD void is_push_heap(Arr& a, int first, int hole, int top, int value) {
    int parent = (hole - 1) / 2;
    while (hole > top && a.less(a.get(first + parent), value)) {
        a.set(first + hole, a.get(first + parent));
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a.set(first + hole, value);
}

template <class Arr>
PGX_HD void is_adjust_heap(Arr& a, int first, int hole, int len, int value) {
    int top = hole, second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (a.less(a.get(first + second), a.get(first + second - 1))) second--;
        a.set(first + hole, a.get(first + second));
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        a.set(first + hole, a.get(first + second - 1));
        hole = second - 1;
    }
    is_push_heap(a, first, hole, top, value);
}

template <class Arr>
PGX_HD void is_heap_sort(Arr& a, int first, int last) {   // __partial_sort(first, last, last)
    int len = last - first;
    if (len >= 2) {                                       // __make_heap
        int parent = (len - 2) / 2;
        while (true) {
            is_adjust_heap(a, first, parent, len, a.get(first + parent));
            if (parent == 0) break;
            parent--;
        }
    }
    while (last - first > 1) {                            // __sort_heap / __pop_heap
        --last;
        int value = a.get(last);
        a.set(last, a.get(first));
        is_adjust_heap(a, first, 0, last - first, value);
    }
}

template <class Arr>
PGX_HD void is_swap(Arr& a, int i, int j) {
    int t = a.get(i);
    a.set(i, a.get(j));
    a.set(j, t);
}

template <class Arr>
PGX_HD int is_partition_pivot(Arr& a, int first, int last) {
    int mid = first + (last - first) / 2;
    // __move_median_to_first(first, first + 1, mid, last - 1)
    int ra = first + 1, rb = mid, rc = last - 1;
    if (a.less(a.get(ra), a.get(rb))) {
        if (a.less(a.get(rb), a.get(rc))) is_swap(a, first, rb);
        else if (a.less(a.get(ra), a.get(rc))) is_swap(a, first, rc);
        else is_swap(a, first, ra);
    } else if (a.less(a.get(ra), a.get(rc))) is_swap(a, first, ra);
    else if (a.less(a.get(rb), a.get(rc))) is_swap(a, first, rc);
    else is_swap(a, first, rb);
    // __unguarded_partition(first + 1, last, first)
    int lo = first + 1, hi = last;
    while (true) {
        while (a.less(a.get(lo), a.get(first))) ++lo;
        --hi;
        while (a.less(a.get(first), a.get(hi))) --hi;
        if (!(lo < hi)) return lo;
        is_swap(a, lo, hi);
        ++lo;
    }
}

template <class Arr>
PGX_HD void is_linear_insert(Arr& a, int last) {         // __unguarded_linear_insert
    int val = a.get(last);
    int next = last - 1;
    while (a.less(val, a.get(next))) {
        a.set(last, a.get(next));
        last = next;
        --next;
    }
    a.set(last, val);
}

template <class Arr>
PGX_HD void is_insertion_sort(Arr& a, int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        if (a.less(a.get(i), a.get(first))) {
            int val = a.get(i);
            for (int k = i; k > first; --k) a.set(k, a.get(k - 1));   // move_backward
            a.set(first, val);
        } else {
            is_linear_insert(a, i);
        }
    }
}

// std::sort(first, last) on a.get/set positions [0, n). The recursion of __introsort_loop
// (right part first, then the loop on the left part) is kept as an explicit stack of
// disjoint ranges, each with the depth limit the recursion would give it.
template <class Arr>
PGX_HD void introsort(Arr& a, int n) {
    if (n <= 1) return;
    int lg = 0;
    while ((2 << lg) <= n) ++lg;                          // floor(log2 n)
    int sf[64], sl[64], sd[64];
    int top = 0;
    sf[0] = 0; sl[0] = n; sd[0] = 2 * lg;
    top = 1;
    while (top > 0) {
        --top;
        int first = sf[top], last = sl[top], depth = sd[top];
        while (last - first > 16) {
            if (depth == 0) {
                is_heap_sort(a, first, last);
                break;
            }
            --depth;
            int cut = is_partition_pivot(a, first, last);
            if (top < 64) { sf[top] = cut; sl[top] = last; sd[top] = depth; ++top; }
            last = cut;
        }
    }
    if (n > 16) {                                         // __final_insertion_sort
        is_insertion_sort(a, 0, 16);
        for (int i = 16; i != n; ++i) is_linear_insert(a, i);
    } else {
        is_insertion_sort(a, 0, n);
    }
}

}  // namespace pgx
