// Host check of diffvg_amd/csrc/dvg_sort.h against the C++ library's std::sort: both
// sort index arrays with a comparator on keys drawn from small alphabets (many ties),
// sorted, reversed and organ-pipe inputs, and keys shaped to exhaust the partition depth
// budget; every permutation must be identical, also when dvg_sort.h sorts (key, index)
// pairs with a comparator on the key, and when the sort runs in the form the device
// rebuild runs it: large ranges partitioned through their stoppers level by level, the
// remaining ranges by the sequential partition phase, the final insertion pass as a
// stable rank.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

#include "../../diffvg_amd/csrc/dvg_sort.h"

namespace {
struct KeyIdx {
    float k;
    int i;
};
struct KeyLess {
    bool operator()(const KeyIdx &x, const KeyIdx &y) const { return x.k < y.k; }
};

// the device rebuild's schedule of dvg::stl::sort, run sequentially: ranges longer than
// `big` (with depth left) are partitioned by partition_by_stoppers one level at a time,
// the others by partition_phase, then every element goes to its stable rank
std::vector<int> two_phase_order(std::vector<KeyIdx> a, int big) {
    using namespace dvg::stl;
    const int n = (int)a.size();
    KeyLess less;
    struct R {
        int lo, hi, depth;
    };
    std::vector<R> cur, next, small;
    auto classify = [&](R r, std::vector<R> &into) {
        if (r.hi - r.lo <= kInsertionLimit) return;
        if (r.hi - r.lo <= big || r.depth == 0)
            small.push_back(r);
        else
            into.push_back(r);
    };
    if (n >= 2) classify(R{0, n, 2 * floor_log2(n)}, cur);
    std::vector<int> Lpos(n + 2), Rpos(n + 2), stack(sort_stack_ints(n));
    while (!cur.empty()) {
        next.clear();
        for (const R &r : cur) {
            const int depth = r.depth - 1;
            median_to(a.data(), r.lo, r.lo + 1, r.lo + (r.hi - r.lo) / 2, r.hi - 1, less);
            int cut = partition_by_stoppers(a.data(), r.lo + 1, r.hi, r.lo, less, Lpos.data(), Rpos.data());
            classify(R{r.lo, cut, depth}, next);
            classify(R{cut, r.hi, depth}, next);
        }
        cur.swap(next);
    }
    for (const R &r : small) partition_phase(a.data(), r.lo, r.hi, r.depth, less, stack.data());
    std::vector<int> out(n);
    for (int i = 0; i < n; i++) {
        int pos = 0;
        for (int j = 0; j < n; j++) pos += (less(a[j], a[i]) || (j < i && !less(a[i], a[j]))) ? 1 : 0;
        out[pos] = a[i].i;
    }
    return out;
}
}  // namespace

extern "C" {
// out[0] = arrays compared, out[1] = arrays whose permutations differ, out[2] = first
// failing n (or -1)
void sort_fuzz(int64_t trials, uint64_t seed, int64_t *out) {
    std::mt19937_64 g(seed);
    int64_t bad = 0, first = -1;
    std::vector<int> a, b, stack;
    std::vector<float> key;
    for (int64_t t = 0; t < trials; t++) {
        int n;
        int r = (int)(g() % 10);
        if (r < 5)
            n = (int)(g() % 40);
        else if (r < 9)
            n = (int)(g() % 600);
        else
            n = (int)(g() % 5000);
        key.resize(n);
        int shape = (int)(g() % 7);
        int alpha = 1 + (int)(g() % 8);
        for (int i = 0; i < n; i++) {
            switch (shape) {
                case 0: key[i] = (float)(g() % alpha); break;                       // heavy ties
                case 1: key[i] = (float)(g() % 1000003) * 0.001f; break;             // few ties
                case 2: key[i] = (float)i; break;                                    // sorted
                case 3: key[i] = (float)(n - i); break;                              // reversed
                case 4: key[i] = (float)std::min(i, n - i); break;                   // organ pipe
                case 5: key[i] = (float)((i * 7919) % (alpha + 3)); break;           // periodic
                default: key[i] = (i & 1) ? (float)i : (float)(n - i % alpha); break;  // interleaved
            }
        }
        a.resize(n);
        std::iota(a.begin(), a.end(), 0);
        b = a;
        auto less = [&](int x, int y) { return key[x] < key[y]; };
        std::sort(a.begin(), a.end(), less);
        stack.assign(dvg::stl::sort_stack_ints(n), 0);
        dvg::stl::sort(b.data(), n, less, stack.data());
        std::vector<KeyIdx> kv(n);
        for (int i = 0; i < n; i++) kv[i] = KeyIdx{key[i], i};
        std::vector<KeyIdx> kv0 = kv;
        dvg::stl::sort(kv.data(), n, KeyLess{}, stack.data());
        bool pairs_same = true;
        for (int i = 0; i < n; i++) pairs_same = pairs_same && kv[i].i == a[i];
        const int big = t % 3 == 0 ? 17 : t % 3 == 1 ? 40 : 192;
        bool two_phase_same = n > 3000 || two_phase_order(kv0, big) == a;
        if (a != b || !pairs_same || !two_phase_same) {
            if (first < 0) first = n;
            bad++;
        }
    }
    out[0] = trials;
    out[1] = bad;
    out[2] = first;
}

// the median-of-three killer sequence for this partition rule makes std::sort fall back
// to heap sort; built adaptively by comparing against gas values (McIlroy's adversary)
void sort_adversary(int n, int64_t *out) {
    std::vector<int> val(n), order(n);
    const int gas = n;
    int nsolid = 0, candidate = 0;
    for (int i = 0; i < n; i++) val[i] = gas;
    auto freeze = [&](int x) { val[x] = nsolid++; };
    auto cmp = [&](int x, int y) {
        if (val[x] == gas && val[y] == gas) {
            if (x == candidate) freeze(x);
            else freeze(y);
        }
        if (val[x] == gas) candidate = x;
        else if (val[y] == gas) candidate = y;
        return val[x] < val[y];
    };
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), cmp);
    // val now holds a fixed key sequence; compare both sorts on it
    std::vector<int> a(n), b(n), stack(dvg::stl::sort_stack_ints(n));
    std::iota(a.begin(), a.end(), 0);
    b = a;
    auto less = [&](int x, int y) { return val[x] < val[y]; };
    std::sort(a.begin(), a.end(), less);
    dvg::stl::sort(b.data(), n, less, stack.data());
    std::vector<KeyIdx> kv(n);
    for (int i = 0; i < n; i++) kv[i] = KeyIdx{(float)val[i], i};
    out[0] = (a == b && (n > 3000 || two_phase_order(kv, 40) == a)) ? 0 : 1;
}
}
