// Host check of diffvg_amd/csrc/dvg_sort.h against the C++ library's std::sort: both
// sort index arrays with a comparator on keys drawn from small alphabets (many ties),
// sorted, reversed and organ-pipe inputs, and keys shaped to exhaust the partition depth
// budget; every permutation must be identical, also when dvg_sort.h sorts (key, index)
// pairs with a comparator on the key.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

#include "../../diffvg_amd/csrc/dvg_sort.h"

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
        struct KeyIdx {
            float k;
            int i;
        };
        std::vector<KeyIdx> kv(n);
        for (int i = 0; i < n; i++) kv[i] = KeyIdx{key[i], i};
        dvg::stl::sort(kv.data(), n, [](const KeyIdx &x, const KeyIdx &y) { return x.k < y.k; }, stack.data());
        bool pairs_same = true;
        for (int i = 0; i < n; i++) pairs_same = pairs_same && kv[i].i == a[i];
        if (a != b || !pairs_same) {
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
    out[0] = a == b ? 0 : 1;
}
}
