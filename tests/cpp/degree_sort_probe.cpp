// Test helper: prints the order GCC's std::sort (mode 0) or std::partial_sort
// over the whole range (mode 1, the heapsort that introsort falls back to)
// leaves items in when compared by key[item] < key[other]. Input on stdin:
// cases, then per case: mode n, n keys, n items. Output: the sorted items of
// each case on one line.
#include <algorithm>
#include <cstdio>
#include <vector>

int main() {
    int cases = 0;
    if (std::scanf("%d", &cases) != 1) return 1;
    for (int c = 0; c < cases; ++c) {
        int mode = 0, n = 0;
        if (std::scanf("%d %d", &mode, &n) != 2) return 1;
        std::vector<int> key(n), items(n);
        for (int& k : key)
            if (std::scanf("%d", &k) != 1) return 1;
        for (int& i : items)
            if (std::scanf("%d", &i) != 1) return 1;
        auto less = [&](int left, int right) { return key[left] < key[right]; };
        if (mode == 0) std::sort(items.begin(), items.end(), less);
        else std::partial_sort(items.begin(), items.end(), items.end(), less);
        for (int i = 0; i < n; ++i) std::printf(i ? " %d" : "%d", items[i]);
        std::printf("\n");
    }
    return 0;
}
