// spans.hpp -- detection of tasks that write memory another task uses
// (host code shared by the batch planner and the polynomial-multiply server).
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <map>
#include <queue>
#include <vector>

namespace bntt {

// A byte range one task reads or writes.
struct Span {
    uintptr_t lo, hi;  // [lo, hi)
    size_t task;
    bool write;
};

// Whether some task writes bytes that another task reads or writes (sweep over
// spans by start; active spans are counted per task, all and writes only).
inline bool cross_task_overlap(std::vector<Span> spans)
{
    std::sort(spans.begin(), spans.end(), [](const Span &a, const Span &b) { return a.lo < b.lo; });
    typedef std::pair<uintptr_t, size_t> End;  // (hi, index)
    std::priority_queue<End, std::vector<End>, std::greater<End>> ends;
    std::map<size_t, size_t> active, active_w;  // task -> spans still open
    auto drop = [](std::map<size_t, size_t> &m, size_t t) {
        auto it = m.find(t);
        if (--it->second == 0) m.erase(it);
    };
    for (size_t i = 0; i < spans.size(); i++) {
        const Span &cur = spans[i];
        while (!ends.empty() && ends.top().first <= cur.lo) {
            const Span &old = spans[ends.top().second];
            drop(active, old.task);
            if (old.write) drop(active_w, old.task);
            ends.pop();
        }
        const std::map<size_t, size_t> &against = cur.write ? active : active_w;
        if (against.size() > 1 || (against.size() == 1 && against.begin()->first != cur.task)) return true;
        active[cur.task]++;
        if (cur.write) active_w[cur.task]++;
        ends.push(End(cur.hi, i));
    }
    return false;
}

}  // namespace bntt
