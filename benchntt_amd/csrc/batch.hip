// batch.hip -- uneven batches of forward NTTs (and pointwise products of their
// outputs) in one persistent work-queue kernel, replayed from a hipGraph.
//
// A plan turns a list of transforms of sizes 2^1..2^24, each with its own
// input, output and twiddle table, plus a list of pointwise products
// c = a * b mod q, into "levels" of equal-size work items:
//   - a single-pass item runs all stages of up to 2048/N whole transforms
//     (N <= 2048) that share one twiddle table;
//   - a multi-pass transform (larger N) contributes one item per 2048-value
//     block to each of its passes (ceil(log2 N / 8) by default; stages split
//     by pass_split), pass p at level p, through a plan-owned intermediate
//     buffer (when the plan has products, the smaller part of the multi-pass
//     work starts one level later, so that products overlap transforms);
//   - a product contributes one item per 2048 elements at the level after the
//     last pass of the transforms whose outputs it reads.
// Each level is one launch of the same persistent kernel (queue.hpp): a grid
// sized to the resident capacity of the chip whose workgroups start with the
// item of their own index and draw the rest from an atomic counter until the
// level's list is exhausted, the last draw rearming the counter for the next
// launch.  Items are ordered largest-first with the memory-bound products
// spread among the compute-bound transform items.
// Kernel boundaries order the levels, so no workgroup ever waits on another.
//
// The radix (points per thread per register sub-pass: 2, 4 or 8, hence the
// LDS tile of 512/1024/2048 values of a 256-thread workgroup) and the number
// of passes are chosen per transform size from a tuning table measured on the
// device (benchntt_amd/autotune.py), or from built-in defaults.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/bntt.h"
#include "internal.hpp"
#include "ntt_core.hpp"

// share of the multi-pass transform work that starts at level 0 when the
// plan has products (the rest starts at level 1)
#ifndef BNTT_STAGGER_SHARE
#define BNTT_STAGGER_SHARE 0.7
#endif
static constexpr double kStaggerShare = BNTT_STAGGER_SHARE;

#include "queue.hpp"
#include "spans.hpp"

namespace bntt {

// Clears the table flags before the twiddle tables are converted (a kernel
// rather than a memset, so that graph replays clear them too).  The punt
// counts need no clearing: each slow launch clears the count it consumed.
__global__ __launch_bounds__(256) void exec_clear_kernel(uint32_t *flags, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) flags[i] = 0;
}

__global__ __launch_bounds__(256) void tw_convert_kernel(const TwChunk *__restrict__ chunks,
                                                         const TwJob *__restrict__ jobs)
{
    const TwChunk ch = chunks[blockIdx.x];
    const TwJob jb = jobs[ch.job];
    bool big = false;
    for (uint32_t e = threadIdx.x; e < ch.count; e += 256) {
        const uint64_t i = ch.off + e;
        const U128 a = ld16(jb.src + i);
        big |= (i != 0) && tw_big(a);  // entry 0 is never read by the network
        jb.dst[i] = to_r31(a);
    }
    if (__syncthreads_or(big) && threadIdx.x == 0 && jb.flag) atomicOr(jb.flag, 1u);
}

}  // namespace bntt

// ------------------------------------------------------------------ host side

using bntt::DevDesc;
using bntt::DevMul;
using bntt::Mod;
using bntt::U128;
using bntt::WorkItem;

struct bntt_batch_plan {
    int dev = -1;
    Mod md{};
    std::vector<size_t> level_items;     // items per level
    std::vector<size_t> level_offset;    // first item of each level in items_d
    std::vector<size_t> level_lds;       // dynamic LDS bytes of each level's launch
    std::vector<bntt::TwArea> level_tw;  // staged-twiddle area of each level's launch
    std::vector<int> level_grid;         // workgroups of each level's launch
    std::vector<int> level_kind;         // fast kernel of each level (QK_FAST or QK_FAST8)
    bool fixed_tw = false;               // digit tables made once, at creation
    WorkItem *items_d = nullptr;
    DevDesc *descs_d = nullptr;
    DevMul *muls_d = nullptr;
    bntt::QueueCtl *ctl_d = nullptr;       // fast launches' counter, then the slow launches'
    uint32_t *punt_d = nullptr;            // punt count of each level, then one index slot per item
    int slow_grid = 0;                     // workgroups of a slow launch over a punt list
    U128 *tmp_d = nullptr;
    size_t tmp_bytes = 0;
    bntt::R31 *twr_d = nullptr;          // digit copies of the distinct tables
    uint32_t *twflag_d = nullptr;        // one flag per distinct table
    bntt::TwJob *twjobs_d = nullptr;
    bntt::TwChunk *twchunks_d = nullptr;
    uint32_t ntwjobs = 0, ntwchunks = 0;
    size_t nitems = 0;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // Executions share the work-queue counters, punt lists, intermediates and
    // digit tables, so they must not overlap on the device: each execution
    // outside a stream capture records `done` on its stream, and the next one
    // on another stream waits for it first.
    hipEvent_t done = nullptr;
    hipStream_t last_stream = nullptr;
    bool executed = false;
    std::mutex mu;
};

namespace {

struct TuneChoice {
    int elog;
    int np;
};

TuneChoice default_choice(int L)
{
    if (L <= 11) return {L < 3 ? L : 3, 1};
    return {3, (L + 7) / 8};
}

bool valid_choice(int L, const TuneChoice &c)
{
    if (c.elog < 1 || c.elog > bntt::kQueueMaxElog || c.np < 1 || c.np > 3) return false;
    const int tile_log = 8 + c.elog;
    if (c.np == 1) return L <= tile_log && L >= c.elog;
    // every pass must fit the tile and hold at least elog stages
    int ks[3];
    return bntt::pass_split(L, c.np, c.elog, tile_log, ks);
}

// Whether two transforms of size 2^L can share the tiles of their last pass
// (it leaves at least two groups per tile).
bool pairable(int L, const TuneChoice &c)
{
    const int tile_log = 8 + c.elog;
    if (c.np == 1) return L < tile_log;
    int ks[3];
    if (!bntt::pass_split(L, c.np, c.elog, tile_log, ks)) return false;
    return ks[c.np - 1] < tile_log;
}

}  // namespace

namespace {

template <typename T>
int upload(T **dst, const std::vector<T> &v)
{
    if (v.empty()) return BNTT_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), v.size() * sizeof(T));
    if (e != hipSuccess) return bntt_set_error(BNTT_ENOMEM, "hipMalloc for a batch plan failed");
    e = hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bntt_set_hip_error(e, "hipMemcpy of a batch plan");
    return BNTT_OK;
}

void free_plan(bntt_batch_plan *p)
{
    if (!p) return;
    if (p->done) (void)hipEventDestroy(p->done);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->items_d);
    (void)hipFree(p->descs_d);
    (void)hipFree(p->muls_d);
    (void)hipFree(p->ctl_d);
    (void)hipFree(p->punt_d);
    (void)hipFree(p->tmp_d);
    (void)hipFree(p->twr_d);
    (void)hipFree(p->twflag_d);
    (void)hipFree(p->twjobs_d);
    (void)hipFree(p->twchunks_d);
    delete p;
}

// Digit copies and large-entry flags of the plan's twiddle tables.
int convert_tables(bntt_batch_plan *p, hipStream_t s)
{
    if (!p->ntwjobs) return BNTT_OK;
    hipLaunchKernelGGL(bntt::exec_clear_kernel, dim3((p->ntwjobs + 255) / 256), dim3(256), 0, s, p->twflag_d,
                       p->ntwjobs);
    hipLaunchKernelGGL(bntt::tw_convert_kernel, dim3(p->ntwchunks), dim3(256), 0, s, p->twchunks_d, p->twjobs_d);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bntt_set_hip_error(e, "twiddle conversion launch");
    return BNTT_OK;
}

int launch_levels(bntt_batch_plan *p, hipStream_t s)
{
    // each level: the fast launch, then a slow launch over the items it left to
    // the literal formulas (usually none); a modulus not of the fast form runs
    // every level on the literal formulas
    const uint32_t nlev = (uint32_t)p->level_items.size();
    if (!p->fixed_tw) {
        const int rc = convert_tables(p, s);
        if (rc) return rc;
    }
    for (size_t l = 0; l < nlev; l++) {
        if (p->level_items[l] == 0) continue;
        const WorkItem *items = p->items_d + p->level_offset[l];
        const uint32_t n = (uint32_t)p->level_items[l];
        if (p->md.pm) {
            const bntt::PuntList punt{p->punt_d + l, p->punt_d + nlev + p->level_offset[l]};
            if (p->level_kind[l] == bntt::QK_FAST8)
                hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_FAST8>, dim3(p->level_grid[l]),
                                   dim3(bntt::kQueueThreads), p->level_lds[l], s, items, n, p->descs_d, p->muls_d,
                                   p->ctl_d, p->level_tw[l], p->md, punt);
            else
                hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_FAST>, dim3(p->level_grid[l]),
                                   dim3(bntt::kQueueThreads), p->level_lds[l], s, items, n, p->descs_d, p->muls_d,
                                   p->ctl_d, p->level_tw[l], p->md, punt);
            hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_SLOW>, dim3(p->slow_grid), dim3(bntt::kQueueThreads),
                               p->level_lds[l], s, items, 0u, p->descs_d, p->muls_d, p->ctl_d + 1, bntt::TwArea{},
                               p->md, punt);
        } else {
            hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_SLOW>, dim3(p->level_grid[l]), dim3(bntt::kQueueThreads),
                               p->level_lds[l], s, items, n, p->descs_d, p->muls_d, p->ctl_d + 1, bntt::TwArea{},
                               p->md, bntt::PuntList{nullptr, nullptr});
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return bntt_set_hip_error(e, "queue kernel launch");
    }
    return BNTT_OK;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int bntt_batch_plan_create(bntt_batch_plan **out, const bntt_ntt_task *tasks, size_t ntasks,
                           const bntt_mul_task *muls, size_t nmuls, const bntt_u128 *m, const bntt_u128 *mu,
                           const bntt_tune_entry *tune, size_t ntune, int flags)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if (!out || !m || !mu) return bntt_set_error(BNTT_EINVAL, "null pointer argument");
    *out = nullptr;
    if ((ntasks && !tasks) || (nmuls && !muls) || (ntune && !tune))
        return bntt_set_error(BNTT_EINVAL, "null task list");
    if (ntasks > 0xffffffffu || nmuls > 0xffffffffu) return bntt_set_error(BNTT_EINVAL, "too many tasks");

    TuneChoice choice[bntt::kMaxLog2 + 1];
    for (int L = 0; L <= bntt::kMaxLog2; L++) choice[L] = default_choice(L < 1 ? 1 : L);
    for (size_t i = 0; i < ntune; i++) {
        const int L = tune[i].log2n;
        if (L < 1 || L > bntt::kMaxLog2) return bntt_set_error(BNTT_EINVAL, "tuning entry log2n out of range");
        const TuneChoice c{tune[i].elog, tune[i].passes};
        if (!valid_choice(L, c)) return bntt_set_error(BNTT_EINVAL, "tuning entry not valid for its size");
        choice[L] = c;
    }

    bntt_batch_plan *p = new bntt_batch_plan();
    if (hipGetDevice(&p->dev) != hipSuccess) {
        delete p;
        return bntt_set_error(BNTT_EHIP, "hipGetDevice");
    }
    p->md = bntt_make_mod(m, mu);
    p->use_graph = (flags & BNTT_PLAN_GRAPH) != 0;
    p->fixed_tw = (flags & BNTT_PLAN_TW_FIXED) != 0;
    // the event only orders executions on one device, which the end of each
    // kernel already releases at device scope: no system-scope fence (that
    // would write back the caches at every execution)
    if (hipEventCreateWithFlags(&p->done, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
        free_plan(p);
        return bntt_set_error(BNTT_EHIP, "hipEventCreateWithFlags");
    }

    // descriptors: one per transform; multi-pass transforms get tmp space
    std::vector<DevDesc> descs(ntasks);
    std::vector<size_t> tmp_off(ntasks, 0);
    size_t tmp_elems = 0;
    for (size_t i = 0; i < ntasks; i++) {
        const bntt_ntt_task &t = tasks[i];
        const int L = t.log2n;
        if (L < 1 || L > bntt::kMaxLog2) {
            free_plan(p);
            return bntt_set_error(BNTT_EINVAL, "task log2n out of range");
        }
        if (!t.x || !t.y || !t.tw || (reinterpret_cast<uintptr_t>(t.x) | reinterpret_cast<uintptr_t>(t.y) |
                                      reinterpret_cast<uintptr_t>(t.tw)) & 15u) {
            free_plan(p);
            return bntt_set_error(BNTT_EINVAL, "task pointers must be non-null and 16-byte aligned");
        }
        descs[i].x = static_cast<const U128 *>(t.x);
        descs[i].y = static_cast<U128 *>(t.y);
        descs[i].tw = static_cast<const U128 *>(t.tw);
        descs[i].tmp = nullptr;
        if (choice[L].np > 1) {
            const size_t n = (size_t)1 << L;
            const char *xb = static_cast<const char *>(t.x), *yb = static_cast<const char *>(t.y);
            if (yb < xb + n * 16 && xb < yb + n * 16) {
                free_plan(p);
                return bntt_set_error(BNTT_EINVAL, "multi-pass transforms need y not overlapping x");
            }
            tmp_off[i] = tmp_elems;
            tmp_elems += n;
        }
    }
    // distinct twiddle tables -> digit copies and flags
    {
        std::unordered_map<const void *, uint32_t> table_id;
        std::vector<uint64_t> table_n;
        std::vector<const U128 *> table_src;
        std::vector<uint32_t> desc_table(ntasks);
        for (size_t i = 0; i < ntasks; i++) {
            auto it = table_id.find(tasks[i].tw);
            const uint64_t n = (uint64_t)1 << tasks[i].log2n;
            if (it == table_id.end()) {
                const uint32_t id = (uint32_t)table_n.size();
                table_id.emplace(tasks[i].tw, id);
                table_n.push_back(n);
                table_src.push_back(static_cast<const U128 *>(tasks[i].tw));
                desc_table[i] = id;
            } else {
                desc_table[i] = it->second;
                if (table_n[it->second] < n) table_n[it->second] = n;  // one table serving several sizes
            }
        }
        const uint32_t nt = (uint32_t)table_n.size();
        std::vector<uint64_t> toff(nt);
        uint64_t tot = 0;
        for (uint32_t t = 0; t < nt; t++) {
            toff[t] = tot;
            tot += table_n[t];
        }
        if (nt) {
            if (hipMalloc(&p->twr_d, tot * sizeof(bntt::R31)) != hipSuccess ||
                hipMalloc(&p->twflag_d, nt * sizeof(uint32_t)) != hipSuccess) {
                free_plan(p);
                return bntt_set_error(BNTT_ENOMEM, "hipMalloc for twiddle digit tables failed");
            }
            std::vector<bntt::TwJob> jobs(nt);
            std::vector<bntt::TwChunk> chunks;
            for (uint32_t t = 0; t < nt; t++) {
                jobs[t] = bntt::TwJob{table_src[t], p->twr_d + toff[t], p->twflag_d + t, table_n[t]};
                for (uint64_t o = 0; o < table_n[t]; o += bntt::kTwChunk)
                    chunks.push_back(
                        bntt::TwChunk{t, (uint32_t)std::min<uint64_t>(bntt::kTwChunk, table_n[t] - o), o});
            }
            if ((rc = upload(&p->twjobs_d, jobs)) || (rc = upload(&p->twchunks_d, chunks))) {
                free_plan(p);
                return rc;
            }
            p->ntwjobs = nt;
            p->ntwchunks = (uint32_t)chunks.size();
            for (size_t i = 0; i < ntasks; i++) {
                descs[i].twr = p->twr_d + toff[desc_table[i]];
                descs[i].twflag = p->twflag_d + desc_table[i];
            }
        }
    }
    if (tmp_elems) {
        p->tmp_bytes = tmp_elems * sizeof(U128);
        if (hipMalloc(&p->tmp_d, p->tmp_bytes) != hipSuccess) {
            free_plan(p);
            return bntt_set_error(BNTT_ENOMEM, "hipMalloc for batch intermediates failed");
        }
        for (size_t i = 0; i < ntasks; i++)
            if (choice[tasks[i].log2n].np > 1) descs[i].tmp = p->tmp_d + tmp_off[i];
    }

    // Products fused into their transforms' last pass: c = y_A * y_B where A
    // and B are transforms of one size and one table whose outputs are exactly
    // the product's operands, each transform in at most one pair, and whose
    // passes all leave room for two groups per tile.
    std::vector<int64_t> partner(ntasks, -1);  // the other transform of a pair
    std::vector<int64_t> pair_mul(ntasks, -1); // the pair's product (on its A)
    std::vector<char> fused(nmuls, 0);
    size_t nunfused = nmuls;
    {
        std::unordered_map<uintptr_t, size_t> by_y;
        for (size_t j = 0; j < ntasks; j++) by_y.emplace(reinterpret_cast<uintptr_t>(tasks[j].y), j);
        for (size_t i = 0; i < nmuls; i++) {
            const uint64_t n = muls[i].n;
            if (n < 2 || (n & (n - 1)) != 0) continue;
            const int L = 63 - __builtin_clzll(n);
            auto fa = by_y.find(reinterpret_cast<uintptr_t>(muls[i].a));
            auto fb = by_y.find(reinterpret_cast<uintptr_t>(muls[i].b));
            if (fa == by_y.end() || fb == by_y.end() || fa->second == fb->second) continue;
            const size_t ia = fa->second, ib = fb->second;
            if (tasks[ia].log2n != L || tasks[ib].log2n != L || tasks[ia].tw != tasks[ib].tw) continue;
            if (partner[ia] >= 0 || partner[ib] >= 0 || !pairable(L, choice[L])) continue;
            partner[ia] = (int64_t)ib;
            partner[ib] = (int64_t)ia;
            pair_mul[ia] = (int64_t)i;
            fused[i] = 1;
            nunfused--;
        }
    }

    // transform items, by level; single-pass transforms of one size and one
    // twiddle table are packed several per item (descriptors made contiguous)
    std::vector<std::vector<WorkItem>> ntt_lv(4), mul_lv(5);
    std::vector<DevDesc> packed;
    packed.reserve(ntasks);
    std::vector<size_t> order(ntasks);
    for (size_t i = 0; i < ntasks; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        if (tasks[a].log2n != tasks[b].log2n) return tasks[a].log2n > tasks[b].log2n;
        return reinterpret_cast<uintptr_t>(tasks[a].tw) < reinterpret_cast<uintptr_t>(tasks[b].tw);
    });
    std::vector<int> out_level(ntasks, 0);  // level at which each transform's output is final
    // With unfused products to run, multi-pass transforms start in two
    // groups: the first kStaggerShare of their work (largest first) at level
    // 0, the rest one level later.  The products of the first group then run
    // beside the last pass of the second (memory-bound items among
    // compute-bound ones) instead of all waiting for a level of their own after
    // every transform.
    std::vector<int> start_level(ntasks, 0);
    bool stagger = false;  // some unfused product reads a multi-pass transform's output
    if (nunfused > 0) {
        std::map<uintptr_t, uintptr_t> multi;  // y -> end, multi-pass transforms
        for (size_t j = 0; j < ntasks; j++)
            if (choice[tasks[j].log2n].np > 1) {
                const uintptr_t y = reinterpret_cast<uintptr_t>(tasks[j].y);
                multi[y] = y + (sizeof(U128) << tasks[j].log2n);
            }
        auto reads_multi = [&](const void *a, uint64_t n) {
            const uintptr_t lo = reinterpret_cast<uintptr_t>(a), hi = lo + sizeof(U128) * n;
            auto it = multi.upper_bound(lo);
            if (it != multi.begin()) --it;
            for (; it != multi.end() && it->first < hi; ++it)
                if (it->second > lo) return true;
            return false;
        };
        for (size_t i = 0; i < nmuls && !stagger; i++)
            stagger = !fused[i] && (reads_multi(muls[i].a, muls[i].n) || reads_multi(muls[i].b, muls[i].n));
    }
    if (stagger) {
        auto work = [&](size_t j) { return (double)tasks[j].log2n * (double)((size_t)1 << tasks[j].log2n); };
        double total = 0, acc = 0;
        for (size_t j : order)
            if (choice[tasks[j].log2n].np > 1) total += work(j);
        for (size_t j : order)
            if (choice[tasks[j].log2n].np > 1) {
                if (acc >= kStaggerShare * total) start_level[j] = 1;
                acc += work(j);
            }
        // the two transforms of a pair start together
        for (size_t j = 0; j < ntasks; j++)
            if (pair_mul[j] >= 0) start_level[(size_t)partner[j]] = start_level[j];
    }
    auto with_prod = [&](size_t j) {
        DevDesc d = descs[j];
        d.prod = pair_mul[j] >= 0 ? static_cast<U128 *>(muls[pair_mul[j]].c) : nullptr;
        return d;
    };
    size_t i0 = 0;
    while (i0 < ntasks) {
        const size_t ti = order[i0];
        const int L = tasks[ti].log2n;
        const TuneChoice c = choice[L];
        if (c.np == 1) {
            // the run of transforms of this size and table
            size_t i1 = i0;
            while (i1 < ntasks && tasks[order[i1]].log2n == L && tasks[order[i1]].tw == tasks[ti].tw) i1++;
            std::vector<size_t> single, firsts;  // unpaired; A of each pair
            for (size_t j = i0; j < i1; j++) {
                const size_t t = order[j];
                if (partner[t] < 0)
                    single.push_back(t);
                else if (pair_mul[t] >= 0)
                    firsts.push_back(t);
            }
            const int Glog = 8 + c.elog - L;  // transforms per item
            WorkItem w{};
            w.kind = bntt::ITEM_NTT;
            w.L = (uint8_t)L;
            w.S = 0;
            w.k = (uint8_t)L;
            w.Glog = (uint8_t)Glog;
            w.elog = (uint8_t)c.elog;
            w.pass = 0;
            w.np = 1;
            w.first = 0;
            for (size_t s0 = 0; s0 < single.size(); s0 += ((size_t)1 << Glog)) {
                const size_t s1 = std::min(single.size(), s0 + ((size_t)1 << Glog));
                w.desc = (uint32_t)packed.size();
                w.nvalid = (uint32_t)(s1 - s0);
                w.flags = 0;
                // transforms back to back in x and in y: one base pointer each
                bool contig = true;
                for (size_t j = s0 + 1; j < s1 && contig; j++)
                    contig = descs[single[j]].x == descs[single[s0]].x + ((j - s0) << L) &&
                             descs[single[j]].y == descs[single[s0]].y + ((j - s0) << L);
                if (contig) w.flags = bntt::IT_CONTIG;
                for (size_t j = s0; j < s1; j++) packed.push_back(descs[single[j]]);
                ntt_lv[0].push_back(w);
            }
            // pairs: up to 2^(Glog-1) per item, descriptors A's then B's
            for (size_t s0 = 0; s0 < firsts.size(); s0 += ((size_t)1 << (Glog - 1))) {
                const size_t s1 = std::min(firsts.size(), s0 + ((size_t)1 << (Glog - 1)));
                w.desc = (uint32_t)packed.size();
                w.nvalid = (uint32_t)(s1 - s0);
                w.flags = bntt::IT_PAIR | bntt::IT_PROD;
                for (size_t j = s0; j < s1; j++) packed.push_back(with_prod(firsts[j]));
                for (size_t j = s0; j < s1; j++) packed.push_back(descs[(size_t)partner[firsts[j]]]);
                ntt_lv[0].push_back(w);
            }
            for (size_t j = i0; j < i1; j++) out_level[order[j]] = 0;
            i0 = i1;
        } else {
            if (partner[ti] >= 0 && pair_mul[ti] < 0) {  // a pair's B: emitted with its A
                i0++;
                continue;
            }
            const bool pair = partner[ti] >= 0;
            const uint32_t di = (uint32_t)packed.size();
            packed.push_back(with_prod(ti));
            if (pair) packed.push_back(descs[(size_t)partner[ti]]);
            int ks[3];
            bntt::pass_split(L, c.np, c.elog, 8 + c.elog, ks);  // valid: checked with the choice
            int S = 0;
            for (int pidx = 0; pidx < c.np; pidx++) {
                const int k = ks[pidx];
                const int Glog = 8 + c.elog - k;
                const size_t groups = (size_t)1 << (L - k);
                // a pair shares the items of its last pass only (that pass
                // reads each group contiguously, so half a tile per transform
                // still moves long runs); earlier passes run each transform
                // on its own
                const bool shared = pair && pidx + 1 == c.np;
                const int plog = Glog - (shared ? 1 : 0);  // groups of one transform per item
                for (int which = 0; which < (pair && !shared ? 2 : 1); which++)
                    for (size_t g0 = 0; g0 < groups; g0 += ((size_t)1 << plog)) {
                        WorkItem w{};
                        w.kind = bntt::ITEM_NTT;
                        w.L = (uint8_t)L;
                        w.S = (uint8_t)S;
                        w.k = (uint8_t)k;
                        w.Glog = (uint8_t)Glog;
                        w.elog = (uint8_t)c.elog;
                        w.pass = (uint8_t)pidx;
                        w.np = (uint8_t)c.np;
                        w.desc = di + (uint32_t)which;
                        w.nvalid = (uint32_t)std::min(groups - g0, (size_t)1 << plog);
                        w.first = g0;
                        if (shared) w.flags = bntt::IT_PAIR | bntt::IT_PROD;
                        ntt_lv[start_level[ti] + pidx].push_back(w);
                    }
                S += k;
            }
            out_level[ti] = start_level[ti] + c.np - 1;
            if (pair) out_level[(size_t)partner[ti]] = out_level[ti];
            i0++;
        }
    }

    // transforms are independent: none may write what another reads or writes;
    // products may read transform outputs and write anything no transform uses
    {
        std::vector<bntt::Span> spans;
        spans.reserve(2 * ntasks + 3 * nmuls);
        for (size_t j = 0; j < ntasks; j++) {
            const uintptr_t bytes = sizeof(U128) << tasks[j].log2n;
            const uintptr_t x = reinterpret_cast<uintptr_t>(tasks[j].x), y = reinterpret_cast<uintptr_t>(tasks[j].y);
            spans.push_back(bntt::Span{x, x + bytes, j, false});
            spans.push_back(bntt::Span{y, y + bytes, j, true});
        }
        for (size_t i = 0; i < nmuls; i++) {
            const uintptr_t c = reinterpret_cast<uintptr_t>(muls[i].c);
            spans.push_back(bntt::Span{c, c + sizeof(U128) * muls[i].n, ntasks + i, true});
        }
        std::vector<bntt::Span> prod;  // products among themselves: operands against results
        prod.reserve(3 * nmuls);
        for (size_t i = 0; i < nmuls; i++) {
            const uintptr_t bytes = sizeof(U128) * muls[i].n;
            const uintptr_t a = reinterpret_cast<uintptr_t>(muls[i].a), b = reinterpret_cast<uintptr_t>(muls[i].b),
                            c = reinterpret_cast<uintptr_t>(muls[i].c);
            prod.push_back(bntt::Span{a, a + bytes, i, false});
            prod.push_back(bntt::Span{b, b + bytes, i, false});
            prod.push_back(bntt::Span{c, c + bytes, i, true});
        }
        if (bntt::cross_task_overlap(spans) || bntt::cross_task_overlap(prod)) {
            free_plan(p);
            return bntt_set_error(BNTT_EINVAL, "a task writes memory another task reads or writes");
        }
    }
    // products: level after the last pass of every transform whose output
    // overlaps one of their operands
    std::map<uintptr_t, std::pair<uintptr_t, int>> outputs;  // y -> (end, level after it)
    for (size_t j = 0; j < ntasks; j++) {
        const uintptr_t y = reinterpret_cast<uintptr_t>(tasks[j].y);
        outputs[y] = std::make_pair(y + (sizeof(U128) << tasks[j].log2n), out_level[j] + 1);
    }
    auto ready_level = [&](const void *a, uint64_t n) {
        const uintptr_t lo = reinterpret_cast<uintptr_t>(a), hi = lo + sizeof(U128) * n;
        int lvl = 0;
        // outputs are disjoint: those overlapping [lo, hi) start below hi and end above lo
        auto it = outputs.upper_bound(lo);
        if (it != outputs.begin()) --it;
        for (; it != outputs.end() && it->first < hi; ++it)
            if (it->second.first > lo) lvl = std::max(lvl, it->second.second);
        return lvl;
    };
    std::vector<DevMul> mv(nmuls);
    for (size_t i = 0; i < nmuls; i++) {
        const bntt_mul_task &t = muls[i];
        if (!t.a || !t.b || !t.c ||
            (reinterpret_cast<uintptr_t>(t.a) | reinterpret_cast<uintptr_t>(t.b) | reinterpret_cast<uintptr_t>(t.c)) &
                15u) {
            free_plan(p);
            return bntt_set_error(BNTT_EINVAL, "product pointers must be non-null and 16-byte aligned");
        }
        mv[i] = DevMul{static_cast<const U128 *>(t.a), static_cast<const U128 *>(t.b), static_cast<U128 *>(t.c),
                       t.n};
        if (fused[i]) continue;  // computed by its pair's last pass
        const int lvl = std::max(ready_level(t.a, t.n), ready_level(t.b, t.n));
        for (uint64_t off = 0; off < t.n; off += bntt::kQueueTile) {
            WorkItem w{};
            w.kind = bntt::ITEM_MUL;
            w.desc = (uint32_t)i;
            w.nvalid = (uint32_t)std::min<uint64_t>(t.n - off, bntt::kQueueTile);
            w.first = off;
            mul_lv[lvl].push_back(w);
        }
    }

    // order each level: largest transform items first, products spread evenly
    std::vector<WorkItem> all;
    int grid_cap = 0;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, p->dev) != hipSuccess) {
            free_plan(p);
            return bntt_set_error(BNTT_EHIP, "hipGetDeviceProperties");
        }
        grid_cap = prop.multiProcessorCount;
    }
    // dynamic LDS above 64 KiB must be allowed before occupancy is queried
    static std::atomic<uint64_t> lds_fast{0}, lds_fast8{0}, lds_slow{0};
    hipError_t attr = bntt_allow_lds(reinterpret_cast<const void *>(&bntt::queue_kernel<bntt::QK_FAST>),
                                     (int)bntt::kQueueMaxLds, lds_fast);
    if (attr == hipSuccess)
        attr = bntt_allow_lds(reinterpret_cast<const void *>(&bntt::queue_kernel<bntt::QK_FAST8>),
                              (int)bntt::kQueueMaxLds, lds_fast8);
    if (attr == hipSuccess)
        attr = bntt_allow_lds(reinterpret_cast<const void *>(&bntt::queue_kernel<bntt::QK_SLOW>),
                              (int)bntt::kQueueMaxLds, lds_slow);
    if (attr != hipSuccess) {
        free_plan(p);
        return bntt_set_hip_error(attr, "hipFuncSetAttribute");
    }
    const size_t nlevels = 5;
    for (size_t l = 0; l < nlevels; l++) {
        std::vector<WorkItem> nt = l < ntt_lv.size() ? ntt_lv[l] : std::vector<WorkItem>();
        std::stable_sort(nt.begin(), nt.end(), [](const WorkItem &x, const WorkItem &y) {
            const int cx = (int)x.k << x.Glog, cy = (int)y.k << y.Glog;  // stages x values ~ work
            return cx > cy;
        });
        const std::vector<WorkItem> &mm = mul_lv[l];
        std::vector<WorkItem> lv;
        lv.reserve(nt.size() + mm.size());
        size_t a_i = 0, b_i = 0;
        while (a_i < nt.size() || b_i < mm.size()) {
            // keep the fraction of products emitted equal to their share of the level
            if (b_i < mm.size() && (a_i >= nt.size() || b_i * nt.size() <= a_i * mm.size()))
                lv.push_back(mm[b_i++]);
            else
                lv.push_back(nt[a_i++]);
        }
        size_t data_slots = 0;
        std::vector<size_t> needs;
        bool all8 = true;  // every transform item has a 2048-value tile
        for (const WorkItem &w : lv)
            if (w.kind == bntt::ITEM_NTT) {
                data_slots = std::max(data_slots, bntt::lds_data_slots(w.k, w.Glog));
                needs.push_back(bntt::tw_slots(w.k, bntt::item_tw_wlog(w.L, w.k, w.S, bntt::item_plog(w))));
                all8 = all8 && w.elog == 3 && w.k >= 3;
            }
        const size_t budget = (flags & BNTT_PLAN_TW_WIDE) ? bntt::kQueueMaxLds
                              : (flags & BNTT_PLAN_TW_LDS) ? bntt::kQueueLds3
                                                           : 0;
        bntt::TwArea area;
        size_t lds = bntt::plan_tw_area(data_slots, needs.data(), needs.size(), budget, &area);
        // the code specialised per stage count when the twiddles come from
        // device memory and every tile is a radix-8 one
        const int kind = (p->md.pm && budget == 0 && all8) ? bntt::QK_FAST8 : bntt::QK_FAST;

        int per_cu = 0;
        const hipError_t oe =
            kind == bntt::QK_FAST8
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bntt::queue_kernel<bntt::QK_FAST8>,
                                                               bntt::kQueueThreads, lds)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bntt::queue_kernel<bntt::QK_FAST>,
                                                               bntt::kQueueThreads, lds);
        if (oe != hipSuccess || per_cu < 1) per_cu = 1;
        p->level_kind.push_back(kind);
        // Items of equal work gain nothing from the dynamic draw: then the
        // level gets a workgroup per item (no counter atomics), otherwise a
        // persistent grid of the resident capacity
        bool equal = !lv.empty();
        for (const WorkItem &w : lv)
            equal = equal && w.kind == bntt::ITEM_NTT && w.k == lv[0].k && w.L == lv[0].L && w.Glog == lv[0].Glog &&
                    w.elog == lv[0].elog && w.np == lv[0].np && w.flags == lv[0].flags;
        const size_t cap = equal ? (size_t)0x7fffffff : (size_t)grid_cap * per_cu;
        const int grid = (int)std::min<size_t>(std::max<size_t>(lv.size(), 1), cap);
        p->level_offset.push_back(all.size());
        p->level_items.push_back(lv.size());
        p->level_lds.push_back(lds);
        p->level_tw.push_back(area);
        p->level_grid.push_back(grid);
        all.insert(all.end(), lv.begin(), lv.end());
    }
    while (!p->level_items.empty() && p->level_items.back() == 0) {
        p->level_kind.pop_back();
        p->level_items.pop_back();
        p->level_offset.pop_back();
        p->level_lds.pop_back();
        p->level_tw.pop_back();
        p->level_grid.pop_back();
    }
    p->nitems = all.size();

    for (size_t lds : p->level_lds)
        if (lds > bntt::kQueueMaxLds) {
            free_plan(p);
            return bntt_set_error(BNTT_EINVAL, "internal: batch item exceeds the LDS budget");
        }
    if ((rc = upload(&p->items_d, all)) || (rc = upload(&p->descs_d, packed)) || (rc = upload(&p->muls_d, mv))) {
        free_plan(p);
        return rc;
    }
    // punt lists: a count per level, then one slot per item of the plan
    const size_t punt_words = p->level_items.size() + p->nitems;
    if (hipMalloc(&p->ctl_d, 2 * sizeof(bntt::QueueCtl)) != hipSuccess ||
        hipMalloc(&p->punt_d, std::max<size_t>(punt_words, 1) * sizeof(uint32_t)) != hipSuccess) {
        free_plan(p);
        return bntt_set_error(BNTT_ENOMEM, "hipMalloc for queue counters failed");
    }
    p->slow_grid = grid_cap;  // one workgroup per CU: the literal formulas are the exception
    if (hipMemset(p->ctl_d, 0, 2 * sizeof(bntt::QueueCtl)) != hipSuccess ||
        hipMemset(p->punt_d, 0, std::max<size_t>(punt_words, 1) * sizeof(uint32_t)) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        free_plan(p);
        return bntt_set_error(BNTT_EHIP, "initialising queue counters failed");
    }
    if (p->fixed_tw) {
        // the tables' digit copies and flags, once
        if ((rc = convert_tables(p, nullptr)) || hipDeviceSynchronize() != hipSuccess) {
            free_plan(p);
            return rc ? rc : bntt_set_error(BNTT_EHIP, "converting the twiddle tables failed");
        }
    }
    *out = p;
    return BNTT_OK;
}

int bntt_batch_plan_execute(bntt_batch_plan *p, void *stream)
{
    if (!p) return bntt_set_error(BNTT_EINVAL, "null plan");
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(p->mu);
    // inside the caller's own stream capture the launches become part of the
    // caller's graph (whose replays the caller orders)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s != nullptr && hipStreamIsCapturing(s, &cs) != hipSuccess)
        return bntt_set_error(BNTT_EHIP, "hipStreamIsCapturing");
    if (cs != hipStreamCaptureStatusNone) return launch_levels(p, s);
    hipError_t e;
    if (p->executed && p->last_stream != s) {
        e = hipStreamWaitEvent(s, p->done, 0);
        if (e != hipSuccess) return bntt_set_hip_error(e, "hipStreamWaitEvent");
    }
    if (!p->use_graph || s == nullptr) {
        const int rc = launch_levels(p, s);
        if (rc) return rc;
    } else {
        if (!p->exec) {
            e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return bntt_set_hip_error(e, "hipStreamBeginCapture");
            const int rc = launch_levels(p, s);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(s, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return bntt_set_hip_error(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                return bntt_set_hip_error(e, "hipGraphInstantiate");
            }
            p->graph = g;
        }
        e = hipGraphLaunch(p->exec, s);
        if (e != hipSuccess) return bntt_set_hip_error(e, "hipGraphLaunch");
    }
    e = hipEventRecord(p->done, s);
    if (e != hipSuccess) return bntt_set_hip_error(e, "hipEventRecord");
    p->last_stream = s;
    p->executed = true;
    return BNTT_OK;
}

int bntt_batch_plan_info(const bntt_batch_plan *p, bntt_batch_info *info)
{
    if (!p || !info) return bntt_set_error(BNTT_EINVAL, "null argument");
    std::memset(info, 0, sizeof *info);
    info->levels = (int)p->level_items.size();
    info->items = p->nitems;
    info->tmp_bytes = p->tmp_bytes;
    for (size_t l = 0; l < p->level_items.size() && l < 8; l++) {
        info->level_items[l] = p->level_items[l];
        info->level_grid[l] = p->level_grid[l];
        info->level_lds_bytes[l] = (uint32_t)p->level_lds[l];
    }
    return BNTT_OK;
}

int bntt_batch_plan_destroy(bntt_batch_plan *p)
{
    if (!p) return BNTT_OK;
    const hipError_t e = hipDeviceSynchronize();
    free_plan(p);
    if (e != hipSuccess) return bntt_set_hip_error(e, "hipDeviceSynchronize");
    return BNTT_OK;
}

int bntt_ntt_launch_each_dev(const bntt_ntt_task *tasks, size_t ntasks, const bntt_u128 *m, const bntt_u128 *mu,
                             void *tmp, void *stream)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if (ntasks && !tasks) return bntt_set_error(BNTT_EINVAL, "null task list");
    if (!m || !mu) return bntt_set_error(BNTT_EINVAL, "null modulus");
    const Mod md = bntt_make_mod(m, mu);
    for (size_t i = 0; i < ntasks; i++) {
        const bntt_ntt_task &t = tasks[i];
        if (t.log2n < bntt::kMinLog2 || t.log2n > bntt::kMaxLog2)
            return bntt_set_error(BNTT_EINVAL, "task log2n out of range");
        if (t.log2n > bntt::kLdsMaxLog2 && !tmp)
            return bntt_set_error(BNTT_EINVAL, "tasks above 2^12 points need tmp");
        const hipError_t e = bntt::launch_ntt_uniform(static_cast<U128 *>(t.y), static_cast<const U128 *>(t.x),
                                                      static_cast<const U128 *>(t.tw), t.log2n, 1, md,
                                                      // one launch for every one-pass size
                                                      t.log2n > bntt::kLdsMaxLog2 ? static_cast<U128 *>(tmp)
                                                                                  : nullptr,
                                                      static_cast<hipStream_t>(stream));
        if (e != hipSuccess) return bntt_set_hip_error(e, "ntt kernel launch");
    }
    return BNTT_OK;
}

}  // extern "C"
#pragma GCC visibility pop
