// serve.hip -- negacyclic polynomial multiplication c = a * b mod (x^N + 1, q)
// with continuous batching: no counterpart in the reference (whose ntt() is a
// single forward transform); built from the same forward network.
//
// Pipeline of one request (N = 2^L, q odd < 2^122 with 2^(L+1) | q - 1), all
// products exact Montgomery products (R = 2^124) so no value leaves [0, q):
//   FA = F(a; psi table), FB = F(b; psi table)      F = the forward network
//   G[k] = FA[brv k] * FB[brv k] * R^-1               (fused into the first load of Y)
//   Y    = F(G; inverse-root table)
//   c[brv i] = Y[i] * S[i] * R^-1                     (fused into the last store of Y)
// with tw_psi[2^s + brv_s(j)] = psi_{2^(s+2)}^(2j+1) (the network is then the
// negacyclic NTT in bit-reversed order), tw_inv[2^s + brv_s(j)] =
// omega_{2^(s+1)}^-j (the cyclic inverse DFT in bit-reversed order), and
// S[i] = N^-1 psi_{2N}^-brv(i) R^2: for a bit-reversed-output network F,
// a_j = N^-1 psi^-j sum_k F(a)[brv k] omega^-jk, i.e. the inverse is F with
// inverse cyclic roots applied to the bit-reversed product, then a twist.
// Twiddle entries are stored as tw*R mod q in digits, so that the network's
// montmul(v, tw*R) = v*tw stays in the ordinary domain; the R^-1 of the
// pointwise product is cancelled by the extra R in S.
//
// A request needs 2N residues of intermediates (FA, FB) when its transforms
// take one pass, 4N (FA, FB and one pass buffer for each of a and b; the
// inverse reuses a's) when they take two, allocated from an arena of a fixed
// HBM budget.  The trace runner admits arrived requests in order while the
// arena has room, launches each batch as four persistent work-queue launches
// (forward pass 0 / 1, inverse pass 0 / 1) on one of several streams, and
// retires batches as their events complete, so batches of different arrivals
// overlap on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "../../include/bntt.h"
#include "internal.hpp"
#include "hostmont.hpp"
#include "queue.hpp"
#include "spans.hpp"

using bntt::DevDesc;
using bntt::Mod;
using bntt::R31;
using bntt::U128;
using bntt::WorkItem;

namespace {

typedef unsigned __int128 u128;
using bntt::HostMont;

uint32_t brv(uint32_t i, int bits) { return bntt::brv_host(i, bits); }

struct SizeTables {
    R31 *fwd = nullptr;    // tw_psi * R, digits
    R31 *inv = nullptr;    // tw_inv * R, digits
    R31 *scale = nullptr;  // N^-1 psi^-brv(i) R^2 by network output index i, then (entries
                           // N..2N-1) the same permuted to output position brv(i), digits
};

// First-fit allocator over [0, size) in 256-byte units.
class Arena {
  public:
    void reset(size_t units)
    {
        free_.clear();
        if (units) free_[0] = units;
        total_ = units;
        used_ = 0;
    }
    bool alloc(size_t units, size_t *off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= units) {
                *off = it->first;
                const size_t rest = it->second - units, start = it->first + units;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += units;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t units)
    {
        auto it = free_.emplace(off, units).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
        used_ -= units;
    }
    size_t total() const { return total_; }

  private:
    std::map<size_t, size_t> free_;
    size_t total_ = 0, used_ = 0;
};

constexpr size_t kUnit = 256;

int passes_for(int L) { return L <= 11 ? 1 : (L + 7) / 8; }
// Passes of a transform in a batch whose items hold 256 << want values: in
// small batches (want < 3) sizes 2^10 and 2^11 take two passes.  In one pass
// such a transform is a single item on one CU (about 12 and 20 us per level);
// in two, each pass spreads it over several, in the levels that the batch's
// larger sizes need anyway.
int passes_in_batch(int L, int want) { return (want < 3 && L >= 10 && L <= 11) ? 2 : passes_for(L); }
// Points per thread per register sub-pass, 2^elog: `want` (3 normally, 2 or 1
// for small batches: more items, each shorter), within what the size allows
// (a one-pass transform must fit the 256 << elog values of an item).
int elog_for(int L, int want, int np)
{
    int e = want < 3 ? want : 3;
    if (np == 1 && e < L - 8) e = L - 8;
    return e < L ? e : L;
}

// Arena units of a request in flight: FA and FB, plus a pass buffer each for
// a and b for sizes that may take two passes.
size_t request_units(int L)
{
    const size_t bufs = L < 10 ? 2 : 4;
    return ((bufs << L) * sizeof(U128) + kUnit - 1) / kUnit;
}

// Butterflies of one request: three transforms of N/2 * log2 N.
uint64_t request_butterflies(int L) { return (uint64_t)3 * ((uint64_t)1 << (L - 1)) * (uint64_t)L; }

// Upper bound of work items one request contributes (three transforms).
size_t items_per_request(int L, int want)
{
    const int np = passes_in_batch(L, want);
    if (np == 1) return 3;
    const int elog = elog_for(L, want, np);
    int ks[3];
    bntt::pass_split(L, np, elog, 8 + elog, ks);
    size_t per_transform = 0;
    for (int p = 0; p < np; p++) {
        const int Glog = 8 + elog - ks[p];
        per_transform += (((size_t)1 << (L - ks[p])) + ((size_t)1 << Glog) - 1) >> Glog;
    }
    return 3 * per_transform;
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bntt::QueueCtl *ctl = nullptr;
    char *batch_d = nullptr;  // a batch's items, then its descriptors (one upload)
    char *stage_h = nullptr;  // pinned staging of the same layout
    size_t cap_items = 0, cap_descs = 0;
    bool busy = false;
    uint64_t launched_ns = 0;                          // trace clock at the batch's launch
    std::vector<size_t> reqs;                          // request indices of the batch
    std::vector<std::pair<size_t, size_t>> allocs;     // arena blocks
    // work lists being built for the slot's next batch: kept across batches
    // so that building one reuses their memory (host time per batch is what
    // bounds the trace runner at middle loads)
    std::vector<std::vector<WorkItem>> lv;
    std::vector<DevDesc> descs, fwd, inv;
};

}  // namespace

struct bntt_pm_server {
    int dev = -1;
    Mod md{};
    HostMont hm;
    int min_l = 0, max_l = 0, max_batch = 0, nslots = 0;
    uint64_t admit_timeout_ns = 0;  // 0: requests wait for admission as long as it takes
    uint64_t max_batch_bfly = 0;    // work cap of a trace batch in butterflies (0: none)
    SizeTables tabs[bntt::kMaxLog2 + 1];
    char *arena_d = nullptr;
    Arena arena;
    std::vector<Slot> slots;  // nslots for traces + 1 for bntt_pm_multiply
    int grid_cap = 0;         // resident workgroups with up to kQueueLds3 of dynamic LDS
    int grid_cap2 = 0;        // ... with up to kQueueMaxLds
    size_t tw_budget = 0;     // twiddles from device memory (staging measured slower)
    bool sparse = false;      // q < 2^121 with 31-bit digits 1 and 2 zero: the shorter Montgomery reduction
};

namespace {

// The serving kernel for the server's modulus; k8: every item of the level
// is a 2048-value radix-8 tile, run by the code specialised per stage count.
template <typename... A>
void launch_serving(const bntt_pm_server *s, bool k8, dim3 grid, size_t lds, hipStream_t st, A... args)
{
    const dim3 block(bntt::kQueueThreads);
    if (s->sparse && k8)
        hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_MONT_S8>, grid, block, lds, st, args...);
    else if (s->sparse)
        hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_MONT_S>, grid, block, lds, st, args...);
    else if (k8)
        hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_MONT8>, grid, block, lds, st, args...);
    else
        hipLaunchKernelGGL(bntt::queue_kernel<bntt::QK_MONT>, grid, block, lds, st, args...);
}

}  // namespace

namespace {

void destroy_server(bntt_pm_server *s)
{
    if (!s) return;
    (void)hipDeviceSynchronize();
    for (auto &t : s->tabs) {
        (void)hipFree(t.fwd);
        (void)hipFree(t.inv);
        (void)hipFree(t.scale);
    }
    for (auto &sl : s->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        (void)hipFree(sl.ctl);
        (void)hipFree(sl.batch_d);
        if (sl.stage_h) (void)hipHostFree(sl.stage_h);
    }
    (void)hipFree(s->arena_d);
    delete s;
}

int build_tables(bntt_pm_server *s, int L)
{
    const HostMont &hm = s->hm;
    U128 w{};
    int v = 0;
    if (!hm.two_adic_root(&w, &v)) return bntt_set_error(BNTT_EINVAL, "no quadratic non-residue below 1000");
    if (L + 1 > v) return bntt_set_error(BNTT_EINVAL, "q - 1 has too few factors of two for this size");
    const size_t n = (size_t)1 << L;
    std::vector<R31> fwd(n), inv(n), scale(2 * n);
    fwd[0] = bntt::to_r31(hm.r1);  // entry 0 is never read
    inv[0] = fwd[0];
    for (int st = 0; st < L; st++) {
        const size_t base = (size_t)1 << st;
        // r: order 2^(st+2); entries r^(2j+1)
        const U128 r = hm.pow(w, (u128)1 << (v - st - 2));
        const U128 r2 = hm.mul(r, r);
        U128 p = r;
        // u: inverse of a root of order 2^(st+1); entries u^j
        const U128 u = hm.pow(w, ((u128)1 << v) - ((u128)1 << (v - st - 1)));
        U128 pu = hm.r1;
        for (size_t j = 0; j < base; j++) {
            const uint32_t at = (uint32_t)base + brv((uint32_t)j, st);
            // the stored operand is value * R mod q: the Montgomery form itself
            fwd[at] = bntt::to_r31(p);
            inv[at] = bntt::to_r31(pu);
            p = hm.mul(p, r2);
            pu = hm.mul(pu, u);
        }
    }
    // S[i] = N^-1 psi^-brv(i) R^2 with psi of order 2^(L+1)
    const U128 psi_inv = hm.pow(w, ((u128)1 << v) - ((u128)1 << (v - L - 1)));
    const U128 ninv = hm.pow(hm.to_m((hm.q + 1) / 2), (u128)L);
    U128 acc = ninv;  // Montgomery form of N^-1 psi^-j
    for (size_t j = 0; j < n; j++) {
        // stored = value * R^2 = montmul(value * R, R^2)
        // position j of c takes network output i = brv(j)
        const R31 sj = bntt::to_r31(bntt::montmul_r31(acc, hm.r2, hm.mq));
        scale[brv((uint32_t)j, L)] = sj;
        scale[n + j] = sj;
        acc = hm.mul(acc, psi_inv);
    }
    SizeTables &t = s->tabs[L];
    const size_t bytes = n * sizeof(R31);
    if (hipMalloc(&t.fwd, bytes) != hipSuccess || hipMalloc(&t.inv, bytes) != hipSuccess ||
        hipMalloc(&t.scale, 2 * bytes) != hipSuccess)
        return bntt_set_error(BNTT_ENOMEM, "hipMalloc for polynomial-multiply tables failed");
    if (hipMemcpy(t.fwd, fwd.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t.inv, inv.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t.scale, scale.data(), 2 * bytes, hipMemcpyHostToDevice) != hipSuccess)
        return bntt_set_error(BNTT_EHIP, "uploading polynomial-multiply tables failed");
    return BNTT_OK;
}

int make_slot(Slot &sl, size_t cap_items, size_t cap_descs)
{
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        // completion only: the results are device memory, which the kernels'
        // own end-of-kernel release makes visible to every stream, so the
        // event needs no system-scope fence (a cache write-back per batch)
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess)
        return bntt_set_error(BNTT_EHIP, "creating a serving stream failed");
    sl.cap_items = cap_items;
    sl.cap_descs = cap_descs;
    if (hipMalloc(&sl.ctl, sizeof(bntt::QueueCtl)) != hipSuccess ||
        hipMalloc(&sl.batch_d, cap_items * sizeof(WorkItem) + cap_descs * sizeof(DevDesc)) != hipSuccess ||
        hipHostMalloc(&sl.stage_h, cap_items * sizeof(WorkItem) + cap_descs * sizeof(DevDesc),
                      hipHostMallocDefault) != hipSuccess)
        return bntt_set_error(BNTT_ENOMEM, "allocating serving buffers failed");
    if (hipMemset(sl.ctl, 0, sizeof(bntt::QueueCtl)) != hipSuccess)
        return bntt_set_error(BNTT_EHIP, "initialising queue counters failed");
    return BNTT_OK;
}

struct ReqView {
    const U128 *a, *b;
    U128 *c;
    int L;
    size_t arena_off;  // in units
};

// Transforms of one size sharing one table, in np passes: packed several per
// item when single-pass.
void add_transforms(std::vector<std::vector<WorkItem>> &lv, int lv0, std::vector<DevDesc> &descs,
                    const std::vector<DevDesc> &group, int L, uint8_t flags, int want, int np)
{
    const int elog = elog_for(L, want, np);
    if (np == 1) {
        const int Glog = 8 + elog - L;
        for (size_t i = 0; i < group.size(); i += ((size_t)1 << Glog)) {
            WorkItem w{};
            w.kind = bntt::ITEM_NTT;
            w.L = w.k = (uint8_t)L;
            w.Glog = (uint8_t)Glog;
            w.elog = (uint8_t)elog;
            w.np = 1;
            w.flags = flags;
            w.desc = (uint32_t)descs.size();
            w.nvalid = (uint32_t)std::min(group.size() - i, (size_t)1 << Glog);
            for (size_t j = i; j < i + w.nvalid; j++) descs.push_back(group[j]);
            lv[lv0].push_back(w);
        }
        return;
    }
    int ks[3];
    bntt::pass_split(L, np, elog, 8 + elog, ks);
    for (const DevDesc &d : group) {
        const uint32_t di = (uint32_t)descs.size();
        descs.push_back(d);
        for (int p = 0, S = 0; p < np; p++) {
            const int k = ks[p];
            const int Glog = 8 + elog - k;
            const size_t groups = (size_t)1 << (L - k);
            for (size_t g0 = 0; g0 < groups; g0 += ((size_t)1 << Glog)) {
                WorkItem w{};
                w.kind = bntt::ITEM_NTT;
                w.L = (uint8_t)L;
                w.S = (uint8_t)S;
                w.k = (uint8_t)k;
                w.Glog = (uint8_t)Glog;
                w.elog = (uint8_t)elog;
                w.pass = (uint8_t)p;
                w.np = (uint8_t)np;
                w.flags = flags;
                w.desc = di;
                w.nvalid = (uint32_t)std::min(groups - g0, (size_t)1 << Glog);
                w.first = g0;
                lv[lv0 + p].push_back(w);
            }
            S += k;
        }
    }
}

int launch_levels(bntt_pm_server *s, Slot &sl, const std::vector<std::vector<WorkItem>> &lv,
                  const std::vector<DevDesc> &descs);

// 2^want points per thread for a batch of items8 radix-8 items: a batch whose
// radix-8 items would leave resident workgroups idle runs 512-value items,
// one that would fill fewer than three rounds 1024-value items
int want_for(const bntt_pm_server *s, size_t items8)
{
    return items8 < (size_t)s->grid_cap ? 1 : items8 < 3 * (size_t)s->grid_cap ? 2 : 3;
}

// Build the four levels of a batch of requests and launch them.
int launch_batch(bntt_pm_server *s, Slot &sl, const std::vector<ReqView> &reqs)
{
    // levels: 0, 1 forward passes; 2, 3 inverse passes (single-pass sizes use 0 and 2)
    constexpr int kLevels = 4;
    std::vector<std::vector<WorkItem>> &lv = sl.lv;
    std::vector<DevDesc> &descs = sl.descs;
    lv.resize(kLevels);
    for (auto &v : lv) v.clear();
    descs.clear();
    U128 *arena = reinterpret_cast<U128 *>(s->arena_d);
    size_t items8 = 0;
    for (const ReqView &r : reqs) items8 += items_per_request(r.L, 3);
    const int want = want_for(s, items8);
    for (int L = s->min_l; L <= s->max_l; L++) {
        const int np = passes_in_batch(L, want);
        std::vector<DevDesc> &fwd = sl.fwd, &inv = sl.inv;
        fwd.clear();
        inv.clear();
        for (const ReqView &r : reqs) {
            if (r.L != L) continue;
            const size_t n = (size_t)1 << L;
            U128 *base = arena + r.arena_off * (kUnit / sizeof(U128));
            U128 *FA = base, *FB = base + n;
            U128 *TA = np > 1 ? base + 2 * n : nullptr, *TB = TA ? base + 3 * n : nullptr;
            DevDesc da{}, db{}, di{};
            da.x = r.a, da.y = FA, da.tmp = TA, da.twr = s->tabs[L].fwd;
            db.x = r.b, db.y = FB, db.tmp = TB, db.twr = s->tabs[L].fwd;
            // inverse: pass 0 gathers FA * FB, the last pass scatters into c
            di.x = FA, di.x2 = FB, di.y = r.c, di.tmp = TA, di.twr = s->tabs[L].inv, di.scale = s->tabs[L].scale;
            fwd.push_back(da);
            fwd.push_back(db);
            inv.push_back(di);
        }
        if (!fwd.empty()) {
            add_transforms(lv, 0, descs, fwd, L, 0, want, np);
            add_transforms(lv, 2, descs, inv, L, bntt::IT_GATHER_MUL | bntt::IT_SCATTER_SCALE, want, np);
        }
    }
    return launch_levels(s, sl, lv, descs);
}

// Copy a batch's levels into the slot's staging buffer, upload them, launch
// one work-queue kernel per non-empty level on the slot's stream (each level
// starts when the previous one ends) and record the slot's event.
int launch_levels(bntt_pm_server *s, Slot &sl, const std::vector<std::vector<WorkItem>> &lv,
                  const std::vector<DevDesc> &descs)
{
    const int kLevels = (int)lv.size();
    size_t nitems = 0;
    for (auto &v : lv) nitems += v.size();
    if (nitems > sl.cap_items || descs.size() > sl.cap_descs)
        return bntt_set_error(BNTT_EINVAL, "internal: batch exceeds the slot capacity");
    // descriptors follow the items directly (WorkItem is 32 bytes, so they
    // stay aligned), and one copy uploads both
    static_assert(sizeof(WorkItem) % alignof(DevDesc) == 0, "descriptor alignment");
    const size_t desc_off = nitems * sizeof(WorkItem);
    WorkItem *it_h = reinterpret_cast<WorkItem *>(sl.stage_h);
    DevDesc *de_h = reinterpret_cast<DevDesc *>(sl.stage_h + desc_off);
    size_t off[8];
    if (kLevels > 8) return bntt_set_error(BNTT_EINVAL, "internal: too many levels");
    size_t pos = 0;
    for (int l = 0; l < kLevels; l++) {
        off[l] = pos;
        std::memcpy(it_h + pos, lv[l].data(), lv[l].size() * sizeof(WorkItem));
        pos += lv[l].size();
    }
    std::memcpy(de_h, descs.data(), descs.size() * sizeof(DevDesc));
    hipError_t e = hipMemcpyAsync(sl.batch_d, sl.stage_h, desc_off + descs.size() * sizeof(DevDesc),
                                  hipMemcpyHostToDevice, sl.stream);
    if (e != hipSuccess) return bntt_set_hip_error(e, "hipMemcpyAsync of a batch");
    for (int l = 0; l < kLevels; l++) {
        if (lv[l].empty()) continue;
        // the largest data tile and staged-table need of the level's items
        size_t data_slots = 0, need = 0;
        bool all8 = true;
        for (const WorkItem &w : lv[l]) {
            data_slots = std::max(data_slots, bntt::lds_data_slots(w.k, w.Glog));
            need = std::max(need, bntt::tw_slots(w.k, bntt::item_tw_wlog(w.L, w.k, w.S, w.Glog)));
            all8 = all8 && w.elog == 3 && w.k >= 3;
        }
        bntt::TwArea area;
        const size_t lds = bntt::plan_tw_area(data_slots, &need, 1, s->tw_budget, &area);
        const int grid = (int)std::min<size_t>(lv[l].size(), (size_t)(lds > bntt::kQueueLds3 ? s->grid_cap2 : s->grid_cap));
        launch_serving(s, all8 && s->tw_budget == 0, dim3(grid), lds, sl.stream, reinterpret_cast<const WorkItem *>(sl.batch_d) + off[l],
                       (uint32_t)lv[l].size(), reinterpret_cast<const DevDesc *>(sl.batch_d + desc_off),
                       static_cast<const bntt::DevMul *>(nullptr), sl.ctl, area, s->md,
                       bntt::PuntList{nullptr, nullptr});
        e = hipGetLastError();
        if (e != hipSuccess) return bntt_set_hip_error(e, "serving kernel launch");
    }
    e = hipEventRecord(sl.done, sl.stream);
    if (e != hipSuccess) return bntt_set_hip_error(e, "hipEventRecord");
    return BNTT_OK;
}

// A request's c may not overlap another request's operands or result.
int check_disjoint(const bntt_pm_request *reqs, size_t n)
{
    std::vector<bntt::Span> spans;
    spans.reserve(3 * n);
    for (size_t i = 0; i < n; i++) {
        const uintptr_t bytes = sizeof(U128) << reqs[i].log2n;
        const uintptr_t a = reinterpret_cast<uintptr_t>(reqs[i].a), b = reinterpret_cast<uintptr_t>(reqs[i].b),
                        c = reinterpret_cast<uintptr_t>(reqs[i].c);
        spans.push_back(bntt::Span{a, a + bytes, i, false});
        spans.push_back(bntt::Span{b, b + bytes, i, false});
        spans.push_back(bntt::Span{c, c + bytes, i, true});
    }
    if (bntt::cross_task_overlap(spans))
        return bntt_set_error(BNTT_EINVAL, "a request's result overlaps another request's operands or result");
    return BNTT_OK;
}

int check_request(const bntt_pm_server *s, const bntt_pm_request &r)
{
    if (r.log2n < s->min_l || r.log2n > s->max_l) return bntt_set_error(BNTT_EINVAL, "request size not served");
    if (!r.a || !r.b || !r.c ||
        ((reinterpret_cast<uintptr_t>(r.a) | reinterpret_cast<uintptr_t>(r.b) | reinterpret_cast<uintptr_t>(r.c)) &
         15u))
        return bntt_set_error(BNTT_EINVAL, "request pointers must be non-null and 16-byte aligned");
    if (request_units(r.log2n) > s->arena.total())
        return bntt_set_error(BNTT_EINVAL, "request needs more than the whole HBM budget");
    return BNTT_OK;
}

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

int bntt_pm_server_create(bntt_pm_server **out, const bntt_u128 *q, const bntt_pm_config *cfg)
{
    int rc = bntt_ensure_init();
    if (rc) return rc;
    if (!out || !q || !cfg) return bntt_set_error(BNTT_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->min_log2n < 1 || cfg->max_log2n > 16 || cfg->min_log2n > cfg->max_log2n || cfg->max_batch < 1 ||
        cfg->max_inflight < 1 || cfg->max_inflight > 8)
        return bntt_set_error(BNTT_EINVAL, "bad server configuration");
    bntt_pm_server *s = new bntt_pm_server();
    if (hipGetDevice(&s->dev) != hipSuccess) {
        delete s;
        return bntt_set_error(BNTT_EHIP, "hipGetDevice");
    }
    if (!s->hm.init(((u128)q->hi << 64) | q->lo)) {
        delete s;
        return bntt_set_error(BNTT_EINVAL, "q must be odd, at least 3 and below 2^122");
    }
    if (!s->hm.is_probable_prime()) {
        delete s;
        return bntt_set_error(BNTT_EINVAL, "q is not prime");
    }
    s->md.m = s->hm.mq.ql;
    s->md.mq = s->hm.mq;
    s->min_l = cfg->min_log2n;
    s->max_l = cfg->max_log2n;
    s->max_batch = cfg->max_batch;
    s->nslots = cfg->max_inflight;
    s->admit_timeout_ns = cfg->admit_timeout_ns;
    s->max_batch_bfly = cfg->max_batch_butterflies;
    for (int L = s->min_l; L <= s->max_l; L++)
        if ((rc = build_tables(s, L))) {
            destroy_server(s);
            return rc;
        }
    const size_t units = cfg->hbm_budget_bytes / kUnit;
    if (hipMalloc(&s->arena_d, units * kUnit) != hipSuccess) {
        destroy_server(s);
        return bntt_set_error(BNTT_ENOMEM, "hipMalloc of the serving arena failed");
    }
    s->arena.reset(units);
    const size_t cap_items = (size_t)s->max_batch * items_per_request(s->max_l, 1) + 64;
    s->slots.resize(s->nslots + 1);
    for (auto &sl : s->slots)
        if ((rc = make_slot(sl, cap_items, 3 * (size_t)s->max_batch))) {
            destroy_server(s);
            return rc;
        }
    // (the sparse product sums a column of the product and of the reduction
    // in one 64-bit accumulator, which needs q < 2^121: digit 3 below 2^28)
    s->sparse = s->md.mq.q.d[1] == 0 && s->md.mq.q.d[2] == 0 && s->md.mq.q.d[3] < (1u << 28);
    const void *kfn = s->sparse ? reinterpret_cast<const void *>(&bntt::queue_kernel<bntt::QK_MONT_S>)
                                : reinterpret_cast<const void *>(&bntt::queue_kernel<bntt::QK_MONT>);
    const void *kfn8 = s->sparse ? reinterpret_cast<const void *>(&bntt::queue_kernel<bntt::QK_MONT_S8>)
                                 : reinterpret_cast<const void *>(&bntt::queue_kernel<bntt::QK_MONT8>);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, s->dev) != hipSuccess ||
        hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bntt::kQueueMaxLds) != hipSuccess ||
        hipFuncSetAttribute(kfn8, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bntt::kQueueMaxLds) != hipSuccess) {
        destroy_server(s);
        return bntt_set_error(BNTT_EHIP, "querying the device for the serving kernel failed");
    }
    // resident workgroups: the smaller of the two kernels' figures
    int per_cu = 1 << 20, per_cu2 = 1 << 20;
    for (const void *f : {kfn, kfn8}) {
        int a = 0, b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, f, bntt::kQueueThreads, bntt::kQueueLds3) != hipSuccess)
            a = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, f, bntt::kQueueThreads, bntt::kQueueMaxLds) != hipSuccess)
            b = 1;
        per_cu = std::min(per_cu, a);
        per_cu2 = std::min(per_cu2, b);
    }
    s->grid_cap = prop.multiProcessorCount * std::max(per_cu, 1);
    s->grid_cap2 = prop.multiProcessorCount * std::max(per_cu2, 1);
    if (hipDeviceSynchronize() != hipSuccess) {
        destroy_server(s);
        return bntt_set_error(BNTT_EHIP, "server set-up failed");
    }
    *out = s;
    return BNTT_OK;
}

int bntt_pm_server_destroy(bntt_pm_server *s)
{
    destroy_server(s);
    return BNTT_OK;
}

int bntt_pm_multiply(bntt_pm_server *s, const bntt_pm_request *reqs, size_t n)
{
    if (!s || (n && !reqs)) return bntt_set_error(BNTT_EINVAL, "null argument");
    if (n > (size_t)s->max_batch) return bntt_set_error(BNTT_EINVAL, "more requests than max_batch");
    Slot &sl = s->slots.back();
    std::vector<ReqView> views;
    std::vector<std::pair<size_t, size_t>> allocs;
    int rc = BNTT_OK;
    for (size_t i = 0; i < n; i++)
        if ((rc = check_request(s, reqs[i]))) return rc;
    if ((rc = check_disjoint(reqs, n))) return rc;
    for (size_t i = 0; i < n && !rc; i++) {
        size_t off;
        const size_t u = request_units(reqs[i].log2n);
        if (!s->arena.alloc(u, &off)) {
            rc = bntt_set_error(BNTT_ENOMEM, "requests exceed the HBM budget");
            break;
        }
        allocs.emplace_back(off, u);
        views.push_back(ReqView{static_cast<const U128 *>(reqs[i].a), static_cast<const U128 *>(reqs[i].b),
                                static_cast<U128 *>(reqs[i].c), reqs[i].log2n, off});
    }
    if (!rc && !views.empty()) {
        rc = launch_batch(s, sl, views);
        if (!rc && hipEventSynchronize(sl.done) != hipSuccess) rc = bntt_set_error(BNTT_EHIP, "batch failed");
        // a launch that failed part-way may have enqueued kernels that write
        // the blocks released below
        if (rc) (void)hipStreamSynchronize(sl.stream);
    }
    for (auto &a : allocs) s->arena.release(a.first, a.second);
    return rc;
}

int bntt_pm_transform(bntt_pm_server *s, int direction, void *y, const void *x, int log2n, size_t count)
{
    if (!s) return bntt_set_error(BNTT_EINVAL, "null argument");
    if (direction != BNTT_PM_FORWARD && direction != BNTT_PM_INVERSE)
        return bntt_set_error(BNTT_EINVAL, "direction must be BNTT_PM_FORWARD or BNTT_PM_INVERSE");
    if (log2n < s->min_l || log2n > s->max_l) return bntt_set_error(BNTT_EINVAL, "transform size not served");
    if (count == 0) return BNTT_OK;
    if (!x || !y || ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15u))
        return bntt_set_error(BNTT_EINVAL, "pointers must be non-null and 16-byte aligned");
    const int L = log2n;
    const size_t n = (size_t)1 << L;
    if (count > (SIZE_MAX / sizeof(U128)) >> L) return bntt_set_error(BNTT_EINVAL, "count too large");
    const uintptr_t xa = reinterpret_cast<uintptr_t>(x), ya = reinterpret_cast<uintptr_t>(y);
    const uintptr_t bytes = (uintptr_t)(count << L) * sizeof(U128);
    if (xa != ya && xa < ya + bytes && ya < xa + bytes)
        return bntt_set_error(BNTT_EINVAL, "y must be x or disjoint from it");
    // transforms per launch: within the slot's descriptors and, for two-pass
    // sizes, one pass buffer each within the HBM budget (every arena block is
    // free between calls)
    size_t chunk = 3 * (size_t)s->max_batch;
    size_t per = 0, off = 0;
    if (passes_for(L) > 1) {
        per = (n * sizeof(U128) + kUnit - 1) / kUnit;
        chunk = std::min(chunk, s->arena.total() / per);
        if (chunk == 0) return bntt_set_error(BNTT_EINVAL, "a transform needs more than the whole HBM budget");
        chunk = std::min(chunk, count);
        if (!s->arena.alloc(chunk * per, &off)) return bntt_set_error(BNTT_ENOMEM, "serving arena in use");
    }
    const bool fwd = direction == BNTT_PM_FORWARD;
    const uint8_t flags = fwd ? bntt::IT_CANON : (bntt::IT_GATHER | bntt::IT_SCATTER_SCALE);
    const SizeTables &tab = s->tabs[L];
    U128 *tmp = reinterpret_cast<U128 *>(s->arena_d) + off * (kUnit / sizeof(U128));
    Slot &sl = s->slots.back();
    int rc = BNTT_OK;
    for (size_t t0 = 0; t0 < count && !rc; t0 += chunk) {
        const size_t m = std::min(chunk, count - t0);
        std::vector<DevDesc> group(m);
        for (size_t t = 0; t < m; t++) {
            DevDesc &d = group[t];
            d.x = static_cast<const U128 *>(x) + ((t0 + t) << L);
            d.y = static_cast<U128 *>(y) + ((t0 + t) << L);
            d.tmp = per ? tmp + t * per * (kUnit / sizeof(U128)) : nullptr;
            d.twr = fwd ? tab.fwd : tab.inv;
            d.scale = fwd ? nullptr : tab.scale;
        }
        std::vector<std::vector<WorkItem>> lv(2);
        std::vector<DevDesc> descs;
        // radix-8 items of the chunk: whole transforms packed 2^Glog per item
        // for one-pass sizes, else blocks of one transform
        const int glog8 = 8 + elog_for(L, 3, 1) - L;
        const size_t items8 = passes_for(L) == 1 ? (m + ((size_t)1 << glog8) - 1) >> glog8
                                                 : m * (items_per_request(L, 3) / 3);
        add_transforms(lv, 0, descs, group, L, flags, want_for(s, items8), passes_for(L));
        rc = launch_levels(s, sl, lv, descs);
        if (!rc && hipEventSynchronize(sl.done) != hipSuccess) rc = bntt_set_error(BNTT_EHIP, "transform failed");
    }
    if (rc) (void)hipStreamSynchronize(sl.stream);  // nothing enqueued may still write the block
    if (per) s->arena.release(off, chunk * per);
    return rc;
}

int bntt_pm_run_trace(bntt_pm_server *s, const bntt_pm_request *reqs, const uint64_t *arrival_ns, size_t n,
                      uint64_t *done_ns, bntt_pm_stats *stats)
{
    if (!s || (n && (!reqs || !arrival_ns || !done_ns))) return bntt_set_error(BNTT_EINVAL, "null argument");
    for (size_t i = 0; i < n; i++) {
        const int rc = check_request(s, reqs[i]);
        if (rc) return rc;
        if (i && arrival_ns[i] < arrival_ns[i - 1]) return bntt_set_error(BNTT_EINVAL, "arrivals must be sorted");
    }
    if (const int rc = check_disjoint(reqs, n)) return rc;
    bntt_pm_stats st{};
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now_ns = [&]() { return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - t0).count(); };
    std::deque<size_t> pending;
    std::vector<char> waited(n, 0);
    size_t next = 0, completed = 0, inflight_reqs = 0;
    std::vector<ReqView> views;
    // on a failure: wait for the batches still in flight, return their arena
    // blocks and free the slots, so the server stays usable
    auto abandon = [&](int rc) {
        for (int k = 0; k < s->nslots; k++) {
            Slot &sl = s->slots[k];
            // a slot whose launch failed part-way is not marked busy but may
            // have kernels enqueued that write its arena blocks: wait for its
            // stream before the blocks are handed out again
            if (sl.busy || !sl.allocs.empty()) (void)hipStreamSynchronize(sl.stream);
            for (auto &a : sl.allocs) s->arena.release(a.first, a.second);
            sl.allocs.clear();
            sl.reqs.clear();
            sl.busy = false;
        }
        return rc;
    };
    uint64_t t_prev = 0;
    while (completed < n) {
        bool progress = false;
        const uint64_t t = now_ns();
        st.max_loop_gap_ns = std::max(st.max_loop_gap_ns, t - t_prev);
        t_prev = t;
        while (next < n && arrival_ns[next] <= t) pending.push_back(next++);
        // retire finished batches
        for (int k = 0; k < s->nslots; k++) {
            Slot &sl = s->slots[k];
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return abandon(bntt_set_hip_error(q, "serving batch failed"));
            const uint64_t tdone = now_ns();
            st.max_batch_ns = std::max(st.max_batch_ns, tdone - sl.launched_ns);
            for (size_t r : sl.reqs) done_ns[r] = tdone;
            for (auto &a : sl.allocs) s->arena.release(a.first, a.second);
            completed += sl.reqs.size();
            inflight_reqs -= sl.reqs.size();
            sl.reqs.clear();
            sl.allocs.clear();
            sl.busy = false;
            progress = true;
        }
        // drop requests that waited too long for admission (the oldest are
        // at the front)
        if (s->admit_timeout_ns)
            while (!pending.empty() && t - arrival_ns[pending.front()] > s->admit_timeout_ns) {
                done_ns[pending.front()] = BNTT_PM_DROPPED;
                pending.pop_front();
                st.dropped++;
                completed++;
                progress = true;
            }
        // admit arrived requests into a free slot while the budget allows
        for (int k = 0; k < s->nslots && !pending.empty(); k++) {
            Slot &sl = s->slots[k];
            if (sl.busy) continue;
            views.clear();
            uint64_t work = 0;  // butterflies of the batch so far
            while (!pending.empty() && (int)views.size() < s->max_batch) {
                const size_t r = pending.front();
                // a batch's duration follows its work: closing batches at a
                // work cap keeps heavy ones from holding every request in them
                const uint64_t w = request_butterflies(reqs[r].log2n);
                if (s->max_batch_bfly && !views.empty() && work + w > s->max_batch_bfly) break;
                const size_t u = request_units(reqs[r].log2n);
                size_t off;
                if (!s->arena.alloc(u, &off)) {
                    if (!waited[r]) st.admission_waits++;
                    waited[r] = 1;
                    break;
                }
                pending.pop_front();
                work += w;
                sl.allocs.emplace_back(off, u);
                sl.reqs.push_back(r);
                views.push_back(ReqView{static_cast<const U128 *>(reqs[r].a), static_cast<const U128 *>(reqs[r].b),
                                        static_cast<U128 *>(reqs[r].c), reqs[r].log2n, off});
            }
            if (views.empty()) break;  // budget full: wait for a batch to retire
            sl.launched_ns = now_ns();
            const int rc = launch_batch(s, sl, views);
            if (rc) return abandon(rc);
            sl.busy = true;
            st.batches++;
            inflight_reqs += views.size();
            st.max_inflight_requests = std::max<uint64_t>(st.max_inflight_requests, inflight_reqs);
            progress = true;
        }
        if (!progress) {
            // nothing to do until the next arrival or completion: sleep through
            // long idle gaps, otherwise poll
            const uint64_t tn = now_ns();
            if (next < n && pending.empty() && inflight_reqs == 0 && arrival_ns[next] > tn + 200000)
                std::this_thread::sleep_for(std::chrono::nanoseconds(arrival_ns[next] - tn - 100000));
            else
                std::this_thread::yield();
        }
    }
    st.requests = n;
    if (stats) *stats = st;
    return BNTT_OK;
}

}  // extern "C"
#pragma GCC visibility pop
