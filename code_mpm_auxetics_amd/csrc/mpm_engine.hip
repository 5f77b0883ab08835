// mpm_engine.hip -- host orchestration and C ABI of the MI355X MPM step.
//
// One handle = one device, one HIP stream, calls serialised by the caller.
// Step pipeline per time step n (2 kernels):
//   k_grid(n):      merge brick arenas(n) -> node momentum, damping, BC,
//                   reaction, platen override (…_MP.cpp:209-265)
//   k_particle(n):  G2P(n) + F/strain/constitutive (…_MP.cpp:271-354), P2G(n+1)
//                   into arenas(n+1) (…_MP.cpp:106-202), platen motion and
//                   override marks for n+1 (…_MP.cpp:360-368)
// The host checks drift flags every cfg.drift_check steps and rebins when a
// point moved past half a cell outside its brick.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/mpm_auxetics.h"
#include "mpm_bin_kernels.hpp"
#include "mpm_step_kernels.hpp"
#include "mpm_contact_kernels.hpp"

using namespace mpm;

namespace {

thread_local std::string g_create_error;

// mpm_constitutive: staging buffer and stream per device, kept until the
// process ends
struct CrStage {
    double *buf = nullptr;
    size_t cap = 0;   // doubles
    hipStream_t stream = nullptr;
};
std::mutex g_cr_mu;
std::map<int, CrStage> g_cr_stage;

struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct mpm_handle {
    // asynchronous step (mpm_step_async): 0 idle, 1 running, 2 finished, not
    // yet collected by mpm_step_wait.  While not idle the worker thread owns
    // the handle and every other call returns MPM_E_STATE.
    std::atomic<int> async_state{0};
    std::thread worker;
    int async_rc = 0;
    int32_t async_finished = 0;
    mpm_config cfg{};
    std::vector<double> tl_t, tl_v;
    std::vector<uint8_t> fixed_host;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t bev[2] = {nullptr, nullptr};   // end of the batch using host buffer set 0 / 1
    std::string err;
    GridParams g{};
    bool fp32 = false;
    size_t rsize = 8;
    long long nnodes = 0;
    int nbricks = 0;

    // points
    int n_total = 0, n_def = 0, n_dc = 0;
    bool uploaded = false;
    std::vector<Buf> pfields[2];   // per buffer: every field allocation
    PArrays<float> pf[2]{};
    PArrays<double> pd[2]{};
    int cur = 0;
    double *xstage[3] = {nullptr, nullptr, nullptr};
    double *dc_x = nullptr, *dc_m = nullptr;
    double *dc_v[2] = {nullptr, nullptr};   // platen velocities, ping-pong by step parity
    // platen fields only the contact step reads (P2G includes platens there)
    double *dc_vol = nullptr, *dc_s = nullptr, *dc_fext = nullptr;
    uint8_t *dc_surf = nullptr;
    uint32_t *dc_id = nullptr;
    uint8_t surf0 = 0;                      // surface flag of original point 0
    bool id0_dc = false;                    // original point 0 is a platen point
    int id0_dc_index = 0;
    std::vector<int32_t> dc_id_host, def_id_host;
    mpm_points dc_saved{};                  // host copy of static platen fields
    std::vector<double> dc_saved_store;

    // grid: no node arrays -- each step's grid update lives in LDS
    uint8_t *fixed = nullptr;
    unsigned long long *marks[2] = {nullptr, nullptr};   // platen override tags, ping-pong

    // bricks
    uint32_t *occ_list = nullptr, *occ_start = nullptr, *occ_count = nullptr;
    int32_t *n_occ_d = nullptr;
    uint32_t *slot_order = nullptr;          // launch order of the brick blocks
    uint8_t *slot_mat = nullptr;             // per slot: shared material index or 0xFF
    int4 *idesc = nullptr;                   // per launch-order item: 2 int4 (k_item_desc)
    std::vector<uint32_t> slot_order_h, occ_count_h;
    int n_occ = 0, occ_cap = 0;
    int n_xcd = 8;                           // XCDs the dispatcher deals workgroups to (MI355X: 8)
    int n_cu = 0;                            // compute units of the device
    uint32_t *queue = nullptr;               // persistent launch: per-step work counters
    int persist_blocks = 0;                  // persistent launch: resident brick blocks
    void *arena[2] = {nullptr, nullptr};
    int8_t *exps[2] = {nullptr, nullptr};   // FP32: per slot, the fixed-point exponents of its arena (8 B)
    int4 *binfo[2] = {nullptr, nullptr};    // per brick (dense): {slot, packed touched node box}
    size_t arena_cap_slots = 0;
    int parity = 0;

    Material *mats = nullptr;
    int n_mats = 0;
    int has_pointE = 0, has_fext = 0;

    // pinned host buffers in two sets (steps_cap entries each), one per batch
    // in flight: step records, reaction history, flags, group history
    StepRec *steps_d = nullptr, *steps_h = nullptr;
    int steps_cap = 0;
    // results of a batch in one device block (out_d): the flags word, the
    // per-step reaction history (steps_cap), the per-group history
    // (steps_cap x n_groups); flags_d, hist_d and ghist_d point into it, and
    // one copy brings a batch's results to out_h (pinned, two sets)
    double *out_d = nullptr, *out_h = nullptr;
    size_t out_stride = 0;                  // doubles per set
    // device addresses of the pinned sets: the plain path's kernels read the
    // step records and write the results there directly (no copies)
    double *out_hd = nullptr;
    StepRec *steps_hd = nullptr;
    int32_t *flags_d = nullptr, *flags_h = nullptr;
    double *kp_part = nullptr;
    double *gpart = nullptr;                // [steps_cap][occ_cap][3] grid-update partials
    size_t kp_cap = 0;
    double *hist_d = nullptr;
    std::vector<double> hist;
    double *checks_d = nullptr;
    double *grid3_d = nullptr;              // last grid update: mass, reaction, active nodes
    // batched worlds: reaction per group (world)
    int n_groups = 0, n_fy = 0;
    int32_t *fy_slot = nullptr, *fy_off = nullptr;   // node -> y-fixed slot; CSR offsets per group
    double *fy_val = nullptr, *ghist_d = nullptr;
    std::vector<double> ghist;              // [step][group] of the last mpm_step call

    // contact step (allocated at its first use)
    struct Contact {
        bool ready = false;
        long long *acc = nullptr;
        double *gkgrad = nullptr, *gkg = nullptr, *gm = nullptr, *gv = nullptr, *gf = nullptr, *ga = nullptr;
        uint32_t *flags = nullptr, *id0 = nullptr;
        unsigned long long *near_d = nullptr, *sg_d = nullptr, *cmarks = nullptr, *bounds = nullptr;
        unsigned long long *near_i = nullptr, *sg_i = nullptr;
        uint8_t *near_surf = nullptr, *fmask = nullptr;
        double *kp = nullptr, *kg = nullptr, *cpart = nullptr, *ppart = nullptr;
        double *kp_id = nullptr, *kg_id = nullptr;   // by original point index
        int nblk_n = 0, nblk_p = 0;         // node blocks; capacity of the point-sum partials
        int n_ppart = 0;                    // point-sum partials the last contact G2P wrote
        bool kernel_valid = false;          // kp/kg hold the last contact step's point kernel
        int bsel = 0;                       // which half of bounds is the coming step's
        bool bnext_valid = false;           // the previous G2P formed the coming step's deformable bounds
        bool kernel_dirty = false;          // kp_id/kg_id not yet rewritten from kp/kg since that step
    } ct;
    int last_path = 0;                      // path of the last step (checks read its partials)
    // hipGraphs of batches of plain steps, keyed by their step count and
    // valid for one brick layout (bumped at every rebin / reallocation)
    struct GraphEntry {
        int steps, parity;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    long long graph_layout = -1, layout = 0;
    bool plain_stale = false;               // a contact step ran: plain-path arenas must be rebuilt

    // rebin scratch
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *radix_hist = nullptr, *radix_off = nullptr;
    std::vector<uint32_t *> scan_scratch;
    uint32_t *heads = nullptr, *headpos = nullptr;
    GatherDesc *gdesc = nullptr;
    size_t gdesc_cap = 0;

    double time = 0.0;
    long long cycle = 0;
    double console_last = 0.0;
    long long rebins = 0;
    double last_ms = 0.0;
    double runtime_s = 0.0;      // device time of all mpm_step calls (d_Runtime_Total)
    std::string log_path;        // report lines are appended here when set
    // profiling (cfg.profile): event pairs around every launch
    std::vector<hipEvent_t> pev;
    double kt[6] = {0, 0, 0, 0, 0, 0};
};

#define HIPCHK(h, call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);                   \
            return MPM_E_HIP;                                                               \
        }                                                                                   \
    } while (0)

#define RCHK(call)                \
    do {                          \
        int r_ = (call);          \
        if (r_ != MPM_OK) return r_; \
    } while (0)

static int dev_alloc(mpm_handle *h, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        h->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return MPM_E_NOMEM;
    }
    return MPM_OK;
}

static void dev_free(void *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T>
static int dalloc(mpm_handle *h, T **p, size_t count)
{
    return dev_alloc(h, (void **)p, count * sizeof(T));
}

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ------------------------------------------------------------------ scans
static int scan_u32(mpm_handle *h, const uint32_t *in, uint32_t *out, uint32_t n, int level,
                    uint32_t *total)
{
    if (n == 0) { if (total) *total = 0; return MPM_OK; }
    const unsigned tiles = cdiv(n, 1024);
    while ((int)h->scan_scratch.size() < 2 * (level + 1)) {
        uint32_t *p = nullptr;
        size_t cap = (size_t)cdiv(std::max<size_t>((size_t)h->n_def, (size_t)h->nbricks) + 1024, 1024) + 64;
        RCHK(dalloc(h, &p, cap));
        h->scan_scratch.push_back(p);
    }
    uint32_t *tsum = h->scan_scratch[2 * level], *toff = h->scan_scratch[2 * level + 1];
    k_scan_tiles<<<tiles, 256, 0, h->stream>>>(in, out, tsum, n);
    HIPCHK(h, hipGetLastError());
    if (tiles > 1) {
        RCHK(scan_u32(h, tsum, toff, tiles, level + 1, nullptr));
        k_scan_add<<<tiles, 256, 0, h->stream>>>(out, toff, n);
        HIPCHK(h, hipGetLastError());
    }
    if (total) {
        uint32_t a = 0, b = 0;
        HIPCHK(h, hipMemcpyAsync(&a, out + n - 1, 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(&b, in + n - 1, 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        *total = a + b;
    }
    return MPM_OK;
}

// stable LSD radix sort of (keys[0], vals[0]) -> result left in keys[r], vals[r]
static int radix_sort(mpm_handle *h, uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1,
                      uint32_t n, uint32_t max_key, uint32_t **kout, uint32_t **vout)
{
    int bits = 0;
    while (bits < 32 && (max_key >> bits) != 0) ++bits;
    int passes = std::max(1, (bits + 7) / 8);
    const unsigned nb = std::max(1u, cdiv(n, kSortTile));
    uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
    for (int ps = 0; ps < passes; ++ps) {
        const int shift = 8 * ps;
        k_radix_hist<<<nb, kBlock, 0, h->stream>>>(ka, n, shift, h->radix_hist, nb);
        HIPCHK(h, hipGetLastError());
        RCHK(scan_u32(h, h->radix_hist, h->radix_off, 256u * nb, 0, nullptr));
        k_radix_scatter<<<nb, kBlock, 0, h->stream>>>(ka, va, kb, vb, n, shift, h->radix_off, nb);
        HIPCHK(h, hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    *kout = ka;
    *vout = va;
    return MPM_OK;
}

static int alloc_sort_scratch(mpm_handle *h, uint32_t n)
{
    for (int b = 0; b < 2; ++b) {
        RCHK(dalloc(h, &h->keys[b], (size_t)n + 1));
        RCHK(dalloc(h, &h->vals[b], (size_t)n + 1));
    }
    const size_t nb = std::max(1u, cdiv(n, kSortTile));
    RCHK(dalloc(h, &h->radix_hist, 256 * nb));
    RCHK(dalloc(h, &h->radix_off, 256 * nb));
    return MPM_OK;
}

// ---------------------------------------------------------------- fields
template <typename real>
static PArrays<real> &parr(mpm_handle *h, int b);
template <>
PArrays<float> &parr<float>(mpm_handle *h, int b) { return h->pf[b]; }
template <>
PArrays<double> &parr<double>(mpm_handle *h, int b) { return h->pd[b]; }

// every particle array with its element size, in a fixed order
template <typename real>
static void field_list(mpm_handle *h, PArrays<real> &p, std::vector<std::pair<void **, int>> &out)
{
    const int rs = (int)sizeof(real);
    for (int a = 0; a < 3; ++a) out.push_back({(void **)&p.v[a], rs});
    for (int a = 0; a < 9; ++a) out.push_back({(void **)&p.F[a], rs});
    for (int a = 0; a < 6; ++a) out.push_back({(void **)&p.s[a], rs});
    for (int a = 0; a < 6; ++a) out.push_back({(void **)&p.ep[a], rs});
    if (h->cfg.track_strain)
        for (int a = 0; a < 6; ++a) out.push_back({(void **)&p.eps[a], rs});
    out.push_back({(void **)&p.alpha, rs});
    out.push_back({(void **)&p.mass, rs});
    out.push_back({(void **)&p.vol0, rs});
    if (h->has_pointE) out.push_back({(void **)&p.E, rs});
    if (h->has_fext)
        for (int a = 0; a < 3; ++a) out.push_back({(void **)&p.fext[a], rs});
    out.push_back({(void **)&p.mat, 1});
    out.push_back({(void **)&p.surf, 1});
    out.push_back({(void **)&p.id, 4});
}

template <typename real>
static int alloc_points(mpm_handle *h, int n)
{
    for (int b = 0; b < 2; ++b) {
        PArrays<real> &p = parr<real>(h, b);
        std::memset(&p, 0, sizeof(p));
        std::vector<std::pair<void **, int>> fl;
        field_list<real>(h, p, fl);
        for (int a = 0; a < 3; ++a) {
            if (sizeof(real) == 4) fl.push_back({(void **)&p.xq[a], 4});
            else fl.push_back({(void **)&p.x[a], 8});
        }
        fl.push_back({(void **)&p.pbrick, 4});
        for (auto &f : fl) {
            RCHK(dev_alloc(h, f.first, (size_t)std::max(n, 1) * f.second));
            h->pfields[b].push_back(Buf{*f.first, (size_t)n * f.second});
        }
    }
    return MPM_OK;
}

static int ensure_slots(mpm_handle *h, int n_occ)
{
    const int want_occ = std::max(1, n_occ + n_occ / 4 + 16);
    if ((size_t)n_occ > h->arena_cap_slots || h->arena[0] == nullptr) {
        for (int b = 0; b < 2; ++b) {
            dev_free(h->arena[b]);
            dev_free((void *&)h->exps[b]);
        }
        dev_free((void *&)h->kp_part);
        dev_free((void *&)h->gpart);
        // the step kernel addresses arena records with 32-bit element
        // offsets; one record of 8 zeros after the last slot is what an
        // uncovered candidate node reads
        const size_t elems = (size_t)want_occ * 8 * h->g.arenaN;
        if (elems + 8 > 0xFFFFFFFFull) {
            h->err = "brick arenas exceed 2^32 elements (use larger bricks)";
            return MPM_E_ARG;
        }
        h->arena_cap_slots = (size_t)want_occ;
        for (int b = 0; b < 2; ++b) {
            RCHK(dev_alloc(h, &h->arena[b], (elems + 8) * h->rsize));
            HIPCHK(h, hipMemsetAsync((char *)h->arena[b] + elems * h->rsize, 0, 8 * h->rsize, h->stream));
            // every arena rebuild writes the exponents of its slots before a step reads them
            RCHK(dalloc(h, &h->exps[b], (size_t)want_occ * 8));
            HIPCHK(h, hipMemsetAsync(h->exps[b], 0x80, (size_t)want_occ * 8, h->stream));
        }
        RCHK(dalloc(h, &h->kp_part, h->arena_cap_slots * 5 * (kPBlock / 64)));
        HIPCHK(h, hipMemsetAsync(h->kp_part, 0, h->arena_cap_slots * 5 * (kPBlock / 64) * 8, h->stream));
        RCHK(dalloc(h, &h->gpart, (size_t)h->steps_cap * h->arena_cap_slots * 3));
        h->occ_cap = want_occ;
        h->layout++;
    }
    return MPM_OK;
}

template <typename real>
static int launch_particle(mpm_handle *h, bool g2p, int step);

// profiling: events recorded around launches, summed after each sync
struct ProfMark {
    int kind;   // 0 particle, 1 grid, 2 rebin
    int ev;     // index of the first event of the pair
};

static int prof_begin(mpm_handle *h, std::vector<ProfMark> &marks, int kind)
{
    if (!h->cfg.profile) return -1;
    const int idx = (int)marks.size() * 2;
    while ((int)h->pev.size() < idx + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        h->pev.push_back(e);
    }
    (void)hipEventRecord(h->pev[idx], h->stream);
    marks.push_back(ProfMark{kind, idx});
    return idx;
}

static void prof_end(mpm_handle *h, int idx)
{
    if (idx >= 0) (void)hipEventRecord(h->pev[idx + 1], h->stream);
}

static void prof_collect(mpm_handle *h, std::vector<ProfMark> &marks)
{
    for (const ProfMark &m : marks) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, h->pev[m.ev], h->pev[m.ev + 1]) == hipSuccess) {
            h->kt[2 * m.kind] += ms;
            h->kt[2 * m.kind + 1] += 1.0;
        }
    }
    marks.clear();
}

// ------------------------------------------------------------------ rebin
template <typename real>
static CArgs<real> contact_args(mpm_handle *h);
// d_Kernel / d3_Kernel_Gradient of the last contact step by original point
// index (readback order), from the per-point values in the current order
template <typename real>
static int kernel_out(mpm_handle *h)
{
    if (!h->ct.kernel_dirty) return MPM_OK;
    h->ct.kernel_dirty = false;
    CArgs<real> ca = contact_args<real>(h);
    if (ca.n_all > 0)
        k_c_kernel_out<real><<<cdiv(ca.n_all, 256), 256, 0, h->stream>>>(ca, h->ct.kp_id, h->ct.kg_id, h->n_total);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));   // readers copy kp_id / kg_id right after
    return MPM_OK;
}

template <typename real>
static int rebin_t(mpm_handle *h, bool from_abs)
{
    RCHK(kernel_out<real>(h));   // before the points are reordered
    const uint32_t n = (uint32_t)h->n_def;
    if (n == 0) {
        h->n_occ = 0;
        RCHK(ensure_slots(h, 0));
        HIPCHK(h, hipMemsetAsync(h->n_occ_d, 0, 4, h->stream));
        return MPM_OK;
    }
    PArrays<real> &src = parr<real>(h, h->cur), &dst = parr<real>(h, 1 - h->cur);
    const unsigned nbk = cdiv(n, kBlock);
    k_rebin_keys<real><<<nbk, kBlock, 0, h->stream>>>(h->g, src, from_abs ? h->xstage[0] : nullptr,
                                                      from_abs ? h->xstage[1] : nullptr,
                                                      from_abs ? h->xstage[2] : nullptr, n, h->keys[0],
                                                      h->vals[0], h->flags_d);
    HIPCHK(h, hipGetLastError());
    const uint32_t max_key = ((uint32_t)(h->nbricks - 1) << h->g.cbits) | ((1u << h->g.cbits) - 1u);
    uint32_t *sk, *perm;
    RCHK(radix_sort(h, h->keys[0], h->vals[0], h->keys[1], h->vals[1], n, max_key, &sk, &perm));
    // brick tables
    k_heads<<<nbk, kBlock, 0, h->stream>>>(sk, n, h->g.cbits, h->heads);
    HIPCHK(h, hipGetLastError());
    uint32_t n_occ = 0;
    RCHK(scan_u32(h, h->heads, h->headpos, n, 0, &n_occ));
    RCHK(ensure_slots(h, (int)n_occ));
    k_occ<<<nbk, kBlock, 0, h->stream>>>(sk, h->heads, h->headpos, n, h->g.cbits, h->occ_list, h->occ_start);
    HIPCHK(h, hipGetLastError());
    k_occ_count<<<cdiv(n_occ, kBlock), kBlock, 0, h->stream>>>(h->occ_start, n_occ, n, h->occ_count);
    HIPCHK(h, hipGetLastError());
    // launch order of the brick blocks.  Workgroups go to the XCDs round
    // robin by block index, so block X k + x takes the k-th slot of group x,
    // the x-th run of consecutive slots (bricks in space order): a brick and
    // most of its neighbours then run on one XCD every step, and the arena
    // records a block reads were written through that XCD's L2.  Group sizes
    // differ by at most one (the first n_occ % X groups take one more), so
    // every round of X blocks but the last visits each group once and the
    // mapping holds to the end.  Inside a group: most point chunks first,
    // ties in slot order (a longest-first schedule).  Every slot still sums
    // in slot order.
    const uint32_t kXcd = (uint32_t)h->n_xcd;
    h->occ_count_h.resize(n_occ);
    h->slot_order_h.clear();
    HIPCHK(h, hipMemcpyAsync(h->occ_count_h.data(), h->occ_count, (size_t)n_occ * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    {
        const uint32_t q = n_occ / kXcd, r = n_occ % kXcd;
        std::vector<std::vector<uint32_t>> grp(kXcd);
        uint32_t k0 = 0;
        for (uint32_t x = 0; x < kXcd; ++x) {
            const uint32_t sz = q + (x < r ? 1u : 0u);
            for (uint32_t k = 0; k < sz; ++k) grp[x].push_back(k0 + k);
            k0 += sz;
        }
        for (auto &gl : grp)
            std::stable_sort(gl.begin(), gl.end(), [h](uint32_t x, uint32_t y) {
                return cdiv(h->occ_count_h[x], kPBlock) > cdiv(h->occ_count_h[y], kPBlock);
            });
        for (uint32_t k = 0; k <= q; ++k)
            for (uint32_t x = 0; x < kXcd; ++x)
                if (k < grp[x].size()) h->slot_order_h.push_back(grp[x][k]);
    }
    HIPCHK(h, hipMemcpyAsync(h->slot_order, h->slot_order_h.data(), (size_t)n_occ * 4, hipMemcpyHostToDevice, h->stream));
    // spread each brick's points over the lanes (permutation inside bricks)
    uint32_t *perm2 = (perm == h->vals[0]) ? h->vals[1] : h->vals[0];
    k_spread<<<nbk, kBlock, 0, h->stream>>>(perm, h->heads, h->headpos, h->occ_start, h->occ_count, n, perm2);
    HIPCHK(h, hipGetLastError());
    perm = perm2;
    // gather all fields
    std::vector<std::pair<void **, int>> fs, fd;
    field_list<real>(h, src, fs);
    field_list<real>(h, dst, fd);
    std::vector<GatherDesc> desc;
    for (size_t k = 0; k < fs.size(); ++k) desc.push_back(GatherDesc{*fs[k].first, *fd[k].first, fs[k].second});
    if (desc.size() > h->gdesc_cap) {
        dev_free((void *&)h->gdesc);
        h->gdesc_cap = desc.size() + 8;
        RCHK(dalloc(h, &h->gdesc, h->gdesc_cap));
    }
    HIPCHK(h, hipMemcpyAsync(h->gdesc, desc.data(), desc.size() * sizeof(GatherDesc), hipMemcpyHostToDevice, h->stream));
    k_gather<<<dim3(nbk, (unsigned)desc.size()), kBlock, 0, h->stream>>>(h->gdesc, perm, n);
    HIPCHK(h, hipGetLastError());
    k_rebin_pos<real><<<nbk, kBlock, 0, h->stream>>>(h->g, src, dst, from_abs ? h->xstage[0] : nullptr,
                                                     from_abs ? h->xstage[1] : nullptr,
                                                     from_abs ? h->xstage[2] : nullptr, perm, sk, n);
    HIPCHK(h, hipGetLastError());
    h->cur = 1 - h->cur;
    h->n_occ = (int)n_occ;
    if (n_occ > 0) {
        k_slot_mat<<<n_occ, 256, 0, h->stream>>>(parr<real>(h, h->cur).mat, h->occ_start, h->occ_count, h->slot_mat);
        HIPCHK(h, hipGetLastError());
        k_item_desc<<<cdiv(n_occ, 256), 256, 0, h->stream>>>(h->g, h->slot_order, h->occ_list, h->occ_start,
                                                              h->occ_count, h->slot_mat, h->mats, h->fp32 ? 1 : 0,
                                                              (int)n_occ, h->idesc);
        HIPCHK(h, hipGetLastError());
    }
    // brick records of both parities: stale slots must read -1; the rebuild
    // below writes this parity's records of the occupied bricks
    for (int b = 0; b < 2; ++b)
        HIPCHK(h, hipMemsetAsync(h->binfo[b], 0xFF, (size_t)h->nbricks * sizeof(int4), h->stream));
    int32_t cnt = (int32_t)n_occ;
    HIPCHK(h, hipMemcpyAsync(h->n_occ_d, &cnt, 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // rebuild the arena the next grid update reads
    RCHK(launch_particle<real>(h, false, 0));
    h->rebins++;
    h->layout++;
    return MPM_OK;
}

static int rebin(mpm_handle *h, bool from_abs)
{
    return h->fp32 ? rebin_t<float>(h, from_abs) : rebin_t<double>(h, from_abs);
}

// ---------------------------------------------------------------- launches
template <typename real>
static int launch_particle(mpm_handle *h, bool g2p, int step)
{
    KpArgs<real> a{};
    a.g = h->g;
    a.p = parr<real>(h, h->cur);
    // a step reads the parity of its step record and writes the other one;
    // a rebuild (P2G only) writes the current parity
    for (int b = 0; b < 2; ++b) {
        a.arena[b] = (real *)h->arena[b];
        a.exps[b] = h->exps[b];
        a.binfo[b] = h->binfo[b];
        a.marks[b] = h->marks[b];
        a.dc_v[b] = h->dc_v[b];
    }
    a.rebuild_par = h->parity;
    a.zero_off = (uint32_t)(h->arena_cap_slots * 8 * h->g.arenaN);
    a.fixed = h->fixed;
    a.occ_list = h->occ_list;
    a.occ_start = h->occ_start;
    a.occ_count = h->occ_count;
    a.n_occ = h->n_occ_d;
    a.slot_order = h->slot_order;
    a.idesc = h->idesc;
    a.slot_mat = h->slot_mat;
    // platen blocks first (a multiple of the XCD count), then the brick blocks
    const bool persist = g2p && h->cfg.persistent && h->persist_blocks > 0;
    const int xcd = std::max(h->n_xcd, 1);
    a.n_pblk = g2p ? (int)(cdiv(cdiv(h->n_dc, kPBlock), xcd) * xcd) : 0;
    a.n_xcd = h->n_xcd;
    a.queue = h->queue;
    const int brick_blocks = persist ? h->persist_blocks : h->n_occ;
    a.mats = h->mats;
    a.steps = h->steps_d;
    a.flags = h->flags_d;
    a.kp_part = h->kp_part;
    a.gpart = h->gpart + (size_t)step * h->occ_cap * 3;
    a.n_dc = h->n_dc;
    a.dc_x = h->dc_x;
    a.fy_slot = (g2p && h->n_fy > 0) ? h->fy_slot : nullptr;
    a.fy_val = h->fy_val ? h->fy_val + (size_t)step * h->n_fy : nullptr;
    const size_t lds = (size_t)kQ * h->g.arenaN * sizeof(long long) + (g2p ? (size_t)6 * h->g.arenaN * h->rsize : 0);
    if (brick_blocks == 0 && h->n_dc == 0) return MPM_OK;
    const unsigned grid = (unsigned)(a.n_pblk + brick_blocks);
    if (grid == 0) return MPM_OK;
    if (persist)
        k_particle<real, true, true><<<grid, kPBlock, lds, h->stream>>>(a, step);
    else if (g2p)
        k_particle<real, true><<<grid, kPBlock, lds, h->stream>>>(a, step);
    else
        k_particle<real, false><<<grid, kPBlock, lds, h->stream>>>(a, step);
    HIPCHK(h, hipGetLastError());
    return MPM_OK;
}

// ---------------------------------------------------------------- contact
static int ensure_contact(mpm_handle *h)
{
    auto &c = h->ct;
    if (c.ready) return MPM_OK;
    const size_t nn = (size_t)h->nnodes;
    const size_t n_all = (size_t)std::max(h->n_def + h->n_dc, 1);
    c.nblk_n = (int)cdiv(nn, 256);
    c.nblk_p = (int)cdiv((size_t)std::max(h->n_def, 1), 256);
    RCHK(dalloc(h, &c.acc, 21 * nn));
    RCHK(dalloc(h, &c.gkgrad, 3 * nn));
    RCHK(dalloc(h, &c.gkg, 3 * nn));
    RCHK(dalloc(h, &c.gm, 2 * nn));
    RCHK(dalloc(h, &c.gv, 6 * nn));
    RCHK(dalloc(h, &c.gf, 6 * nn));
    RCHK(dalloc(h, &c.ga, 6 * nn));
    RCHK(dalloc(h, &c.fmask, n_all));
    RCHK(dalloc(h, &c.flags, nn));
    RCHK(dalloc(h, &c.near_i, nn));
    RCHK(dalloc(h, &c.sg_i, nn));
    RCHK(dalloc(h, &c.id0, 2));
    RCHK(dalloc(h, &c.near_d, nn));
    RCHK(dalloc(h, &c.sg_d, nn));
    RCHK(dalloc(h, &c.cmarks, nn));
    RCHK(dalloc(h, &c.bounds, (size_t)2 * kCB));   // this step's and the next step's
    RCHK(dalloc(h, &c.near_surf, nn));
    RCHK(dalloc(h, &c.kp, n_all));
    RCHK(dalloc(h, &c.kg, 3 * n_all));
    RCHK(dalloc(h, &c.cpart, (size_t)h->steps_cap * c.nblk_n * 3));
    RCHK(dalloc(h, &c.ppart, (size_t)c.nblk_p * 5));
    RCHK(dalloc(h, &c.kp_id, (size_t)std::max(h->n_total, 1)));
    RCHK(dalloc(h, &c.kg_id, 3 * (size_t)std::max(h->n_total, 1)));
    HIPCHK(h, hipMemsetAsync(c.cmarks, 0, nn * 8, h->stream));
    HIPCHK(h, hipMemsetAsync(c.ppart, 0, (size_t)c.nblk_p * 5 * 8, h->stream));
    c.ready = true;
    return MPM_OK;
}

template <typename real>
static CArgs<real> contact_args(mpm_handle *h)
{
    auto &c = h->ct;
    CArgs<real> a{};
    a.g = h->g;
    a.p = parr<real>(h, h->cur);
    a.n_def = h->n_def;
    a.n_dc = h->n_dc;
    a.n_all = h->n_def + h->n_dc;
    a.surf = a.p.surf;
    a.dc_x = h->dc_x;
    a.dc_x_out = h->dc_x;
    a.dc_v = h->dc_v[h->parity];
    a.dc_m = h->dc_m;
    a.dc_vol = h->dc_vol;
    a.dc_s = h->dc_s;
    a.dc_fext = h->dc_fext;
    a.dc_surf = h->dc_surf;
    a.dc_id = h->dc_id;
    a.def_id0 = c.id0;
    a.surf0 = h->surf0;
    a.nn = h->nnodes;
    a.fixed = h->fixed;
    a.acc = c.acc;
    a.gkgrad = c.gkgrad;
    a.gkg = c.gkg;
    a.gm = c.gm;
    a.gv = c.gv;
    a.gf = c.gf;
    a.flags = c.flags;
    a.near_d = c.near_d;
    a.sg_d = c.sg_d;
    a.near_i = c.near_i;
    a.sg_i = c.sg_i;
    a.near_surf = c.near_surf;
    a.fmask = c.fmask;
    a.ga = c.ga;
    a.cmarks = c.cmarks;
    a.bounds = c.bounds + (size_t)kCB * c.bsel;
    a.bounds_next = c.bounds + (size_t)kCB * (c.bsel ^ 1);
    a.def_bounds = c.bnext_valid ? 0 : 1;
    a.kp = c.kp;
    a.kg = c.kg;
    a.mats = h->mats;
    a.steps = h->steps_d;
    a.gflags = h->flags_d;
    a.ppart = c.ppart;
    a.slot_order = h->slot_order;
    a.occ_start = h->occ_start;
    a.occ_count = h->occ_count;
    a.n_bb = h->n_def > 0 ? h->n_occ : 0;
    a.n_xcd = std::max(h->n_xcd, 1);
    int H = 2;
    while ((1ll << (H - 2)) <= (long long)a.n_all) ++H;   // 2^(H-2) > points touching a node
    a.H = H;
    a.mu = h->cfg.friction;
    return a;
}

// one contact step (…_MP_Contact.cpp:30-646): a fixed sequence of launches
template <typename real>
static int launch_contact(mpm_handle *h, int k)
{
    auto &c = h->ct;
    CArgs<real> a = contact_args<real>(h);
    a.cpart = c.cpart + (size_t)k * c.nblk_n * 3;
    // point passes: one block per occupied brick (all its points), then
    // blocks of 256 platen points; G2P: blocks of 256 deformable points
    const unsigned gp = (unsigned)a.n_bb + cdiv((size_t)a.n_dc, 256), gn = (unsigned)c.nblk_n;
    hipStream_t st = h->stream;
    k_c_clear<real><<<gn, 256, 0, st>>>(a);
    // H >= 11 headroom bits keep every scaled contribution below 2^51 (the
    // short round-to-int64 of c_toll)
    const bool small = a.H >= 11;
    if (gp > 0) {   // kernel density, active nodes, bounds
        if (small) k_c_kernel<real, true><<<gp, 256, 0, st>>>(a);
        else k_c_kernel<real, false><<<gp, 256, 0, st>>>(a);
    }
    k_c_kgrad<real><<<gn, 256, 0, st>>>(a);
    if (gp > 0) {
        k_c_point1<real><<<gp, 256, 0, st>>>(a);
        k_c_point2<real><<<cdiv((size_t)a.n_all, 256), 256, 0, st>>>(a);   // blocks of 256 points
        k_c_publish<real><<<gn, 256, 0, st>>>(a);
        if (small) {   // mass with detection and the field mask, then momentum and force
            k_c_p2g<real, true, true><<<gp, 256, 0, st>>>(a);
            k_c_p2g<real, false, true><<<gp, 256, 0, st>>>(a);
        } else {
            k_c_p2g<real, true, false><<<gp, 256, 0, st>>>(a);
            k_c_p2g<real, false, false><<<gp, 256, 0, st>>>(a);
        }
    }
    if (a.n_dc > 0) k_c_dcmarks<real><<<cdiv(a.n_dc, 256), 256, 0, st>>>(a, k);
    k_c_grid<real><<<gn, 256, 0, st>>>(a, k);
    if (a.n_def > 0) k_c_g2p<real><<<(unsigned)c.nblk_p, 256, 0, st>>>(a, k);
    c.n_ppart = a.n_def > 0 ? c.nblk_p : 0;
    if (a.n_dc > 0) k_c_dcmove<real><<<cdiv(a.n_dc, 256), 256, 0, st>>>(a, k);
    HIPCHK(h, hipGetLastError());
    c.kernel_valid = true;
    // the G2P maxed the next step's deformable bounds over the new state
    c.bsel ^= 1;
    c.bnext_valid = a.n_def > 0;
    return MPM_OK;
}

// where original point 0 lives now (platen index, or its sorted position)
template <typename real>
static int locate_point0(mpm_handle *h)
{
    uint32_t v[2] = {h->id0_dc ? 1u : 0u, h->id0_dc ? (uint32_t)h->id0_dc_index : 0u};
    HIPCHK(h, hipMemcpyAsync(h->ct.id0, v, sizeof(v), hipMemcpyHostToDevice, h->stream));
    if (!h->id0_dc && h->n_def > 0)
        k_find_id<<<cdiv(h->n_def, 256), 256, 0, h->stream>>>(parr<real>(h, h->cur).id, (uint32_t)h->n_def, 0u, h->ct.id0);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MPM_OK;
}

// TimeLine::getVelocity (TimeLine.h:33-46), evaluated on the host per step
static void timeline_velocity(const mpm_handle *h, double t, double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    const int n = (int)h->tl_t.size();
    for (int i = 0; i < n - 1; ++i) {
        if (h->tl_t[i] <= t && t < h->tl_t[i + 1]) {
            const double sc = (t - h->tl_t[i]) / (h->tl_t[i + 1] - h->tl_t[i]);
            for (int a = 0; a < 3; ++a) out[a] = (1.0 - sc) * h->tl_v[3 * i + a] + sc * h->tl_v[3 * (i + 1) + a];
        }
    }
}

static int check_flags(mpm_handle *h, bool *need_rebin, bool copied = false)
{
    if (!copied) HIPCHK(h, hipMemcpyAsync(h->flags_h, h->flags_d, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int f = *h->flags_h;
    *need_rebin = (f & FLAG_REBIN) != 0;
    if (f & (FLAG_OUT_OF_GRID | FLAG_ARENA | FLAG_NAN)) {
        if (f & FLAG_NAN) { h->err = "non-finite point state"; return MPM_E_NAN; }
        if (f & FLAG_OUT_OF_GRID) { h->err = "a material point left the grid box"; return MPM_E_OUT_OF_GRID; }
        h->err = "a point moved more than one cell beyond its brick between drift checks "
                 "(lower mpm_config.drift_check)";
        return MPM_E_ARENA;
    }
    if (*need_rebin) HIPCHK(h, hipMemsetAsync(h->flags_d, 0, 4, h->stream));
    return MPM_OK;
}

static int checks_impl(mpm_handle *h, double out[12]);

// Script(double, p) of Definitions.h:43-56: scientific, p digits, explicit sign
static std::string script(double v, int p)
{
    char b[64];
    std::snprintf(b, sizeof b, "%+.*e", p, v);
    return b;
}

// reportConsole (PhysicsEngine.cpp:33-120): the description, then time,
// runtime, the displacement monitor (the first platen point, as the worlds
// mark it: …PolygonCell.cpp:189), the reaction Force_y (the sum of the last
// step's Force_Temp) and sum m|v|^2 over all points.  The stress monitor is
// empty in the reference worlds (…PolygonCell.cpp:230 is commented out), so
// its fields are not printed.
static int report_line(mpm_handle *h, const char *desc, double runtime_s, std::string &out)
{
    double c[12];
    int rc = checks_impl(h, c);
    if (rc) return rc;
    out = desc ? desc : "";
    out += "\ttime: " + script(h->time, 6);
    out += "\tRuntime: " + script(runtime_s, 3);
    if (h->n_dc > 0) {
        double y = 0.0;
        HIPCHK(h, hipMemcpyAsync(&y, h->dc_x + 1, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        out += "\tPosition_y: " + script(y, 6);
    }
    out += "\tForce_y: " + script(c[1], 6);
    out += "\tKinetic Energy: " + script(c[2], 6);
    out += "\n";
    return MPM_OK;
}

// the console (cfg.verbose) and the log file (_STR_LOGFILE, Definitions.h:22)
static void report_emit(mpm_handle *h, const std::string &line)
{
    if (h->cfg.verbose) {
        std::fputs(line.c_str(), stdout);
        std::fflush(stdout);
    }
    if (!h->log_path.empty()) {
        if (FILE *f = std::fopen(h->log_path.c_str(), "a")) {
            std::fputs(line.c_str(), f);
            std::fclose(f);
        }
    }
}

static void report_console(mpm_handle *h, double runtime_s)
{
    if (!h->cfg.verbose && h->log_path.empty()) return;
    std::string line;
    if (report_line(h, nullptr, runtime_s, line) == MPM_OK) report_emit(h, line);
}

// batched worlds: clear this batch's y-fixed slots (uncovered nodes stay 0)
static int group_pre(mpm_handle *h, int nsteps)
{
    if (h->n_fy == 0) return MPM_OK;
    const size_t n = (size_t)nsteps * h->n_fy;
    k_fill_f64<<<cdiv(n, 256), 256, 0, h->stream>>>(h->fy_val, n, 0.0);
    HIPCHK(h, hipGetLastError());
    return MPM_OK;
}
// ... and reduce them per (step, group) in a fixed order into dst (the
// group part of a result set)
static int group_post(mpm_handle *h, int nsteps, double *dst)
{
    if (h->n_groups == 0) return MPM_OK;
    if (h->n_fy > 0) {
        k_group_sums<<<(unsigned)(nsteps * h->n_groups), 256, 0, h->stream>>>(h->fy_val, h->n_fy, h->fy_off, h->n_groups,
                                                                            dst);
    } else {
        const size_t n = (size_t)nsteps * h->n_groups;
        k_fill_f64<<<cdiv(n, 256), 256, 0, h->stream>>>(dst, n, 0.0);
    }
    HIPCHK(h, hipGetLastError());
    return MPM_OK;
}

static void drop_graphs(mpm_handle *h)
{
    for (auto &e : h->graphs) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.graph);
    }
    h->graphs.clear();
}

// A batch of nsteps plain steps with the step records of host set `slot`
// and its results written to result set `slot`: the records in, the fused
// step kernels, the per-step reaction sums with the flags word, the last
// step's grid totals and the per-group reactions.  Kernels only, with the
// pinned sets accessed in place.  Parities come from the step records, so
// the same sequence serves every starting parity.
template <typename real>
static int batch_ops(mpm_handle *h, int nsteps, int slot, std::vector<ProfMark> *marks)
{
    double *out = h->out_hd + (size_t)slot * h->out_stride;
    k_load_steps<<<1, 64, 0, h->stream>>>(reinterpret_cast<const uint32_t *>(h->steps_hd + (size_t)slot * h->steps_cap),
                                         reinterpret_cast<uint32_t *>(h->steps_d),
                                         (int)(nsteps * sizeof(StepRec) / 4));
    HIPCHK(h, hipGetLastError());
    if (h->cfg.persistent && h->persist_blocks > 0) {
        k_fill_u32<<<cdiv((size_t)nsteps * kQueueStride, 256), 256, 0, h->stream>>>(h->queue, (size_t)nsteps * kQueueStride, 0u);
        HIPCHK(h, hipGetLastError());
    }
    RCHK(group_pre(h, nsteps));
    for (int k = 0; k < nsteps; ++k) {
        const int pe = marks ? prof_begin(h, *marks, 0) : -1;
        RCHK(launch_particle<real>(h, true, k));
        prof_end(h, pe);
    }
    k_batch_sums<<<(unsigned)nsteps + 1, 256, 0, h->stream>>>(h->gpart, h->occ_cap, h->n_occ_d, nsteps, out,
                                                              h->grid3_d, h->flags_d);
    HIPCHK(h, hipGetLastError());
    RCHK(group_post(h, nsteps, out + 1 + h->steps_cap));
    return MPM_OK;
}

// Bytes of a batch's result block that the host reads: the flags word, the
// reaction history and (batched worlds) the per-group history.
static size_t out_bytes(const mpm_handle *h, int nsteps)
{
    return sizeof(double) * (h->n_groups > 0 ? 1 + (size_t)h->steps_cap + (size_t)nsteps * h->n_groups
                                             : 1 + (size_t)nsteps);
}

// The hipGraph of batch_ops(nsteps, slot) for the current brick layout,
// captured on first use.  No memset nodes.  upload and every plain-path rebin
// capture the full batch length for both host sets at once, so that no
// capture falls inside a run of steps.
template <typename real>
static int batch_graph(mpm_handle *h, int nsteps, int slot, hipGraphExec_t *out)
{
    if (h->graph_layout != h->layout) {
        drop_graphs(h);
        h->graph_layout = h->layout;
    }
    for (const auto &e : h->graphs)
        if (e.steps == nsteps && e.parity == slot) { *out = e.exec; return MPM_OK; }
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_ops<real>(h, nsteps, slot, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (rc != MPM_OK || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        h->err = "hipGraph capture failed";
        return MPM_E_HIP;
    }
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        h->err = "hipGraphInstantiate failed";
        return MPM_E_HIP;
    }
    if (h->graphs.size() >= 4) drop_graphs(h);
    h->graphs.push_back({nsteps, slot, g, exec});
    *out = exec;
    return MPM_OK;
}

// steps of a full plain-path batch (the drift check interval)
static int batch_len(const mpm_handle *h)
{
    return std::max(1, std::min(h->cfg.drift_check > 0 ? h->cfg.drift_check : 8, h->steps_cap));
}

static bool graphs_on(const mpm_handle *h) { return h->cfg.launch == MPM_LAUNCH_GRAPH && !h->cfg.profile; }

// capture the full-length batch for the current layout now (outside any run
// of steps)
static int prepare_graph(mpm_handle *h)
{
    if (!graphs_on(h) || h->n_occ == 0) return MPM_OK;
    hipGraphExec_t e = nullptr;
    for (int slot = 0; slot < 2; ++slot)
        RCHK(h->fp32 ? batch_graph<float>(h, batch_len(h), slot, &e) : batch_graph<double>(h, batch_len(h), slot, &e));
    return MPM_OK;
}

template <typename real>
static int step_impl(mpm_handle *h, double dt_total, int path, int32_t *finished)
{
    if (path == MPM_PATH_CONTACT) {
        RCHK(ensure_contact(h));
        RCHK(locate_point0<real>(h));
        // the first contact step of a call forms every bound from the point
        // state as it is now (uploads and plain steps may have changed it)
        h->ct.bnext_valid = false;
    } else if (h->plain_stale) {
        // contact steps moved the points: re-bin, rebuild the arenas and the
        // platen tags of the current step
        RCHK(rebin(h, false));
        if (h->n_dc > 0) {
            k_dc_marks<<<cdiv(h->n_dc, 256), 256, 0, h->stream>>>(h->g, h->dc_x, h->n_dc, (uint32_t)h->cycle,
                                                                 h->marks[h->parity]);
            HIPCHK(h, hipGetLastError());
        }
        h->plain_stale = false;
        RCHK(prepare_graph(h));
    }
    h->hist.clear();
    h->ghist.clear();
    for (double &v : h->kt) v = 0.0;
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    // Plain-path batches are pipelined: batch i+1 is enqueued before the
    // host reads batch i's flags, so the device never waits for the host
    // between batches.  Step records, history and flags of the two batches
    // use two sets of pinned host buffers (slot), and a rebin that batch i
    // asks for runs after batch i+1.  The contact path and profiling runs
    // consume every batch before the next.
    const bool pipeline = path == MPM_PATH_PLAIN && !h->cfg.profile;
    const int K = batch_len(h);
    const size_t cap = (size_t)h->steps_cap;
    struct Batch {
        bool on = false;
        int n = 0, slot = 0;
    };
    Batch pend;   // enqueued, results not yet read
    int slot = 0, in_batch = 0;
    std::vector<ProfMark> marks;
    auto recs = [&](int sl) { return h->steps_h + (size_t)sl * cap; };

    // enqueue the in_batch steps whose records are in recs(slot)
    auto launch = [&]() -> int {
        StepRec *r = recs(slot);
        for (int k = 0; k < in_batch; ++k) {
            r[k].par = (h->parity + k) & 1;
            r[k].skip_sums = (k + 1 < in_batch) ? 1 : 0;   // sums are read after the batch's last step
        }
        if (path == MPM_PATH_PLAIN && h->n_occ > 0) {
            if (graphs_on(h) && in_batch == K) {
                hipGraphExec_t exec = nullptr;
                RCHK(batch_graph<real>(h, in_batch, slot, &exec));
                HIPCHK(h, hipGraphLaunch(exec, h->stream));
            } else {
                RCHK(batch_ops<real>(h, in_batch, slot, h->cfg.profile ? &marks : nullptr));
            }
            h->last_path = 0;
        } else if (path == MPM_PATH_CONTACT) {
            HIPCHK(h, hipMemcpyAsync(h->steps_d, r, sizeof(StepRec) * in_batch, hipMemcpyHostToDevice, h->stream));
            for (int k = 0; k < in_batch; ++k) RCHK(launch_contact<real>(h, k));
            // the point kernel by original index is written when it is read
            // (readback) or before a rebin reorders the points
            h->ct.kernel_dirty = true;
            // per-step platen reaction of the batch
            k_c_step_sums<<<(unsigned)in_batch, 256, 0, h->stream>>>(h->ct.cpart, h->ct.nblk_n, h->hist_d);
            k_sum3<<<1, 256, 0, h->stream>>>(h->ct.cpart + (size_t)(in_batch - 1) * h->ct.nblk_n * 3, nullptr,
                                            h->ct.nblk_n, h->grid3_d);
            HIPCHK(h, hipGetLastError());
            if (h->n_groups > 0)
                k_fill_f64<<<cdiv((size_t)in_batch * h->n_groups, 256), 256, 0, h->stream>>>(
                    h->ghist_d, (size_t)in_batch * h->n_groups, 0.0);
            HIPCHK(h, hipMemcpyAsync(h->out_h + (size_t)slot * h->out_stride, h->out_d, out_bytes(h, in_batch),
                                     hipMemcpyDeviceToHost, h->stream));
            h->plain_stale = true;
            h->last_path = 1;
        } else {
            // no deformable points: platen blocks only, no sums
            HIPCHK(h, hipMemcpyAsync(h->steps_d, r, sizeof(StepRec) * in_batch, hipMemcpyHostToDevice, h->stream));
            for (int k = 0; k < in_batch; ++k) RCHK(launch_particle<real>(h, true, k));
            k_fill_f64<<<cdiv(in_batch, 256), 256, 0, h->stream>>>(h->hist_d, (size_t)in_batch, 0.0);
            k_fill_f64<<<1, 256, 0, h->stream>>>(h->grid3_d, 3, 0.0);
            HIPCHK(h, hipGetLastError());
            if (h->n_groups > 0)
                k_fill_f64<<<cdiv((size_t)in_batch * h->n_groups, 256), 256, 0, h->stream>>>(
                    h->ghist_d, (size_t)in_batch * h->n_groups, 0.0);
            HIPCHK(h, hipMemcpyAsync(h->out_h + (size_t)slot * h->out_stride, h->out_d, out_bytes(h, in_batch),
                                     hipMemcpyDeviceToHost, h->stream));
            h->last_path = 0;
        }
        HIPCHK(h, hipEventRecord(h->bev[slot], h->stream));
        h->parity ^= (in_batch & 1);
        pend.on = true;
        pend.n = in_batch;
        pend.slot = slot;
        slot ^= 1;
        in_batch = 0;
        return MPM_OK;
    };
    // wait for a launched batch and read its results; *need is set when it
    // asks for a rebin
    auto collect = [&](Batch &b, bool *need) -> int {
        if (!b.on) return MPM_OK;
        b.on = false;
        HIPCHK(h, hipEventSynchronize(h->bev[b.slot]));
        const double *o = h->out_h + (size_t)b.slot * h->out_stride;
        const int f = *reinterpret_cast<const int32_t *>(o);
        for (int k = 0; k < b.n; ++k) h->hist.push_back(o[1 + k]);
        if (h->n_groups > 0) {
            const double *gh = o + 1 + cap;
            for (int k = 0; k < b.n; ++k)
                for (int gr = 0; gr < h->n_groups; ++gr) h->ghist.push_back(gh[(size_t)k * h->n_groups + gr]);
        }
        if (f & (FLAG_OUT_OF_GRID | FLAG_ARENA | FLAG_NAN)) {
            (void)hipStreamSynchronize(h->stream);
            if (f & FLAG_NAN) { h->err = "non-finite point state"; return MPM_E_NAN; }
            if (f & FLAG_OUT_OF_GRID) { h->err = "a material point left the grid box"; return MPM_E_OUT_OF_GRID; }
            h->err = "a point moved more than one cell beyond its brick between drift checks "
                     "(lower mpm_config.drift_check)";
            return MPM_E_ARENA;
        }
        *need = *need || (f & FLAG_REBIN) != 0;
        return MPM_OK;
    };
    // launch the steps gathered so far; read the previous batch (or, with
    // drain, every batch); rebin when asked
    auto flush = [&](bool drain) -> int {
        Batch prev = pend;
        pend.on = false;
        if (in_batch > 0) RCHK(launch());
        else pend = prev, prev.on = false;
        bool need = false;
        RCHK(collect(prev, &need));
        if (need || drain || !pipeline) RCHK(collect(pend, &need));
        if (h->cfg.profile) prof_collect(h, marks);
        if (need) {
            HIPCHK(h, hipMemsetAsync(h->flags_d, 0, 4, h->stream));
            const int pe = prof_begin(h, marks, 2);
            RCHK(rebin(h, false));
            prof_end(h, pe);
            HIPCHK(h, hipStreamSynchronize(h->stream));
            if (h->cfg.profile) prof_collect(h, marks);
            if (path == MPM_PATH_CONTACT) RCHK(locate_point0<real>(h));
            else RCHK(prepare_graph(h));
        }
        return MPM_OK;
    };
    int rc = MPM_OK;
    double acc = 0.0;
    while (acc < dt_total) {
        double dt = h->cfg.dt_max;
        if (h->cfg.dt_max > (dt_total - acc)) dt = dt_total - acc;
        acc += dt;
        StepRec &r = recs(slot)[in_batch];
        r.dt = dt;
        r.t = h->time;
        r.has_tl = h->tl_t.empty() ? 0 : 1;
        timeline_velocity(h, h->time, r.vtl);
        r.epoch = (uint32_t)h->cycle;
        ++in_batch;
        // reportConsole every console_interval (…_MP.cpp:374-379); the batch is
        // cut there only when the line goes somewhere
        bool report = h->cfg.console_interval > 0.0 && (h->time - h->console_last > h->cfg.console_interval);
        if (report && !h->cfg.verbose && h->log_path.empty()) {
            h->console_last = h->time;
            report = false;
        }
        if (in_batch == K || report) {
            rc = flush(report);
            if (rc != MPM_OK) break;
        }
        if (report) {
            h->console_last = h->time;
            float ms_now = 0.f;
            HIPCHK(h, hipEventRecord(h->ev1, h->stream));
            HIPCHK(h, hipEventSynchronize(h->ev1));
            HIPCHK(h, hipEventElapsedTime(&ms_now, h->ev0, h->ev1));
            report_console(h, h->runtime_s + ms_now * 1e-3);
        }
        h->time += dt;
        h->cycle++;
    }
    if (rc == MPM_OK) rc = flush(true);
    if (rc != MPM_OK) {
        (void)hipStreamSynchronize(h->stream);
        return rc;
    }
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
    h->runtime_s += ms * 1e-3;
    if (finished) *finished = (h->time < h->cfg.time_end) ? 0 : 1;
    if (finished && *finished) {
        h->console_last = h->time;
        report_console(h, h->runtime_s);
        if (h->cfg.verbose) std::printf("Analysis complete ...\n");
    }
    return MPM_OK;
}

static int checks_impl(mpm_handle *h, double out[12])
{
    if (h->last_path == 1)
        k_finalize<<<1, 256, 0, h->stream>>>(h->ct.ppart, nullptr, h->ct.n_ppart, h->grid3_d, h->dc_v[h->parity],
                                           h->dc_m, h->n_dc, h->checks_d);
    else
        k_finalize<<<1, 256, 0, h->stream>>>(h->kp_part, h->n_occ_d, 0, h->grid3_d, h->dc_v[h->parity], h->dc_m,
                                           h->n_dc, h->checks_d, kPBlock / 64);
    HIPCHK(h, hipGetLastError());
    double r[8] = {0};
    HIPCHK(h, hipMemcpyAsync(r, h->checks_d, 8 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < 7; ++k) out[k] = r[k];
    out[7] = h->time;
    out[8] = (double)h->cycle;
    out[9] = (double)h->rebins;
    out[10] = (double)h->n_occ;
    out[11] = r[7];
    return MPM_OK;
}

// ================================================================== C ABI
extern "C" {

int mpm_abi_version(void) { return MPM_ABI_VERSION; }

// every call on a handle with an asynchronous step in flight is refused
#define BUSYCHK(h)                                                                  \
    do {                                                                            \
        if ((h) && (h)->async_state.load(std::memory_order_acquire) != 0) return MPM_E_STATE; \
    } while (0)

static const char *kBusy = "an asynchronous step is in flight on this handle; call mpm_step_wait first";

const char *mpm_last_error(const mpm_handle *h)
{
    if (h && h->async_state.load(std::memory_order_acquire) != 0) return kBusy;
    return h ? h->err.c_str() : g_create_error.c_str();
}

static void destroy_all(mpm_handle *h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    for (int b = 0; b < 2; ++b) {
        for (auto &f : h->pfields[b]) dev_free(f.p);
        h->pfields[b].clear();
        dev_free(h->arena[b]);
        dev_free((void *&)h->exps[b]);
        dev_free((void *&)h->keys[b]);
        dev_free((void *&)h->vals[b]);
    }
    for (int a = 0; a < 3; ++a) dev_free((void *&)h->xstage[a]);
    {
        auto &c = h->ct;
        void **cp[] = {(void **)&c.acc, (void **)&c.gkgrad, (void **)&c.gkg, (void **)&c.gm, (void **)&c.gv,
                       (void **)&c.gf, (void **)&c.flags, (void **)&c.near_i, (void **)&c.sg_i, (void **)&c.id0,
                       (void **)&c.near_d, (void **)&c.sg_d, (void **)&c.cmarks, (void **)&c.bounds,
                       (void **)&c.near_surf, (void **)&c.kp, (void **)&c.kg, (void **)&c.cpart, (void **)&c.ppart,
                       (void **)&c.kp_id, (void **)&c.kg_id, (void **)&c.ga, (void **)&c.fmask,
                       (void **)&h->dc_vol, (void **)&h->dc_s, (void **)&h->dc_fext, (void **)&h->dc_surf,
                       (void **)&h->dc_id};
        for (void **q : cp) dev_free(*q);
    }
    void **ptrs[] = {(void **)&h->dc_x, (void **)&h->dc_v[0], (void **)&h->dc_v[1], (void **)&h->dc_m,
                     (void **)&h->fixed, (void **)&h->marks[0], (void **)&h->marks[1], (void **)&h->occ_list,
                     (void **)&h->occ_start, (void **)&h->occ_count, (void **)&h->n_occ_d,
                     (void **)&h->slot_order, (void **)&h->slot_mat, (void **)&h->idesc, (void **)&h->binfo[0], (void **)&h->binfo[1], (void **)&h->mats,
                     (void **)&h->steps_d,
                     (void **)&h->out_d, (void **)&h->queue, (void **)&h->kp_part, (void **)&h->gpart, (void **)&h->grid3_d,
                     (void **)&h->fy_slot, (void **)&h->fy_off, (void **)&h->fy_val,
                     (void **)&h->checks_d, (void **)&h->radix_hist, (void **)&h->radix_off, (void **)&h->heads,
                     (void **)&h->headpos, (void **)&h->gdesc};
    for (void **p : ptrs) dev_free(*p);
    for (auto *p : h->scan_scratch) { void *q = p; dev_free(q); }
    h->scan_scratch.clear();
    if (h->steps_h) (void)hipHostFree(h->steps_h);
    if (h->flags_h) (void)hipHostFree(h->flags_h);
    if (h->out_h) (void)hipHostFree(h->out_h);
    h->out_h = nullptr;
    h->flags_d = nullptr;
    h->hist_d = h->ghist_d = nullptr;
    h->steps_h = nullptr;
    h->flags_h = nullptr;
    for (hipEvent_t e : h->pev) (void)hipEventDestroy(e);
    h->pev.clear();
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    for (hipEvent_t &e : h->bev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

void mpm_destroy(mpm_handle *h)
{
    if (!h) return;
    if (h->worker.joinable()) h->worker.join();
    destroy_all(h);
    delete h;
}

int mpm_create(const mpm_config *cfg, mpm_handle **out)
{
    if (!cfg || !out) { g_create_error = "null argument"; return MPM_E_ARG; }
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (!(cfg->length[a] > 0.0) || cfg->cells[a] < 1) { g_create_error = "grid length and cells must be positive"; return MPM_E_ARG; }
    // cell indices enter the wave reductions of the contact windows, which
    // take values in [0, 2^20]
    for (int a = 0; a < 3; ++a)
        if (cfg->cells[a] > (1 << 20) - 2) { g_create_error = "at most 2^20 - 2 cells per axis"; return MPM_E_ARG; }
    if (!(cfg->dt_max > 0.0)) { g_create_error = "dt_max must be positive"; return MPM_E_ARG; }
    if (cfg->precision != MPM_FP64 && cfg->precision != MPM_FP32) { g_create_error = "unknown precision"; return MPM_E_ARG; }
    if (cfg->n_timeline < 0 || (cfg->n_timeline > 0 && (!cfg->timeline_t || !cfg->timeline_v))) {
        g_create_error = "bad timeline";
        return MPM_E_ARG;
    }
    // FP32 state holds positions as Q6.26 cells relative to the brick origin
    // (|u| < 32); a point stays within [-1, S + 1] cells of it plus one
    // step's motion before the drift check flags it
    for (int a = 0; a < 3; ++a) {
        const int S = std::min(cfg->brick[a] > 0 ? cfg->brick[a] : (a == 2 ? 10 : 4), std::max(1, cfg->cells[a]));
        if (cfg->precision == MPM_FP32 && S > kMaxBrickFp32) {
            g_create_error = "FP32 state: brick size must be at most " + std::to_string(kMaxBrickFp32) +
                             " cells per axis (Q6.26 positions relative to the brick origin)";
            return MPM_E_ARG;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return MPM_E_NO_DEVICE; }
    if (cfg->device < 0 || cfg->device >= ndev) { g_create_error = "bad device ordinal"; return MPM_E_ARG; }
    mpm_handle *h = new mpm_handle();
    h->cfg = *cfg;
    h->device = cfg->device;
    h->fp32 = cfg->precision == MPM_FP32;
    h->rsize = h->fp32 ? 4 : 8;
    h->tl_t.assign(cfg->timeline_t, cfg->timeline_t + cfg->n_timeline);
    h->tl_v.assign(cfg->timeline_v, cfg->timeline_v + 3 * cfg->n_timeline);
    h->cfg.timeline_t = nullptr;
    h->cfg.timeline_v = nullptr;
    h->cfg.node_fixed = nullptr;
    h->time = cfg->time_start;
    // d_TimeConsole_Last = 0 after the world's first report (…PolygonCell.cpp:311)
    h->console_last = cfg->time_start;
    auto fail = [&](int code) {
        g_create_error = h->err;
        destroy_all(h);
        delete h;
        return code;
    };
    if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return fail(MPM_E_HIP); }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { h->err = "stream"; return fail(MPM_E_HIP); }
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&h->bev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->bev[1], hipEventDisableTiming) != hipSuccess) {
        h->err = "event";
        return fail(MPM_E_HIP);
    }

    GridParams &g = h->g;
    for (int a = 0; a < 3; ++a) {
        g.L[a] = cfg->length[a];
        g.cells[a] = cfg->cells[a];
        g.nodes[a] = cfg->cells[a] + 1;
        g.h[a] = cfg->length[a] / (double)cfg->cells[a];
        g.hf[a] = (float)g.h[a];
        g.ihf[a] = (float)(1.0 / g.h[a]);
        g.ih[a] = 1.0 / g.h[a];
        int S = cfg->brick[a] > 0 ? cfg->brick[a] : (a == 2 ? 10 : 4);
        S = std::min(S, std::max(1, cfg->cells[a]));
        // an arena window reaches one neighbour per side only when S >= 3
        if (S < 3 && S < cfg->cells[a]) { h->err = "brick size must be >= 3 cells (or span the axis)"; return fail(MPM_E_ARG); }
        g.S[a] = S;
        g.nb[a] = g.nodes[a] / S + ((g.nodes[a] % S) ? 1 : 0);
        g.A[a] = S + 3;
    }
    g.arenaN = g.A[0] * g.A[1] * g.A[2];
    {
        // dynamic LDS of the step kernel: fixed-point arena + node window
        // (plus about 1 KB of static LDS)
        const size_t need = (size_t)kQ * g.arenaN * 8 + (size_t)6 * g.arenaN * h->rsize;
        int maxlds = 0;
        if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) != hipSuccess)
            maxlds = 64 * 1024;
        if (need + 1024 > (size_t)maxlds) { h->err = "brick too large for the LDS arena and node window"; return fail(MPM_E_ARG); }
        hipError_t e1, e2;
        if (h->fp32) {
            e1 = hipFuncSetAttribute((const void *)k_particle<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
            e2 = hipFuncSetAttribute((const void *)k_particle<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
            if (e1 == hipSuccess)
                e1 = hipFuncSetAttribute((const void *)k_particle<float, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
        } else {
            e1 = hipFuncSetAttribute((const void *)k_particle<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
            e2 = hipFuncSetAttribute((const void *)k_particle<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
            if (e1 == hipSuccess)
                e1 = hipFuncSetAttribute((const void *)k_particle<double, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
        }
        if (e1 != hipSuccess || e2 != hipSuccess) { h->err = "hipFuncSetAttribute (dynamic LDS) failed"; return fail(MPM_E_HIP); }
    }
    {
        // the XCDs workgroups are dealt to round robin (brick launch order)
        int nx = 0;
        if (hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, h->device) == hipSuccess && nx >= 1 &&
            nx <= kQueueStride)
            h->n_xcd = nx;
        (void)hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, h->device);
    }
    const int ncb = g.S[0] * g.S[1] * g.S[2];
    g.cbits = 0;
    while ((1 << g.cbits) < ncb) ++g.cbits;
    h->nbricks = g.nb[0] * g.nb[1] * g.nb[2];
    if (((unsigned long long)h->nbricks << g.cbits) >= (1ull << 32)) { h->err = "grid too large for 32-bit brick keys"; return fail(MPM_E_ARG); }
    g.m_min = cfg->mass_min;
    g.damping = cfg->damping;
    g.strain_form = h->fp32 ? 1 : cfg->strain_form;
    g.neohookean = cfg->neohookean;
    g.track_strain = cfg->track_strain;
    h->nnodes = (long long)g.nodes[0] * g.nodes[1] * g.nodes[2];

    int rc;
    if ((rc = dalloc(h, &h->fixed, (size_t)h->nnodes))) return fail(rc);
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(h, &h->marks[b], (size_t)h->nnodes))) return fail(rc);
        (void)hipMemsetAsync(h->marks[b], 0, (size_t)h->nnodes * 8, h->stream);
    }
    h->fixed_host.assign((size_t)h->nnodes, 0);
    if (cfg->node_fixed) h->fixed_host.assign(cfg->node_fixed, cfg->node_fixed + h->nnodes);
    if (cfg->node_fixed) {
        if (hipMemcpyAsync(h->fixed, h->fixed_host.data(), (size_t)h->nnodes, hipMemcpyHostToDevice, h->stream) != hipSuccess) {
            h->err = "copy node_fixed";
            return fail(MPM_E_HIP);
        }
    } else {
        (void)hipMemsetAsync(h->fixed, 0, (size_t)h->nnodes, h->stream);
    }
    const size_t nb = (size_t)h->nbricks;
    if ((rc = dalloc(h, &h->occ_list, nb)) || (rc = dalloc(h, &h->occ_start, nb)) || (rc = dalloc(h, &h->occ_count, nb)) ||
        (rc = dalloc(h, &h->n_occ_d, 1)) || (rc = dalloc(h, &h->slot_order, nb)) || (rc = dalloc(h, &h->slot_mat, nb)) || (rc = dalloc(h, &h->idesc, 2 * nb)) ||
        (rc = dalloc(h, &h->binfo[0], nb)) || (rc = dalloc(h, &h->binfo[1], nb)) ||
        (rc = dalloc(h, &h->checks_d, 16)) || (rc = dalloc(h, &h->grid3_d, 3)))
        return fail(rc);
    (void)hipMemsetAsync(h->n_occ_d, 0, 4, h->stream);
    (void)hipMemsetAsync(h->grid3_d, 0, 24, h->stream);
    h->steps_cap = 64;
    if ((rc = dalloc(h, &h->steps_d, h->steps_cap)) || (rc = dalloc(h, &h->queue, (size_t)h->steps_cap * kQueueStride)))
        return fail(rc);
    if (cfg->persistent) {
        // resident brick blocks: blocks per CU at the step kernel's LDS, a
        // multiple of the XCD count
        const size_t need = (size_t)kQ * g.arenaN * 8 + (size_t)6 * g.arenaN * h->rsize;
        int per_cu = 0;
        const hipError_t e = h->fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_particle<float, true, true>, kPBlock, need)
                                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_particle<double, true, true>, kPBlock, need);
        if (e != hipSuccess || per_cu < 1 || h->n_cu < 1) { h->err = "occupancy query for the persistent launch failed"; return fail(MPM_E_HIP); }
        h->persist_blocks = std::max(h->n_xcd, (per_cu * h->n_cu) / h->n_xcd * h->n_xcd);
    }
    if (hipHostMalloc((void **)&h->steps_h, 2 * sizeof(StepRec) * h->steps_cap,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&h->steps_hd, h->steps_h, 0) != hipSuccess ||
        hipHostMalloc((void **)&h->flags_h, 16, hipHostMallocDefault) != hipSuccess) {
        h->err = "hipHostMalloc";
        return fail(MPM_E_NOMEM);
    }
    // batched worlds: y-fixed nodes of every group, in group order
    if (cfg->node_group && cfg->n_groups > 0) {
        const int ng = cfg->n_groups;
        std::vector<int32_t> cnt(ng + 1, 0), slot((size_t)h->nnodes, -1);
        for (long long I = 0; I < h->nnodes; ++I) {
            if (cfg->node_group[I] >= ng) { h->err = "node_group id >= n_groups"; return fail(MPM_E_ARG); }
            if (cfg->node_fixed && (cfg->node_fixed[I] & MPM_FIX_Y)) cnt[cfg->node_group[I] + 1]++;
        }
        for (int gr = 0; gr < ng; ++gr) cnt[gr + 1] += cnt[gr];
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        for (long long I = 0; I < h->nnodes; ++I)
            if (cfg->node_fixed && (cfg->node_fixed[I] & MPM_FIX_Y)) slot[I] = fill[cfg->node_group[I]]++;
        h->n_groups = ng;
        h->n_fy = cnt[ng];
        if ((rc = dalloc(h, &h->fy_slot, (size_t)h->nnodes)) || (rc = dalloc(h, &h->fy_off, (size_t)ng + 1)) ||
            (rc = dalloc(h, &h->fy_val, (size_t)h->steps_cap * std::max(h->n_fy, 1))))
            return fail(rc);
        if (hipMemcpy(h->fy_slot, slot.data(), sizeof(int32_t) * h->nnodes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->fy_off, cnt.data(), sizeof(int32_t) * (ng + 1), hipMemcpyHostToDevice) != hipSuccess) {
            h->err = "group tables";
            return fail(MPM_E_HIP);
        }
    }
    // the result block of a batch and its two host sets
    h->out_stride = 1 + (size_t)h->steps_cap * (1 + (size_t)h->n_groups);
    if ((rc = dalloc(h, &h->out_d, h->out_stride))) return fail(rc);
    if (hipHostMalloc((void **)&h->out_h, 2 * sizeof(double) * h->out_stride,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&h->out_hd, h->out_h, 0) != hipSuccess) {
        h->err = "hipHostMalloc";
        return fail(MPM_E_NOMEM);
    }
    h->flags_d = reinterpret_cast<int32_t *>(h->out_d);
    h->hist_d = h->out_d + 1;
    h->ghist_d = h->n_groups > 0 ? h->out_d + 1 + h->steps_cap : nullptr;
    (void)hipMemsetAsync(h->out_d, 0, sizeof(double) * h->out_stride, h->stream);
    if (hipStreamSynchronize(h->stream) != hipSuccess) { h->err = "init sync"; return fail(MPM_E_HIP); }
    *out = h;
    return MPM_OK;
}

}  // extern "C"

template <typename real>
static int upload_t(mpm_handle *h, const mpm_points *pts)
{
    const int n = pts->n;
    std::vector<int> def, dc;
    for (int i = 0; i < n; ++i) ((pts->dc && pts->dc[i]) ? dc : def).push_back(i);
    h->n_total = n;
    h->n_def = (int)def.size();
    h->n_dc = (int)dc.size();
    h->def_id_host.assign(def.begin(), def.end());
    h->dc_id_host.assign(dc.begin(), dc.end());
    // materials: dedupe (type, nu, sy, C0, C1[, E])
    auto mkey = [&](int i, bool withE) {
        return std::make_tuple(pts->type ? pts->type[i] : 0, pts->nu ? pts->nu[i] : 0.0, pts->sy ? pts->sy[i] : 0.0,
                               pts->C0 ? pts->C0[i] : 0.0, pts->C1 ? pts->C1[i] : 0.0,
                               withE && pts->E ? pts->E[i] : 0.0);
    };
    std::map<std::tuple<int, double, double, double, double, double>, int> mm;
    bool withE = true;
    for (int i : def) { mm.emplace(mkey(i, true), (int)mm.size()); if (mm.size() > 255) break; }
    if (mm.size() > 255) {
        withE = false;
        mm.clear();
        for (int i : def) { mm.emplace(mkey(i, false), (int)mm.size()); if (mm.size() > 255) break; }
        if (mm.size() > 255) { h->err = "more than 255 distinct materials"; return MPM_E_ARG; }
    }
    h->has_pointE = withE ? 0 : 1;
    h->has_fext = 0;
    if (pts->fext)
        for (int i : def) if (pts->fext[3 * i] != 0.0 || pts->fext[3 * i + 1] != 0.0 || pts->fext[3 * i + 2] != 0.0) { h->has_fext = 1; break; }
    h->g.has_pointE = h->has_pointE;
    h->g.has_fext = h->has_fext;
    std::vector<Material> mats(std::max<size_t>(mm.size(), 1));
    for (auto &kv : mm) {
        Material m{};
        m.type = std::get<0>(kv.first);
        m.nu = std::get<1>(kv.first);
        m.sy = std::get<2>(kv.first);
        m.C0 = std::get<3>(kv.first);
        m.C1 = std::get<4>(kv.first);
        m.E = std::get<5>(kv.first);
        m.C = m.E / (1.0 + m.nu) / (1.0 - 2.0 * m.nu);
        m.G = 0.5 * m.E / (1.0 + m.nu);
        m.k0 = m.sy / std::sqrt(3.0);
        m.invE = 1.0 / m.E;
        m.invG = 1.0 / m.G;
        m.Cn = 1.0 / ((1.0 + m.nu) * (1.0 - 2.0 * m.nu));
        m.hs = std::sqrt(2.0 / 3.0) * m.C1;
        mats[kv.second] = m;
    }
    RCHK(dalloc(h, &h->mats, mats.size()));
    HIPCHK(h, hipMemcpyAsync(h->mats, mats.data(), mats.size() * sizeof(Material), hipMemcpyHostToDevice, h->stream));
    h->n_mats = (int)mats.size();

    RCHK(alloc_points<real>(h, h->n_def));
    RCHK(alloc_sort_scratch(h, (uint32_t)std::max(h->n_def, 1)));
    RCHK(dalloc(h, &h->heads, (size_t)std::max(h->n_def, 1)));
    RCHK(dalloc(h, &h->headpos, (size_t)std::max(h->n_def, 1)));
    for (int a = 0; a < 3; ++a) RCHK(dalloc(h, &h->xstage[a], (size_t)std::max(h->n_def, 1)));

    // host staging in input order, converted to the device precision
    const int nd = h->n_def;
    PArrays<real> &P = parr<real>(h, h->cur);
    std::vector<real> tmp(std::max(nd, 1));
    std::vector<double> tmpd(std::max(nd, 1));
    auto put = [&](real *dst, auto getter) -> int {
        for (int k = 0; k < nd; ++k) tmp[k] = (real)getter(def[k]);
        HIPCHK(h, hipMemcpy(dst, tmp.data(), sizeof(real) * nd, hipMemcpyHostToDevice));
        return MPM_OK;
    };
    for (int a = 0; a < 3; ++a) {
        for (int k = 0; k < nd; ++k) tmpd[k] = pts->x[3 * def[k] + a];
        HIPCHK(h, hipMemcpy(h->xstage[a], tmpd.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
        RCHK(put(P.v[a], [&](int i) { return pts->v ? pts->v[3 * i + a] : 0.0; }));
        if (h->has_fext) RCHK(put(P.fext[a], [&](int i) { return pts->fext[3 * i + a]; }));
    }
    for (int a = 0; a < 9; ++a)
        RCHK(put(P.F[a], [&](int i) {   // stored form: F - I in FP32 state
            const double F = pts->F ? pts->F[9 * i + a] : ((a % 4) == 0 ? 1.0 : 0.0);
            return (double)fstore<real>(F, a);
        }));
    for (int a = 0; a < 6; ++a) {
        RCHK(put(P.s[a], [&](int i) { return pts->stress ? pts->stress[6 * i + a] : 0.0; }));
        RCHK(put(P.ep[a], [&](int i) { return pts->eps_p ? pts->eps_p[6 * i + a] : 0.0; }));
        if (h->cfg.track_strain) RCHK(put(P.eps[a], [&](int i) { return pts->strain ? pts->strain[6 * i + a] : 0.0; }));
    }
    RCHK(put(P.alpha, [&](int i) { return pts->alpha ? pts->alpha[i] : 0.0; }));
    RCHK(put(P.mass, [&](int i) { return pts->mass[i]; }));
    RCHK(put(P.vol0, [&](int i) { return pts->vol0 ? pts->vol0[i] : pts->vol[i]; }));
    if (h->has_pointE) RCHK(put(P.E, [&](int i) { return pts->E[i]; }));
    {
        std::vector<uint8_t> mt(std::max(nd, 1));
        std::vector<uint32_t> id(std::max(nd, 1));
        for (int k = 0; k < nd; ++k) {
            mt[k] = (uint8_t)mm.at(mkey(def[k], withE));
            id[k] = (uint32_t)def[k];
        }
        HIPCHK(h, hipMemcpy(P.mat, mt.data(), nd, hipMemcpyHostToDevice));
        for (int k = 0; k < nd; ++k) mt[k] = (pts->surface && pts->surface[def[k]]) ? 1 : 0;
        HIPCHK(h, hipMemcpy(P.surf, mt.data(), nd, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(P.id, id.data(), 4 * (size_t)nd, hipMemcpyHostToDevice));
    }
    // platens
    RCHK(dalloc(h, &h->dc_x, 3 * (size_t)std::max(h->n_dc, 1)));
    for (int b = 0; b < 2; ++b) RCHK(dalloc(h, &h->dc_v[b], 3 * (size_t)std::max(h->n_dc, 1)));
    RCHK(dalloc(h, &h->dc_m, (size_t)std::max(h->n_dc, 1)));
    RCHK(dalloc(h, &h->dc_vol, (size_t)std::max(h->n_dc, 1)));
    RCHK(dalloc(h, &h->dc_s, 6 * (size_t)std::max(h->n_dc, 1)));
    RCHK(dalloc(h, &h->dc_fext, 3 * (size_t)std::max(h->n_dc, 1)));
    RCHK(dalloc(h, &h->dc_surf, (size_t)std::max(h->n_dc, 1)));
    RCHK(dalloc(h, &h->dc_id, (size_t)std::max(h->n_dc, 1)));
    h->surf0 = (n > 0 && pts->surface && pts->surface[0]) ? 1 : 0;
    h->id0_dc = false;
    for (int j = 0; j < h->n_dc; ++j)
        if (dc[j] == 0) { h->id0_dc = true; h->id0_dc_index = j; }
    if (h->n_dc > 0) {
        std::vector<double> vol(h->n_dc), st(6 * (size_t)h->n_dc, 0.0), fe(3 * (size_t)h->n_dc, 0.0);
        std::vector<uint8_t> sf(h->n_dc);
        std::vector<uint32_t> id(h->n_dc);
        for (int j = 0; j < h->n_dc; ++j) {
            const int i = dc[j];
            vol[j] = pts->vol ? pts->vol[i] : pts->vol0[i];
            for (int a = 0; a < 6; ++a) st[6 * j + a] = pts->stress ? pts->stress[6 * i + a] : 0.0;
            for (int a = 0; a < 3; ++a) fe[3 * j + a] = pts->fext ? pts->fext[3 * i + a] : 0.0;
            sf[j] = (pts->surface && pts->surface[i]) ? 1 : 0;
            id[j] = (uint32_t)i;
        }
        HIPCHK(h, hipMemcpy(h->dc_vol, vol.data(), sizeof(double) * h->n_dc, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dc_s, st.data(), sizeof(double) * 6 * h->n_dc, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dc_fext, fe.data(), sizeof(double) * 3 * h->n_dc, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dc_surf, sf.data(), h->n_dc, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dc_id, id.data(), sizeof(uint32_t) * h->n_dc, hipMemcpyHostToDevice));
    }
    {
        std::vector<double> x(3 * h->n_dc + 1), v(3 * h->n_dc + 1), m(h->n_dc + 1);
        for (int j = 0; j < h->n_dc; ++j) {
            const int i = dc[j];
            for (int a = 0; a < 3; ++a) {
                x[3 * j + a] = pts->x[3 * i + a];
                v[3 * j + a] = pts->v ? pts->v[3 * i + a] : 0.0;
            }
            m[j] = pts->mass[i];
        }
        HIPCHK(h, hipMemcpy(h->dc_x, x.data(), sizeof(double) * 3 * h->n_dc, hipMemcpyHostToDevice));
        for (int b = 0; b < 2; ++b)
            HIPCHK(h, hipMemcpy(h->dc_v[b], v.data(), sizeof(double) * 3 * h->n_dc, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->dc_m, m.data(), sizeof(double) * h->n_dc, hipMemcpyHostToDevice));
        // static platen fields for readback (F, stress, ... never change)
        h->dc_saved_store.assign((size_t)h->n_dc * 36, 0.0);
        for (int j = 0; j < h->n_dc; ++j) {
            const int i = dc[j];
            double *o = &h->dc_saved_store[(size_t)36 * j];
            for (int a = 0; a < 9; ++a) o[a] = pts->F ? pts->F[9 * i + a] : ((a % 4) == 0 ? 1.0 : 0.0);
            for (int a = 0; a < 6; ++a) o[9 + a] = pts->stress ? pts->stress[6 * i + a] : 0.0;
            for (int a = 0; a < 6; ++a) o[15 + a] = pts->strain ? pts->strain[6 * i + a] : 0.0;
            for (int a = 0; a < 6; ++a) o[21 + a] = pts->eps_p ? pts->eps_p[6 * i + a] : 0.0;
            o[27] = pts->alpha ? pts->alpha[i] : 0.0;
            o[28] = pts->vol ? pts->vol[i] : (pts->vol0 ? pts->vol0[i] : 0.0);
        }
    }
    for (int b = 0; b < 2; ++b) HIPCHK(h, hipMemsetAsync(h->marks[b], 0, (size_t)h->nnodes * 8, h->stream));
    h->parity = 0;
    if (h->n_dc > 0) {
        k_dc_marks<<<cdiv(h->n_dc, 256), 256, 0, h->stream>>>(h->g, h->dc_x, h->n_dc, (uint32_t)h->cycle,
                                                             h->marks[h->parity]);
        HIPCHK(h, hipGetLastError());
    }
    HIPCHK(h, hipMemsetAsync(h->flags_d, 0, 4, h->stream));
    RCHK(rebin(h, true));
    bool need = false;
    RCHK(check_flags(h, &need));
    HIPCHK(h, hipMemsetAsync(h->flags_d, 0, 4, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->rebins = 0;
    RCHK(prepare_graph(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MPM_OK;
}

extern "C" {

int mpm_upload_points(mpm_handle *h, const mpm_points *pts)
{
    BUSYCHK(h);
    if (!h || !pts || pts->n < 0 || !pts->x || !pts->mass) { if (h) h->err = "null points"; return MPM_E_ARG; }
    if (h->uploaded) { h->err = "points already uploaded"; return MPM_E_STATE; }
    if (!pts->vol0 && !pts->vol) { h->err = "vol0 or vol required"; return MPM_E_ARG; }
    if ((uint32_t)pts->n > kMaxPoints) { h->err = "more than 2^28 points"; return MPM_E_ARG; }
    {   // every given real input must be finite: nothing non-finite reaches a kernel
        const size_t n = (size_t)pts->n;
        const struct { const double *p; size_t w; const char *name; } in[] = {
            {pts->x, 3, "x"}, {pts->v, 3, "v"}, {pts->fext, 3, "fext"}, {pts->F, 9, "F"}, {pts->strain, 6, "strain"},
            {pts->stress, 6, "stress"}, {pts->eps_p, 6, "eps_p"}, {pts->alpha, 1, "alpha"}, {pts->mass, 1, "mass"},
            {pts->vol, 1, "vol"}, {pts->vol0, 1, "vol0"}, {pts->E, 1, "E"}, {pts->nu, 1, "nu"}, {pts->sy, 1, "sy"},
            {pts->C0, 1, "C0"}, {pts->C1, 1, "C1"}};
        for (const auto &f : in) {
            if (!f.p) continue;
            for (size_t k = 0; k < n * f.w; ++k)
                if (!std::isfinite(f.p[k])) {
                    h->err = std::string("non-finite ") + f.name + " of point " + std::to_string(k / f.w);
                    return MPM_E_ARG;
                }
        }
    }
    HIPCHK(h, hipSetDevice(h->device));
    int rc = h->fp32 ? upload_t<float>(h, pts) : upload_t<double>(h, pts);
    if (rc == MPM_OK) h->uploaded = true;
    return rc;
}

static int step_args(mpm_handle *h, double dt_total, int32_t path)
{
    if (!h->uploaded) { h->err = "mpm_step before mpm_upload_points"; return MPM_E_STATE; }
    if (path != MPM_PATH_PLAIN && path != MPM_PATH_CONTACT) { h->err = "unknown step path"; return MPM_E_ARG; }
    if (!(dt_total >= 0.0)) { h->err = "dt_total must be >= 0"; return MPM_E_ARG; }
    return MPM_OK;
}

static int step_run(mpm_handle *h, double dt_total, int32_t path, int32_t *finished)
{
    HIPCHK(h, hipSetDevice(h->device));
    return h->fp32 ? step_impl<float>(h, dt_total, path, finished) : step_impl<double>(h, dt_total, path, finished);
}

int mpm_step(mpm_handle *h, double dt_total, int32_t path, int32_t *finished)
{
    if (!h) return MPM_E_ARG;
    BUSYCHK(h);
    RCHK(step_args(h, dt_total, path));
    return step_run(h, dt_total, path, finished);
}

int mpm_step_async(mpm_handle *h, double dt_total, int32_t path)
{
    if (!h) return MPM_E_ARG;
    BUSYCHK(h);
    RCHK(step_args(h, dt_total, path));
    if (h->worker.joinable()) h->worker.join();
    h->async_rc = MPM_OK;
    h->async_finished = 0;
    h->async_state.store(1, std::memory_order_release);
    try {
        h->worker = std::thread([h, dt_total, path]() {
            int32_t fin = 0;
            const int rc = step_run(h, dt_total, path, &fin);
            h->async_rc = rc;
            h->async_finished = fin;
            h->async_state.store(2, std::memory_order_release);
        });
    } catch (...) {
        h->async_state.store(0, std::memory_order_release);
        h->err = "could not start the step thread";
        return MPM_E_STATE;
    }
    return MPM_OK;
}

int mpm_step_poll(mpm_handle *h, int32_t *done)
{
    if (!h || !done) return MPM_E_ARG;
    const int st = h->async_state.load(std::memory_order_acquire);
    if (st == 0) { h->err = "no asynchronous step on this handle"; return MPM_E_STATE; }
    *done = st == 2 ? 1 : 0;
    return MPM_OK;
}

int mpm_step_wait(mpm_handle *h, int32_t *finished)
{
    if (!h) return MPM_E_ARG;
    if (h->async_state.load(std::memory_order_acquire) == 0) { h->err = "no asynchronous step on this handle"; return MPM_E_STATE; }
    h->worker.join();
    const int rc = h->async_rc;
    if (finished) *finished = h->async_finished;
    h->async_state.store(0, std::memory_order_release);
    return rc;
}

}  // extern "C"

template <typename real>
static int readback_t(mpm_handle *h, uint32_t mask, mpm_points *out)
{
    const int n = h->n_total;
    constexpr int NF = 9;   // x v F stress strain eps_p alpha vol j2 (doubles)
    double *st[NF] = {nullptr};
    const int widths[NF] = {3, 3, 9, 6, 6, 6, 1, 1, 2};
    const uint32_t bits[NF] = {MPM_F_X, MPM_F_V, MPM_F_F, MPM_F_STRESS, MPM_F_STRAIN, MPM_F_EPSP,
                               MPM_F_ALPHA, MPM_F_VOL, MPM_F_J2};
    double *dst[NF] = {out->x, out->v, out->F, out->stress, out->strain, out->eps_p, out->alpha, out->vol, out->j2};
    for (int k = 0; k < NF; ++k) {
        if (!(mask & bits[k]) || !dst[k]) continue;
        RCHK(dalloc(h, &st[k], (size_t)widths[k] * std::max(n, 1)));
    }
    int32_t *cell_d = nullptr;
    uint32_t *order_d = nullptr;
    const bool want_cell = (mask & MPM_F_CELL) && out->cell && out->order;
    if (want_cell) {
        RCHK(dalloc(h, &cell_d, 3 * (size_t)std::max(n, 1)));
        RCHK(dalloc(h, &order_d, (size_t)std::max(n, 1)));
    }
    uint32_t m = 0;
    for (int k = 0; k < NF; ++k) if (st[k]) m |= bits[k];
    if (want_cell) m |= MPM_F_CELL;
    ReadbackOut o{st[0], st[1], st[2], st[3], st[4], st[5], st[6], st[7], st[8], cell_d, order_d};
    if (h->n_def > 0) {
        k_readback<real><<<cdiv(h->n_def, kBlock), kBlock, 0, h->stream>>>(h->g, parr<real>(h, h->cur), (uint32_t)h->n_def, o, m);
        HIPCHK(h, hipGetLastError());
    }
    std::vector<double> hx(3 * (size_t)h->n_dc + 1), hv(3 * (size_t)h->n_dc + 1);
    if (h->n_dc > 0) {
        HIPCHK(h, hipMemcpyAsync(hx.data(), h->dc_x, sizeof(double) * 3 * h->n_dc, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(hv.data(), h->dc_v[h->parity], sizeof(double) * 3 * h->n_dc, hipMemcpyDeviceToHost, h->stream));
    }
    for (int k = 0; k < NF; ++k)
        if (st[k] && h->n_def > 0) {
            // copy whole array; platen entries are overwritten below
            std::vector<double> buf((size_t)widths[k] * n);
            HIPCHK(h, hipMemcpyAsync(buf.data(), st[k], sizeof(double) * widths[k] * n, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            for (int i : h->def_id_host)
                for (int w = 0; w < widths[k]; ++w) dst[k][(size_t)widths[k] * i + w] = buf[(size_t)widths[k] * i + w];
        }
    if (want_cell && h->n_def > 0) {
        std::vector<int32_t> cb(3 * (size_t)n);
        std::vector<uint32_t> ob((size_t)n);
        HIPCHK(h, hipMemcpyAsync(cb.data(), cell_d, sizeof(int32_t) * 3 * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(ob.data(), order_d, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (int i : h->def_id_host) {
            for (int a = 0; a < 3; ++a) out->cell[3 * (size_t)i + a] = cb[3 * (size_t)i + a];
            out->order[i] = ob[i];
        }
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int j = 0; j < h->n_dc; ++j) {
        const int i = h->dc_id_host[j];
        const double *sv = &h->dc_saved_store[(size_t)36 * j];
        if (st[0]) for (int a = 0; a < 3; ++a) out->x[3 * i + a] = hx[3 * j + a];
        if (st[1]) for (int a = 0; a < 3; ++a) out->v[3 * i + a] = hv[3 * j + a];
        if (st[2]) for (int a = 0; a < 9; ++a) out->F[9 * i + a] = sv[a];
        if (st[3]) for (int a = 0; a < 6; ++a) out->stress[6 * i + a] = sv[9 + a];
        if (st[4]) for (int a = 0; a < 6; ++a) out->strain[6 * i + a] = sv[15 + a];
        if (st[5]) for (int a = 0; a < 6; ++a) out->eps_p[6 * i + a] = sv[21 + a];
        if (st[6]) out->alpha[i] = sv[27];
        if (st[7]) out->vol[i] = sv[28];
        if (st[8]) {   // platen points keep their initial stress and plastic strain
            out->j2[2 * (size_t)i] = j2_of(sv + 9);
            out->j2[2 * (size_t)i + 1] = j2_of(sv + 21);
        }
        if (want_cell) {   // platens are not binned: the cell of trunc(x / h), no order
            for (int a = 0; a < 3; ++a) out->cell[3 * (size_t)i + a] = (int32_t)(hx[3 * j + a] / h->g.h[a]);
            out->order[i] = 0xFFFFFFFFu;
        }
    }
    for (int k = 0; k < NF; ++k) dev_free((void *&)st[k]);
    dev_free((void *&)cell_d);
    dev_free((void *&)order_d);
    // d_Kernel / d3_Kernel_Gradient of the last contact step (zero before one)
    if (mask & MPM_F_KERNEL) {
        RCHK(kernel_out<real>(h));
        if (out->kernel) {
            if (h->ct.kernel_valid) HIPCHK(h, hipMemcpy(out->kernel, h->ct.kp_id, sizeof(double) * n, hipMemcpyDeviceToHost));
            else std::fill(out->kernel, out->kernel + n, 0.0);
        }
        if (out->kernel_grad) {
            if (h->ct.kernel_valid) HIPCHK(h, hipMemcpy(out->kernel_grad, h->ct.kg_id, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
            else std::fill(out->kernel_grad, out->kernel_grad + 3 * (size_t)n, 0.0);
        }
    }
    return MPM_OK;
}

// mpm_grid_nodes: node fields into dense device arrays, then to the caller
template <typename real>
static int grid_nodes_t(mpm_handle *h, mpm_nodes *out)
{
    const size_t nn = (size_t)h->nnodes;
    uint8_t *act = nullptr;
    double *buf = nullptr;   // mass 1, v 3, f 3, ft 3, k 1, kg 3 per node
    RCHK(dalloc(h, &act, nn));
    if (int rc = dalloc(h, &buf, 14 * nn)) { dev_free((void *&)act); return rc; }
    NodeOut o{act, buf, buf + nn, buf + 4 * nn, buf + 7 * nn, buf + 10 * nn, buf + 11 * nn};
    int rc = MPM_OK;
    auto run = [&]() -> int {
        HIPCHK(h, hipMemsetAsync(act, 0, nn, h->stream));
        HIPCHK(h, hipMemsetAsync(buf, 0, 14 * nn * sizeof(double), h->stream));
        if (h->last_path == 1 && h->ct.ready) {
            k_c_nodes<real><<<cdiv(nn, 256), 256, 0, h->stream>>>(contact_args<real>(h), o);
            HIPCHK(h, hipGetLastError());
        } else if (h->n_occ > 0) {
            // the next step's record: current time, dt_max, epoch = step count
            StepRec st{};
            st.dt = h->cfg.dt_max;
            st.t = h->time;
            st.epoch = (uint32_t)h->cycle;
            st.par = h->parity;
            k_grid_state<real><<<(unsigned)h->n_occ, 256, 0, h->stream>>>(
                h->g, (const real *)h->arena[h->parity], h->binfo[h->parity],
                (uint32_t)(h->arena_cap_slots * 8 * h->g.arenaN), h->fixed, h->marks[h->parity], h->dc_v[h->parity],
                h->occ_list, h->n_occ_d, st, o);
            HIPCHK(h, hipGetLastError());
        }
        std::vector<double> hb(14 * nn);
        std::vector<uint8_t> ha(nn);
        HIPCHK(h, hipMemcpyAsync(hb.data(), buf, 14 * nn * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(ha.data(), act, nn, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        auto put = [&](double *dst, size_t off, int w) {
            if (dst) std::memcpy(dst, hb.data() + off, sizeof(double) * w * nn);
        };
        put(out->mass, 0, 1);
        put(out->v, nn, 3);
        put(out->f, 4 * nn, 3);
        put(out->f_temp, 7 * nn, 3);
        put(out->kernel, 10 * nn, 1);
        put(out->kernel_grad, 11 * nn, 3);
        if (out->active) std::memcpy(out->active, ha.data(), nn);
        if (out->fixed) std::memcpy(out->fixed, h->fixed_host.data(), nn);
        if (out->x) {   // GridPoint_Factory.h:40-42: (i L) / n per axis
            const int nx = h->g.nodes[0], ny = h->g.nodes[1];
            for (size_t I = 0; I < nn; ++I) {
                const int i[3] = {(int)(I % nx), (int)((I / nx) % ny), (int)(I / ((size_t)nx * ny))};
                for (int a = 0; a < 3; ++a) out->x[3 * I + a] = ((double)i[a] * h->g.L[a]) / (double)h->g.cells[a];
            }
        }
        return MPM_OK;
    };
    rc = run();
    dev_free((void *&)act);
    dev_free((void *&)buf);
    return rc;
}

extern "C" {

int mpm_grid_nodes(mpm_handle *h, mpm_nodes *out)
{
    BUSYCHK(h);
    if (!h || !out) return MPM_E_ARG;
    if (!h->uploaded) { h->err = "grid readback before upload"; return MPM_E_STATE; }
    if (out->n != h->nnodes) { h->err = "grid readback node count mismatch"; return MPM_E_ARG; }
    HIPCHK(h, hipSetDevice(h->device));
    return h->fp32 ? grid_nodes_t<float>(h, out) : grid_nodes_t<double>(h, out);
}

int mpm_readback(mpm_handle *h, uint32_t field_mask, mpm_points *out)
{
    BUSYCHK(h);
    if (!h || !out) return MPM_E_ARG;
    if (!h->uploaded) { h->err = "readback before upload"; return MPM_E_STATE; }
    if (out->n != h->n_total) { h->err = "readback point count mismatch"; return MPM_E_ARG; }
    if ((field_mask & MPM_F_STRAIN) && !h->cfg.track_strain) { h->err = "strain not tracked (mpm_config.track_strain=0)"; return MPM_E_ARG; }
    HIPCHK(h, hipSetDevice(h->device));
    return h->fp32 ? readback_t<float>(h, field_mask, out) : readback_t<double>(h, field_mask, out);
}

int mpm_checks(mpm_handle *h, double out[12])
{
    BUSYCHK(h);
    if (!h || !out) return MPM_E_ARG;
    if (!h->uploaded) { h->err = "checks before upload"; return MPM_E_STATE; }
    HIPCHK(h, hipSetDevice(h->device));
    return checks_impl(h, out);
}

int mpm_reaction_history(mpm_handle *h, double *out, int32_t cap, int32_t *n)
{
    BUSYCHK(h);
    if (!h || (!out && cap > 0) || !n) return MPM_E_ARG;
    const int m = std::min<int>((int)h->hist.size(), cap);
    for (int k = 0; k < m; ++k) out[k] = h->hist[k];
    *n = m;
    return MPM_OK;
}

int mpm_kernel_times(mpm_handle *h, double out[6])
{
    BUSYCHK(h);
    if (!h || !out) return MPM_E_ARG;
    for (int k = 0; k < 6; ++k) out[k] = h->kt[k];
    return MPM_OK;
}

int mpm_group_reaction_history(mpm_handle *h, double *out, int32_t cap, int32_t *n)
{
    BUSYCHK(h);
    if (!h || (!out && cap > 0) || !n) return MPM_E_ARG;
    if (h->n_groups == 0) { *n = 0; return MPM_OK; }
    const int steps = (int)(h->ghist.size() / h->n_groups);
    const int m = std::min<int>(steps, cap);
    for (size_t k = 0; k < (size_t)m * h->n_groups; ++k) out[k] = h->ghist[k];
    *n = m;
    return MPM_OK;
}

int mpm_get_time(mpm_handle *h, double *time, double *time_end, double *dt_max)
{
    BUSYCHK(h);
    if (!h) return MPM_E_ARG;
    if (time) *time = h->time;
    if (time_end) *time_end = h->cfg.time_end;
    if (dt_max) *dt_max = h->cfg.dt_max;
    return MPM_OK;
}

int mpm_last_step_ms(mpm_handle *h, double *ms)
{
    BUSYCHK(h);
    if (!h || !ms) return MPM_E_ARG;
    *ms = h->last_ms;
    return MPM_OK;
}

int mpm_set_log(mpm_handle *h, const char *path)
{
    BUSYCHK(h);
    if (!h) return MPM_E_ARG;
    h->log_path = path ? path : "";
    return MPM_OK;
}

int mpm_report(mpm_handle *h, const char *description, char *buf, int32_t cap, int32_t *len)
{
    BUSYCHK(h);
    if (!h || (buf && cap < 1)) return MPM_E_ARG;
    if (h->n_total == 0) {
        h->err = "report before upload";
        return MPM_E_STATE;
    }
    HIPCHK(h, hipSetDevice(h->device));
    std::string line;
    int rc = report_line(h, description, h->runtime_s, line);
    if (rc) return rc;
    report_emit(h, line);
    if (len) *len = (int32_t)line.size();
    if (buf) {
        const size_t k = std::min<size_t>(line.size(), (size_t)cap - 1);
        std::memcpy(buf, line.data(), k);
        buf[k] = '\0';
    }
    return MPM_OK;
}

int mpm_sort_keys(const uint32_t *keys_host, int32_t n, uint32_t *sorted_host, uint32_t *perm_host, int32_t device)
{
    if (n < 0 || (n > 0 && (!keys_host || !sorted_host || !perm_host))) { g_create_error = "bad sort args"; return MPM_E_ARG; }
    if (n == 0) return MPM_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return MPM_E_NO_DEVICE; }
    mpm_handle tmp;
    mpm_handle *h = &tmp;
    h->device = device;
    h->n_def = n;
    h->nbricks = 1;
    int rc = MPM_OK;
    auto run = [&]() -> int {
        HIPCHK(h, hipSetDevice(device));
        HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        RCHK(alloc_sort_scratch(h, (uint32_t)n));
        uint32_t mx = 0;
        for (int i = 0; i < n; ++i) mx = std::max(mx, keys_host[i]);
        std::vector<uint32_t> iota(n);
        for (int i = 0; i < n; ++i) iota[i] = (uint32_t)i;
        HIPCHK(h, hipMemcpy(h->keys[0], keys_host, 4 * (size_t)n, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->vals[0], iota.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
        uint32_t *ko, *vo;
        RCHK(radix_sort(h, h->keys[0], h->vals[0], h->keys[1], h->vals[1], (uint32_t)n, mx, &ko, &vo));
        HIPCHK(h, hipMemcpyAsync(sorted_host, ko, 4 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(perm_host, vo, 4 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return MPM_OK;
    };
    rc = run();
    if (rc != MPM_OK) g_create_error = h->err;
    destroy_all(h);
    return rc;
}

int mpm_constitutive(int32_t law, const mpm_cr_batch *b, int32_t device)
{
    if (!b || b->n < 0) { g_create_error = "bad constitutive batch"; return MPM_E_ARG; }
    if (law == MPM_LAW_VISCOELASTIC || law == MPM_LAW_IDEALGAS) {
        g_create_error = "the ViscoElastic and IdealGas laws are not dispatched by the reference step and not provided";
        return MPM_E_ARG;
    }
    if (!cr_law_known(law)) {
        g_create_error = "unknown constitutive law";
        return MPM_E_ARG;
    }
    if (b->n == 0) return MPM_OK;
    CrPart parts[kCrParts];
    const char *why = nullptr;
    const long long total = cr_layout(law, b, parts, &why);
    if (total < 0) { g_create_error = why; return MPM_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return MPM_E_NO_DEVICE; }
    if (device < 0 || device >= ndev) { g_create_error = "bad device ordinal"; return MPM_E_ARG; }
    // one staging buffer and stream per device, kept for the process and
    // grown when a batch needs more; calls are serialised per device
    std::lock_guard<std::mutex> lk(g_cr_mu);
    CrStage &cs = g_cr_stage[device];
    if (hipSetDevice(device) != hipSuccess) { g_create_error = "hipSetDevice"; return MPM_E_HIP; }
    if (!cs.stream && hipStreamCreateWithFlags(&cs.stream, hipStreamNonBlocking) != hipSuccess) {
        cs.stream = nullptr;
        g_create_error = "hipStreamCreate";
        return MPM_E_HIP;
    }
    if ((size_t)total > cs.cap) {
        if (cs.buf) (void)hipFree(cs.buf);
        cs.buf = nullptr;
        cs.cap = 0;
        const size_t want = std::max<size_t>((size_t)total, 1u << 16);
        if (hipMalloc((void **)&cs.buf, want * sizeof(double)) != hipSuccess) {
            cs.buf = nullptr;
            g_create_error = "hipMalloc (constitutive staging)";
            return MPM_E_NOMEM;
        }
        cs.cap = want;
    }
    int rc = MPM_OK;
    for (int k = 0; k < kCrParts && rc == MPM_OK; ++k)
        if (parts[k].src && hipMemcpyAsync(cs.buf + parts[k].off, parts[k].src, parts[k].len * sizeof(double),
                                           hipMemcpyHostToDevice, cs.stream) != hipSuccess)
            rc = MPM_E_HIP;
    if (rc == MPM_OK) {
        const CrBatch cb = cr_views(b->n, cs.buf, parts);
        k_constitutive<<<cdiv((size_t)b->n, 256), 256, 0, cs.stream>>>(law, cb);
        if (hipGetLastError() != hipSuccess) rc = MPM_E_HIP;
    }
    for (int k = 0; k < kCrParts && rc == MPM_OK; ++k)
        if (parts[k].dst && hipMemcpyAsync(parts[k].dst, cs.buf + parts[k].off, parts[k].len * sizeof(double),
                                           hipMemcpyDeviceToHost, cs.stream) != hipSuccess)
            rc = MPM_E_HIP;
    if (hipStreamSynchronize(cs.stream) != hipSuccess) rc = MPM_E_HIP;
    if (rc != MPM_OK) g_create_error = "constitutive batch: HIP error";
    return rc;
}

}  // extern "C"
// ---- mpm_constitutive_noalloc: the same batch through a fixed device array.
// `scratch` is a file-scope device array (one copy per GPU the library runs
// on, held until the process ends).  In host code its name is only a
// stand-in: kernels and copies use the address hipGetSymbolAddress returns,
// looked up once per device with that device current.
constexpr size_t kNoAllocInputs = 8192;                 // inputs per call
constexpr size_t kNoAllocDoubles = 41 * kNoAllocInputs;  // 41 = the most doubles one input uses
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
double *g_na_ptr[kMaxDevices] = {nullptr};
}  // namespace

extern "C" int mpm_constitutive_noalloc(int32_t law, const mpm_cr_batch *b, int32_t device)
{
    if (!b || b->n < 0) { g_create_error = "bad constitutive batch"; return MPM_E_ARG; }
    if (law == MPM_LAW_VISCOELASTIC || law == MPM_LAW_IDEALGAS) {
        g_create_error = "the ViscoElastic and IdealGas laws are not dispatched by the reference step and not provided";
        return MPM_E_ARG;
    }
    if (!cr_law_known(law)) { g_create_error = "unknown constitutive law"; return MPM_E_ARG; }
    if ((size_t)b->n > kNoAllocInputs) {
        g_create_error = "mpm_constitutive_noalloc: at most " + std::to_string(kNoAllocInputs) +
                         " inputs per call (use mpm_constitutive for larger batches)";
        return MPM_E_ARG;
    }
    if (b->n == 0) return MPM_OK;
    CrPart parts[kCrParts];
    const char *why = nullptr;
    const long long total = cr_layout(law, b, parts, &why);
    if (total < 0) { g_create_error = why; return MPM_E_ARG; }
    if ((size_t)total > kNoAllocDoubles) { g_create_error = "mpm_constitutive_noalloc: batch exceeds the scratch array"; return MPM_E_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device"; return MPM_E_NO_DEVICE; }
    if (device < 0 || device >= ndev || device >= kMaxDevices) { g_create_error = "bad device ordinal"; return MPM_E_ARG; }
    // calls on one device share its array: serialise them in this process
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (hipSetDevice(device) != hipSuccess) { g_create_error = "hipSetDevice"; return MPM_E_HIP; }
    if (!g_na_ptr[device]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return MPM_E_HIP;
        }
        g_na_ptr[device] = p;
    }
    double *const buf = g_na_ptr[device];
    int rc = MPM_OK;
    for (int k = 0; k < kCrParts && rc == MPM_OK; ++k)
        if (parts[k].src && hipMemcpy(buf + parts[k].off, parts[k].src, parts[k].len * sizeof(double),
                                      hipMemcpyHostToDevice) != hipSuccess)
            rc = MPM_E_HIP;
    if (rc == MPM_OK) {
        const CrBatch cb = cr_views(b->n, buf, parts);
        k_constitutive<<<cdiv((size_t)b->n, 256), 256, 0, 0>>>(law, cb);
        if (hipGetLastError() != hipSuccess) rc = MPM_E_HIP;
    }
    for (int k = 0; k < kCrParts && rc == MPM_OK; ++k)
        if (parts[k].dst && hipMemcpy(parts[k].dst, buf + parts[k].off, parts[k].len * sizeof(double),
                                      hipMemcpyDeviceToHost) != hipSuccess)
            rc = MPM_E_HIP;
    if (hipDeviceSynchronize() != hipSuccess) rc = MPM_E_HIP;
    if (rc != MPM_OK) g_create_error = "constitutive batch: HIP error";
    return rc;
}
